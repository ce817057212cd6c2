"""Distilled head: oracle pins (CPU) and HIP parity (GPU).

CPU: the numpy network oracle against torch autograd in float64 (an
independent implementation: energies, dE/db, and the force-matching
parameter gradient through double backward) and central differences; the
per-row-coefficient slot gradients of the MTP oracle against central
differences of the descriptor energy; Jacobian linearity.
GPU: every head kernel and the composed head evaluation (with exact
fallback) against the oracle at 1e-10 relative (FP64 throughout).
"""
import numpy as np
import pytest

from conftest import oracle_for

S12, SP = 29, 3


class _M:
    """Minimal HeadModel-like holder for the oracle."""

    def __init__(self, theta, mean, inv_std, lo, hi, e_shift, e_scale, lin=None, lin_sc=None):
        self.theta, self.mean, self.inv_std, self.lo, self.hi = theta, mean, inv_std, lo, hi
        self.e_shift, self.e_scale, self.lin, self.lin_sc = e_shift, e_scale, lin, lin_sc


def _as_loaded(jm, p):
    """The coefficients a potential opened from MTPData actually holds (its .mtp text)."""
    return jm.parse_mtp_text(jm.format_mtp(p))


def _model(rng, S=S12, Sp=SP, box=False):
    import head_oracle as HO

    theta = rng.normal(0, 0.3, HO.N_THETA)
    W1 = theta[:HO.IN * HO.H].reshape(HO.IN, HO.H)
    W1[S + Sp:] = 0.0
    mean = rng.normal(0, 1, S)
    inv_std = rng.uniform(0.5, 2.0, S)
    lo = hi = None
    if box:
        lo = np.full((Sp, S), -1.5)
        hi = np.full((Sp, S), 1.5)
    return _M(theta, mean, inv_std, lo, hi, rng.normal(0, 1, Sp), 0.7)


def _torch_net(model, basis, itypes):
    import torch
    import head_oracle as HO

    th = torch.tensor(model.theta, requires_grad=True)
    b = torch.tensor(basis, requires_grad=True)
    S = basis.shape[1]
    Sp = len(model.e_shift)
    o = 0
    W1 = th[o:o + HO.IN * HO.H].reshape(HO.IN, HO.H); o += HO.IN * HO.H
    b1 = th[o:o + HO.H]; o += HO.H
    W2 = th[o:o + HO.H * HO.H].reshape(HO.H, HO.H); o += HO.H * HO.H
    b2 = th[o:o + HO.H]; o += HO.H
    w3 = th[o:o + HO.H]; o += HO.H
    b3 = th[o]
    ts = HO.wrap_species(itypes, Sp)
    zb = (b - torch.tensor(model.mean)) * torch.tensor(model.inv_std)
    oh = torch.zeros((len(basis), HO.IN - S), dtype=torch.float64)
    oh[np.arange(len(basis)), ts] = 1.0
    z = torch.cat([zb, oh], 1)
    h1 = torch.nn.functional.silu(z @ W1 + b1)
    h2 = torch.nn.functional.silu(h1 @ W2 + b2)
    E = torch.tensor(model.e_shift)[ts] + model.e_scale * (h2 @ w3 + b3)
    return th, b, E


def test_network_matches_torch_autograd():
    import torch
    import head_oracle as HO

    rng = np.random.default_rng(1)
    model = _model(rng)
    A = 40
    basis = rng.normal(0, 1, (A, S12)) * 2 + model.mean
    itypes = rng.integers(-1, SP, A)
    E, xi, _ = HO.eval_head(model, basis, itypes)
    th, b, Et = _torch_net(model, basis, itypes)
    np.testing.assert_allclose(E, Et.detach().numpy(), rtol=1e-13, atol=1e-13)
    (dEdb,) = torch.autograd.grad(Et.sum(), b, create_graph=True)
    np.testing.assert_allclose(xi, dEdb.detach().numpy(), rtol=1e-12, atol=1e-13)
    # force-matching gradient: d/dtheta sum alpha E + v . dE/db  (double backward)
    alpha = rng.normal(0, 1, A)
    v = rng.normal(0, 1, (A, S12))
    phi = (torch.tensor(alpha) * Et).sum() + (torch.tensor(v) * dEdb).sum()
    (gt,) = torch.autograd.grad(phi, th)
    g = HO.param_grad(model, basis, itypes, alpha, v)
    np.testing.assert_allclose(g, gt.numpy(), rtol=1e-11, atol=1e-12 * np.abs(gt.numpy()).max())


def test_param_grad_central_differences():
    import head_oracle as HO

    rng = np.random.default_rng(2)
    model = _model(rng)
    A = 12
    basis = rng.normal(0, 1, (A, S12)) + model.mean
    itypes = rng.integers(0, SP, A)
    alpha = rng.normal(0, 1, A)
    v = rng.normal(0, 1, (A, S12))
    g = HO.param_grad(model, basis, itypes, alpha, v)

    def phi():
        E, xi, _ = HO.eval_head(model, basis, itypes)
        return alpha @ E + np.sum(v * xi)

    h = 1e-4
    for q in [0, 5, HO.IN * HO.H + 3, HO.IN * HO.H + HO.H + 77, HO.N_THETA - 2, HO.N_THETA - 1]:
        x0 = model.theta[q]
        vals = []
        for t in (-2, -1, 1, 2):
            model.theta[q] = x0 + t * h
            vals.append(phi())
        model.theta[q] = x0
        fd = (vals[0] - 8 * vals[1] + 8 * vals[2] - vals[3]) / (12 * h)
        assert g[q] == pytest.approx(fd, rel=1e-7, abs=1e-9 * np.abs(g).max())


def _level12_case(jm, seed=0, natoms=24, cell=7.5):
    w = jm.configs.small_cluster(12, SP, natoms, seed, cell=cell, cutoff=5.0, max_nbr=64, radial_basis_size=8)
    nb = w.neighbors()
    return w, nb


def test_oracle_per_row_coefficients(jm, O):
    w, nb = _level12_case(jm, seed=3, natoms=12)
    orc = oracle_for(O, w.params)
    A, K = nb["js"].shape
    rng = np.random.default_rng(3)
    xi = rng.normal(0, 1, (A, S12))
    Ei, g, basis = orc.compact_xi(nb["itypes"], nb["rij"], nb["jtypes"], xi)
    # energy is species_coeff + xi . basis
    sc = orc._keep["species"][nb["itypes"]]
    np.testing.assert_allclose(Ei, sc + np.sum(xi * basis, 1), rtol=1e-13)
    # the shared vector reproduces the plain evaluation
    E0, g0 = orc.compact(nb["itypes"], nb["rij"], nb["jtypes"])
    E1, g1, _ = orc.compact_xi(nb["itypes"], nb["rij"], nb["jtypes"], orc._keep["moment"])
    np.testing.assert_array_equal(E0, E1)
    np.testing.assert_array_equal(g0, g1)
    # slot gradients: central differences of E_i (degree-unbounded, so a 4th-order stencil)
    h = 1e-5
    rij = nb["rij"].copy()
    for i, k, c in [(0, 0, 0), (3, 5, 1), (7, nb["counts"][7] - 1, 2)]:
        vals = []
        for t in (-2, -1, 1, 2):
            r = rij.copy()
            r[i, k, c] += t * h
            vals.append(orc.compact_xi(nb["itypes"], r, nb["jtypes"], xi, grad=False)[0][i])
        fd = (vals[0] - 8 * vals[1] + 8 * vals[2] - vals[3]) / (12 * h)
        assert g[i, k, c] == pytest.approx(fd, rel=1e-7, abs=1e-9 * np.abs(g).max())
    # Jacobian linearity
    J = orc.basis_jacobian(nb["itypes"], nb["rij"], nb["jtypes"])
    np.testing.assert_allclose(np.einsum("as,askc->akc", xi, J), g, rtol=1e-11, atol=1e-12 * np.abs(g).max())


def test_pair_invariants_descriptor(jm, O):
    """Descriptor with pair moments M_{mu,0} = sum_k f_mu(r_k) (mu = 3, 4, 5)
    appended to the level-12 basis: its first 29 basis values are those of the
    plain level-12 descriptor (the moment renumbering is right), and each new
    one equals the single basis value of a level-2 potential whose radial
    function 0 is f_mu (a pair sum by construction)."""
    rng = np.random.default_rng(21)
    base = jm.configs.random_params(12, SP, 8, 2.0, 5.0, rng)
    aug = jm.distill.with_pair_invariants(jm.parse_mtp_text(jm.format_mtp(base)), (3, 4, 5))
    aug.radial_coeffs = rng.normal(0, 0.3, np.asarray(aug.radial_coeffs).shape)
    base.radial_coeffs = aug.radial_coeffs[:, :, :3, :].copy()
    w, nb = _level12_case(jm, seed=21, natoms=20, cell=8.0)
    args = (nb["itypes"], nb["rij"], nb["jtypes"])
    _, _, b_base = oracle_for(O, base).compact_xi(*args, np.zeros(S12), grad=False)
    _, _, b_aug = oracle_for(O, aug).compact_xi(*args, np.zeros(S12 + 3), grad=False)
    np.testing.assert_allclose(b_aug[:, :S12], b_base, rtol=1e-13, atol=1e-13 * np.abs(b_base).max())
    for q, mu in enumerate((3, 4, 5)):
        pair = jm.configs.random_params(2, SP, 8, 2.0, 5.0, rng)
        pair.radial_coeffs = aug.radial_coeffs[:, :, mu:mu + 1, :].copy()
        _, _, b_pair = oracle_for(O, pair).compact_xi(*args, np.zeros(1), grad=False)
        np.testing.assert_allclose(b_aug[:, S12 + q], b_pair[:, 0], rtol=1e-13, atol=1e-13 * np.abs(b_pair).max())


def test_head_model_roundtrip(jm, tmp_path):
    rng = np.random.default_rng(4)
    desc = jm.configs.random_params(12, SP, 8, 2.0, 5.0, rng)
    m = jm.head.HeadModel(desc, jm.head.init_theta(S12, SP, rng), rng.normal(size=S12), rng.uniform(1, 2, S12),
                          -np.ones((SP, S12)), np.ones((SP, S12)), rng.normal(size=SP), 0.3)
    m.save(tmp_path / "h.npz")
    m2 = jm.head.HeadModel.load(tmp_path / "h.npz")
    for k in ("theta", "mean", "inv_std", "lo", "hi", "e_shift"):
        np.testing.assert_array_equal(getattr(m, k), getattr(m2, k))
    assert m2.e_scale == m.e_scale
    np.testing.assert_array_equal(m2.desc.radial_coeffs, desc.radial_coeffs)
    assert len(jm.head.init_theta(S12, SP, rng)) == jm.head.N_THETA


# ---------------------------------------------------------------------------
# GPU
# ---------------------------------------------------------------------------
def _rel(a, r):
    a = np.asarray(a)
    r = np.asarray(r)
    return float(np.max(np.abs(a - r)) / max(np.max(np.abs(r)), 1e-300))


@pytest.mark.gpu
@pytest.mark.parametrize("pairs", [False, True])
def test_gpu_descriptor_jacobian_gather(jm, O, pairs):
    """Level-12 descriptor; with pairs, the level-12 basis plus the pair moments of mu = 3, 4, 5 (six
    radial functions, a non-standard alpha table)."""
    import torch

    w, nb = _level12_case(jm, seed=5, natoms=40, cell=8.2)
    if pairs:
        w.params = jm.distill.with_pair_invariants(w.params, (3, 4, 5))
        w.params.radial_coeffs = np.random.default_rng(55).normal(0, 0.2, np.asarray(w.params.radial_coeffs).shape)
    S = len(w.params.moment_coeffs)
    print(f"[case] level-12 descriptor natoms={w.natoms} pairs={pairs} S={S}", flush=True)
    orc = oracle_for(O, w.params)
    pot = jm.MTPPotential(w.params, max_atoms=w.natoms, max_nbr=w.max_nbr)
    A = w.natoms
    rng = np.random.default_rng(5)
    xi = rng.normal(0, 1, (A, S))
    _, g_ref, b_ref = orc.compact_xi(nb["itypes"], nb["rij"], nb["jtypes"], xi)
    b = pot.descriptor(nb["itypes"], nb["rij"], nb["jtypes"]).cpu().numpy()
    assert _rel(b, b_ref) <= 1e-12
    g = pot.slot_grad_xi(nb["itypes"], nb["rij"], nb["jtypes"], xi).cpu().numpy()
    assert _rel(g, g_ref) <= 1e-10
    J = pot.basis_jacobian(nb["itypes"], nb["rij"], nb["jtypes"]).cpu().numpy()
    J_ref = orc.basis_jacobian(nb["itypes"], nb["rij"], nb["jtypes"])
    assert _rel(J, J_ref) <= 1e-10
    F = pot.gather_forces(nb["js"], torch.as_tensor(g_ref, device="cuda"), A).cpu().numpy()
    np.testing.assert_allclose(F, O.Oracle.forces(g_ref, nb["js"], A), rtol=0, atol=1e-12 * np.abs(F).max())
    pot.close()


@pytest.mark.gpu
@pytest.mark.parametrize("residual", [False, True])
@pytest.mark.parametrize("level", [12, 14])
def test_gpu_head_network_and_gradient(jm, O, residual, level):
    """Level 12: 29 basis inputs (32 with the species, half the input tiles
    unused); level 14: 52 (55), every input tile in use."""
    import torch
    import head_oracle as HO

    rng = np.random.default_rng(6)
    desc = jm.configs.random_params(level, SP, 8, 2.0, 5.0, rng)
    S = len(desc.moment_coeffs)
    model = _model(rng, S=S, box=True)
    if residual:
        dl = _as_loaded(jm, desc)
        model.lin, model.lin_sc = np.asarray(dl.moment_coeffs), np.asarray(dl.species_coeffs)
    hm = jm.head.HeadModel(desc, model.theta, model.mean, model.inv_std, model.lo, model.hi, model.e_shift,
                           model.e_scale, residual)
    A = 3000
    head = jm.head.DistilledHead(hm, None, max_atoms=A, max_nbr=16)
    basis = rng.normal(0, 1, (A, S)) + model.mean
    itypes = rng.integers(-1, SP, A)
    dev = lambda x, dt=torch.float64: torch.as_tensor(x, dtype=dt, device="cuda").contiguous()
    E, xi = head.eval_network(dev(itypes, torch.int32), dev(basis))
    E_ref, xi_ref, _ = HO.eval_head(model, basis, itypes)
    assert _rel(E.cpu().numpy(), E_ref) <= 1e-12
    assert _rel(xi.cpu().numpy(), xi_ref) <= 1e-12
    alpha = rng.normal(0, 1, A)
    v = rng.normal(0, 1, (A, S))
    g = head.param_grad(dev(itypes, torch.int32), dev(basis), dev(alpha), dev(v)).cpu().numpy()
    g_ref = HO.param_grad(model, basis, itypes, alpha, v)
    print(f"  head param grad rel err {_rel(g, g_ref):.3e}", flush=True)
    assert _rel(g, g_ref) <= 1e-11
    head.close()


@pytest.mark.gpu
@pytest.mark.parametrize("exact_level,desc_variant", [(16, 3), (24, 3), (24, 1), (24, 2), (24, 6)])
def test_gpu_head_compute_with_fallback(jm, O, exact_level, desc_variant):
    """Descriptor variants of 64, 128 and 256 threads reuse the descriptor
    pass's moments in the adjoint pass; 512 threads recompute them."""
    import head_oracle as HO

    w, nb = _level12_case(jm, seed=7, natoms=48, cell=8.6)
    print(f"[case] head compute natoms={w.natoms} exact level {exact_level} descriptor variant {desc_variant}",
          flush=True)
    rng = np.random.default_rng(7)
    model = _model(rng, box=False)
    orc_d = oracle_for(O, w.params)
    # box from the configuration's own inputs, narrowed so that some rows fall outside
    _, _, b = orc_d.compact_xi(nb["itypes"], nb["rij"], nb["jtypes"], np.zeros(S12), grad=False)
    z = (b - model.mean) * model.inv_std
    lo = np.tile(np.quantile(z, 0.03, axis=0), (SP, 1))
    hi = np.tile(np.quantile(z, 0.97, axis=0), (SP, 1))
    model.lo, model.hi = lo, hi
    exact_p = jm.configs.random_params(exact_level, SP, 8, 2.0, 5.0, rng)
    orc_x = oracle_for(O, exact_p)
    A, K = nb["js"].shape
    n_e = n_f = w.natoms - 3  # masks exercised
    E_r, F_r, S_r, eat_r, fl_r, _ = HO.head_compute(model, orc_d, orc_x, nb["itypes"], nb["js"], nb["rij"],
                                                     nb["jtypes"], 3, w.volume, n_e, n_f)
    assert 0 < fl_r.sum() < A
    model.lin, model.lin_sc = orc_d._keep["moment"], orc_d._keep["species"]  # residual form
    E_r, F_r, S_r, eat_r, fl_r, _ = HO.head_compute(model, orc_d, orc_x, nb["itypes"], nb["js"], nb["rij"],
                                                     nb["jtypes"], 3, w.volume, n_e, n_f)
    exact = jm.MTPPotential(exact_p, max_atoms=A, max_nbr=K)
    hm = jm.head.HeadModel(w.params, model.theta, model.mean, model.inv_std, lo, hi, model.e_shift, model.e_scale,
                           True)
    head = jm.head.DistilledHead(hm, exact, max_atoms=A, max_nbr=K)
    head.desc.set_variant(desc_variant)
    E, F, St, extra = head.calc_energy_forces_stress_padded(nb["itypes"], nb["js"], nb["rij"], nb["jtypes"], 3,
                                                            w.volume, n_e, n_f, return_all=True)
    assert float(E) == pytest.approx(E_r, rel=1e-10)
    assert _rel(F.cpu().numpy(), F_r) <= 1e-10
    assert _rel(St.cpu().numpy(), S_r) <= 1e-10
    assert _rel(extra["eatom"].cpu().numpy(), eat_r) <= 1e-10
    np.testing.assert_array_equal(extra["flags"].cpu().numpy(), fl_r)
    assert extra["n_fallback"] == int(((fl_r != 0) & (nb["itypes"] >= 0)).sum())
    assert float(extra["status"]) == 0.0
    head.close()
    exact.close()


@pytest.mark.gpu
def test_gpu_head_graph_replay(jm, O):
    """Repeated head evaluations with the same arrays replay a recorded HIP
    graph; results stay equal to the first (eager) call and follow positions
    changed in place, against the oracle composition."""
    import torch
    import head_oracle as HO

    w, nb = _level12_case(jm, seed=7, natoms=48, cell=8.6)
    rng = np.random.default_rng(7)
    model = _model(rng, box=False)
    orc_d = oracle_for(O, w.params)
    _, _, b = orc_d.compact_xi(nb["itypes"], nb["rij"], nb["jtypes"], np.zeros(S12), grad=False)
    z = (b - model.mean) * model.inv_std
    lo = np.tile(np.quantile(z, 0.03, axis=0), (SP, 1))
    hi = np.tile(np.quantile(z, 0.97, axis=0), (SP, 1))
    model.lo, model.hi = lo, hi
    model.lin, model.lin_sc = orc_d._keep["moment"], orc_d._keep["species"]
    exact_p = jm.configs.random_params(24, SP, 8, 2.0, 5.0, rng)
    orc_x = oracle_for(O, exact_p)
    A, K = nb["js"].shape
    exact = jm.MTPPotential(exact_p, max_atoms=A, max_nbr=K)
    hm = jm.head.HeadModel(w.params, model.theta, model.mean, model.inv_std, lo, hi, model.e_shift, model.e_scale,
                           True)
    head = jm.head.DistilledHead(hm, exact, max_atoms=A, max_nbr=K)
    dev = torch.device("cuda")
    it = torch.as_tensor(nb["itypes"], dtype=torch.int32, device=dev)
    js = torch.as_tensor(nb["js"], dtype=torch.int32, device=dev)
    rij = torch.as_tensor(nb["rij"], dtype=torch.float64, device=dev)
    jt = torch.as_tensor(nb["jtypes"], dtype=torch.int32, device=dev)
    out = torch.empty(14, dtype=torch.float64, device=dev)
    F = torch.empty((A, 3), dtype=torch.float64, device=dev)
    nfb = torch.zeros(1, dtype=torch.int32, device=dev)
    L = jm.capi.lib()
    st = torch.cuda.current_stream().cuda_stream

    def call():
        jm.capi.check(L.mtp_head_compute_padded_device(head.h, A, K, it.data_ptr(), js.data_ptr(), rij.data_ptr(),
                                                       jt.data_ptr(), 3, w.volume, A, A, out.data_ptr(),
                                                       F.data_ptr(), None, None, nfb.data_ptr(), st), head=head.h)
        torch.cuda.synchronize()
        return float(out[0].item()), F.cpu().numpy().copy(), int(nfb.item())

    def check(res, rij_np, tag):
        E_r, F_r, S_r, eat_r, fl_r, _ = HO.head_compute(model, orc_d, orc_x, nb["itypes"], nb["js"], rij_np,
                                                         nb["jtypes"], 3, w.volume, A, A)
        print(f"[case] head graph replay {tag}: rel F {_rel(res[1], F_r):.3e}", flush=True)
        assert _rel(res[1], F_r) <= 1e-10, tag
        assert res[0] == pytest.approx(E_r, rel=1e-10), tag
        assert res[2] == int(((fl_r != 0) & (nb["itypes"] >= 0)).sum()), tag

    first = call()
    check(first, nb["rij"], "eager")
    for q in range(3):
        again = call()
        assert again[0] == first[0] and np.array_equal(again[1], first[1]) and again[2] == first[2], q
    moved = nb["rij"] + np.where(nb["jtypes"][..., None] >= 0, rng.normal(0, 0.02, nb["rij"].shape), 0.0)
    rij.copy_(torch.as_tensor(moved, dtype=torch.float64, device=dev))
    check(call(), moved, "positions changed in place")
    check(call(), moved, "positions changed in place, replayed")
    head.close()
    exact.close()


@pytest.mark.gpu
def test_gpu_head_training_gradient_and_fit(jm, O):
    """The autograd Function's gradient of the distillation loss against the
    oracle composition, then a short fit that lowers the loss."""
    import torch
    import head_oracle as HO

    rng = np.random.default_rng(8)
    cells = jm.training.random_cells(4, rng, natoms_range=(20, 40), species=SP, kinds=("fcc", "random"))
    b = jm.training.pack_configs(cells, cutoff=5.0, max_nbr=64)
    desc_p = jm.configs.random_params(12, SP, 8, 1.6, 5.0, rng)
    teacher_p = jm.configs.random_params(16, SP, 8, 1.6, 5.0, rng, sigma=0.05)
    print(f"[case] head training natoms={b.natoms}", flush=True)
    orc_t = oracle_for(O, teacher_p)
    E_site, g_t = orc_t.compact(b.itypes, b.rij, b.jtypes)
    F_t = O.Oracle.forces(g_t, b.js, b.natoms)
    desc = jm.MTPPotential(desc_p, max_atoms=b.natoms, max_nbr=b.nslots)
    tr = jm.head.HeadTrainer(desc, b, E_site, F_t, weights=(1.0, 0.5), seed=0)
    L = tr.loss()
    L.backward()
    g_gpu = tr.theta.grad.detach().cpu().numpy()
    # oracle: dL/dE, dL/dF -> v = J^T Delta -> parameter gradient
    orc_d = oracle_for(O, desc_p)
    tm = tr.model
    assert tm.residual
    m = _M(tm.theta, tm.mean, tm.inv_std, tm.lo, tm.hi, tm.e_shift, tm.e_scale, orc_d._keep["moment"],
           orc_d._keep["species"])
    _, _, basis = orc_d.compact_xi(b.itypes, b.rij, b.jtypes, np.zeros(S12), grad=False)
    E, xi, _ = HO.eval_head(m, basis, b.itypes)
    _, g, _ = orc_d.compact_xi(b.itypes, b.rij, b.jtypes, xi)
    F = O.Oracle.forces(g, b.js, b.natoms)
    N = b.natoms
    gE = 2 * 1.0 * (E - E_site) / N
    gF = 2 * 0.5 * (F - F_t) / (3 * N)
    J = orc_d.basis_jacobian(b.itypes, b.rij, b.jtypes)
    gFx = np.vstack([gF, np.zeros((1, 3))])
    js = np.where(b.js < N, b.js, N)
    delta = gF[:, None, :] - gFx[js]
    v = np.einsum("askc,akc->as", J, delta)
    g_ref = HO.param_grad(m, basis, b.itypes, gE, v)
    err = _rel(g_gpu, g_ref)
    print(f"  distillation gradient rel err {err:.3e}", flush=True)
    assert err <= 1e-10
    l0 = float(tr.loss().detach())
    tr.fit(steps=150, lr=3e-3)
    l1 = float(tr.loss().detach())
    print(f"  distillation loss {l0:.4e} -> {l1:.4e}", flush=True)
    assert l1 < 0.5 * l0
    tr.head.close()
    desc.close()


@pytest.mark.gpu
def test_gpu_head_input_width_limit(jm):
    """A descriptor whose basis values and species exceed the 64 input lanes
    (level 16, 3 species: 92 + 3) is refused at open."""
    rng = np.random.default_rng(11)
    desc = jm.configs.random_params(16, SP, 8, 2.0, 5.0, rng)
    S = len(desc.moment_coeffs)
    assert S + SP > 64
    hm = jm.head.HeadModel(desc, jm.head.init_theta(min(S, 60), SP, rng), np.zeros(S), np.ones(S), None, None,
                           np.zeros(SP), 1.0)
    with pytest.raises(jm.capi.MTPError) as e:
        jm.head.DistilledHead(hm, None, max_atoms=8, max_nbr=16)
    assert e.value.code == jm.capi.MTP_ERR_UNSUPPORTED


@pytest.mark.gpu
def test_gpu_distillation_gain(jm):
    """The trained network against the method without it: on a held-out cell
    the head's force RMSE (vs the level-24 teacher) is well below that of the
    level-12 MTP alone (same inputs, least-squares coefficients)."""
    # decay 0.5: a teacher whose levels above 12 carry a visible share of the forces
    res = jm.distill.run(n_env=16384, infer_ncell=(6, 6, 6), head_steps=1000, decay=0.5, seed=0,
                         log=lambda s: print(s, flush=True), desc_level=12)
    print({k: res[k] for k in ("force_rmse_head_meV_per_A", "force_rmse_desc_meV_per_A",
                               "force_rms_exact_eV_per_A", "n_fallback", "speedup")}, flush=True)
    assert res["force_rmse_head_meV_per_A"] < 0.8 * res["force_rmse_desc_meV_per_A"]
    assert res["force_rmse_head_meV_per_A"] < 0.15 * 1e3 * res["force_rms_exact_eV_per_A"]
