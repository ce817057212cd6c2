"""Training path: packed batches, the loss, and the hand-written parameter gradient.

CPU part pins the exact parameter-gradient oracle (5-point stencil of the loss
linearisation, exact because the linearisation is a polynomial of degree <= 4
along each coordinate) against plain finite differences of the loss itself.
GPU part compares the HIP batch forward and parameter backward with it
(norm-wise relative error <= 1e-10) and runs a few optimizer steps.
"""
import numpy as np
import pytest

from conftest import oracle_for


def _batch(jm, nconf=4, seed=0, level=16, species=2, natoms=(8, 40)):
    rng = np.random.default_rng(seed)
    cells = jm.training.random_cells(nconf, rng, natoms_range=natoms, species=species)
    batch = jm.training.pack_configs(cells, cutoff=5.0, max_nbr=96)
    params = jm.configs.random_params(level, species, 8, 1.6, 5.0, rng)
    return batch, params, rng


def _oracle_outputs(O, orc, b):
    return O.batch_outputs(orc, b.itypes, b.js, b.rij, b.jtypes, b.conf_off, b.volumes)


def test_pack_configs_layout(jm):
    b, p, rng = _batch(jm, nconf=5, seed=1)
    n = np.diff(b.conf_off)
    assert b.natoms == n.sum() and len(b.volumes) == 5
    assert np.all(b.conf_of == np.repeat(np.arange(5), n))
    # js point inside the own configuration or to the empty-slot marker N
    for c in range(5):
        rows = slice(b.conf_off[c], b.conf_off[c + 1])
        j = b.js[rows][b.jtypes[rows] >= 0]
        assert np.all((j >= b.conf_off[c]) & (j < b.conf_off[c + 1]))
    assert np.all(b.js[b.jtypes < 0] == b.natoms)


def test_linearisation_degree_at_most_4(jm, O):
    b, p, rng = _batch(jm, nconf=2, seed=2, level=12)
    orc = oracle_for(O, p)
    C, N = len(b.volumes), b.natoms
    gE, gF, gS = rng.normal(size=C), rng.normal(size=(N, 3)), rng.normal(size=(C, 6))
    ts = np.arange(-3, 4) * 0.3
    for which, q in [("radial", 5), ("radial", 17), ("moment", 3)]:
        v = O.J_along(orc, b.itypes, b.js, b.rij, b.jtypes, b.conf_off, b.volumes, gE, gF, gS, which, q, ts)
        d5 = np.diff(v, 5)
        assert np.max(np.abs(d5)) <= 1e-9 * np.max(np.abs(v)), (which, q, d5)


def test_exact_gradient_matches_loss_finite_differences(jm, O):
    b, p, rng = _batch(jm, nconf=3, seed=3, level=10)
    orc = oracle_for(O, p)
    E, F, S, g = _oracle_outputs(O, orc, b)
    E_ref = E + rng.normal(0, 0.1, E.shape)
    F_ref = F + rng.normal(0, 0.1, F.shape)
    S_ref = S + rng.normal(0, 0.01, S.shape)
    wE, wF, wS = jm.training.LOSS_WEIGHTS

    def loss():
        E, F, S, _ = _oracle_outputs(O, orc, b)
        return wE * np.sum((E - E_ref) ** 2) + wF * np.sum((F - F_ref) ** 2) + wS * np.sum((S - S_ref) ** 2)

    gE, gF, gS = 2 * wE * (E - E_ref), 2 * wF * (F - F_ref), 2 * wS * (S - S_ref)
    grad = O.param_grad_exact(orc, b.itypes, b.js, b.rij, b.jtypes, b.conf_off, b.volumes, gE, gF, gS)
    nS, sp = len(p.moment_coeffs), p.species_count
    assert grad.shape == (nS + sp + p.radial_coeffs.size,)
    arrays = [orc._keep["moment"], orc._keep["species"], orc._keep["radial"].reshape(-1)]
    offsets = [0, nS, nS + sp]
    # the loss is a degree-8 polynomial in the radial coefficients with large energies, so a plain
    # central difference needs a 4th-order stencil to get below 1e-6
    h = 1e-6
    for a, off in zip(arrays, offsets):
        for q in list(range(0, a.size, max(1, a.size // 6)))[:6]:
            x0 = a[q]
            v = []
            for t in (-2, -1, 1, 2):
                a[q] = x0 + t * h
                v.append(loss())
            a[q] = x0
            fd = (v[0] - 8 * v[1] + 8 * v[2] - v[3]) / (12 * h)
            assert grad[off + q] == pytest.approx(fd, rel=1e-6, abs=1e-7 * np.max(np.abs(grad)))


@pytest.mark.gpu
# (16, 8): 256 radial-coefficient columns per row, one per thread of the
# kernel's final reduction; the others split slots over thread groups
@pytest.mark.parametrize("level,species", [(2, 1), (6, 3), (10, 1), (12, 3), (16, 2), (16, 8), (20, 2)])
def test_gpu_batch_forward_and_param_grad(jm, O, level, species):
    import torch

    b, p, rng = _batch(jm, nconf=5, seed=level, level=level, species=species, natoms=(8, 48))
    print(f"[case] training level={level} species={species} natoms={b.natoms}", flush=True)
    orc = oracle_for(O, p)
    E, F, S, g = _oracle_outputs(O, orc, b)
    pot = jm.MTPPotential(p, max_atoms=b.natoms, max_nbr=b.nslots)
    db = b.to_device()
    Eg, Fg, Sg = jm.training.batch_forward(pot, db)
    torch.cuda.synchronize()
    rel = lambda a, r: float(np.max(np.abs(a - r)) / np.max(np.abs(r)))
    assert rel(Eg.cpu().numpy(), E) <= 1e-10
    assert rel(Fg.cpu().numpy(), F) <= 1e-10
    assert rel(Sg.cpu().numpy(), S) <= 1e-10
    C, N = len(b.volumes), b.natoms
    gE, gF, gS = rng.normal(size=C), rng.normal(size=(N, 3)), rng.normal(size=(C, 6))
    ref = O.param_grad_exact(orc, b.itypes, b.js, b.rij, b.jtypes, b.conf_off, b.volumes, gE, gF, gS)
    dev = lambda x: torch.as_tensor(x, dtype=torch.float64, device="cuda")
    grad = jm.training.batch_param_grad(pot, db, dev(gE), dev(gF), dev(gS)).cpu().numpy()
    nS, sp = len(p.moment_coeffs), p.species_count
    for name, sl in [("moment", slice(0, nS)), ("species", slice(nS, nS + sp)), ("radial", slice(nS + sp, None))]:
        err = rel(grad[sl], ref[sl])
        print(f"  grad {name}: rel err {err:.3e}", flush=True)
        assert err <= 1e-10, name
    pot.close()


@pytest.mark.gpu
def test_gpu_training_reduces_loss(jm):
    """Labels from a seeded higher-level teacher; a few L-BFGS steps on the hand-written gradient."""
    import torch

    rng = np.random.default_rng(5)
    cells = jm.training.random_cells(8, rng, natoms_range=(8, 32), species=2)
    b = jm.training.pack_configs(cells, cutoff=5.0, max_nbr=96)
    teacher = jm.configs.random_params(12, 2, 8, 1.6, 5.0, rng, sigma=0.05)
    tp = jm.MTPPotential(teacher, max_atoms=b.natoms, max_nbr=b.nslots)
    E_ref, F_ref, S_ref = [t.clone() for t in jm.training.batch_forward(tp, b.to_device())]
    tp.close()
    student = jm.configs.random_params(10, 2, 8, 1.6, 5.0, rng, sigma=0.05)
    pot = jm.MTPPotential(student, max_atoms=b.natoms, max_nbr=b.nslots)
    tr0 = jm.training.Trainer(pot, b, E_ref, F_ref, S_ref, lls=False)
    l_init = float(tr0.loss().detach())
    tr = jm.training.Trainer(pot, b, E_ref, F_ref, S_ref, lls=True)
    l0 = float(tr.loss().detach())
    losses = [tr.step() for _ in range(8)]
    print("L-BFGS losses", losses, flush=True)
    l1 = float(tr.loss().detach())
    print(f"loss {l_init:.6e} -> LLS {l0:.6e} -> L-BFGS {l1:.6e}", flush=True)
    assert l0 < l_init
    assert l1 < l0
    pot.close()


@pytest.mark.gpu
def test_gpu_device_params_match_host(jm):
    import torch

    b, p, rng = _batch(jm, nconf=4, seed=11, level=16, species=2)
    p1 = jm.configs.random_params(16, 2, 8, 1.6, 5.0, rng)
    db = b.to_device()
    A = jm.MTPPotential(p, max_atoms=b.natoms, max_nbr=b.nslots)
    B = jm.MTPPotential(p, max_atoms=b.natoms, max_nbr=b.nslots)
    A.set_params(p1.moment_coeffs, p1.species_coeffs, p1.radial_coeffs)
    B.set_params_device(torch.tensor(jm.training.theta_of(p1), dtype=torch.float64, device="cuda"))
    for x, y in zip(jm.training.batch_forward(A, db), jm.training.batch_forward(B, db)):
        assert torch.equal(x, y)
    A.close()
    B.close()


@pytest.mark.gpu
def test_gpu_graphed_training_evaluation(jm):
    """Loss and gradient from the captured graph equal the eager autograd path on
    every replay, with eager GPU work between replays."""
    import torch

    rng = np.random.default_rng(12)
    cells = jm.training.random_cells(6, rng, natoms_range=(8, 40), species=2)
    b = jm.training.pack_configs(cells, cutoff=5.0, max_nbr=96)
    teacher = jm.configs.random_params(12, 2, 8, 1.6, 5.0, rng, sigma=0.05)
    tp = jm.MTPPotential(teacher, max_atoms=b.natoms, max_nbr=b.nslots)
    E_ref, F_ref, S_ref = [t.clone() for t in jm.training.batch_forward(tp, b.to_device())]
    tp.close()
    student = jm.configs.random_params(10, 2, 8, 1.6, 5.0, rng, sigma=0.05)
    pot = jm.MTPPotential(student, max_atoms=b.natoms, max_nbr=b.nslots)
    tr_e = jm.training.Trainer(pot, b, E_ref, F_ref, S_ref, lls=False, graph=False)
    tr_g = jm.training.Trainer(pot, b, E_ref, F_ref, S_ref, lls=False, graph=True)
    theta0 = tr_e.theta.detach().clone()
    for rep in range(4):
        th = theta0 + 0.01 * rep * torch.randn_like(theta0)
        Lg, gg = tr_g.value_and_grad(th)
        junk = torch.randn(4096, device="cuda").sum()  # eager work between replays
        Le, ge = tr_e.value_and_grad(th)
        print(f"  replay {rep}: loss graph {Lg:.12e} eager {Le:.12e}", flush=True)
        assert Lg == pytest.approx(Le, rel=1e-12)
        assert float((gg - ge).abs().max() / ge.abs().max()) <= 1e-12
        del junk
    # a short L-BFGS run on the graphed evaluation still lowers the loss
    l0 = tr_g.value_and_grad(tr_g.theta.detach())[0]
    for _ in range(3):
        l1 = tr_g.step()
    assert l1 < l0
    pot.close()


@pytest.mark.gpu
def test_gpu_graphed_force_evaluation(jm):
    """mtp_compute_padded_device captured in a graph: each replay equals the eager result."""
    import torch

    w = jm.configs.fcc256_level10(0)
    nb = w.neighbors()
    pot = jm.MTPPotential(w.params, max_atoms=w.natoms, max_nbr=w.max_nbr)
    dev = torch.device("cuda")
    rij = torch.as_tensor(nb["rij"], dtype=torch.float64, device=dev)
    it = torch.as_tensor(nb["itypes"], dtype=torch.int32, device=dev)
    js = torch.as_tensor(nb["js"], dtype=torch.int32, device=dev)
    jt = torch.as_tensor(nb["jtypes"], dtype=torch.int32, device=dev)
    static_r = rij.clone()
    args = lambda r: (it, js, r, jt, 3, w.volume, w.natoms, w.natoms)
    pot.calc_energy_forces_stress_padded(*args(static_r), sync=False)  # warm-up
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        E_s, F_s, S_s = pot.calc_energy_forces_stress_padded(*args(static_r), sync=False)
    gen = torch.Generator(device="cuda").manual_seed(0)
    for rep in range(4):
        r = rij + 0.01 * rep * torch.randn(rij.shape, dtype=torch.float64, device=dev, generator=gen)
        static_r.copy_(r)
        g.replay()
        E_g, F_g, S_g = E_s.clone(), F_s.clone(), S_s.clone()
        E_e, F_e, S_e = pot.calc_energy_forces_stress_padded(*args(r))
        assert torch.equal(E_g, E_e) and torch.equal(F_g, F_e) and torch.equal(S_g, S_e), rep
    pot.close()
