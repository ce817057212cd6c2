"""GPU parity: the HIP path through the C ABI against the FP64 CPU oracle.

Tolerance (FP64 path): norm-wise relative error <= 1e-10 for the energy,
the forces (max-abs error / max-abs reference force) and the stress.
FP32-radial variant: its error against the FP64 oracle is reported and
bounded by 1e-4 (norm-wise) on forces.
"""
import numpy as np
import pytest

from conftest import oracle_for

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _rel(a, b):
    a = np.asarray(a, dtype=float); b = np.asarray(b, dtype=float)
    den = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / den) if den > 0 else float(np.max(np.abs(a - b)))


def _run_gpu(jm, w, nb, radial_fp32=False, n_e=None, n_f=None, rows=None, variant=None):
    A = nb["js"].shape[0]
    pot = jm.MTPPotential(w.params, max_atoms=A, max_nbr=nb["js"].shape[1], radial_fp32=radial_fp32)
    if variant is not None:
        pot.set_variant(variant)
    n_e = w.natoms if n_e is None else n_e
    n_f = w.natoms if n_f is None else n_f
    E, F, s6, extra = pot.calc_energy_forces_stress_padded(nb["itypes"], nb["js"], nb["rij"], nb["jtypes"], 3,
                                                           w.volume, n_e, n_f, return_all=True)
    out = dict(E=float(E.item()), F=F.cpu().numpy(), s6=s6.cpu().numpy(), v6=extra["virial6"].cpu().numpy(),
               eatom=extra["eatom"].cpu().numpy(), g=extra["g"].cpu().numpy(), status=float(extra["status"].item()))
    pot.close()
    return out


def _run_oracle(O, w, nb, n_e=None, n_f=None):
    orc = oracle_for(O, w.params)
    n_e = w.natoms if n_e is None else n_e
    n_f = w.natoms if n_f is None else n_f
    E, F, s6, Ei, g = orc.calc_energy_forces_stress_padded(nb["itypes"], nb["js"], nb["rij"], nb["jtypes"], 3,
                                                           w.volume, n_e, n_f)
    return dict(E=E, F=F, s6=s6, eatom=Ei, g=g)


def _run_potential(pot, w, nb):
    E, F, s6, extra = pot.calc_energy_forces_stress_padded(nb["itypes"], nb["js"], nb["rij"], nb["jtypes"], 3,
                                                           w.volume, w.natoms, w.natoms, return_all=True)
    return dict(E=float(E.item()), F=F.cpu().numpy(), s6=s6.cpu().numpy(), eatom=extra["eatom"].cpu().numpy(),
                g=extra["g"].cpu().numpy(), status=float(extra["status"].item()))


def _check(gpu, ref, tol=TOL):
    print(f"E {gpu['E']:.17g} ref {ref['E']:.17g} | rel F {_rel(gpu['F'], ref['F']):.3e} "
          f"| rel s6 {_rel(gpu['s6'], ref['s6']):.3e} | rel Ei {_rel(gpu['eatom'], ref['eatom']):.3e}", flush=True)
    assert gpu["status"] == 0.0
    assert abs(gpu["E"] - ref["E"]) <= tol * max(abs(ref["E"]), np.max(np.abs(ref["eatom"])))
    assert _rel(gpu["eatom"], ref["eatom"]) <= tol
    assert _rel(gpu["g"], ref["g"]) <= tol
    assert _rel(gpu["F"], ref["F"]) <= tol
    assert _rel(gpu["s6"], ref["s6"]) <= tol


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4, 5, 6])
@pytest.mark.parametrize("level", [2, 4, 6, 8, 10, 12, 14, 16, 18, 20, 22, 24])
def test_levels_random_cells(jm, O, level, variant):
    w = jm.configs.small_cluster(level, 3, 24, seed=100 + level, cell=8.0, cutoff=5.0, max_nbr=80)
    nb = w.neighbors()
    print(f"[case] level={level} variant={variant} natoms={w.natoms} maxn={nb['counts'].max()}", flush=True)
    _check(_run_gpu(jm, w, nb, variant=variant), _run_oracle(O, w, nb))


def _wide_mu_params(jm, seed):
    """A potential whose basic moments carry eight radial functions at degree 0 and seven at degree 1
    (every shipped level has at most six): the widest monomial-column shape of the 256- and
    512-thread kernels (eight mu per monomial).  Scalars: the eight degree-0 moments and four dot
    products of degree-1 moments, one of them a square."""
    rng = np.random.default_rng(seed)
    p = jm.configs.random_params(2, 2, 8, 2.0, 5.0, rng)
    R = 8
    basic = [[mu, 0, 0, 0] for mu in range(8)]
    for mu in range(7):
        basic += [[mu, 1, 0, 0], [mu, 0, 1, 0], [mu, 0, 0, 1]]
    vec = lambda mu: 8 + 3 * mu  # first component of M_{mu,1}
    pairs = [(0, 1), (2, 3), (4, 6), (5, 5)]
    times = []
    for q, (a, b) in enumerate(pairs):
        t = len(basic) + q
        for c in range(3):
            times.append([vec(a) + c, vec(b) + c, 1, t])
    p.radial_funcs_count = R
    p.alpha_index_basic = np.array(basic, dtype=np.int64)
    p.alpha_index_basic_count = len(basic)
    p.alpha_index_times = np.array(times, dtype=np.int64)
    p.alpha_index_times_count = len(times)
    p.alpha_moments_count = len(basic) + len(pairs)
    p.alpha_moment_mapping = np.array(list(range(8)) + [len(basic) + q for q in range(len(pairs))], dtype=np.int64)
    p.alpha_scalar_moments = len(p.alpha_moment_mapping)
    p.moment_coeffs = rng.normal(0.0, 0.1, p.alpha_scalar_moments)
    p.radial_coeffs = rng.normal(0.0, 0.1, (2, 2, R, 8))
    return p


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4, 5, 6])
def test_eight_radial_functions(jm, O, variant):
    w = jm.configs.small_cluster(2, 2, 24, seed=71, cell=8.0, cutoff=5.0, max_nbr=80)
    w.params = _wide_mu_params(jm, 5)
    nb = w.neighbors()
    print(f"[case] eight radial functions variant={variant} natoms={w.natoms} maxn={nb['counts'].max()}", flush=True)
    _check(_run_gpu(jm, w, nb, variant=variant), _run_oracle(O, w, nb))


def test_fcc256_level10(jm, O):
    w = jm.configs.fcc256_level10(seed=0)
    nb = w.neighbors()
    print(f"[case] fcc256 level10 maxn={nb['counts'].max()}", flush=True)
    gpu, ref = _run_gpu(jm, w, nb), _run_oracle(O, w, nb)
    _check(gpu, ref)
    # periodic cell: total force vanishes
    assert np.abs(gpu["F"].sum(axis=0)).max() <= 1e-10 * np.abs(gpu["F"]).max()


def test_bcc2889_level16(jm, O):
    w = jm.configs.bcc2889_level16(seed=0)
    nb = w.neighbors()
    print(f"[case] bcc2889 level16 maxn={nb['counts'].max()}", flush=True)
    _check(_run_gpu(jm, w, nb), _run_oracle(O, w, nb))


@pytest.mark.parametrize("variant", [1, 2, 4, 5])
@pytest.mark.parametrize("rows", [24, 25])
def test_masks_and_padding_rows(jm, O, rows, variant):
    """natoms_energy below the row count and natoms_force different from it; an odd row count leaves the
    paired-row kernels' last workgroup with one row.

    The reference's padded entry hands natoms_energy to calc_energy_forces_stress's natoms_force parameter
    (jax_backend.py:414-418), so accumulate_forces (:279-301) masks rows and Newton-3 targets by
    natoms_energy and natoms_force changes nothing.  The expected forces are built here from the oracle's
    slot gradients with that mask, apart from the oracle's padded entry, and a second call with another
    natoms_force must give the same bits."""
    w = jm.configs.small_cluster(12, 2, 20, seed=9, cell=7.5, cutoff=5.0, max_nbr=64)
    nb = jm.neighbors.padded_neighbors(w.pos, w.types, w.cell, w.cutoff, w.max_nbr, rows=rows)
    print(f"[case] masks rows={rows} natoms=20 n_e=17 n_f=15 variant={variant}", flush=True)
    gpu = _run_gpu(jm, w, nb, n_e=17, n_f=15, variant=variant)
    orc = oracle_for(O, w.params)
    Ei, g = orc.compact(nb["itypes"], nb["rij"], nb["jtypes"])
    js = nb["js"]
    A, K = js.shape
    F = np.zeros((A, 3))
    for i in range(17):  # F_a = [a < n_e] sum_k g_ak - sum_{i < n_e, k: j_ik = a < n_e} g_ik
        F[i] += g[i].sum(axis=0)
        for k in range(K):
            if js[i, k] < 17:
                F[js[i, k]] -= g[i, k]
    s6, _ = O.Oracle.stress(nb["rij"], g, w.volume, 3)
    ref = dict(E=float(Ei[:17].sum()), F=F, s6=s6, eatom=Ei, g=g)
    _check(gpu, ref)
    assert np.all(gpu["F"][17:] == 0.0)
    other = _run_gpu(jm, w, nb, n_e=17, n_f=rows, variant=variant)  # natoms_force has no effect
    np.testing.assert_array_equal(other["F"], gpu["F"])
    assert other["E"] == gpu["E"]


def test_nonfinite_slot_is_an_error(jm, O):
    """A NaN or Inf component in one neighbor vector makes the reference's energy and forces NaN
    (jax_backend.py:83-95, 232).  Here the host-rows entry returns MTP_ERR_NONFINITE and the device entry
    sets status bit 2 with a NaN energy; the next clean call is clean again (the flag is cleared)."""
    w = jm.configs.small_cluster(10, 2, 16, seed=3, cell=7.0, cutoff=5.0, max_nbr=48)
    nb = w.neighbors()
    pot = jm.MTPPotential(w.params, max_atoms=w.natoms, max_nbr=48)
    for bad in (np.nan, np.inf):
        rij = nb["rij"].copy()
        rij[3, 0, 1] = bad  # slot 0 of row 3 holds a typed neighbor
        assert nb["jtypes"][3, 0] >= 0
        print(f"[case] non-finite slot {bad}", flush=True)
        with pytest.raises(jm.capi.MTPError) as e:
            pot.compute_rows(w.natoms, rij.reshape(w.natoms, -1), nb["itypes"], nb["js"], nb["counts"],
                             nb["jtypes"], w.volume)
        assert e.value.code == jm.capi.MTP_ERR_NONFINITE
        E, F, s6, extra = pot.calc_energy_forces_stress_padded(nb["itypes"], nb["js"], rij, nb["jtypes"], 3,
                                                               w.volume, w.natoms, w.natoms, return_all=True)
        assert int(extra["status"].item()) & 2 and np.isnan(float(E.item()))
    E, F, s6, extra = pot.calc_energy_forces_stress_padded(nb["itypes"], nb["js"], nb["rij"], nb["jtypes"], 3,
                                                           w.volume, w.natoms, w.natoms, return_all=True)
    assert float(extra["status"].item()) == 0.0
    ref = _run_oracle(O, w, nb)
    assert abs(float(E.item()) - ref["E"]) <= TOL * np.max(np.abs(ref["eatom"]))
    pot.close()


def test_repeated_calls_graph_replay(jm, O):
    """Repeated calls with the same arrays run as a recorded HIP graph (the
    second call records it, later ones replay it).  Replays read the arrays'
    current contents and parameters: positions changed in place, new moment
    coefficients and a new scaling all show up, each against the oracle."""
    import copy
    import ctypes

    import torch

    w = jm.configs.small_cluster(14, 2, 24, seed=41, cell=8.0, cutoff=5.0, max_nbr=64)
    nb = jm.neighbors.padded_neighbors(w.pos, w.types, w.cell, w.cutoff, w.max_nbr, rows=24)
    A, K = nb["js"].shape
    dev = torch.device("cuda")
    it = torch.as_tensor(nb["itypes"], dtype=torch.int32, device=dev)
    js = torch.as_tensor(nb["js"], dtype=torch.int32, device=dev)
    rij = torch.as_tensor(nb["rij"], dtype=torch.float64, device=dev)
    jt = torch.as_tensor(nb["jtypes"], dtype=torch.int32, device=dev)
    out = torch.empty(16, dtype=torch.float64, device=dev)
    F = torch.empty((A, 3), dtype=torch.float64, device=dev)
    pot = jm.MTPPotential(w.params, max_atoms=A, max_nbr=K)
    L = jm.capi.lib()
    stream = torch.cuda.current_stream().cuda_stream

    def call():
        jm.capi.check(L.mtp_compute_padded_device(pot.ctx, A, K, it.data_ptr(), js.data_ptr(), rij.data_ptr(),
                                                  jt.data_ptr(), 3, w.volume, A, A, out.data_ptr(), F.data_ptr(),
                                                  None, stream), pot.ctx)
        torch.cuda.synchronize()
        return float(out[0].item()), F.cpu().numpy().copy()

    def check(res, params, rij_np, tag):
        n2 = dict(nb, rij=rij_np)
        ref = _run_oracle(O, jm.configs.Workload(w.name, params, w.pos, w.types, w.cell, w.cutoff, w.max_nbr), n2)
        print(f"[case] graph replay {tag}: rel F {_rel(res[1], ref['F']):.3e}", flush=True)
        assert _rel(res[1], ref["F"]) <= TOL, tag
        assert abs(res[0] - ref["E"]) <= TOL * np.max(np.abs(ref["eatom"])), tag

    first = call()
    check(first, w.params, nb["rij"], "eager")
    for q in range(3):  # records, then replays
        again = call()
        assert again[0] == first[0] and np.array_equal(again[1], first[1]), q
    rng = np.random.default_rng(3)
    moved = nb["rij"] + np.where(nb["jtypes"][..., None] >= 0, rng.normal(0, 0.02, nb["rij"].shape), 0.0)
    rij.copy_(torch.as_tensor(moved, dtype=torch.float64, device=dev))
    check(call(), w.params, moved, "positions changed in place")
    p2 = copy.deepcopy(w.params)
    p2.moment_coeffs = np.asarray(p2.moment_coeffs) * 1.5 + 0.01
    pot.set_params(moment_coeffs=p2.moment_coeffs)
    check(call(), p2, moved, "moment coefficients")
    p2.scaling = 1.3
    pot.set_params(scaling=1.3)
    check(call(), p2, moved, "scaling")
    check(call(), p2, moved, "scaling, replayed")
    pot.close()


def test_reverse_neighbor_reuse(jm, O):
    """One potential over a sequence of calls: the reverse-neighbor CSR of the
    force gather is reused when js, the force mask (natoms_energy) and the
    shape repeat, and rebuilt when any of them changes (a different neighbor
    list of the same shape, a smaller natoms_energy)."""
    wa = jm.configs.small_cluster(12, 2, 24, seed=31, cell=8.0, cutoff=5.0, max_nbr=64)
    wb = jm.configs.small_cluster(12, 2, 24, seed=32, cell=8.0, cutoff=5.0, max_nbr=64)
    wb.params = wa.params
    na = jm.neighbors.padded_neighbors(wa.pos, wa.types, wa.cell, wa.cutoff, wa.max_nbr, rows=24)
    nb = jm.neighbors.padded_neighbors(wb.pos, wb.types, wb.cell, wb.cutoff, wb.max_nbr, rows=24)
    assert na["js"].shape == nb["js"].shape and not np.array_equal(na["js"], nb["js"])
    pot = jm.MTPPotential(wa.params, max_atoms=24, max_nbr=64)
    seq = [(wa, na, 24), (wa, na, 24), (wb, nb, 24), (wa, na, 24), (wa, na, 19), (wa, na, 19), (wb, nb, 19)]
    for q, (w, nbr, n_e) in enumerate(seq):
        print(f"[case] reuse step {q} n_e={n_e}", flush=True)
        E, F, s6, extra = pot.calc_energy_forces_stress_padded(nbr["itypes"], nbr["js"], nbr["rij"], nbr["jtypes"], 3,
                                                               w.volume, n_e, 24, return_all=True)
        ref = _run_oracle(O, w, nbr, n_e=n_e, n_f=24)
        assert _rel(F.cpu().numpy(), ref["F"]) <= TOL, q
        assert abs(float(E.item()) - ref["E"]) <= TOL * np.max(np.abs(ref["eatom"])), q
    pot.close()


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4, 5, 6])
def test_fcc32k_level24_full_size(jm, O, variant):
    """The 32,000-atom level-24 workload at full size against the oracle."""
    w = jm.configs.fcc32k_level24(seed=0)
    nb = w.neighbors()
    print(f"[case] fcc32k level24 variant={variant} maxn={nb['counts'].max()}", flush=True)
    gpu = _run_gpu(jm, w, nb, variant=variant)
    ref = _fcc32k_reference(O, w, nb)
    _check(gpu, ref)
    assert np.abs(gpu["F"].sum(axis=0)).max() <= 1e-9 * np.abs(gpu["F"]).max()


_REF_CACHE = {}


def _fcc32k_reference(O, w, nb):
    if "fcc32k" not in _REF_CACHE:
        _REF_CACHE["fcc32k"] = _run_oracle(O, w, nb)
    return _REF_CACHE["fcc32k"]


def test_radial_fp32_variant_error(jm, O):
    w = jm.configs.fcc32k_level24(seed=0, ncell=6)
    nb = w.neighbors()
    print(f"[case] fp32-radial fcc{w.natoms} level24", flush=True)
    gpu = _run_gpu(jm, w, nb, radial_fp32=True)
    ref = _run_oracle(O, w, nb)
    errF = _rel(gpu["F"], ref["F"])
    errE = abs(gpu["E"] - ref["E"]) / np.max(np.abs(ref["eatom"]))
    print(f"FP32-radial: force err {errF:.3e}, energy err {errE:.3e}", flush=True)
    assert errF < 1e-4 and errE < 1e-4


def test_host_rows_api_matches_device_api(jm):
    """mtp_compute (pair-style row contract) equals the device entry bit for bit."""
    w = jm.configs.small_cluster(16, 2, 30, seed=4, cell=8.5, cutoff=5.0, max_nbr=64)
    nb = w.neighbors()
    pot = jm.MTPPotential(w.params, max_atoms=w.natoms, max_nbr=64)
    E, F, s6, v6 = pot.compute_rows(w.natoms, nb["rij"].reshape(w.natoms, -1), nb["itypes"], nb["js"], nb["counts"],
                                    nb["jtypes"], w.volume)
    E2, F2, s62, extra = pot.calc_energy_forces_stress_padded(nb["itypes"], nb["js"], nb["rij"], nb["jtypes"], 3,
                                                              w.volume, w.natoms, w.natoms, return_all=True)
    assert E == float(E2.item())
    np.testing.assert_array_equal(F, F2.cpu().numpy())
    np.testing.assert_array_equal(s6, s62.cpu().numpy())
    np.testing.assert_array_equal(v6, extra["virial6"].cpu().numpy())
    # LAMMPS virial = -V * stress, reordered xx yy zz xy xz yz
    np.testing.assert_allclose(v6, -w.volume * s6[[0, 1, 2, 5, 4, 3]], rtol=1e-12, atol=1e-14)
    pot.close()


def test_per_atom_energy_and_virial(jm, O):
    """mtp_last_atom_values (the pair style's eatom / vatom): E_i against the oracle, the per-atom
    virial against -sym(sum_k r_ik (x) g_ik) of the oracle's slot gradients, and both summing to the
    totals of the same mtp_compute."""
    w = jm.configs.small_cluster(24, 3, 30, seed=6, cell=8.5, cutoff=5.0, max_nbr=64)
    nb = w.neighbors()
    pot = jm.MTPPotential(w.params, max_atoms=w.natoms, max_nbr=64)
    E, F, s6, v6 = pot.compute_rows(w.natoms, nb["rij"].reshape(w.natoms, -1), nb["itypes"], nb["js"], nb["counts"],
                                    nb["jtypes"], w.volume)
    eatom, vatom = pot.last_atom_values(w.natoms)
    part_e, _ = pot.last_atom_values(w.natoms // 2)  # a prefix of the same rows
    np.testing.assert_array_equal(part_e, eatom[: w.natoms // 2])
    ref = _run_oracle(O, w, nb)
    assert _rel(eatom, ref["eatom"]) <= TOL
    W = np.einsum("ika,ikb->iab", nb["rij"], ref["g"])  # per-row virial of the oracle's gradients
    vref = -np.stack([W[:, 0, 0], W[:, 1, 1], W[:, 2, 2], 0.5 * (W[:, 0, 1] + W[:, 1, 0]),
                      0.5 * (W[:, 0, 2] + W[:, 2, 0]), 0.5 * (W[:, 1, 2] + W[:, 2, 1])], axis=1)
    assert _rel(vatom, vref) <= TOL
    np.testing.assert_allclose(eatom.sum(), E, rtol=1e-12, atol=1e-12 * np.abs(eatom).max())
    np.testing.assert_allclose(vatom.sum(axis=0), v6, rtol=1e-12, atol=1e-12 * np.abs(vatom).max())
    with pytest.raises(RuntimeError):
        pot.last_atom_values(w.natoms + 1)
    # a device entry point overwrites the per-row buffers: the values are no longer offered
    pot.calc_energy_forces_stress_padded(nb["itypes"], nb["js"], nb["rij"], nb["jtypes"], 3, w.volume,
                                         w.natoms, w.natoms)
    with pytest.raises(RuntimeError):
        pot.last_atom_values(1)
    pot.close()


def test_deterministic(jm):
    w = jm.configs.small_cluster(20, 2, 40, seed=8, cell=9.0, cutoff=5.0, max_nbr=80)
    nb = w.neighbors()
    a = _run_gpu(jm, w, nb)
    b = _run_gpu(jm, w, nb)
    assert a["E"] == b["E"]
    np.testing.assert_array_equal(a["F"], b["F"])
    np.testing.assert_array_equal(a["s6"], b["s6"])


@pytest.mark.parametrize("variant", [1, 2, 3])
def test_widest_rows(jm, O, variant):
    """K = 256 slots (the largest row the kernels take), 195-237 neighbors per
    atom: four slot chunks for the 64-thread kernel, one slot per thread for
    256.  (The paired-row variants do not fit this row width in LDS and are
    reported unavailable.)"""
    rng = np.random.default_rng(12)
    n, cell = 150, 10.0
    pos = rng.uniform(0, cell, (n, 3))
    types = rng.integers(0, 2, n)
    params = jm.configs.random_params(12, 2, 8, 0.5, 7.0, rng, sigma=0.02)
    w = jm.configs.Workload("dense150", params, pos, types, np.array([cell] * 3), 7.0, 256)
    nb = jm.neighbors.padded_neighbors(pos, types, w.cell, 7.0, 256)
    print(f"[case] widest rows K=256 maxn={nb['counts'].max()} variant={variant}", flush=True)
    assert nb["counts"].max() > 192
    _check(_run_gpu(jm, w, nb, variant=variant), _run_oracle(O, w, nb))


@pytest.mark.parametrize("variant", [3, 1, 2])
def test_empty_and_single_slot(jm, O, variant):
    """No rows (energy = 0, nothing launched but the totals), rows with a single neighbor and rows with
    none, on the 64-thread kernel (one monomial per thread) and the 256- and 512-thread kernels
    (monomial columns: a quad with one slot pair or none)."""
    import torch

    w = jm.configs.small_cluster(10, 1, 6, seed=5, cell=6.0, cutoff=5.0, max_nbr=40)
    pot = jm.MTPPotential(w.params, max_atoms=8, max_nbr=8)
    dev = torch.device("cuda")
    z = lambda shape, dt: torch.zeros(shape, dtype=dt, device=dev)
    E, F, s6 = pot.calc_energy_forces_stress_padded(z((0,), torch.int32), z((0, 1), torch.int32),
                                                    z((0, 1, 3), torch.float64), z((0, 1), torch.int32), 3,
                                                    w.volume, 0, 0)
    assert float(E.item()) == 0.0 and F.shape[0] == 0
    pot.close()
    # rows of two slots (the smallest width the pair style allows), one neighbor and one empty slot
    nb = jm.neighbors.padded_neighbors(w.pos, w.types, w.cell, 5.0, 40)
    assert nb["counts"].max() < 40  # slot 39 is empty in every row
    keep = np.argmax(nb["jtypes"] >= 0, axis=1)
    A = nb["js"].shape[0]
    cols = np.stack([keep, np.full(A, 39)], axis=1)
    cols[::3, 0] = 39  # every third row has no neighbor at all
    rows = np.arange(A)[:, None]
    one = dict(nb, js=nb["js"][rows, cols].copy(), jtypes=nb["jtypes"][rows, cols].copy(),
               rij=nb["rij"][rows, cols].copy())
    w1 = jm.configs.Workload(w.name, w.params, w.pos, w.types, w.cell, w.cutoff, 2)
    print(f"[case] single and empty slots, variant {variant}", flush=True)
    _check(_run_gpu(jm, w1, one, variant=variant), _run_oracle(O, w1, one))


def test_capacity_errors(jm):
    w = jm.configs.small_cluster(10, 1, 8, seed=1, cell=6.0, cutoff=5.0, max_nbr=40)
    nb = w.neighbors()
    pot = jm.MTPPotential(w.params, max_atoms=4, max_nbr=40)
    with pytest.raises(jm.capi.MTPError) as e:
        pot.calc_energy_forces_stress_padded(nb["itypes"], nb["js"], nb["rij"], nb["jtypes"], 3, w.volume, 8, 8)
    assert e.value.code == jm.capi.MTP_ERR_CAPACITY
    pot.close()


def test_level26_and_level28(jm, O):
    """Levels 26 and 28 (alpha tables of the reference's untrained_mtps/26.mtp, 28.mtp; no JSON trees exist
    for them, so the oracle is the compact form).  Level 26's moments take more than 64 KB of LDS, so its
    contraction streams address operands by 8-byte slot index; the stream kernels (one workgroup per CU at
    512 threads) are the default, and every single-row variant, the per-target kernel included, is
    checked.  Level 28's moments and adjoints (17,366 of each, 290 KB) exceed the LDS: the per-target
    kernel keeps them in a global slice per resident workgroup, and a fixed grid of workgroups walks the
    rows.  Checked on a small cluster and on a 1,372-atom FCC cell, more rows than the grid, so that
    workgroups reuse their slices."""
    rng = np.random.default_rng(26)
    p = jm.mtp_io.untrained(26)
    p.species_count, p.radial_basis_size, p.min_dist, p.max_dist = 2, 8, 1.5, 5.0
    p.species_coeffs = rng.normal(0, 0.1, 2)
    p.moment_coeffs = rng.normal(0, 0.1, len(p.alpha_moment_mapping))
    p.radial_coeffs = rng.normal(0, 0.1, (2, 2, p.radial_funcs_count, 8))
    w = jm.configs.small_cluster(10, 2, 12, seed=26, cell=7.0, cutoff=5.0, max_nbr=80)
    w.params = p
    nb = w.neighbors()
    ref = _run_oracle(O, w, nb)
    pot = jm.MTPPotential(p, max_atoms=w.natoms, max_nbr=80)
    assert pot.variant == 2 and pot.info()["level"] == 26
    for variant in (2, 1, 3, 0):
        print(f"[case] level 26, {w.natoms} atoms, variant {variant}", flush=True)
        pot.set_variant(variant)
        _check(_run_potential(pot, w, nb), ref)
    pot.close()

    q = jm.mtp_io.untrained(28)
    q.species_count, q.radial_basis_size, q.min_dist, q.max_dist = 2, 8, 1.5, 5.0
    q.species_coeffs = rng.normal(0, 0.1, 2)
    q.moment_coeffs = rng.normal(0, 0.1, len(q.alpha_moment_mapping))
    q.radial_coeffs = rng.normal(0, 0.1, (2, 2, q.radial_funcs_count, 8))
    pos, cell = jm.configs.fcc(3.6, (7, 7, 7))
    cases = [jm.configs.small_cluster(10, 2, 12, seed=28, cell=7.0, cutoff=5.0, max_nbr=80),
             jm.configs.Workload("fcc1372_level28", None, pos + rng.normal(0.0, 0.05, pos.shape),
                                 rng.integers(0, 2, len(pos)).astype(np.int32), cell, 5.0, 80)]
    for w in cases:
        w.params = q
        nb = w.neighbors()
        print(f"[case] level 28, {w.natoms} atoms", flush=True)
        pot = jm.MTPPotential(q, max_atoms=w.natoms, max_nbr=80)
        assert pot.variant == 0 and pot.info()["level"] == 28
        gpu = _run_potential(pot, w, nb)
        pot.close()
        _check(gpu, _run_oracle(O, w, nb))
