"""Parity at the benchmark's own sizes (GPU against the FP64 CPU oracle).

* The packed training step: the 512 configurations (8-256 atoms, FCC/BCC/
  random-dense, 2 species; 70,037 atoms) of the benchmark's training leg.
  Forward energies, forces and stresses of the level-16 student and of the
  level-20 teacher, and the hand-written parameter gradient of
  J = sum gE.E + sum gF.F + sum gS.S along random directions: along any
  direction J is a polynomial of degree <= 5 in the step (linear in the
  moment and species coefficients, degree <= 4 in the radial ones), so the
  6-point central stencil over J at t = +-h, +-2h, +-3h is exact up to
  rounding; the test checks that degree bound on the same 7 values.
  Reference: motep_jax_train_function.py:67-100, 191.
* The distilled head on the 100,000-atom 3-species FCC cell with exact
  level-24 re-evaluation of out-of-box atoms: energy, forces, stress,
  per-atom energies, flags and the fallback count against the oracle
  composition (head_oracle.head_compute).

Tolerance: norm-wise relative 1e-10 (FP64), 1e-9 for the directional
derivative (one number summed over 70k atoms and 222 parameters).
"""
import numpy as np
import pytest

from conftest import oracle_for

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a = np.asarray(a, dtype=float); b = np.asarray(b, dtype=float)
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


_BATCH = {}


def _bench_batch(jm):
    """The benchmark's training batch and potentials (bench.py training_leg, seed 0)."""
    if "b" not in _BATCH:
        rng = np.random.default_rng(0)
        cells = jm.training.random_cells(512, rng, natoms_range=(8, 256), species=2)
        b = jm.training.pack_configs(cells, cutoff=5.0, max_nbr=64)
        teacher = jm.configs.random_params(20, 2, 8, 1.6, 5.0, rng, sigma=0.05)
        student = jm.configs.random_params(16, 2, 8, 1.6, 5.0, rng, sigma=0.05)
        _BATCH.update(b=b, teacher=teacher, student=student)
    return _BATCH["b"], _BATCH["teacher"], _BATCH["student"]


@pytest.mark.parametrize("which", ["student", "teacher"])
def test_training_batch_forward_full_size(jm, O, which):
    import torch

    b, teacher, student = _bench_batch(jm)
    p = student if which == "student" else teacher
    print(f"[case] training forward full size: {which} level {p.level}, {b.nconf} configurations, "
          f"{b.natoms} atoms, K={b.nslots}", flush=True)
    assert b.nconf == 512 and b.natoms == 70037
    orc = oracle_for(O, p)
    E, F, S, _ = O.batch_outputs(orc, b.itypes, b.js, b.rij, b.jtypes, b.conf_off, b.volumes)
    pot = jm.MTPPotential(p, max_atoms=b.natoms, max_nbr=b.nslots)
    Eg, Fg, Sg = jm.training.batch_forward(pot, b.to_device())
    torch.cuda.synchronize()
    print(f"  rel E {_rel(Eg.cpu().numpy(), E):.2e} F {_rel(Fg.cpu().numpy(), F):.2e} "
          f"S {_rel(Sg.cpu().numpy(), S):.2e}", flush=True)
    assert _rel(Eg.cpu().numpy(), E) <= 1e-10
    assert _rel(Fg.cpu().numpy(), F) <= 1e-10
    assert _rel(Sg.cpu().numpy(), S) <= 1e-10
    pot.close()


def test_training_param_grad_full_size(jm, O):
    import torch

    b, _, p = _bench_batch(jm)
    print(f"[case] parameter gradient full size: {b.natoms} atoms", flush=True)
    orc = oracle_for(O, p)
    rng = np.random.default_rng(11)
    C, N = b.nconf, b.natoms
    gE, gF, gS = rng.normal(size=C), rng.normal(size=(N, 3)), rng.normal(size=(C, 6))
    pot = jm.MTPPotential(p, max_atoms=N, max_nbr=b.nslots)
    dev = lambda x: torch.as_tensor(x, dtype=torch.float64, device="cuda")
    grad = jm.training.batch_param_grad(pot, b.to_device(), dev(gE), dev(gF), dev(gS)).cpu().numpy()
    pot.close()
    nS, sp, nR = len(p.moment_coeffs), p.species_count, np.asarray(p.radial_coeffs).size
    assert grad.shape == (nS + sp + nR,)
    scale = {"moment": 0.05, "species": 0.05, "radial": 0.05}
    masks = {"all": np.ones(grad.size, bool),
             "radial only": np.r_[np.zeros(nS + sp, bool), np.ones(nR, bool)],
             "moment and species": np.r_[np.ones(nS + sp, bool), np.zeros(nR, bool)]}
    h = 1.0
    ts = np.arange(-3, 4) * h
    for name, mask in masks.items():
        v = np.where(mask, rng.normal(0.0, scale["radial"], grad.size), 0.0)
        J = O.J_direction(orc, b.itypes, b.js, b.rij, b.jtypes, b.conf_off, b.volumes, gE, gF, gS, v, ts)
        # degree <= 5 along the line: the 6th difference of the 7 values vanishes
        assert abs(np.diff(J, 6)[0]) <= 1e-9 * np.max(np.abs(J)), (name, np.diff(J, 6))
        fd = (45.0 * (J[4] - J[2]) - 9.0 * (J[5] - J[1]) + (J[6] - J[0])) / (60.0 * h)
        dd = float(grad @ v)
        print(f"  direction {name}: GPU {dd:.15e} oracle {fd:.15e} rel {abs(dd - fd) / abs(fd):.2e}", flush=True)
        assert dd == pytest.approx(fd, rel=1e-9)


def test_head_100k_with_fallback(jm, O):
    """The 100,000-atom head evaluation with level-24 fallback against the oracle composition."""
    import head_oracle as HO

    w = jm.distill.inference_cell((25, 25, 40), seed=2)
    nb = w.neighbors()
    A, K = nb["js"].shape
    print(f"[case] head 100k: {A} atoms, K={K}, max n {nb['counts'].max()}", flush=True)
    assert A == 100000
    rng = np.random.default_rng(21)
    desc = jm.configs.random_params(12, 3, 8, 2.0, w.cutoff, rng, sigma=0.1)
    exact_p = jm.configs.random_params(24, 3, 10, 2.0, w.cutoff, rng, sigma=0.1)
    orc_d = oracle_for(O, desc)
    orc_x = oracle_for(O, exact_p)
    S = len(desc.moment_coeffs)
    _, _, bas = orc_d.compact_xi(nb["itypes"], nb["rij"], nb["jtypes"], np.zeros(S), grad=False)
    theta = rng.normal(0, 0.3, HO.N_THETA)
    theta[:HO.IN * HO.H].reshape(HO.IN, HO.H)[S + 3:] = 0.0
    mean, inv_std = bas.mean(0), 1.0 / (bas.std(0) + 1e-12)
    z = (bas - mean) * inv_std
    # per-species box from the cell's own inputs, cut so that a few per cent of the atoms fall outside
    ts = HO.wrap_species(nb["itypes"], 3)
    lo = np.stack([np.quantile(z[ts == s], 0.001, axis=0) for s in range(3)])
    hi = np.stack([np.quantile(z[ts == s], 0.999, axis=0) for s in range(3)])
    e_shift, e_scale = rng.normal(0, 1, 3), 0.7

    class _M:
        pass

    model = _M()
    model.theta, model.mean, model.inv_std, model.lo, model.hi = theta, mean, inv_std, lo, hi
    model.e_shift, model.e_scale = e_shift, e_scale
    model.lin, model.lin_sc = orc_d._keep["moment"], orc_d._keep["species"]  # residual form
    E_r, F_r, S_r, eat_r, fl_r, _ = HO.head_compute(model, orc_d, orc_x, nb["itypes"], nb["js"], nb["rij"],
                                                     nb["jtypes"], 3, w.volume, A, A)
    n_fb = int(((fl_r != 0) & (nb["itypes"] >= 0)).sum())
    print(f"  oracle: {n_fb} atoms re-evaluated exactly", flush=True)
    assert 100 < n_fb < A // 4
    exact = jm.MTPPotential(exact_p, max_atoms=A, max_nbr=K)
    hm = jm.head.HeadModel(desc, theta, mean, inv_std, lo, hi, e_shift, e_scale, True)
    head = jm.head.DistilledHead(hm, exact, max_atoms=A, max_nbr=K)
    E, F, St, extra = head.calc_energy_forces_stress_padded(nb["itypes"], nb["js"], nb["rij"], nb["jtypes"], 3,
                                                            w.volume, A, A, return_all=True)
    print(f"  rel E {abs(float(E) - E_r) / abs(E_r):.2e} F {_rel(F.cpu().numpy(), F_r):.2e} "
          f"Ei {_rel(extra['eatom'].cpu().numpy(), eat_r):.2e} fallback {extra['n_fallback']}", flush=True)
    assert float(E) == pytest.approx(E_r, rel=1e-10)
    assert _rel(F.cpu().numpy(), F_r) <= 1e-10
    assert _rel(St.cpu().numpy(), S_r) <= 1e-10
    assert _rel(extra["eatom"].cpu().numpy(), eat_r) <= 1e-10
    np.testing.assert_array_equal(extra["flags"].cpu().numpy(), fl_r)
    assert extra["n_fallback"] == n_fb
    assert float(extra["status"]) == 0.0
    head.close()
    exact.close()
