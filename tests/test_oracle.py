"""Pins the FP64 CPU oracle (oracle/) before it is used as the parity checker.

Each check is chosen so that a plausible mistake (dropped term, wrong sign or
index, transposed operand) fails it:

* two independent formulations of the same basis — the reference's JSON
  full-tensor contraction tree (jax_backend.py:162-247) and the .mtp alpha
  tables (conversion.py:39-102) — each with its own hand-derived adjoint,
  must agree on energies and per-slot gradients;
* the reference's own coefficient-matching check at its test vectors
  (conversion.py:112-168, utils.py:3-13) must map every basis function;
* closed forms: no neighbors, level-2 pair potential against numpy's
  Chebyshev series (a textbook routine, not the oracle's recurrence), single
  neighbor (every scalar contraction is a product of radial values);
* central finite differences of the total energy for forces (positions) and
  the virial (homogeneous strain);
* invariances: rotation, slot permutation, extra padded slots, sum F = 0.
"""
import numpy as np
import pytest

from conftest import oracle_for


def _cluster(jm, level, species=2, natoms=12, seed=0, cell=7.0, max_nbr=80):
    return jm.configs.small_cluster(level, species, natoms, seed=seed, cell=cell, cutoff=5.0, max_nbr=max_nbr)


@pytest.mark.parametrize("level", [2, 4, 6, 8, 10, 12, 14, 16, 18, 20, 22, 24])
def test_compact_matches_tree(jm, O, level):
    w = _cluster(jm, level, seed=level)
    nb = w.neighbors()
    orc = oracle_for(O, w.params, with_tree=True)
    assert orc.remap_missing == []
    Ec, gc = orc.compact(nb["itypes"], nb["rij"], nb["jtypes"])
    Et, gt = orc.tree_eval(nb["itypes"], nb["rij"], nb["jtypes"])
    assert np.max(np.abs(Ec - Et)) <= 1e-12 * max(1.0, np.max(np.abs(Ec)))
    assert np.max(np.abs(gc - gt)) <= 1e-11 * max(1e-300, np.max(np.abs(gc)))


def test_remap_check_every_level(jm, O):
    # conversion.py:158-161 raises if any MLIP basis value is unmatched
    for level in [2, 4, 6, 8, 10, 12, 14, 16, 18, 20, 22, 24]:
        rng = np.random.default_rng(level)
        p = jm.configs.random_params(level, 1, 8, 2.0, 5.0, rng)
        orc = oracle_for(O, p, with_tree=True)
        assert orc.remap_missing == []
        # the remap is a permutation of the coefficients
        assert sorted(orc.xi_tree.tolist()) == sorted(orc.mtp.moment_coeffs.tolist())


def test_no_neighbors_gives_species_coeff(jm, O):
    p = jm.configs.random_params(16, 3, 8, 2.0, 5.0, np.random.default_rng(1))
    orc = oracle_for(O, p)
    K = 8
    rij = np.full((3, K, 3), 2000.0)
    jt = np.full((3, K), -1, np.int32)
    Ei, g = orc.compact(np.array([0, 1, 2]), rij, jt)
    np.testing.assert_array_equal(Ei, orc.mtp.species_coeffs)
    assert np.all(g == 0.0)


def test_level2_pair_potential_closed_form(jm, O):
    """Level 2: E_i = c_ti + xi_0 sum_k scaling (rc-r)^2 sum_n c[ti,tk,0,n] T_n(x)."""
    rng = np.random.default_rng(7)
    p = jm.configs.random_params(2, 2, 8, 1.7, 5.0, rng, scaling=0.8)
    orc = oracle_for(O, p)
    K = 6
    rij = rng.normal(0, 2.5, (4, K, 3))
    jt = rng.integers(-1, 2, (4, K)).astype(np.int32)
    it = np.array([0, 1, 1, 0], np.int32)
    Ei, _ = orc.compact(it, rij, jt)
    for i in range(4):
        e = p.species_coeffs[it[i]]
        for k in range(K):
            r = np.linalg.norm(rij[i, k])
            if jt[i, k] < 0 or r >= p.max_dist:
                continue
            x = (2 * r - (p.min_dist + p.max_dist)) / (p.max_dist - p.min_dist)
            series = np.polynomial.chebyshev.chebval(x, p.radial_coeffs[it[i], jt[i, k], 0])
            e += p.moment_coeffs[0] * p.scaling * (p.max_dist - r) ** 2 * series
        assert Ei[i] == pytest.approx(e, rel=1e-13, abs=1e-15)


def test_single_neighbor_products(jm, O):
    """With one neighbor every fully contracted scalar is a product of radial values."""
    level = 16
    p = jm.configs.random_params(level, 1, 8, 2.0, 5.0, np.random.default_rng(3))
    orc = oracle_for(O, p, with_tree=True)
    u = np.array([[0.3, -0.5, 0.81]])
    u /= np.linalg.norm(u)
    rb = np.array([[1.3], [0.7], [-1.9], [0.45], [2.2], [0.9]])[: p.radial_funcs_count]
    tree_vals = orc.tree_basis_rb(u, rb)
    for val, key in zip(tree_vals, orc.tree_info.scalar_contractions):
        leaves = []
        O._flatten_leaves(key, leaves)
        assert val == pytest.approx(np.prod([rb[m[0], 0] for m in leaves]), rel=1e-12)
    mlip_vals = orc.compact_basis_rb(u, rb)
    np.testing.assert_allclose(np.sort(mlip_vals), np.sort(tree_vals), rtol=1e-12)


def _total_energy(orc, pos, types, cell, cutoff, K, jm):
    nb = jm.neighbors.padded_neighbors(pos, types, cell, cutoff, K)
    Ei, g = orc.compact(nb["itypes"], nb["rij"], nb["jtypes"])
    return Ei.sum(), nb, g


@pytest.mark.parametrize("level", [10, 16])
def test_forces_match_finite_differences(jm, O, level):
    w = _cluster(jm, level, natoms=10, seed=11 + level, cell=6.5)
    orc = oracle_for(O, w.params)
    E0, nb, g = _total_energy(orc, w.pos, w.types, w.cell, w.cutoff, w.max_nbr, jm)
    F = orc.forces(g, nb["js"], w.natoms)
    h = 1e-5
    for a in range(0, w.natoms, 3):
        for q in range(3):
            pp = w.pos.copy(); pp[a, q] += h
            pm = w.pos.copy(); pm[a, q] -= h
            Ep = _total_energy(orc, pp, w.types, w.cell, w.cutoff, w.max_nbr, jm)[0]
            Em = _total_energy(orc, pm, w.types, w.cell, w.cutoff, w.max_nbr, jm)[0]
            fd = -(Ep - Em) / (2 * h)
            assert F[a, q] == pytest.approx(fd, rel=2e-6, abs=2e-6 * np.abs(F).max())


def test_virial_matches_strain_derivative(jm, O):
    w = _cluster(jm, 12, natoms=10, seed=5, cell=6.5)
    orc = oracle_for(O, w.params)
    nb = w.neighbors()
    Ei, g = orc.compact(nb["itypes"], nb["rij"], nb["jtypes"])
    s6, W = orc.stress(nb["rij"], g, w.volume)
    valid = nb["jtypes"] >= 0
    h = 1e-6
    for a in range(3):
        for b in range(3):
            eps = np.zeros((3, 3)); eps[a, b] = h
            rp = nb["rij"].copy(); rm = nb["rij"].copy()
            rp[valid] = nb["rij"][valid] @ (np.eye(3) + eps).T
            rm[valid] = nb["rij"][valid] @ (np.eye(3) - eps).T
            dE = (orc.compact(nb["itypes"], rp, nb["jtypes"], grad=False)[0].sum()
                  - orc.compact(nb["itypes"], rm, nb["jtypes"], grad=False)[0].sum()) / (2 * h)
            assert W[b, a] == pytest.approx(dE, rel=1e-6, abs=1e-7 * np.abs(W).max())
    # Voigt order [xx, yy, zz, yz, xz, xy] of (W + W^T)/(2V)  (jax_backend.py:357-360)
    S = (W + W.T) / (2 * w.volume)
    np.testing.assert_allclose(s6, [S[0, 0], S[1, 1], S[2, 2], S[1, 2], S[0, 2], S[0, 1]], rtol=1e-14)
    s_bad, _ = orc.stress(nb["rij"], g, w.volume, cell_rank=2)
    assert np.all(np.isnan(s_bad))


def test_invariances(jm, O):
    w = _cluster(jm, 16, natoms=12, seed=21, cell=7.0)
    orc = oracle_for(O, w.params)
    nb = w.neighbors()
    E0, g0 = orc.compact(nb["itypes"], nb["rij"], nb["jtypes"])
    # rotation
    rng = np.random.default_rng(0)
    Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    valid = nb["jtypes"] >= 0
    rr = nb["rij"].copy(); rr[valid] = nb["rij"][valid] @ Q.T
    E1, g1 = orc.compact(nb["itypes"], rr, nb["jtypes"])
    np.testing.assert_allclose(E1, E0, rtol=1e-11, atol=1e-12 * np.abs(E0).max())
    np.testing.assert_allclose(g1[valid], g0[valid] @ Q.T, atol=1e-10 * np.abs(g0).max())
    # slot permutation
    perm = rng.permutation(nb["rij"].shape[1])
    E2, g2 = orc.compact(nb["itypes"], nb["rij"][:, perm], nb["jtypes"][:, perm])
    np.testing.assert_allclose(E2, E0, rtol=1e-12, atol=1e-13 * np.abs(E0).max())
    np.testing.assert_allclose(g2, g0[:, perm], atol=1e-12 * np.abs(g0).max())
    # extra padded slots (t = -1 with the reference sentinels)
    pad_r = np.concatenate([nb["rij"], np.full((w.natoms, 5, 3), 20.0)], axis=1)
    pad_t = np.concatenate([nb["jtypes"], np.full((w.natoms, 5), -1, np.int32)], axis=1)
    E3, _ = orc.compact(nb["itypes"], pad_r, pad_t)
    np.testing.assert_array_equal(E3, E0)
    # total force vanishes in a periodic cell
    F = orc.forces(g0, nb["js"], w.natoms)
    assert np.abs(F.sum(axis=0)).max() <= 1e-11 * np.abs(F).max()


def test_force_mask_semantics(jm, O):
    """jax_backend.py:279-301: rows >= n_f and targets >= n_f are dropped."""
    rng = np.random.default_rng(2)
    A, K = 5, 4
    g = rng.normal(size=(A, K, 3))
    js = rng.integers(0, A + 1, (A, K)).astype(np.int32)
    n = 3
    F = O.Oracle.forces(g, js, n)
    ref = np.zeros((A, 3))
    for i in range(A):
        for k in range(K):
            if i < n:
                ref[i] += g[i, k]
                if js[i, k] < n:
                    ref[js[i, k]] -= g[i, k]
    np.testing.assert_allclose(F, ref, rtol=0, atol=1e-15)
    assert O.Oracle.energy(np.arange(5.0), 3) == 3.0
