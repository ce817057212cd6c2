"""The MD that generates the distillation snapshots (jax-mtp_amd/distill.py):
the Morse forces against finite differences of the Morse energy, Newton's
third law, and the padded rows of the cells against a brute-force search.
CPU only (torch CPU tensors)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")


def _morse_energy(D, x, box, cutoff):
    d_, al, r0 = D.MORSE
    d = D._min_image(x, box)
    r = torch.sqrt((d * d).sum(-1))
    n = x.shape[1]
    iu = torch.triu_indices(n, n, 1)
    r = r[:, iu[0], iu[1]]

    def v(rr):
        e = torch.exp(-al * (rr - r0))
        return d_ * (e * e - 2.0 * e)

    def dv(rr):
        e = torch.exp(-al * (rr - r0))
        return 2.0 * al * d_ * (e - e * e)

    rc = torch.tensor(cutoff, dtype=torch.float64)
    # force-shifted form: V(r) - V(rc) - (r - rc) V'(rc) inside the cutoff
    e = torch.where(r < cutoff, v(r) - v(rc) - (r - rc) * dv(rc), torch.zeros_like(r))
    return e.sum()


def test_morse_forces_are_minus_the_energy_gradient(jm):
    D = jm.distill
    rng = np.random.default_rng(0)
    base, cell = jm.configs.fcc(D.A_FCC, (4, 4, 4))
    x = torch.tensor(base[None] + rng.normal(0.0, 0.12, (1,) + base.shape), dtype=torch.float64)
    box = torch.tensor(np.asarray(cell, float), dtype=torch.float64)
    F = D._morse_forces(x, box, D.CUTOFF)
    assert float(F.sum(1).abs().max()) < 1e-10  # Newton's third law
    h = 1e-5
    for (i, q) in ((0, 0), (17, 1), (200, 2)):
        xp, xm = x.clone(), x.clone()
        xp[0, i, q] += h
        xm[0, i, q] -= h
        g = (_morse_energy(D, xp, box, D.CUTOFF) - _morse_energy(D, xm, box, D.CUTOFF)) / (2 * h)
        assert abs(float(F[0, i, q]) + float(g)) < 1e-6 * max(1.0, abs(float(g)))


def test_cell_rows_match_brute_force(jm):
    D = jm.distill
    rng = np.random.default_rng(1)
    base, cell = jm.configs.fcc(D.A_FCC, (4, 4, 4))
    C, n = 2, len(base)
    xn = base[None] + rng.normal(0.0, 0.1, (C,) + base.shape)
    types = torch.tensor(rng.integers(0, 3, (C, n)), dtype=torch.int64)
    x = torch.tensor(xn, dtype=torch.float64)
    box = torch.tensor(np.asarray(cell, float), dtype=torch.float64)
    it, js, rij, jt = D._cell_rows(x, types, box, D.CUTOFF, D.MAX_NBR)
    L = np.asarray(cell, float)
    for c, i in ((0, 0), (1, 77), (1, 255)):
        row = c * n + i
        d = xn[c] - xn[c, i]
        d -= L * np.round(d / L)
        r = np.linalg.norm(d, axis=1)
        r[i] = np.inf
        want = sorted((float(r[j]), c * n + j) for j in np.nonzero(r < D.CUTOFF)[0])
        valid = jt[row].numpy() >= 0
        got = sorted((float(np.linalg.norm(v)), int(j)) for v, j in zip(rij[row].numpy()[valid], js[row].numpy()[valid]))
        assert [j for _, j in got] == [j for _, j in want]
        np.testing.assert_allclose([g for g, _ in got], [w for w, _ in want], rtol=0, atol=1e-12)
        np.testing.assert_array_equal(jt[row].numpy()[valid], types[c].numpy()[js[row].numpy()[valid] - c * n])
        assert np.all(js[row].numpy()[~valid] == C * n) and np.all(rij[row].numpy()[~valid] == 2000.0)
    assert int(it[n + 5]) == int(types[1, 5])
