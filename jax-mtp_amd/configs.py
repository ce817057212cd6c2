"""Seeded synthetic workloads (lattices, displacements, species, parameters).

The workloads are the named configurations of the evaluation path:

* ``fcc256_level10``   256-atom FCC (a=4.05 A, 4x4x4 cells), sigma=0.05 A,
                       1 species, level 10, r_c=5.0, r_min=2.0, N=8, K=60.
* ``bcc2889_level16``  2-species BCC, a=3.2 A, sigma=0.08 A, level 16,
                       r_c=5.0, K=60.  2889 atoms = 1444 full 2-atom cells
                       plus one extra site is not a lattice, so the cell is
                       built as 11x11x12 BCC cells (2904 sites) with the
                       last 15 sites removed (recorded as ``n_removed``).
* ``fcc32k_level24``   32,000-atom FCC (a=3.6 A, 20^3 cells), sigma=0.1 A,
                       3 species 40/30/30, level 24, r_c=5.5, r_min=2.0,
                       N=10, K=80, scaling=1.0.

All parameters (moment, species, radial coefficients) are drawn N(0, 0.1)
from the seed.  ``bcc2889_nve_params`` builds the level-16 potential of the
NVE run on the same cell (physically scaled; see its docstring).  The level's alpha tables come from the untrained potential
of that level with species_count and radial_basis_size overridden.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .mtp_io import MTPData, untrained
from .neighbors import padded_neighbors


def fcc(a: float, n: tuple[int, int, int]):
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    return _lattice(a, n, base)


def bcc(a: float, n: tuple[int, int, int]):
    base = np.array([[0, 0, 0], [0.5, 0.5, 0.5]])
    return _lattice(a, n, base)


def _lattice(a, n, base):
    g = np.stack(np.meshgrid(*[np.arange(k) for k in n], indexing="ij"), axis=-1).reshape(-1, 3)
    pos = (g[:, None, :] + base[None, :, :]).reshape(-1, 3) * a
    cell = np.array(n, dtype=float) * a
    return pos, cell


def random_params(level: int, species_count: int, radial_basis_size: int, min_dist: float, max_dist: float,
                  rng: np.random.Generator, sigma: float = 0.1, scaling: float = 1.0) -> MTPData:
    p = untrained(level)
    p.species_count = species_count
    p.radial_basis_size = radial_basis_size
    p.min_dist = float(min_dist)
    p.max_dist = float(max_dist)
    p.scaling = float(scaling)
    p.potential_name = f"synthetic_level{level}"
    p.moment_coeffs = rng.normal(0.0, sigma, p.alpha_scalar_moments)
    p.species_coeffs = rng.normal(0.0, sigma, species_count)
    p.radial_coeffs = rng.normal(0.0, sigma, (species_count, species_count, p.radial_funcs_count, radial_basis_size))
    return p


def species_by_fraction(n: int, fractions, rng) -> np.ndarray:
    counts = [int(round(f * n)) for f in fractions]
    counts[-1] = n - sum(counts[:-1])
    t = np.concatenate([np.full(c, s, np.int32) for s, c in enumerate(counts)])
    return t[rng.permutation(n)]


@dataclass
class Workload:
    name: str
    params: MTPData
    pos: np.ndarray
    types: np.ndarray
    cell: np.ndarray
    cutoff: float
    max_nbr: int
    nbr: dict | None = None

    @property
    def natoms(self) -> int:
        return len(self.pos)

    @property
    def volume(self) -> float:
        return float(np.prod(self.cell))

    def neighbors(self):
        if self.nbr is None:
            self.nbr = padded_neighbors(self.pos, self.types, self.cell, self.cutoff, self.max_nbr)
        return self.nbr


def fcc256_level10(seed: int = 0) -> Workload:
    rng = np.random.default_rng(seed)
    pos, cell = fcc(4.05, (4, 4, 4))
    pos = pos + rng.normal(0.0, 0.05, pos.shape)
    types = np.zeros(len(pos), np.int32)
    params = random_params(10, 1, 8, 2.0, 5.0, rng)
    return Workload("fcc256_level10", params, pos, types, cell, 5.0, 60)


def bcc2889_level16(seed: int = 0) -> Workload:
    rng = np.random.default_rng(seed)
    pos, cell = bcc(3.2, (11, 11, 12))
    pos = pos[:2889]
    pos = pos + rng.normal(0.0, 0.08, pos.shape)
    types = species_by_fraction(len(pos), [0.5, 0.5], rng)
    params = random_params(16, 2, 8, 2.0, 5.0, rng)
    return Workload("bcc2889_level16", params, pos, types, cell, 5.0, 60)


def fcc32k_level24(seed: int = 0, ncell: int = 20) -> Workload:
    rng = np.random.default_rng(seed)
    pos, cell = fcc(3.6, (ncell, ncell, ncell))
    pos = pos + rng.normal(0.0, 0.1, pos.shape)
    types = species_by_fraction(len(pos), [0.4, 0.3, 0.3], rng)
    params = random_params(24, 3, 10, 2.0, 5.5, rng)
    return Workload(f"fcc{len(pos)}_level24", params, pos, types, cell, 5.5, 80)


def small_cluster(level: int, species: int, natoms: int, seed: int, cell=8.0, cutoff=5.0, max_nbr=64,
                  min_sep=1.6, radial_basis_size=8) -> Workload:
    """Random periodic cell with a minimum separation (for parity tests)."""
    rng = np.random.default_rng(seed)
    L = np.array([cell, cell, cell], dtype=float)
    pts = []
    while len(pts) < natoms:
        p = rng.uniform(0, cell, 3)
        ok = True
        for q in pts:
            d = p - q
            d -= L * np.round(d / L)
            if np.dot(d, d) < min_sep * min_sep:
                ok = False
                break
        if ok:
            pts.append(p)
    pos = np.array(pts)
    types = rng.integers(0, species, natoms).astype(np.int32)
    params = random_params(level, species, radial_basis_size, 1.5, cutoff, rng)
    return Workload(f"cluster{natoms}_level{level}", params, pos, types, L, cutoff, max_nbr)


def _morse_chebyshev(D, alpha, r0, rmin, rmax, N, scaling=1.0):
    """Chebyshev coefficients c_n with scaling (rmax - r)^2 sum_n c_n T_n(x) ~ shifted-force Morse."""
    r = np.linspace(rmin, rmax - 1e-3, 400)

    def phi(x):
        e = np.exp(-alpha * (x - r0))
        return D * (e * e - 2.0 * e)

    def dphi(x):
        e = np.exp(-alpha * (x - r0))
        return D * (-2.0 * alpha * e * e + 2.0 * alpha * e)

    target = phi(r) - phi(rmax) - dphi(rmax) * (r - rmax)
    g = target / (scaling * (rmax - r) ** 2)
    x = (2.0 * r - (rmin + rmax)) / (rmax - rmin)
    V = np.polynomial.chebyshev.chebvander(x, N - 1)
    return np.linalg.lstsq(V, g, rcond=None)[0]


def bcc2889_nve_params(seed: int = 0, many_body: float = 0.1) -> MTPData:
    """Level-16, 2-species potential for 1 fs NVE dynamics on the 2889-atom BCC cell.

    The N(0, 0.1) parameters of ``bcc2889_level16`` give RMS forces of ~4e3
    eV/A, which no 1 fs velocity-Verlet run can follow.  Here the mu = 0
    radial function of every species pair is a least-squares Chebyshev fit of
    a shifted-force Morse pair potential (D 0.30/0.35/0.40 eV, alpha
    1.5/A, r0 = 2.77 A, the BCC nearest-neighbour distance), carried by the
    level-2 basis function with coefficient 1/2 (each pair is seen from both
    atoms).  Every other basis function gets a random coefficient normalised
    by its RMS on the starting cell, scaled so that these many-body terms
    carry ``many_body`` times the RMS pair force.  Needs the GPU (basis values
    and forces for the normalisation)."""
    import torch

    from .potential import MTPPotential

    rng = np.random.default_rng(seed)
    w = bcc2889_level16(seed)
    p = w.params
    D = np.array([[0.30, 0.35], [0.35, 0.40]])
    rc = np.array(p.radial_coeffs)
    for a in range(2):
        for b in range(2):
            rc[a, b, 0] = _morse_chebyshev(D[a, b], 1.5, 2.77, p.min_dist, p.max_dist, p.radial_basis_size,
                                           p.scaling)
    p.radial_coeffs = rc
    p.species_coeffs = np.zeros(2)
    S = len(p.moment_coeffs)
    nb = w.neighbors()
    pot = MTPPotential(p, max_atoms=w.natoms, max_nbr=w.max_nbr)
    pair = np.zeros(S)
    pair[0] = 0.5
    pot.set_params(moment_coeffs=pair)
    args = (nb["itypes"], nb["js"], nb["rij"], nb["jtypes"], 3, w.volume, w.natoms, w.natoms)
    f_pair = float(torch.sqrt((pot.calc_energy_forces_stress_padded(*args)[1] ** 2).mean()))
    b = pot.descriptor(nb["itypes"], nb["rij"], nb["jtypes"])
    rms = torch.sqrt((b ** 2).mean(0)).cpu().numpy()
    mb = np.where(rms > 0, rng.normal(0.0, 1.0, S) / np.maximum(rms, 1e-300), 0.0)
    mb[0] = 0.0
    pot.set_params(moment_coeffs=mb)
    f_mb = float(torch.sqrt((pot.calc_energy_forces_stress_padded(*args)[1] ** 2).mean()))
    pot.close()
    p.moment_coeffs = pair + mb * (many_body * f_pair / max(f_mb, 1e-300))
    p.potential_name = "bcc2889_level16_nve"
    return p
