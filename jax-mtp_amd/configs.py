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
from the seed.  The level's alpha tables come from the untrained potential
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
