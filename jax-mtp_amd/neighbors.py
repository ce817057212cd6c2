"""Periodic padded neighbor lists (host side, numpy/scipy).

Produces the padded per-row arrays of the device contract
(calc_energy_forces_stress_padded, jax_backend.py:374-421): itypes[A],
js[A,K], rij[A,K,3], jtypes[A,K], with the staging sentinels of
zero_overhead.cpp:76-92 (j = natoms, type = -1, r = 2000) in unused slots.

Unlike the pair style's single minimum image (pair_jax_mtp.cpp:128), every
periodic image inside the cutoff is listed, so cells smaller than twice the
cutoff (the small training cells) are handled; for cells at least twice the
cutoff the two agree.  Neighbors of an atom are ordered by (j, image).
"""
from __future__ import annotations

import itertools

import numpy as np

SENTINEL_R = 2000.0


class NeighborOverflow(RuntimeError):
    pass


def _pairs_min_image(pos, cell, cutoff):
    from scipy.spatial import cKDTree

    L = np.asarray(cell, dtype=float)
    wrapped = np.mod(pos, L)
    tree = cKDTree(wrapped, boxsize=L)
    pairs = tree.query_pairs(cutoff, output_type="ndarray")
    i = np.concatenate([pairs[:, 0], pairs[:, 1]])
    j = np.concatenate([pairs[:, 1], pairs[:, 0]])
    d = pos[j] - pos[i]
    d -= L * np.round(d / L)
    r2 = np.einsum("ij,ij->i", d, d)
    keep = r2 < cutoff * cutoff
    i, j, d = i[keep], j[keep], d[keep]
    img = np.zeros(len(i), dtype=np.int64)
    return i, j, d, img


def _pairs_images(pos, cell, cutoff):
    L = np.asarray(cell, dtype=float)
    nmax = [int(np.ceil(cutoff / L[q])) for q in range(3)]
    shifts = np.array(list(itertools.product(*[range(-n, n + 1) for n in nmax])), dtype=float)
    n = len(pos)
    I, J, D, IMG = [], [], [], []
    for s_idx, s in enumerate(shifts):
        d = pos[None, :, :] + (s * L)[None, None, :] - pos[:, None, :]  # [i, j, 3]
        r2 = np.einsum("ijk,ijk->ij", d, d)
        m = r2 < cutoff * cutoff
        if not np.any(s):
            m[np.arange(n), np.arange(n)] = False
        ii, jj = np.nonzero(m)
        I.append(ii); J.append(jj); D.append(d[ii, jj]); IMG.append(np.full(len(ii), s_idx))
    return np.concatenate(I), np.concatenate(J), np.concatenate(D), np.concatenate(IMG)


def neighbor_pairs(pos, cell, cutoff):
    """All (i, j, r_j + shift - r_i) with |.| < cutoff over periodic images."""
    pos = np.asarray(pos, dtype=float)
    L = np.asarray(cell, dtype=float)
    if np.all(L >= 2.0 * cutoff + 1e-9) and len(pos) > 64:
        i, j, d, img = _pairs_min_image(pos, L, cutoff)
    else:
        i, j, d, img = _pairs_images(pos, L, cutoff)
    order = np.lexsort((img, j, i))
    return i[order], j[order], d[order]


def padded_neighbors(pos, types, cell, cutoff, max_nbr, rows=None, on_overflow="raise"):
    """Returns dict(itypes, js, rij, jtypes, counts) with K = max_nbr slots.

    on_overflow: "raise" (default) or "truncate" (pair_jax_mtp.cpp:132 behaviour).
    """
    pos = np.asarray(pos, dtype=float)
    n = len(pos)
    A = n if rows is None else rows
    i, j, d = neighbor_pairs(pos, cell, cutoff)
    counts_all = np.bincount(i, minlength=n)
    if counts_all.size and counts_all.max() > max_nbr:
        if on_overflow == "raise":
            raise NeighborOverflow(f"an atom has {counts_all.max()} neighbors > max_nbr={max_nbr}")
    starts = np.concatenate([[0], np.cumsum(counts_all)[:-1]])
    slot = np.arange(len(i)) - starts[i]
    keep = slot < max_nbr
    i, j, d, slot = i[keep], j[keep], d[keep], slot[keep]
    types = np.asarray(types, dtype=np.int32)
    itypes = np.full(A, -1, np.int32)
    itypes[:n] = types
    js = np.full((A, max_nbr), n, np.int32)
    jtypes = np.full((A, max_nbr), -1, np.int32)
    rij = np.full((A, max_nbr, 3), SENTINEL_R)
    js[i, slot] = j
    jtypes[i, slot] = types[j]
    rij[i, slot] = d
    counts = np.zeros(A, np.int32)
    counts[:n] = np.minimum(counts_all, max_nbr)
    return dict(itypes=itypes, js=js, rij=rij, jtypes=jtypes, counts=counts)
