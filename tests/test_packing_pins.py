"""Pins of the pair style's packing and staging (pair_jax_mtp.cpp:96-173,
zero_overhead.cpp:61-117) from geometry, independent of any restatement of
the loop: a simple cubic lattice whose neighbor shells are known exactly.

* strict cutoff: r^2 < cutoff^2 (pair_jax_mtp.cpp:132), so a cutoff equal to
  the nearest-neighbor distance keeps nobody and one a hair larger keeps the
  six nearest neighbors;
* order and truncation: the first MAX_ATOMS_PER_NEIGHBOR in-cutoff entries
  of the list, in list order, the rest dropped (and counted here);
* self entries skipped, minimum-image vectors of +-a along the axes, ghosts
  folded onto the local atom with their tag, an unknown tag -> nlocal with
  type -1, types shifted to 0-based;
* sentinels: (20, 20, 20) / nlocal / -1 in the packed rows, then
  2000 / nlocal / -1 and itype -1 after staging.
Both the library (mtp_pack_full_list, mtp_stage_host) and the oracle's
restatement are checked against these expectations.
"""
import ctypes

import numpy as np
import pytest

A_LAT, N_CELL = 2.5, 4
L_BOX = A_LAT * N_CELL


def _lattice():
    """Simple cubic locals (tags 1..64) in a periodic 10 A box, a ghost copy of each one box length along x
    (same tag; toward the nearer face, so that one minimum-image shift recovers the nearest vector, as for
    LAMMPS ghosts), and one ghost with an unknown tag at the cube center (1.25, 1.25, -1.25)."""
    g = np.stack(np.meshgrid(*[np.arange(N_CELL)] * 3, indexing="ij"), -1).reshape(-1, 3) * A_LAT
    n = len(g)
    types = (np.arange(n) % 2 + 1).astype(np.int32)  # 1-based, alternating
    tags = np.arange(1, n + 1, dtype=np.int32)
    shift = np.where(g[:, :1] >= 0.5 * L_BOX, -L_BOX, L_BOX) * [1.0, 0.0, 0.0]
    x = np.vstack([g, g + shift, [[1.25, 1.25, -1.25]]])
    t = np.concatenate([types, types, [2]]).astype(np.int32)
    tag = np.concatenate([tags, tags, [999]]).astype(np.int32)
    return x, t, tag, n


def _min_image(d):
    """Domain::minimum_image for an orthogonal periodic box: one shift when |d| > L/2."""
    return np.where(np.abs(d) > 0.5 * L_BOX, d - np.sign(d) * L_BOX, d)


def _lib_pack(jm, x, t, tag, nlocal, neigh, cutoff, max_nbr, ilist=None):
    L = jm.capi.lib()
    ilist = np.arange(nlocal, dtype=np.int32) if ilist is None else np.asarray(ilist, np.int32)
    A = nlocal
    pos = np.zeros((A, 3 * max_nbr)); nbr = np.zeros((A, max_nbr), np.int32); nt = np.zeros((A, max_nbr), np.int32)
    cnt = np.zeros(A, np.int32); at = np.zeros(A, np.int32)
    rows = [np.ascontiguousarray(r, dtype=np.int32) for r in neigh]
    ptrs = (ctypes.c_void_p * len(rows))(*[r.ctypes.data for r in rows])
    numneigh = np.array([len(r) for r in rows], np.int32)
    xc = np.ascontiguousarray(x, dtype=np.float64)
    trunc = ctypes.c_longlong()
    vp = ctypes.c_void_p
    prd = np.array([L_BOX] * 3)
    tot = L.mtp_pack_full_list(nlocal, xc.ctypes.data_as(vp), t.ctypes.data_as(vp), tag.ctypes.data_as(vp),
                               len(ilist), ilist.ctypes.data_as(vp), numneigh.ctypes.data_as(vp),
                               ctypes.cast(ptrs, vp), float(cutoff), prd.ctypes.data_as(vp),
                               np.array([1, 1, 1], np.int32).ctypes.data_as(vp), A, max_nbr, pos.ctypes.data_as(vp),
                               nbr.ctypes.data_as(vp), nt.ctypes.data_as(vp), cnt.ctypes.data_as(vp),
                               at.ctypes.data_as(vp), ctypes.byref(trunc))
    return pos, nbr, nt, cnt, at, tot, trunc.value


def _oracle_pack(O, x, t, tag, nlocal, neigh, cutoff, max_nbr):
    pos, nbr, nt, cnt, at, tot = O.pack_full_list(x, t, tag, np.arange(nlocal, dtype=np.int32), neigh, nlocal,
                                                  cutoff, [L_BOX] * 3, [1, 1, 1], nlocal, max_nbr)
    return pos, nbr, nt, cnt, at, tot, None


def _lists(n, rng=None, unknown=False):
    """Each local's list holds every local once (itself included), as the local index or as its ghost copy
    (alternating), plus the unknown-tag ghost when asked; optionally shuffled."""
    out = []
    for i in range(n):
        js = [j if (j == i or (i + j) % 2) else n + j for j in range(n)] + ([2 * n] if unknown else [])
        js = np.array(js, np.int32)
        out.append(rng.permutation(js).astype(np.int32) if rng is not None else js)
    return out


@pytest.mark.parametrize("packer", ["library", "oracle"])
def test_strict_cutoff_and_shells(jm, O, packer):
    x, t, tag, n = _lattice()
    neigh = _lists(n)
    pack = (lambda *a: _lib_pack(jm, *a)) if packer == "library" else (lambda *a: _oracle_pack(O, *a))
    # cutoff exactly the nearest-neighbor distance: r^2 == cutoff^2 is not < cutoff^2
    *_, cnt, at, tot, _ = pack(x, t, tag, n, neigh, A_LAT, 32)
    assert tot == 0 and np.all(cnt == 0)
    # a hair more: the six nearest neighbors (4x4x4 periodic: +-a along each axis)
    pos, nbr, nt, cnt, at, tot, _ = pack(x, t, tag, n, neigh, A_LAT * (1 + 1e-12), 32)
    assert np.all(cnt == 6) and tot == 6 * n
    for i in range(n):
        v = pos[i, :18].reshape(6, 3)
        assert sorted(map(tuple, np.round(v, 12))) == sorted(
            [(s * A_LAT * (q == 0), s * A_LAT * (q == 1), s * A_LAT * (q == 2)) for q in range(3) for s in (-1, 1)])
        # the neighbors are locals (ghost copies fold onto their owner's index), types 0-based
        assert np.all(nbr[i, :6] < n)
        np.testing.assert_array_equal(nt[i, :6], t[nbr[i, :6]] - 1)
        # padding of the packed rows
        assert np.all(pos[i, 18:] == 20.0) and np.all(nbr[i, 6:] == n) and np.all(nt[i, 6:] == -1)
    np.testing.assert_array_equal(at, t[:n] - 1)
    # second shell included below a*sqrt(2) * (1 + eps): 6 + 12
    *_, cnt, _, _, _ = pack(x, t, tag, n, neigh, A_LAT * np.sqrt(2) * (1 + 1e-12), 32)
    assert np.all(cnt == 18)


@pytest.mark.parametrize("packer", ["library", "oracle"])
def test_list_order_truncation_and_unknown_tag(jm, O, packer):
    x, t, tag, n = _lattice()
    rng = np.random.default_rng(3)
    neigh = _lists(n, rng, unknown=True)
    pack = (lambda *a: _lib_pack(jm, *a)) if packer == "library" else (lambda *a: _oracle_pack(O, *a))
    cut, K = A_LAT * (1 + 1e-12), 4
    pos, nbr, nt, cnt, at, tot, dropped = pack(x, t, tag, n, neigh, cut, K)
    n_in = []
    for i in range(n):
        d = _min_image(x[neigh[i]] - x[i])
        keep = (neigh[i] != i) & (np.sum(d * d, axis=1) < cut * cut)
        first = neigh[i][keep][:K]  # list order decides who is kept
        n_in.append(int(keep.sum()))
        assert cnt[i] == len(first)
        np.testing.assert_array_equal(pos[i, :3 * len(first)].reshape(-1, 3), d[keep][:K])
        unknown = tag[first] > n
        np.testing.assert_array_equal(nbr[i, :len(first)], np.where(unknown, n, tag[first] - 1))
        np.testing.assert_array_equal(nt[i, :len(first)], np.where(unknown, -1, t[first] - 1))
    # 6 nearest neighbors each; the cube-center ghost is in range (1.25 * sqrt(3) A) of its 8 corners
    assert sorted(n_in) == [6] * (n - 8) + [7] * 8
    assert np.all(cnt == K) and tot == K * n
    assert sum(1 for i in range(n) if np.any(nbr[i, :K] == n)) > 0  # the unknown tag was kept somewhere
    if dropped is not None:
        assert dropped == sum(c - K for c in n_in)


@pytest.mark.parametrize("stager", ["library", "oracle"])
def test_staging_sentinels(jm, O, stager):
    """zero_overhead.cpp:76-117: every slot 2000 / natoms / -1 and itype -1, then the valid prefixes."""
    rows, K, natoms = 5, 3, 3
    counts = np.array([2, 0, 3], np.int32)
    r = np.arange(natoms * 3 * K, dtype=np.float64).reshape(natoms, 3 * K) + 0.5
    js = np.array([[1, 2, 7], [9, 9, 9], [0, 1, 2]], np.int32)
    jt = np.array([[0, 1, 5], [4, 4, 4], [1, 0, 1]], np.int32)
    types = np.array([1, 0, 1], np.int32)
    if stager == "library":
        L = jm.capi.lib()
        vp = ctypes.c_void_p
        arr = lambda a: ctypes.cast((vp * natoms)(*[a[k].ctypes.data for k in range(natoms)]), vp)
        it_o = np.zeros(rows, np.int32); js_o = np.zeros((rows, K), np.int32); r_o = np.zeros((rows, K, 3))
        jt_o = np.zeros((rows, K), np.int32)
        rc = L.mtp_stage_host(rows, K, natoms, arr(r), types.ctypes.data_as(vp), arr(js), counts.ctypes.data_as(vp),
                              arr(jt), it_o.ctypes.data_as(vp), js_o.ctypes.data_as(vp), r_o.ctypes.data_as(vp),
                              jt_o.ctypes.data_as(vp))
        assert rc == 0
    else:
        it_o, js_o, r_o, jt_o = O.stage(rows, K, natoms, r, types, js, counts, jt)
    np.testing.assert_array_equal(it_o, [1, 0, 1, -1, -1])
    np.testing.assert_array_equal(js_o, [[1, 2, 3], [3, 3, 3], [0, 1, 2], [3, 3, 3], [3, 3, 3]])
    np.testing.assert_array_equal(jt_o, [[0, 1, -1], [-1, -1, -1], [1, 0, 1], [-1, -1, -1], [-1, -1, -1]])
    np.testing.assert_array_equal(r_o[0, :2].ravel(), r[0, :6])
    assert np.all(r_o[0, 2] == 2000.0) and np.all(r_o[1] == 2000.0) and np.all(r_o[3:] == 2000.0)
    np.testing.assert_array_equal(r_o[2].ravel(), r[2])
