"""mtp_compute_full_list: the pair style's packing (pair_jax_mtp.cpp:96-173)
and staging (zero_overhead.cpp:61-117) done on the GPU from the uploaded
LAMMPS arrays.  The rows are bit for bit the host packer's, so every output
equals the host path's (mtp_pack_full_list + mtp_compute) exactly, on mock
full lists with ghosts carrying owner tags, a ghost of unknown tag, self
entries, out-of-cutoff entries, a shuffled ilist, and rows cut at
MAX_ATOMS_PER_NEIGHBOR (truncation counted the same)."""
import ctypes

import numpy as np
import pytest

from test_host import _mock_lammps

pytestmark = pytest.mark.gpu


def _host_path(jm, pot, x, t, tag, nlocal, ilist, neigh, prd, max_nbr):
    L = jm.capi.lib()
    A = nlocal
    pos = np.zeros((A, 3 * max_nbr)); nbr = np.zeros((A, max_nbr), np.int32); nt = np.zeros((A, max_nbr), np.int32)
    cnt = np.zeros(A, np.int32); at = np.zeros(A, np.int32)
    rows = [np.ascontiguousarray(n, dtype=np.int32) for n in neigh]
    ptrs = (ctypes.c_void_p * len(rows))(*[r.ctypes.data for r in rows])
    numneigh = np.array([len(r) for r in rows], np.int32)
    xc = np.ascontiguousarray(x, dtype=np.float64)
    trunc = ctypes.c_longlong()
    vp = ctypes.c_void_p
    tot = L.mtp_pack_full_list(nlocal, xc.ctypes.data_as(vp), t.ctypes.data_as(vp), tag.ctypes.data_as(vp),
                               len(ilist), ilist.ctypes.data_as(vp), numneigh.ctypes.data_as(vp),
                               ctypes.cast(ptrs, vp), 5.0, prd.ctypes.data_as(vp),
                               np.array([1, 1, 1], np.int32).ctypes.data_as(vp), A, max_nbr, pos.ctypes.data_as(vp),
                               nbr.ctypes.data_as(vp), nt.ctypes.data_as(vp), cnt.ctypes.data_as(vp),
                               at.ctypes.data_as(vp), ctypes.byref(trunc))
    assert tot >= 0
    E, F, s6, v6 = pot.compute_rows(nlocal, pos, at, nbr, cnt, nt, float(np.prod(prd)))
    return E, F, s6, v6, trunc.value


@pytest.mark.parametrize("max_nbr,n_cells,shuffle", [(80, 2, False), (20, 2, False), (64, 3, True)])
def test_full_list_bitexact_vs_host_packing(jm, max_nbr, n_cells, shuffle):
    x, t, tag, nlocal, neigh, prd = _mock_lammps(seed=n_cells, n_cells=n_cells)
    print(f"[case] full list nlocal={nlocal} nall={len(x)} K={max_nbr} shuffle={shuffle}", flush=True)
    rng = np.random.default_rng(5)
    params = jm.configs.random_params(10, 2, 8, 1.5, 5.0, rng)
    pot = jm.MTPPotential(params, max_atoms=nlocal, max_nbr=max_nbr, cutoff=5.0)
    ilist = np.arange(nlocal, dtype=np.int32)
    if shuffle:
        ilist = rng.permutation(ilist).astype(np.int32)
    ref = _host_path(jm, pot, x, t, tag, nlocal, ilist, neigh, prd, max_nbr)
    got = pot.compute_full_list(nlocal, x, t, tag, ilist, neigh, prd)
    assert got[0] == ref[0]
    for a, b in zip(got[1:4], ref[1:4]):
        np.testing.assert_array_equal(a, b)
    assert got[4] == ref[4]
    assert (got[4] > 0) == (max_nbr == 20)
    # per-atom values are offered after the full-list call as after mtp_compute
    eatom, vatom = pot.last_atom_values(nlocal)
    np.testing.assert_allclose(eatom.sum(), got[0], rtol=1e-12)
    pot.close()


def test_full_list_rejects_bad_lists(jm):
    x, t, tag, nlocal, neigh, prd = _mock_lammps(seed=1)
    params = jm.configs.random_params(6, 2, 8, 1.5, 5.0, np.random.default_rng(1))
    pot = jm.MTPPotential(params, max_atoms=nlocal, max_nbr=64, cutoff=5.0)
    bad = [n.copy() for n in neigh]
    bad[3] = np.append(bad[3], len(x) + 5).astype(np.int32)  # past nall
    with pytest.raises(jm.capi.MTPError) as e:
        pot.compute_full_list(nlocal, x, t, tag, np.arange(nlocal, dtype=np.int32), bad, prd)
    assert e.value.code == jm.capi.MTP_ERR_ARG
    with pytest.raises(jm.capi.MTPError) as e:
        pot.compute_full_list(nlocal, x, t, tag, np.array([0, nlocal], np.int32), neigh, prd)
    assert e.value.code == jm.capi.MTP_ERR_ARG
    pot2 = jm.MTPPotential(params, max_atoms=nlocal - 1, max_nbr=64, cutoff=5.0)
    with pytest.raises(jm.capi.MTPError) as e:
        pot2.compute_full_list(nlocal, x, t, tag, np.arange(nlocal, dtype=np.int32), neigh, prd)
    assert e.value.code == jm.capi.MTP_ERR_CAPACITY
    pot.close()
    pot2.close()
