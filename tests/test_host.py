"""Host-side logic without a GPU: the C ABI library loads and exports every
declared symbol, the .mtp format round-trips and parses identically in the
three readers (C++ library, package, oracle), neighbor packing and staging
are bit-exact against the oracle's restatement of the pair style, and error
paths return the documented codes."""
import ctypes
import re

import numpy as np
import pytest

from conftest import ROOT


def test_single_hip_runtime_in_process(jm):
    """The library binds to PyTorch's HIP runtime; two runtimes in one process hide the GPU from PyTorch."""
    jm.capi.lib()
    paths = {line.split()[-1] for line in open("/proc/self/maps") if "libamdhip64" in line}
    assert len(paths) == 1, paths


def test_library_exports_every_declared_symbol(jm):
    header = (ROOT / "include" / "mtp_hip.h").read_text()
    declared = set(re.findall(r"\b(mtp_[a-z_0-9]+)\s*\(", header))
    L = jm.capi.lib()
    missing = [s for s in sorted(declared) if not hasattr(L, s)]
    assert not missing, missing
    assert declared == set(jm.capi.SIGNATURES), declared ^ set(jm.capi.SIGNATURES)


@pytest.mark.parametrize("level", [2, 6, 10, 16, 24, 28])
def test_mtp_readers_agree(jm, O, tmp_path, level):
    p = jm.configs.random_params(level, 2, 8, 1.9, 5.0, np.random.default_rng(level))
    path = tmp_path / "p.mtp"
    jm.write_mtp(path, p)
    a = jm.read_mtp(path)
    b = O.read_mtp(path)
    for key in ["alpha_index_basic", "alpha_index_times", "alpha_moment_mapping", "species_coeffs",
                "moment_coeffs", "radial_coeffs"]:
        np.testing.assert_array_equal(np.asarray(getattr(a, key)), np.asarray(getattr(b, key)), err_msg=key)
    for key in ["scaling", "min_dist", "max_dist", "species_count", "radial_basis_size", "radial_funcs_count",
                "alpha_moments_count"]:
        assert getattr(a, key) == getattr(b, key), key
    # the C++ reader inside the library
    v = [ctypes.c_int() for _ in range(6)]
    lds = ctypes.c_longlong()
    rc = jm.capi.lib().mtp_check_file(str(path).encode(), 64, *[ctypes.byref(x) for x in v], ctypes.byref(lds))
    assert rc == 0, jm.capi.lib().mtp_global_error()
    assert [x.value for x in v] == [level, 2, a.alpha_moments_count, len(a.alpha_index_basic),
                                    len(a.alpha_index_times), len(a.alpha_moment_mapping)]
    assert lds.value > 0


def test_write_read_roundtrip_is_stable(jm, tmp_path):
    p = jm.read_mtp(jm.mtp_io.POTENTIALS / "Ni3Al-12g.mtp")
    t1 = jm.format_mtp(p)
    t2 = jm.format_mtp(jm.parse_mtp_text(t1))
    assert t1 == t2
    assert p.level == 12 and p.species_count == 2


def test_parameter_vector_groups(jm):
    """The flat parameter vector is (scaling,) moment, species, radial in that order (data.py:60-102);
    the setter inverts the getter, and initialize draws only the missing groups within their ranges."""
    p = jm.mtp_io.untrained(10)
    p.species_count = 2
    p.initialize(np.random.default_rng(1))
    assert np.all(np.abs(p.moment_coeffs) <= 5) and np.all(np.abs(p.species_coeffs) <= 5)
    assert p.radial_coeffs.shape == (2, 2, p.radial_funcs_count, p.radial_basis_size)
    assert np.all(np.abs(p.radial_coeffs) <= 0.1)
    flat = p.parameters
    S, R = len(p.moment_coeffs), p.radial_coeffs.size
    np.testing.assert_array_equal(flat[:S], p.moment_coeffs)
    np.testing.assert_array_equal(flat[S:S + 2], p.species_coeffs)
    np.testing.assert_array_equal(flat[S + 2:], p.radial_coeffs.ravel())
    q = p.copy()
    q.optimized = ["scaling"] + q.optimized
    q.parameters = np.concatenate([[1.25], flat * 2])
    assert q.scaling == 1.25
    np.testing.assert_array_equal(q.radial_coeffs, 2 * p.radial_coeffs)
    keep = p.moment_coeffs.copy()
    p.initialize(np.random.default_rng(2))  # nothing missing: nothing redrawn
    np.testing.assert_array_equal(p.moment_coeffs, keep)
    # a count declared again after the radial block is rejected by the Python reader as well
    with pytest.raises(ValueError):
        jm.parse_mtp_text(jm.format_mtp(p) + "radial_basis_size = 3\n")


def test_untrained_file_rejected_by_library(jm):
    path = jm.mtp_io.POTENTIALS / "untrained_10.mtp"
    rc = jm.capi.lib().mtp_check_file(str(path).encode(), 64, None, None, None, None, None, None, None)
    assert rc == jm.capi.MTP_ERR_FILE
    assert b"radial_coeffs" in jm.capi.lib().mtp_global_error()


def test_malformed_files(jm, tmp_path):
    p = jm.configs.random_params(10, 1, 8, 2.0, 5.0, np.random.default_rng(0))
    good = jm.format_mtp(p)
    L = jm.capi.lib()
    cases = {
        "missing": None,
        "bad_count": good.replace("alpha_index_times_count = 42", "alpha_index_times_count = 43"),
        # moment 25 is read (entry {0, 25, 1, 34}) before the entry that builds it
        "out_of_order": good.replace("{0, 25, 1, 34}, ", "").replace("{0, 0, 1, 25}", "{0, 25, 1, 34}, {0, 0, 1, 25}"),
        "bad_number": good.replace("min_dist = ", "min_dist = x"),
        # a count declared again after the radial block leaves the table short of the final sizes
        "late_radial_count": good + "radial_basis_size = 9\n",
        "late_species_count": good.replace("species_count = 1", "species_count = 1") + "species_count = 2\n",
    }
    for name, text in cases.items():
        path = tmp_path / f"{name}.mtp"
        if text is not None:
            path.write_text(text)
        rc = L.mtp_check_file(str(path).encode(), 64, None, None, None, None, None, None, None)
        assert rc == jm.capi.MTP_ERR_FILE, name


def test_header_variant_count_matches_library(jm):
    """MTP_NUM_VARIANTS sizes mtp_autotune's ms_per_variant: it must equal the library's variant count."""
    header = (ROOT / "include" / "mtp_hip.h").read_text()
    n = int(re.search(r"#define\s+MTP_NUM_VARIANTS\s+(\d+)", header).group(1))
    assert n == len(jm.MTPPotential.VARIANTS)
    for v in jm.MTPPotential.VARIANTS:
        assert re.search(rf"\b{v}\b", header.split("#define MTP_NUM_VARIANTS")[0].rsplit("/*", 1)[1]), v


def test_per_atom_values_need_a_handle(jm):
    """mtp_last_atom_values rejects a null handle before touching the GPU."""
    assert jm.capi.lib().mtp_last_atom_values(None, 1, None, None) == -1


def test_open_argument_checks_without_gpu(jm, tmp_path):
    """pair_jax_mtp.cpp:241-243, 268-269: positive sizes, max_neighbors > 1."""
    p = jm.configs.random_params(10, 1, 8, 2.0, 5.0, np.random.default_rng(0))
    path = tmp_path / "p.mtp"
    jm.write_mtp(path, p)
    L = jm.capi.lib()
    ctx = ctypes.c_void_p()
    for args in [(0, 60, 5.0), (10, 1, 5.0), (10, 60, 0.0), (10, 60, -1.0)]:
        assert L.mtp_open(str(path).encode(), args[0], args[1], args[2], 0, ctypes.byref(ctx)) == jm.capi.MTP_ERR_ARG
    assert L.mtp_open(str(path).encode(), 10, 300, 5.0, 0, ctypes.byref(ctx)) == jm.capi.MTP_ERR_UNSUPPORTED
    assert L.mtp_open(b"/nonexistent.mtp", 10, 60, 5.0, 0, ctypes.byref(ctx)) == jm.capi.MTP_ERR_FILE


def _mock_lammps(seed=0, n_cells=2, a=3.6, ghosts=True):
    """Local atoms of an FCC box plus ghost copies (tags repeat), with a full
    neighbor list including self-images and out-of-cutoff entries."""
    rng = np.random.default_rng(seed)
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    g = np.stack(np.meshgrid(*[np.arange(n_cells)] * 3, indexing="ij"), -1).reshape(-1, 3)
    xl = ((g[:, None] + base[None]).reshape(-1, 3) * a + rng.normal(0, 0.05, (len(g) * 4, 3)))
    nlocal = len(xl)
    L = n_cells * a
    tags = np.arange(1, nlocal + 1)
    types = rng.integers(1, 3, nlocal)
    xs, ts, tg = [xl], [types], [tags]
    if ghosts:
        for s in [(1, 0, 0), (0, 1, 0), (0, 0, -1)]:
            xs.append(xl + np.array(s) * L); ts.append(types); tg.append(tags)
    x = np.concatenate(xs); t = np.concatenate(ts).astype(np.int32); tag = np.concatenate(tg).astype(np.int32)
    # a ghost with an unknown tag maps to nlocal (pair_jax_mtp.cpp:137-145)
    x = np.vstack([x, xl[0] + 0.9]); t = np.append(t, 1).astype(np.int32); tag = np.append(tag, 10_000).astype(np.int32)
    neigh = []
    for i in range(nlocal):
        d = np.linalg.norm(x - x[i], axis=1)
        js = np.nonzero(d < 6.0)[0]
        js = rng.permutation(js)  # includes i itself; the pair style skips it
        neigh.append(js.astype(np.int32))
    return x, t, tag, nlocal, neigh, np.array([L, L, L])


@pytest.mark.parametrize("max_nbr", [80, 20])
def test_pack_bitexact_vs_oracle(jm, O, max_nbr):
    x, t, tag, nlocal, neigh, prd = _mock_lammps()
    ilist = np.arange(nlocal, dtype=np.int32)
    ref = O.pack_full_list(x, t, tag, ilist, neigh, nlocal, 5.0, prd, [1, 1, 1], nlocal + 3, max_nbr)
    # library
    L = jm.capi.lib()
    A = nlocal + 3
    pos = np.zeros((A, 3 * max_nbr)); nbr = np.zeros((A, max_nbr), np.int32); nt = np.zeros((A, max_nbr), np.int32)
    cnt = np.zeros(A, np.int32); at = np.zeros(A, np.int32)
    rows = [np.ascontiguousarray(n) for n in neigh]
    ptrs = (ctypes.c_void_p * len(rows))(*[r.ctypes.data for r in rows])
    numneigh = np.array([len(r) for r in rows], np.int32)
    xc = np.ascontiguousarray(x, dtype=np.float64)
    trunc = ctypes.c_longlong()
    vp = ctypes.c_void_p
    tot = L.mtp_pack_full_list(nlocal, xc.ctypes.data_as(vp), t.ctypes.data_as(vp), tag.ctypes.data_as(vp), nlocal,
                               ilist.ctypes.data_as(vp), numneigh.ctypes.data_as(vp), ctypes.cast(ptrs, vp), 5.0,
                               prd.ctypes.data_as(vp), np.array([1, 1, 1], np.int32).ctypes.data_as(vp), A, max_nbr,
                               pos.ctypes.data_as(vp), nbr.ctypes.data_as(vp), nt.ctypes.data_as(vp),
                               cnt.ctypes.data_as(vp), at.ctypes.data_as(vp), ctypes.byref(trunc))
    assert tot == ref[5]
    for a, b in zip((pos, nbr, nt, cnt, at), ref[:5]):
        np.testing.assert_array_equal(a[:nlocal], b[:nlocal])
    np.testing.assert_array_equal(at, ref[4])
    if max_nbr == 20:
        assert trunc.value > 0  # silent truncation in the reference is counted here
    else:
        assert trunc.value == 0
    assert (nbr[:nlocal] == nlocal).any()  # the unknown-tag ghost


def test_pack_capacity_error(jm):
    x, t, tag, nlocal, neigh, prd = _mock_lammps()
    L = jm.capi.lib()
    z = np.zeros(1)
    vp = ctypes.c_void_p
    rc = L.mtp_pack_full_list(nlocal, z.ctypes.data_as(vp), z.ctypes.data_as(vp), z.ctypes.data_as(vp), 0,
                              z.ctypes.data_as(vp), z.ctypes.data_as(vp), z.ctypes.data_as(vp), 5.0,
                              z.ctypes.data_as(vp), z.ctypes.data_as(vp), nlocal - 1, 10, z.ctypes.data_as(vp),
                              z.ctypes.data_as(vp), z.ctypes.data_as(vp), z.ctypes.data_as(vp),
                              z.ctypes.data_as(vp), None)
    assert rc == jm.capi.MTP_ERR_CAPACITY


def test_stage_bitexact_vs_oracle(jm, O):
    x, t, tag, nlocal, neigh, prd = _mock_lammps(seed=3)
    ilist = np.arange(nlocal, dtype=np.int32)
    K = 64
    pos, nbr, nt, cnt, at, tot = O.pack_full_list(x, t, tag, ilist, neigh, nlocal, 5.0, prd, [1, 1, 1], nlocal, K)
    ref = O.stage(nlocal + 5, K, nlocal, pos, at, nbr, cnt, nt)
    L = jm.capi.lib()
    rows = nlocal + 5
    vp = ctypes.c_void_p
    it = np.zeros(rows, np.int32); js = np.zeros((rows, K), np.int32); rij = np.zeros((rows, K, 3))
    jt = np.zeros((rows, K), np.int32)

    def rp(a):
        return (ctypes.c_void_p * nlocal)(*[a[i].ctypes.data for i in range(nlocal)])

    pos_c = np.ascontiguousarray(pos); nbr_c = np.ascontiguousarray(nbr); nt_c = np.ascontiguousarray(nt)
    rc = L.mtp_stage_host(rows, K, nlocal, ctypes.cast(rp(pos_c), vp), at.ctypes.data_as(vp),
                          ctypes.cast(rp(nbr_c), vp), cnt.ctypes.data_as(vp), ctypes.cast(rp(nt_c), vp),
                          it.ctypes.data_as(vp), js.ctypes.data_as(vp), rij.ctypes.data_as(vp),
                          jt.ctypes.data_as(vp))
    assert rc == 0
    for a, b in zip((it, js, rij, jt), ref):
        np.testing.assert_array_equal(a, b)
    # a count above max_nbr is a capacity error
    cnt2 = cnt.copy(); cnt2[0] = K + 1
    rc = L.mtp_stage_host(rows, K, nlocal, ctypes.cast(rp(pos_c), vp), at.ctypes.data_as(vp),
                          ctypes.cast(rp(nbr_c), vp), cnt2.ctypes.data_as(vp), ctypes.cast(rp(nt_c), vp),
                          it.ctypes.data_as(vp), js.ctypes.data_as(vp), rij.ctypes.data_as(vp),
                          jt.ctypes.data_as(vp))
    assert rc == jm.capi.MTP_ERR_CAPACITY


def test_padded_neighbors_periodic_images(jm):
    """Small cells list every image inside the cutoff; large cells agree with the minimum image."""
    rng = np.random.default_rng(0)
    pos = rng.uniform(0, 4.5, (8, 3))
    nb = jm.neighbors.padded_neighbors(pos, np.zeros(8, np.int32), [4.5, 4.5, 4.5], 5.0, 200)
    # brute force count over images
    for i in range(8):
        c = 0
        for s in np.array(np.meshgrid(*[range(-2, 3)] * 3)).reshape(3, -1).T:
            d = pos + s * 4.5 - pos[i]
            r = np.linalg.norm(d, axis=1)
            c += np.sum((r < 5.0) & (r > 0))
        assert nb["counts"][i] == c
    w = jm.configs.fcc256_level10()
    nb1 = jm.neighbors.padded_neighbors(w.pos, w.types, w.cell, w.cutoff, w.max_nbr)
    i, j, d, img = jm.neighbors._pairs_images(w.pos, w.cell, w.cutoff)
    assert nb1["counts"].sum() == len(i)
