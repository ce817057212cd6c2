"""mtp_gather_forces_noalloc (the Newton-3 force gather on the library's
static per-device scratch, for callers that cannot allocate) against the
context entry mtp_gather_forces_device: bit-identical forces on two calls in
a row, the second with a smaller input, and the capacity limit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(jm, natoms, seed, K):
    w = jm.configs.small_cluster(10, 2, natoms, seed=seed, cell=(natoms / 0.08) ** (1 / 3), cutoff=5.0, max_nbr=K)
    nb = jm.neighbors.padded_neighbors(w.pos, w.types, w.cell, w.cutoff, w.max_nbr, rows=natoms + 3)
    rng = np.random.default_rng(seed)
    g = np.where(nb["jtypes"][..., None] >= 0, rng.normal(size=nb["rij"].shape), 0.0)
    return w, nb, g


def test_noalloc_gather_matches_context_gather(jm):
    import torch

    L = jm.capi.lib()
    dev = torch.device("cuda")
    w0, nb0, g0 = _case(jm, 120, 1, 64)
    pot = jm.MTPPotential(w0.params, max_atoms=nb0["js"].shape[0], max_nbr=64)
    stream = torch.cuda.current_stream().cuda_stream
    for natoms, seed in [(120, 1), (40, 2)]:  # the second call is smaller: stale scratch must not leak in
        w, nb, g = _case(jm, natoms, seed, 64)
        A, K = nb["js"].shape
        n_f = natoms
        print(f"[case] noalloc gather rows={A} slots={K} n_f={n_f}", flush=True)
        js = torch.as_tensor(nb["js"], dtype=torch.int32, device=dev)
        gd = torch.as_tensor(g, dtype=torch.float64, device=dev)
        F_ref = torch.full((A, 3), np.nan, dtype=torch.float64, device=dev)
        F_new = torch.full((A, 3), np.nan, dtype=torch.float64, device=dev)
        jm.capi.check(L.mtp_gather_forces_device(pot.ctx, A, K, js.data_ptr(), gd.data_ptr(), n_f, F_ref.data_ptr(),
                                                 stream), pot.ctx)
        rc = L.mtp_gather_forces_noalloc(A, K, js.data_ptr(), gd.data_ptr(), n_f, F_new.data_ptr(), stream)
        assert rc == 0, L.mtp_global_error()
        torch.cuda.synchronize()
        a, b = F_ref.cpu().numpy(), F_new.cpu().numpy()
        np.testing.assert_array_equal(a, b)
        # and the gather itself: F_a = [a < n_f] sum_k g_ak - sum_{i < n_f, k: j_ik = a < n_f} g_ik
        ref = np.zeros((A, 3))
        ref[:n_f] = g[:n_f].sum(axis=1)
        jj = nb["js"][:n_f].ravel()
        keep = jj < n_f
        np.subtract.at(ref, jj[keep], g[:n_f].reshape(-1, 3)[keep])
        np.testing.assert_allclose(b, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    pot.close()


def test_noalloc_gather_capacity(jm):
    L = jm.capi.lib()
    assert L.mtp_gather_forces_noalloc(131073, 2, None, None, 0, None, None) == jm.capi.MTP_ERR_CAPACITY
    assert L.mtp_gather_forces_noalloc(65536, 65, None, None, 0, None, None) == jm.capi.MTP_ERR_CAPACITY
    assert b"static scratch" in L.mtp_global_error()
    assert L.mtp_gather_forces_noalloc(0, 8, None, None, 0, None, None) == 0
