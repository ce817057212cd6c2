"""Python host interface of the HIP evaluator.

``MTPPotential.calc_energy_forces_stress_padded`` mirrors the reference's
entry point of the same name (jax_backend.py:374-460): the same 8 dynamic
arguments, the same padding/masking semantics, and the same outputs
(energy, forces[A,3], stress6 in the reference's Voigt order), in FP64.
As in the reference, both the energy and the forces are masked by
natoms_energy (jax_backend.py:414-418 passes it into the force mask's slot);
natoms_force is accepted and has no effect.
PyTorch only supplies device memory and the stream; all arithmetic runs in
the HIP kernels of libmtp_hip.so.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import capi
from .mtp_io import MTPData, format_mtp, read_mtp


class _DevArray:
    """Minimal __cuda_array_interface__ view of a device pointer."""

    def __init__(self, ptr: int, shape, typestr: str):
        self.__cuda_array_interface__ = {"shape": tuple(shape), "typestr": typestr, "data": (ptr, False),
                                         "version": 2, "strides": None}


def _i32(t, dev):
    import torch

    return torch.as_tensor(t, dtype=torch.int32, device=dev).contiguous()


def _f64(t, dev):
    import torch

    return torch.as_tensor(t, dtype=torch.float64, device=dev).contiguous()


class MTPPotential:
    """One potential on one GPU (an ``mtp_ctx``)."""

    def __init__(self, params, max_atoms: int, max_nbr: int, cutoff: float | None = None, debug_level: int = 0,
                 radial_fp32: bool = False):
        if isinstance(params, (str, bytes)) or hasattr(params, "__fspath__"):
            params = read_mtp(params)
        self.params: MTPData = params
        self.max_atoms = int(max_atoms)
        self.max_nbr = int(max_nbr)
        L = capi.lib()
        ctx = ctypes.c_void_p()
        cut = float(cutoff if cutoff is not None else params.max_dist)
        capi.check(L.mtp_open_text(format_mtp(params).encode(), self.max_atoms, self.max_nbr, cut, int(debug_level),
                                   ctypes.byref(ctx)))
        self.ctx = ctx
        self.set_precision(radial_fp32)

    def close(self):
        if getattr(self, "ctx", None):
            capi.lib().mtp_close(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_precision(self, radial_fp32: bool):
        self.radial_fp32 = bool(radial_fp32)
        capi.check(capi.lib().mtp_set_precision(self.ctx, int(self.radial_fp32)), self.ctx)

    VARIANTS = {0: "per-target, 256 threads", 1: "streams, 256 threads", 2: "streams, 512 threads",
                3: "streams, 64 threads", 4: "paired rows, 1024 threads", 5: "paired rows, 512 threads",
                6: "streams, 128 threads"}

    @property
    def variant(self) -> int:
        v = ctypes.c_int()
        capi.check(capi.lib().mtp_get_variant(self.ctx, ctypes.byref(v), None), self.ctx)
        return v.value

    def set_variant(self, variant: int):
        capi.check(capi.lib().mtp_set_variant(self.ctx, int(variant)), self.ctx)

    def autotune(self, itypes, all_js, all_rijs, all_jtypes, reps: int = 3):
        """Time every kernel variant on these (device) arrays and keep the fastest; returns (best, ms list)."""
        import torch

        dev = torch.device("cuda")
        it, js, rij, jt = _i32(itypes, dev), _i32(all_js, dev), _f64(all_rijs, dev), _i32(all_jtypes, dev)
        A, K = js.shape
        best = ctypes.c_int()
        ms = (ctypes.c_double * len(self.VARIANTS))()
        capi.check(capi.lib().mtp_autotune(self.ctx, A, K, it.data_ptr(), js.data_ptr(), rij.data_ptr(),
                                           jt.data_ptr(), int(reps), ctypes.byref(best), ms), self.ctx)
        return best.value, list(ms)

    def info(self) -> dict:
        v = [ctypes.c_int() for _ in range(8)]
        capi.check(capi.lib().mtp_info(self.ctx, *[ctypes.byref(x) for x in v]), self.ctx)
        keys = ["level", "species_count", "radial_funcs_count", "radial_basis_size", "alpha_moments_count",
                "basic_count", "times_count", "scalar_count"]
        return {k: x.value for k, x in zip(keys, v)}

    def set_params(self, moment_coeffs=None, species_coeffs=None, radial_coeffs=None, scaling=None):
        def ptr(a):
            if a is None:
                return None, None
            arr = np.ascontiguousarray(a, dtype=np.float64)
            return arr, arr.ctypes.data_as(ctypes.c_void_p)

        m, mp = ptr(moment_coeffs)
        s, sp = ptr(species_coeffs)
        r, rp = ptr(radial_coeffs)
        if scaling is not None and not (np.isfinite(scaling) and scaling > 0):
            raise ValueError("scaling must be a positive finite number (None keeps the current one)")
        capi.check(capi.lib().mtp_set_params(self.ctx, mp, sp, rp, float("nan") if scaling is None else float(scaling)),
                   self.ctx)
        if m is not None:
            self.params.moment_coeffs = m.copy()
        if s is not None:
            self.params.species_coeffs = s.copy()
        if r is not None:
            self.params.radial_coeffs = r.reshape(self.params.radial_coeffs.shape).copy()
        if scaling is not None:
            self.params.scaling = float(scaling)

    def set_params_device(self, theta, stream=None):
        """Parameters from a device tensor theta = [moment, species, radial] (the
        training order), stream-ordered (capturable in a graph).  self.params is
        not updated."""
        import torch

        t = theta.detach()
        if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
            raise ValueError("theta must be a contiguous float64 device tensor")
        S = len(self.params.moment_coeffs)
        sp = self.params.species_count
        if t.numel() != S + sp + int(np.prod(self.params.radial_coeffs.shape)):
            raise ValueError("theta has the wrong length")
        base = t.data_ptr()
        st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        capi.check(capi.lib().mtp_set_params_device(self.ctx, base, base + 8 * S, base + 8 * (S + sp), st), self.ctx)

    # ------------------------------------------------------------------
    def calc_energy_forces_stress_padded(self, itypes, all_js, all_rijs, all_jtypes, cell_rank, volume,
                                         natoms_energy, natoms_force, return_all: bool = False, sync: bool = True):
        """jax_backend.py:374-460 on the GPU.

        Inputs may be numpy arrays or torch tensors (device tensors are used
        in place).  Returns (energy, forces[A,3], stress6) as torch float64
        tensors on the device; with ``return_all`` also a dict holding
        ``virial6`` (LAMMPS order), ``eatom`` [A], ``status`` (bit 1: non-finite
        energy or virial, bit 2: a non-finite r_ij slot, energy then NaN) and the
        per-slot gradients ``g`` [A,K,3].
        """
        import torch

        dev = torch.device("cuda")
        it = _i32(itypes, dev)
        js = _i32(all_js, dev)
        rij = _f64(all_rijs, dev)
        jt = _i32(all_jtypes, dev)
        A, K = js.shape
        if rij.shape != (A, K, 3) or jt.shape != (A, K) or it.shape != (A,):
            raise ValueError("inconsistent padded shapes")
        out = torch.empty(14, dtype=torch.float64, device=dev)
        F = torch.empty((A, 3), dtype=torch.float64, device=dev)
        eatom = torch.empty(A, dtype=torch.float64, device=dev) if return_all else None
        stream = torch.cuda.current_stream().cuda_stream
        L = capi.lib()
        capi.check(L.mtp_compute_padded_device(
            self.ctx, A, K, it.data_ptr(), js.data_ptr(), rij.data_ptr(), jt.data_ptr(), int(cell_rank),
            float(volume), int(natoms_energy), int(natoms_force), out.data_ptr(), F.data_ptr(),
            eatom.data_ptr() if eatom is not None else None, stream), self.ctx)
        if sync:
            capi.check(L.mtp_sync(self.ctx, stream), self.ctx)
        E, stress = out[0], out[1:7]
        if not return_all:
            return E, F, stress
        gptr = L.mtp_last_slot_gradients(self.ctx)
        g = torch.as_tensor(_DevArray(gptr, (A, K, 3), "<f8"), device=dev).clone()
        return E, F, stress, dict(virial6=out[7:13], eatom=eatom, status=out[13], g=g)

    # ------------------------------------------------------------------
    # Pieces of the atom kernel (device arrays in, torch tensors out); the
    # distilled head and its training are built from these.
    def _rows(self, itypes, all_rijs, all_jtypes):
        import torch

        dev = torch.device("cuda")
        it, rij, jt = _i32(itypes, dev), _f64(all_rijs, dev), _i32(all_jtypes, dev)
        A, K = jt.shape
        if rij.shape != (A, K, 3) or it.shape != (A,):
            raise ValueError("inconsistent padded shapes")
        return it, rij, jt, A, K, torch.cuda.current_stream().cuda_stream

    def descriptor(self, itypes, all_rijs, all_jtypes):
        """Basis values b[A, S] (m[alpha_moment_mapping]) of every row."""
        import torch

        it, rij, jt, A, K, st = self._rows(itypes, all_rijs, all_jtypes)
        S = len(self.params.moment_coeffs)
        b = torch.empty((A, S), dtype=torch.float64, device=rij.device)
        capi.check(capi.lib().mtp_descriptor_device(self.ctx, A, K, it.data_ptr(), rij.data_ptr(), jt.data_ptr(),
                                                    b.data_ptr(), st), self.ctx)
        return b

    def slot_grad_xi(self, itypes, all_rijs, all_jtypes, xi, with_virial=False):
        """Slot gradients g[A, K, 3] of xi_row . b_row for per-row xi [A, S] (or one xi [S])."""
        import torch

        it, rij, jt, A, K, st = self._rows(itypes, all_rijs, all_jtypes)
        x = _f64(xi, rij.device)
        S = len(self.params.moment_coeffs)
        stride = 0 if x.dim() == 1 else S
        if x.shape[-1] != S or (x.dim() == 2 and x.shape[0] != A):
            raise ValueError("xi must be [S] or [A, S]")
        g = torch.empty((A, K, 3), dtype=torch.float64, device=rij.device)
        W = torch.empty((A, 9), dtype=torch.float64, device=rij.device) if with_virial else None
        capi.check(capi.lib().mtp_slot_grad_xi_device(
            self.ctx, A, K, it.data_ptr(), rij.data_ptr(), jt.data_ptr(), x.data_ptr(), stride, g.data_ptr(),
            W.data_ptr() if W is not None else None, st), self.ctx)
        return (g, W) if with_virial else g

    def basis_jacobian(self, itypes, all_rijs, all_jtypes):
        """Dense J[A, S, K, 3] = d b / d r."""
        import torch

        it, rij, jt, A, K, st = self._rows(itypes, all_rijs, all_jtypes)
        S = len(self.params.moment_coeffs)
        J = torch.empty((A, S, K, 3), dtype=torch.float64, device=rij.device)
        capi.check(capi.lib().mtp_basis_jacobian_device(self.ctx, A, K, it.data_ptr(), rij.data_ptr(),
                                                        jt.data_ptr(), J.data_ptr(), st), self.ctx)
        return J

    def gather_forces(self, all_js, g, natoms_force):
        """Newton-3 gather (jax_backend.py:263-303) of slot gradients g[A, K, 3]."""
        import torch

        dev = torch.device("cuda")
        js, gg = _i32(all_js, dev), _f64(g, dev)
        A, K = js.shape
        F = torch.empty((A, 3), dtype=torch.float64, device=dev)
        capi.check(capi.lib().mtp_gather_forces_device(self.ctx, A, K, js.data_ptr(), gg.data_ptr(),
                                                       int(natoms_force), F.data_ptr(),
                                                       torch.cuda.current_stream().cuda_stream), self.ctx)
        return F

    # ------------------------------------------------------------------
    def compute_rows(self, natoms, rij_rows, types, nbr_rows, nbr_counts, nbr_type_rows, volume):
        """The pair style's call (zero_overhead.hpp:58-64) from numpy row arrays.

        rij_rows [natoms, 3*K], nbr_rows/nbr_type_rows [natoms, K].
        Returns (energy, forces[natoms,3], stress6, virial6).
        """
        K = self.max_nbr
        rij_rows = np.ascontiguousarray(rij_rows, dtype=np.float64).reshape(natoms, 3 * K)
        nbr_rows = np.ascontiguousarray(nbr_rows, dtype=np.int32).reshape(natoms, K)
        nbr_type_rows = np.ascontiguousarray(nbr_type_rows, dtype=np.int32).reshape(natoms, K)
        types = np.ascontiguousarray(types, dtype=np.int32)
        counts = np.ascontiguousarray(nbr_counts, dtype=np.int32)

        def rows(a):
            return (ctypes.c_void_p * max(natoms, 1))(*[a[i].ctypes.data for i in range(natoms)])

        pr, pn, pt = rows(rij_rows), rows(nbr_rows), rows(nbr_type_rows)
        E = ctypes.c_double()
        F = np.zeros((natoms, 3))
        s6 = np.zeros(6)
        v6 = np.zeros(6)
        vp = ctypes.c_void_p
        capi.check(capi.lib().mtp_compute(
            self.ctx, natoms, int(counts.sum()), ctypes.cast(pr, vp), types.ctypes.data_as(vp), ctypes.cast(pn, vp),
            counts.ctypes.data_as(vp), ctypes.cast(pt, vp), float(volume), ctypes.byref(E),
            F.ctypes.data_as(vp), s6.ctypes.data_as(vp), v6.ctypes.data_as(vp)), self.ctx)
        return E.value, F, s6, v6

    def compute_full_list(self, nlocal, x, types, tags, ilist, neighbors, prd, periodic=(1, 1, 1), volume=None,
                          list_is_new=True):
        """The pair style's force call with packing and staging on the GPU (mtp_compute_full_list).

        x [nall, 3], types [nall] (1-based), tags [nall]; ilist [inum] and neighbors (a list of
        index arrays, one per local atom) form the full neighbor list.  Returns (energy,
        forces[nlocal, 3], stress6, virial6, n_truncated)."""
        vp = ctypes.c_void_p
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        types = np.ascontiguousarray(types, dtype=np.int32)
        tags = np.ascontiguousarray(tags, dtype=np.int32)
        ilist = np.ascontiguousarray(ilist, dtype=np.int32)
        rows = [np.ascontiguousarray(r, dtype=np.int32) for r in neighbors]
        numneigh = np.array([len(r) for r in rows], dtype=np.int32)
        first = (vp * max(len(rows), 1))(*[r.ctypes.data if len(r) else None for r in rows])
        prd = np.ascontiguousarray(prd, dtype=np.float64)
        per = np.ascontiguousarray(periodic, dtype=np.int32)
        vol = float(np.prod(prd)) if volume is None else float(volume)
        E = ctypes.c_double()
        F = np.zeros((nlocal, 3))
        s6, v6 = np.zeros(6), np.zeros(6)
        trunc = ctypes.c_longlong()
        capi.check(capi.lib().mtp_compute_full_list(
            self.ctx, int(nlocal), len(x), x.ctypes.data_as(vp), types.ctypes.data_as(vp), tags.ctypes.data_as(vp),
            len(ilist), ilist.ctypes.data_as(vp), numneigh.ctypes.data_as(vp), ctypes.cast(first, vp),
            int(bool(list_is_new)), prd.ctypes.data_as(vp), per.ctypes.data_as(vp), vol, ctypes.byref(E), F.ctypes.data_as(vp),
            s6.ctypes.data_as(vp), v6.ctypes.data_as(vp), ctypes.byref(trunc)), self.ctx)
        return E.value, F, s6, v6, trunc.value

    def last_atom_values(self, natoms):
        """Per-atom energies [natoms] and LAMMPS virials [natoms, 6] (xx yy zz xy xz yz) of the last
        compute_rows call; the virial rows sum to its virial6."""
        eatom = np.zeros(natoms)
        vatom = np.zeros((natoms, 6))
        vp = ctypes.c_void_p
        capi.check(capi.lib().mtp_last_atom_values(self.ctx, natoms, eatom.ctypes.data_as(vp),
                                                   vatom.ctypes.data_as(vp)), self.ctx)
        return eatom, vatom
