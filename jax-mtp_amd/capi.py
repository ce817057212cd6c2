"""ctypes binding of libmtp_hip.so (the C ABI declared in include/mtp_hip.h).

The library must be built in-tree (``python jax-mtp_amd/build.py``); there is
no fallback: loading fails loudly if it is missing.
"""
from __future__ import annotations

import ctypes
import pathlib

LIB_PATH = pathlib.Path(__file__).resolve().parent / "libmtp_hip.so"
HEADER = pathlib.Path(__file__).resolve().parent.parent / "include" / "mtp_hip.h"

MTP_OK = 0
MTP_ERR_ARG = -1
MTP_ERR_FILE = -2
MTP_ERR_CAPACITY = -3
MTP_ERR_HIP = -4
MTP_ERR_NONFINITE = -5
MTP_ERR_UNSUPPORTED = -6

_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_LL = ctypes.c_longlong

SIGNATURES = {
    "mtp_open": (_I, [ctypes.c_char_p, _I, _I, _D, _I, ctypes.POINTER(_P)]),
    "mtp_open_text": (_I, [ctypes.c_char_p, _I, _I, _D, _I, ctypes.POINTER(_P)]),
    "mtp_close": (None, [_P]),
    "mtp_last_error": (ctypes.c_char_p, [_P]),
    "mtp_global_error": (ctypes.c_char_p, []),
    "mtp_set_precision": (_I, [_P, _I]),
    "mtp_info": (_I, [_P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "mtp_set_params": (_I, [_P, _P, _P, _P, _D]),
    "mtp_set_params_device": (_I, [_P, _P, _P, _P, _P]),
    "mtp_compute": (_I, [_P, _I, _I, _P, _P, _P, _P, _P, _D, _P, _P, _P, _P]),
    "mtp_compute_full_list": (_I, [_P, _I, _I, _P, _P, _P, _I, _P, _P, _P, _I, _P, _P, _D, _P, _P, _P, _P, _P]),
    "mtp_compute_padded_device": (_I, [_P, _I, _I, _P, _P, _P, _P, _I, _D, _I, _I, _P, _P, _P, _P]),
    "mtp_last_slot_gradients": (_P, [_P]),
    "mtp_last_atom_values": (_I, [_P, _I, _P, _P]),
    "mtp_sync": (_I, [_P, _P]),
    "mtp_set_variant": (_I, [_P, _I]),
    "mtp_get_variant": (_I, [_P, _P, _P]),
    "mtp_autotune": (_I, [_P, _I, _I, _P, _P, _P, _P, _I, _P, _P]),
    "mtp_profile": (_I, [_P, _I]),
    "mtp_profile_read": (_I, [_P, _P, _P]),
    "mtp_batch_forward": (_I, [_P, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "mtp_batch_param_grad": (_I, [_P, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "mtp_descriptor_device": (_I, [_P, _I, _I, _P, _P, _P, _P, _P]),
    "mtp_slot_grad_xi_device": (_I, [_P, _I, _I, _P, _P, _P, _P, _I, _P, _P, _P]),
    "mtp_basis_jacobian_device": (_I, [_P, _I, _I, _P, _P, _P, _P, _P]),
    "mtp_jacobian_apply_device": (_I, [_I, _I, _I, _P, _P, _P, _P]),
    "mtp_jacobian_tapply_device": (_I, [_I, _I, _I, _I, _P, _P, _P, _P, _P]),
    "mtp_gather_forces_device": (_I, [_P, _I, _I, _P, _P, _I, _P, _P]),
    "mtp_gather_forces_noalloc": (_I, [_I, _I, _P, _P, _I, _P, _P]),
    "mtp_head_theta_size": (_I, []),
    "mtp_head_open": (_I, [_P, _P, _P, _P, _P, _P, _P, _P, _D, _I, ctypes.POINTER(_P)]),
    "mtp_head_close": (None, [_P]),
    "mtp_head_last_error": (ctypes.c_char_p, [_P]),
    "mtp_head_set_theta": (_I, [_P, _P, _I, _P]),
    "mtp_head_compute_padded_device": (_I, [_P, _I, _I, _P, _P, _P, _P, _I, _D, _I, _I, _P, _P, _P, _P, _P, _P]),
    "mtp_head_eval_device": (_I, [_P, _I, _P, _P, _P, _P, _P, _P]),
    "mtp_head_param_grad_device": (_I, [_P, _I, _P, _P, _P, _P, _P, _P]),
    "mtp_check_file": (_I, [ctypes.c_char_p, _I, _P, _P, _P, _P, _P, _P, _P]),
    "mtp_pack_full_list": (_LL, [_I, _P, _P, _P, _I, _P, _P, _P, _D, _P, _P, _I, _I, _P, _P, _P, _P, _P, _P]),
    "mtp_stage_host": (_I, [_I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
}

_lib = None


class MTPError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def _load_torch_hip_runtime() -> None:
    """Load PyTorch (and with it its HIP runtime) before the library.

    PyTorch-ROCm ships its own libamdhip64 (SONAME libamdhip64.so.7).  If the
    library were loaded first, the dynamic linker would bind it to the system
    ROCm copy and the process would end up with two HIP runtimes, after which
    PyTorch no longer sees the GPU.  Loading PyTorch first makes the library's
    libamdhip64.so.7 dependency resolve to the runtime PyTorch already uses.
    """
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python jax-mtp_amd/build.py` "
                               "(the HIP extension is required; there is no fallback)")
        _load_torch_hip_runtime()
        L = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int, ctx=None, head=None) -> int:
    if rc < 0:
        L = lib()
        if head:
            msg = L.mtp_head_last_error(head)
        else:
            msg = L.mtp_last_error(ctx) if ctx else L.mtp_global_error()
        raise MTPError(rc, (msg or b"").decode())
    return rc
