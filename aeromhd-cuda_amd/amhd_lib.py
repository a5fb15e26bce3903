"""ctypes binding of libaeromhd_hip.so (the C ABI of include/aeromhd.h).

The library is the only compute path: if it is missing or no GPU is
visible, every entry point raises; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("AMHD_LIB", os.path.join(HERE, "libaeromhd_hip.so"))

AMHD_F64 = 0
AMHD_F32 = 1

_vp = C.c_void_p
_d = C.c_double
_i = C.c_int
_l = C.c_long
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class AmhdParams(C.Structure):
    """Mirror of ``amhd_params`` (include/aeromhd.h)."""

    _fields_ = [
        ("gamma", _d), ("dx", _d), ("dy", _d), ("x_min", _d), ("y_min", _d), ("glm_alpha", _d),
        ("bc_x_periodic", _i), ("bc_y_periodic", _i), ("powell", _i), ("reserved", _i),
    ]


# name -> (restype, argtypes); every function returns int status
SIGNATURES = {
    "amhd_last_error": (C.c_char_p, []),
    "amhd_abi_version": (_i, []),
    "amhd_create": (_i, [C.POINTER(_vp), _i, _i, _i, C.POINTER(AmhdParams)]),
    "amhd_destroy": (_i, [_vp]),
    "amhd_set_rows": (_i, [_vp, _i]),
    "amhd_get_rows": (_i, [_vp]),
    "amhd_set_packing": (_i, [_vp, _i]),
    "amhd_set_balance": (_i, [_vp, _i]),
    "amhd_upload_state": (_i, [_vp, _vp, _i]),
    "amhd_download_state": (_i, [_vp, _vp, _i]),
    "amhd_compute_dt": (_i, [_vp, _d, _dp, _dp]),
    "amhd_step_ssprk3": (_i, [_vp, _d, _d, _d]),
    "amhd_run": (_i, [_vp, _i, _d, _d, _i, _i, _dp, _ip, _dp]),
    "amhd_run_profiled": (_i, [_vp, _i, _d, _d, _i, _dp, _ip, _dp, _dp, _dp]),
    "amhd_boundary_energy": (_i, [_vp, _dp, _dp]),
    "amhd_reset_boundary_energy": (_i, [_vp, _d]),
    "amhd_compute_rhs": (_i, [_vp, _vp, _vp, _d]),
    "amhd_global_sums": (_i, [_vp, _dp]),
    "amhd_concentration_profile": (_i, [_vp, _dp, _dp]),
    "amhd_flow_diagnostics": (_i, [_vp, _dp, _dp, _dp]),
    "amhd_set_halo": (_i, [_vp, _vp, _vp]),
    "amhd_stage": (_i, [_vp, _i, _d, _d, _d]),
    "amhd_edge_rows": (_i, [_vp, _i, _i, _vp]),
    "amhd_cfl_maxima": (_i, [_vp, _dp]),
    "amhd_dt_from_maxima": (_i, [_dp, _d, _d, _d, _dp, _dp]),
    "amhd_synchronize": (_i, [_vp]),
    "amhd_state_check": (_i, [_vp, _dp]),
    "amhd_state_ptr": (_i, [_vp, C.POINTER(_vp)]),
    "amhd_get_stream": (_i, [_vp, C.POINTER(_vp)]),
    "amhd_cons_to_prim": (_i, [_vp, _vp, _l, _d, _vp]),
    "amhd_prim_to_cons": (_i, [_vp, _vp, _l, _d, _vp]),
    "amhd_muscl": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp]),
    "amhd_hlld_flux": (_i, [_vp, _vp, _vp, _vp, _dp, _l, _i, _d, _d, _d, _d, _i, _vp]),
    "amhd_apply_bc": (_i, [_vp, _i, _i, _i, _i, _vp]),
    "amhd_enforce_scalar_bounds": (_i, [_vp, _l, _vp]),
    "amhd_compute_dt_w": (_i, [_vp, _l, _d, _d, _d, _d, _dp, _dp, _vp]),
    "amhd_compute_dt_w_noalloc": (_i, [_vp, _l, _d, _d, _d, _d, _dp, _dp, _vp]),
    "amhd_batch_create": (_i, [C.POINTER(_vp), _i, _ip, _ip, _i, C.POINTER(AmhdParams)]),
    "amhd_batch_destroy": (_i, [_vp]),
    "amhd_batch_set_rows": (_i, [_vp, _i]),
    "amhd_batch_info": (_i, [_vp, _ip, _ip, _ip]),
    "amhd_batch_upload": (_i, [_vp, _i, _vp, _i]),
    "amhd_batch_download": (_i, [_vp, _i, _vp, _i]),
    "amhd_batch_state_ptrs": (_i, [_vp, C.POINTER(_vp), C.POINTER(_vp)]),
    "amhd_batch_last_dt": (_i, [_vp, _dp]),
    "amhd_batch_run": (_i, [_vp, _i, _dp, _dp, _ip, _i, _dp, _ip]),
    "amhd_conv3x3_kpad": (_i, [_i]),
    "amhd_conv3x3_fwd": (_i, [_vp, _i, _i, _i, _i, _vp, _vp, _i, _i, _vp, _i, _vp]),
    "amhd_conv3x3_bwd_data": (_i, [_vp, _i, _i, _i, _i, _vp, _i, _vp, _vp, _vp]),
    "amhd_conv3x3_bwd_weight": (_i, [_vp, _vp, _i, _i, _i, _i, _i, _vp, _vp, _vp, _i, _vp]),
    "amhd_mfma_probe": (_i, [_vp, _vp, _vp, _vp]),
    "amhd_closure_features": (_i, [C.POINTER(_vp), _i, C.POINTER(_i), _i, _i, _i, C.POINTER(C.c_float),
                                   C.POINTER(C.c_float), _vp, _vp]),
    "amhd_closure_source": (_i, [_vp, C.POINTER(_vp), C.POINTER(_vp), _i, C.POINTER(_i), _i, _i, _i,
                                 C.POINTER(C.c_float), _d, _d, _dp, _d, _d, _vp, C.POINTER(_vp), _i, _vp]),
    "amhd_closure_check": (_i, [C.POINTER(_vp), C.POINTER(_vp), _dp, _i, _i, _i, _i, _d, _d, _d, _vp, _dp, _vp]),
    "amhd_closure_apply": (_i, [C.POINTER(_vp), C.POINTER(_vp), _dp, _i, _i, _i, _i, _vp]),
    "amhd_closure_sharpen": (_i, [C.POINTER(_vp), C.POINTER(_vp), _i, _i, _i, _i, _d, _vp, _vp]),
}

_lib = None
_lock = threading.Lock()


class AmhdError(RuntimeError):
    pass


def load(path: str | None = None) -> C.CDLL:
    """Load the library (once) and declare every signature."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        p = path or LIB_PATH
        if not os.path.exists(p):
            raise AmhdError(
                f"{p} is missing: build it first (python aeromhd-cuda_amd/build_lib.py). "
                "The HIP library is the only compute path.")
        lib = C.CDLL(p)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = _lib.amhd_last_error().decode(errors="replace") if _lib is not None else ""
        raise AmhdError(f"{what or 'aeromhd'} failed (code {rc}): {msg}")


def call(name: str, *args) -> None:
    lib = load()
    check(getattr(lib, name)(*args), name)


def require_gpu():
    """Return torch with a visible GPU, or raise: no silent CPU path."""
    import torch

    if not torch.cuda.is_available():
        raise AmhdError("no GPU visible: the aeromhd HIP path needs a ROCm device")
    load()
    return torch


def stream_handle(torch_mod) -> C.c_void_p:
    return C.c_void_p(torch_mod.cuda.current_stream().cuda_stream)
