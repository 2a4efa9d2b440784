"""libcc_amd — closed-shell CCSD(T) triples correction on AMD MI355X (gfx950).

The product is the C-ABI shared library ``libcc_amd/libcc_mi355x.so`` (declared in
``include/libcc_mi355x.h``): hand-written HIP kernels behind a C host layer.  This
Python package only loads it with ctypes and mirrors the reference's driver
interface for tests and benchmarks (``libcc_amd.triples``).  There is no CPU
fallback: if the library is missing, loading raises.
"""

from __future__ import annotations

import ctypes
import os
import sys

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcc_mi355x.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "libcc_mi355x.h")

_LIB = None


class CcxError(RuntimeError):
    """A libcc_mi355x call returned a non-zero status."""

    def __init__(self, status: int, what: str):
        self.status = status
        msg = lib().ccx_error_string(status).decode() if _LIB is not None else str(status)
        super().__init__(f"{what}: status {status} ({msg})")


def lib() -> ctypes.CDLL:
    """Load libcc_mi355x.so (built by ``make -C libcc_amd``); raises if it is absent."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run `make -C libcc_amd` (no CPU fallback exists)")
    # PyTorch bundles its own HIP runtime under the same soname.  Loaded first, it is the one
    # runtime of the process and the library binds to it; loaded after the library has
    # initialised the system runtime, torch finds no GPU.  So torch, where installed, goes first.
    if "torch" not in sys.modules:
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
    L = ctypes.CDLL(LIB_PATH)
    dp, i64, vp = ctypes.POINTER(ctypes.c_double), ctypes.c_int64, ctypes.c_void_p
    ctxp = ctypes.c_void_p
    L.ccx_error_string.argtypes = [ctypes.c_int]
    L.ccx_error_string.restype = ctypes.c_char_p
    L.ccx_last_status.argtypes = []
    L.ccx_last_status.restype = ctypes.c_int
    L.ccx_create.argtypes = [i64, i64, ctypes.POINTER(ctxp)]
    L.ccx_create.restype = ctypes.c_int
    L.ccx_destroy.argtypes = [ctxp]
    L.ccx_destroy.restype = ctypes.c_int
    L.ccx_load_host.argtypes = [ctxp] + [dp] * 7 + [ctypes.c_int]
    L.ccx_load_host.restype = ctypes.c_int
    L.ccx_load_device.argtypes = [ctxp] + [vp] * 7 + [ctypes.c_int]
    L.ccx_load_device.restype = ctypes.c_int
    L.ccx_energy.argtypes = [ctxp, ctypes.c_int, dp, dp, vp, ctypes.c_int]
    L.ccx_energy.restype = ctypes.c_int
    L.ccx_energy_tuples.argtypes = [ctxp, ctypes.c_int, i64, ctypes.POINTER(i64), dp, dp, vp, ctypes.c_int]
    L.ccx_energy_tuples.restype = ctypes.c_int
    L.ccx_tuple_w.argtypes = [ctxp, i64, i64, i64, dp]
    L.ccx_tuple_w.restype = ctypes.c_int
    L.ccx_w_abc.argtypes = [ctxp, i64, i64, i64, dp]
    L.ccx_w_abc.restype = ctypes.c_int
    L.ccx_energy_abc.argtypes = [ctxp, dp]
    L.ccx_energy_abc.restype = ctypes.c_int
    L.ccx_last_timing.argtypes = [ctxp, dp, dp]
    L.ccx_last_timing.restype = ctypes.c_int
    L.ccx_work_stats.argtypes = [ctxp, ctypes.POINTER(i64), dp]
    L.ccx_work_stats.restype = ctypes.c_int
    L.ccx_gradient.argtypes = [ctxp, dp, dp] + [vp] * 5 + [ctypes.c_int]
    L.ccx_gradient.restype = ctypes.c_int
    L.ccx_gradient_form.argtypes = [ctxp, ctypes.c_int, dp, dp] + [vp] * 5 + [ctypes.c_int]
    L.ccx_gradient_form.restype = ctypes.c_int
    L.ccx_set_stream.argtypes = [ctxp, vp]
    L.ccx_set_stream.restype = ctypes.c_int
    L.ccx_energy_static.argtypes = [i64, i64] + [vp] * 9 + [vp]
    L.ccx_energy_static.restype = ctypes.c_int
    pi64 = ctypes.POINTER(i64)
    L.common_cc_init_.argtypes = [pi64, pi64]
    L.common_cc_init_.restype = None
    L.ccx_register_vvvo_.argtypes = [pi64, dp, pi64]
    L.ccx_register_vvvo_.restype = None
    L.ccx_set_d_vvvo_.argtypes = [pi64]
    L.ccx_set_d_vvvo_.restype = None
    L.cc_triples_acc_.argtypes = [dp] * 9 + [pi64, dp, dp]
    L.cc_triples_acc_.restype = None
    L.cc_triples_.argtypes = [dp] * 12
    L.cc_triples_.restype = None
    L.ccx_get_ets_.argtypes = []
    L.ccx_get_ets_.restype = ctypes.c_double
    L.ccx_get_etd_.argtypes = []
    L.ccx_get_etd_.restype = ctypes.c_double
    L.ijk_tuple_.argtypes = [pi64, pi64, pi64] + [dp] * 13
    L.ijk_tuple_.restype = None
    L.t1wt3_ijk_.argtypes = [pi64] * 5 + [dp] * 10
    L.t1wt3_ijk_.restype = None
    L.ddcc_t_ijk_acc_.argtypes = [pi64] * 5 + [dp] * 10
    L.ddcc_t_ijk_acc_.restype = None
    for name in ("ddcc_t_ijj_acc_", "ddcc_t_iij_acc_"):
        getattr(L, name).argtypes = [pi64] * 4 + [dp] * 12
        getattr(L, name).restype = None
    for name in ("trant3_1_", "trant3_3_", "trant3_4_", "trant3_5_"):
        getattr(L, name).argtypes = [pi64, dp]
        getattr(L, name).restype = None
    for name in ("trant3_acc_", "trant3_smp_"):
        getattr(L, name).argtypes = [dp, pi64, pi64]
        getattr(L, name).restype = None
    L.tranmd_23_.argtypes = [dp, pi64, pi64, pi64, pi64]
    L.tranmd_23_.restype = None
    for name in ("trant3_1_async_", "trant3_4_async_"):
        getattr(L, name).argtypes = [pi64, pi64, dp]
        getattr(L, name).restype = None
    L.tranmd_23_acc_.argtypes = [pi64, dp, pi64, pi64, pi64, pi64]
    L.tranmd_23_acc_.restype = None
    L.tranmd_smp_.argtypes = [dp, pi64, pi64, pi64, pi64, pi64]
    L.tranmd_smp_.restype = None
    L.ddcc_t_ijk_big_.argtypes = [pi64] * 5 + [dp] * 13
    L.ddcc_t_ijk_big_.restype = None
    L.ijk_gpu_init_.argtypes = [pi64, pi64] + [dp] * 6
    L.ijk_gpu_init_.restype = None
    L.ijk_gpu_driver_.argtypes = [pi64] * 5 + [dp] * 7
    L.ijk_gpu_driver_.restype = None
    L.ijk_gpu_finalize_.argtypes = [pi64, pi64, dp]
    L.ijk_gpu_finalize_.restype = None
    cp = ctypes.c_char_p
    L.ccx_restart_dims.argtypes = [cp, ctypes.POINTER(i64), ctypes.POINTER(i64)]
    L.ccx_restart_dims.restype = ctypes.c_int
    L.ccx_read_restart.argtypes = [cp, i64, i64] + [dp] * 7
    L.ccx_read_restart.restype = ctypes.c_int
    L.ccx_write_restart.argtypes = [cp, i64, i64] + [dp] * 7 + [ctypes.c_int]
    L.ccx_write_restart.restype = ctypes.c_int
    L.ccx_load_restart.argtypes = [ctxp, cp]
    L.ccx_load_restart.restype = ctypes.c_int
    _LIB = L
    return L
