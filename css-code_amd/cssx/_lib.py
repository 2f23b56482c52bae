"""ctypes binding of libcssx.so (the C ABI declared in include/cssx.h).

The library is built in-tree by ``make -C css-code_amd/csrc`` (``__graft_entry__.build()``).
There is no fallback: if the library is missing or fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcssx.so")

CSSX_OK = 0
CSSX_ENOMEM = 12
CSSX_EINVAL = 22
CSSX_ENAN = 33
CSSX_EHIP = 100
CSSX_ESIZE = 27
PREC_F64 = 0
PREC_F32 = 1
AXIS_R, AXIS_THETA, AXIS_PHI = 0, 1, 2
PHASES = ("deriv_r", "deriv_theta", "deriv_phi", "pointwise", "rk_update", "other", "sld", "rhs_main",
          "rhs_final")

_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)


class CssxError(RuntimeError):
    def __init__(self, code: int, where: str):
        self.code = code
        super().__init__(f"{where}: cssx error {code} ({strerror(code)})")


class Patch(ctypes.Structure):
    """cssx_patch (include/cssx.h)."""
    _fields_ = [("nr", ctypes.c_int), ("nth", ctypes.c_int), ("nphi", ctypes.c_int), ("in_", ctypes.c_void_p),
                ("d_r", ctypes.c_void_p), ("d_th", ctypes.c_void_p), ("d_ph", ctypes.c_void_p),
                ("bdry", ctypes.c_int), ("dri", ctypes.c_double), ("dti", ctypes.c_double), ("dpi", ctypes.c_double)]


class Grid(ctypes.Structure):
    """cssx_grid (include/cssx.h)."""
    _fields_ = [("nr", ctypes.c_int), ("nth", ctypes.c_int), ("nphi", ctypes.c_int),
                ("r_lo_wall", ctypes.c_int), ("r_hi_wall", ctypes.c_int),
                ("th_lo_wall", ctypes.c_int), ("th_hi_wall", ctypes.c_int),
                ("dri", ctypes.c_double), ("dti", ctypes.c_double), ("dpi", ctypes.c_double)]


_lib = None

# exported symbols and their ctypes signatures (checked against include/cssx.h by the tests)
SIGNATURES = {
    "cssx_version": ([], ctypes.c_int),
    "cssx_strerror": ([ctypes.c_int], ctypes.c_char_p),
    "cssx_ctx_create": ([ctypes.POINTER(Grid), ctypes.POINTER(_vp)], ctypes.c_int),
    "cssx_ctx_destroy": ([_vp], None),
    "cssx_deriv": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, _vp, _vp],
                   ctypes.c_int),
    "cssx_compact_fd6": ([ctypes.c_int, ctypes.c_long, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                          _vp, _vp, _vp, _vp, _vp, _vp], ctypes.c_int),
    "cssx_compact_fd6_noalloc": ([ctypes.c_int, ctypes.c_long, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                  _vp, _vp, _vp, _vp, _vp, _vp], ctypes.c_int),
    "cssx_mhd_create": ([_vp, _vp, _vp, ctypes.POINTER(_vp)], ctypes.c_int),
    "cssx_mhd_destroy": ([_vp], None),
    "cssx_rhs": ([_vp, ctypes.c_int, _vp, _vp, _vp, _vp], ctypes.c_int),
    "cssx_bound": ([_vp, ctypes.c_int, _vp, _vp], ctypes.c_int),
    "cssx_rk4": ([_vp, ctypes.c_int, _vp, ctypes.c_double, ctypes.c_int, _vp], ctypes.c_int),
    "cssx_cfl": ([_vp, _vp, ctypes.c_double, ctypes.POINTER(ctypes.c_double), _vp], ctypes.c_int),
    "cssx_conservation": ([_vp, _vp, _dp, _vp], ctypes.c_int),
    "cssx_diffusion": ([_vp, ctypes.c_int, _vp, _vp, _vp, _vp], ctypes.c_int),
    "cssx_rk3_tvd": ([_vp, ctypes.c_int, _vp, ctypes.c_double, ctypes.c_int, _vp], ctypes.c_int),
    "cssx_rk2_tvd": ([_vp, ctypes.c_int, _vp, ctypes.c_double, ctypes.c_int, _vp], ctypes.c_int),
    "cssx_timing_enable": ([ctypes.c_int], ctypes.c_int),
    "cssx_timing_read": ([ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_long),
                          ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "cssx_plan_tables": ([ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                          _dp, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    # batched patches
    "cssx_patchset_create": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)], ctypes.c_int),
    "cssx_patchset_destroy": ([_vp], None),
    "cssx_patchset_run": ([_vp, ctypes.c_int, _vp], ctypes.c_int),
    "cssx_patchset_autotune": ([_vp, ctypes.c_int, _vp, ctypes.POINTER(ctypes.c_int), _fp], ctypes.c_int),
    "cssx_patch_fill": ([_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_ulonglong, _vp],
                        ctypes.c_int),
    "cssx_normal_fill": ([_vp, ctypes.c_int, ctypes.c_long, ctypes.c_ulonglong, _vp], ctypes.c_int),
    # learned closure
    "cssx_mlp_create": ([ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_ulonglong,
                         ctypes.POINTER(_vp)], ctypes.c_int),
    "cssx_mlp_destroy": ([_vp], None),
    "cssx_mlp_num_params": ([_vp], ctypes.c_long),
    "cssx_mlp_set_params": ([_vp, _fp], ctypes.c_int),
    "cssx_mlp_get_params": ([_vp, _fp, _fp], ctypes.c_int),
    "cssx_mlp_set_norm": ([_vp, _fp, _fp, _fp, _fp], ctypes.c_int),
    "cssx_mlp_forward": ([_vp, _vp, ctypes.c_int, _vp, _vp], ctypes.c_int),
    "cssx_mlp_grad": ([_vp, _vp, _vp, ctypes.c_int, _vp, _vp], ctypes.c_int),
    "cssx_mlp_train_step": ([_vp, _vp, _vp, ctypes.c_int, ctypes.c_float, _vp, _vp], ctypes.c_int),
    "cssx_gemm_bf16": ([_vp, ctypes.c_long, ctypes.c_long, _vp, ctypes.c_long, ctypes.c_long, ctypes.c_int,
                        ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_long, _vp], ctypes.c_int),
    "cssx_box_filter": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                         _vp, _vp], ctypes.c_int),
    "cssx_closure_features": ([_vp, _vp, ctypes.c_int, ctypes.c_int, _vp, _vp, ctypes.c_int, ctypes.c_int,
                               ctypes.c_int, ctypes.c_long, ctypes.c_long, _vp, _vp, _vp], ctypes.c_int),
    "cssx_closure_targets": ([_vp, _vp, ctypes.c_int, ctypes.c_long, _vp, _vp, _vp, _vp, _vp], ctypes.c_int),
    "cssx_mhd_set_closure": ([_vp, _vp], ctypes.c_int),
    "cssx_closure_tendency": ([_vp, _vp, ctypes.c_int, ctypes.c_int, _vp, _vp, ctypes.c_int, ctypes.c_int,
                               ctypes.c_int, _vp, _vp], ctypes.c_int),
    "cssx_colstats": ([_vp, ctypes.c_long, ctypes.c_int, _dp, _dp, _vp], ctypes.c_int),
    "cssx_field_stats": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                          _vp, _vp, _dp, _dp, _vp], ctypes.c_int),
    "cssx_gather_rows": ([_vp, ctypes.c_long, _vp, ctypes.c_long, _vp, _vp], ctypes.c_int),
    "cssx_standardize_rows": ([_vp, ctypes.c_long, ctypes.c_int, _vp, _vp, _vp, _vp, _vp], ctypes.c_int),
}


def lib():
    """Load libcssx.so once; raises if it is absent (no CPU fallback exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"libcssx.so not built at {LIB_PATH}: run __graft_entry__.build() "
                              "(make -C css-code_amd/csrc)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def strerror(code: int) -> str:
    try:
        return lib().cssx_strerror(code).decode()
    except Exception:  # library missing: the code itself is still meaningful
        return "unknown"


def check(code: int, where: str) -> None:
    if code != CSSX_OK:
        raise CssxError(code, where)


def timing_enable(on: bool = True) -> None:
    """Bracket every library kernel launch with hipEvents (cssx_timing_enable)."""
    check(lib().cssx_timing_enable(1 if on else 0), "cssx_timing_enable")


def timing_read() -> dict:
    """{phase: (ms, launches, algorithmic bytes)} since the last read; synchronizes."""
    n = len(PHASES)
    ms = (ctypes.c_double * n)()
    cnt = (ctypes.c_long * n)()
    by = (ctypes.c_double * n)()
    check(lib().cssx_timing_read(ms, cnt, by), "cssx_timing_read")
    return {PHASES[i]: (ms[i], cnt[i], by[i]) for i in range(n)}
