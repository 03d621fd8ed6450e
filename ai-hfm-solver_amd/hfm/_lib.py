"""ctypes binding of libhfm.so (the C ABI declared in include/hfm.h).

torch is imported first so that the HIP runtime torch ships is the one the
library binds to (both carry the soname libamdhip64.so.7); the library then
launches on torch's streams.  With HFM_NO_TORCH=1 in the environment torch is
not imported (the library binds to the system HIP runtime; stream_handle is
unavailable) -- for torch-free drivers such as counter-collection runs.
There is no fallback: if the library is missing every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np
if os.environ.get("HFM_NO_TORCH") != "1":
    import torch  # noqa: F401  (loads the HIP runtime before libhfm.so)

from ._build import LIB_PATH

HFM_OK = 0

STORE = {"f32": 0, "bf16": 1, "f16": 2}
FORM = {"apply": 0, "resid": 1, "jacobi": 2, "subtract": 3, "axpy": 4, "jacobi_mul": 5}
FORM_ZERO_OUTLET = 16
CORRECT_ZERO_OUTLET = 1
CORRECT_RECIPROCAL = 2
SUM_SOURCE_BC = 1
BC = {"top_copy": 0, "pressure": 1, "source": 2, "indicator": 3, "zero_outlet": 4}
SWEEP = {"zero": 0, "plain": 1, "correct": 2}
FIN = {"init": 0, "beta": 1, "alpha": 2, "check": 3}
STATUS = {0: "running", 1: "converged", 2: "maxit", 3: "breakdown"}
METHOD = {"fcg": 0, "refine": 1}

TASK_DTYPE = np.dtype([
    ("D", "<i4"), ("H", "<i4"), ("W", "<i4"), ("inst", "<i4"),
    ("blk0", "<i4"), ("nblk", "<i4"), ("slot0", "<i4"), ("tile_rows", "<i4"),
    ("magW", "<u8"), ("magH", "<u8"),
    ("Dc", "<i4"), ("Hc", "<i4"), ("Wc", "<i4"), ("tile_planes", "<i4"),
    ("magWc", "<u8"), ("magHc", "<u8"),
    ("c", "<u8", (6,)), ("xin", "<u8"), ("xout", "<u8"), ("b", "<u8"), ("ec", "<u8"),
    ("x64", "<u8"), ("r64", "<u8"), ("rlow", "<u8"), ("b32", "<u8"),
    ("o", "<u8", (6,)),
])

SCALARS_DTYPE = np.dtype([
    ("bnorm2", "<f8"), ("rz", "<f8"), ("alpha", "<f8"), ("beta", "<f8"),
    ("rr", "<f8"), ("rel", "<f8"), ("tol", "<f8"), ("pap", "<f8"),
    ("iter", "<i4"), ("maxit", "<i4"), ("active", "<i4"), ("status", "<i4"),
    ("slot0", "<i4"), ("nslot", "<i4"), ("xlag", "<i4"), ("method", "<i4"),
    ("alpha_prev", "<f8"),
    ("sexp", "<i4"), ("zexp", "<i4"), ("sbase", "<i4"), ("scaled", "<i4"),
])

_vp = C.c_void_p
_i64 = C.c_int64
_int = C.c_int
_f32 = C.c_float

# name -> (restype, argtypes)
SIGNATURES = {
    "hfm_last_error": (C.c_char_p, []),
    "hfm_version": (_int, []),
    "hfm_struct_sizes": (_int, [C.POINTER(_i64), C.POINTER(_i64)]),
    "hfm_conv27_f32": (_int, [_vp, _vp, _vp, _vp, _f32, _i64, _i64, _i64, _int, _vp]),
    "hfm_var6_f32": (_int, [_vp, _vp, _vp, _vp, _vp, _i64, _i64, _i64, _int, _vp]),
    "hfm_restrict_mean_f32": (_int, [_vp, _vp, _i64, _i64, _i64, _int, _int, _int, _vp]),
    "hfm_resid_restrict_f32": (_int, [_vp, _vp, _vp, _vp, _i64, _i64, _i64, _int, _vp]),
    "hfm_prolong_nearest_f32": (_int, [_vp, _vp, _i64, _i64, _i64, _int, _int, _int, _vp]),
    "hfm_prolong_upsum_f32": (_int, [_vp, _vp, _i64, _i64, _i64, _int, _int, _int, _vp]),
    "hfm_sum_f64": (_int, [_vp, _i64, _vp, _vp]),
    "hfm_prolong_jacobi_f32": (_int, [_vp, _vp, _vp, _vp, _f32, _vp, _vp, _i64, _i64, _i64,
                                      _int, _int, _int, _vp]),
    "hfm_correct_jacobi_f32": (_int, [_vp, _vp, _vp, _vp, _vp, _f32, _vp, _vp, _i64, _i64, _i64,
                                      _int, _vp]),
    "hfm_prolong_correct_jacobi_f32": (_int, [_vp, _vp, _vp, _vp, _vp, _f32, _vp, _vp, _i64, _i64,
                                              _i64, _int, _int, _int, _int, _vp]),
    "hfm_advect27_f32": (_int, [_vp, _vp, _vp, _vp, _int, _vp, _f32, _i64, _i64, _i64, _vp]),
    "hfm_advect27_post_f32": (_int, [_vp, _vp, _vp, _vp, _int, _vp, _f32, _vp, _f32, _int,
                                     _i64, _i64, _i64, _vp]),
    "hfm_grad3_f32": (_int, [_vp, _vp, _vp, _vp, _vp, _vp, _f32, _int, _i64, _i64, _i64, _vp]),
    "hfm_bc_velocity_bluff_inplace_f32": (_int, [_vp, _vp, _vp, _vp, _i64, _i64, _i64, _f32, _vp,
                                                 _vp, _f32, _vp]),
    "hfm_advect27_dis_f32": (_int, [_vp, _vp, _vp, _vp, _vp, _int, _vp, _f32, _i64, _i64, _i64,
                                    _vp]),
    "hfm_pg_coef_f32": (_int, [_vp, _vp, _vp, _vp, _f32, _f32, _i64, _i64, _i64, _vp]),
    "hfm_pg_flux_f32": (_int, [_vp, _vp, _vp, _vp, _i64, _i64, _i64, _vp]),
    "hfm_conv27_sum_f32": (_int, [_vp, _vp, _int, _vp, _f32, _i64, _i64, _i64, _int, _vp]),
    "hfm_bc_f32": (_int, [_vp, _vp, _i64, _i64, _i64, _int, _vp]),
    "hfm_bc_velocity_f32": (_int, [_vp, _vp, _vp, _vp, _vp, _vp, _i64, _i64, _i64, _f32, _vp,
                                   _vp]),
    "hfm_bc_velocity_inplace_f32": (_int, [_vp, _vp, _vp, _vp, _i64, _i64, _i64, _f32, _vp, _vp]),
    "hfm_bc_velocity_inplace_noalloc_f32": (_int, [_vp, _vp, _vp, _i64, _i64, _i64, _f32, _vp,
                                                   _vp]),
    "hfm_bluff_body_f32": (_int, [_vp, _vp, _int, _vp, _f32, _i64, _vp]),
    "hfm_mg_sweep": (_int, [_vp, _vp, _i64, _vp, _int, _int, _int, _f32, _f32, _int, _vp, _vp]),
    "hfm_mg_resid_restrict": (_int, [_vp, _vp, _i64, _vp, _int, _int, _vp]),
    "hfm_mg_smooth_restrict": (_int, [_vp, _vp, _i64, _vp, _int, _f32, _vp]),
    "hfm_mg_smooth2": (_int, [_vp, _vp, _i64, _vp, _int, _int, _f32, _f32, _vp, _vp]),
    "hfm_mg_tile": (_int, [_i64, _i64, _i64, C.c_int32, C.POINTER(C.c_int32),
                           C.POINTER(C.c_int32), C.POINTER(_i64)]),
    "hfm_mg_coarse_solve": (_int, [_vp, _i64, _vp, _int, _vp]),
    "hfm_mg_coarse_cycle": (_int, [_vp, _vp, _i64, _vp, _int, _int, _int, _f32, _f32, _vp]),
    "hfm_fcg_init": (_int, [_vp, _vp, _i64, _vp, _int, _vp, _vp]),
    "hfm_fcg_dir": (_int, [_vp, _vp, _i64, _vp, _int, _vp, _vp]),
    "hfm_fcg_update": (_int, [_vp, _vp, _i64, _vp, _int, _int, _vp, _vp]),
    "hfm_fcg_init_sweep": (_int, [_vp, _vp, _i64, _vp, _int, _f32, _vp, _vp]),
    "hfm_fcg_update_sweep": (_int, [_vp, _vp, _i64, _vp, _int, _int, _f32, _vp, _vp]),
    "hfm_fcg_flush": (_int, [_vp, _vp, _i64, _vp, _int, _vp]),
    "hfm_fcg_finalize": (_int, [_vp, _i64, _vp, _int, _vp]),
    "hfm_mg_faces": (_int, [_vp, _vp, _vp, _vp, _i64, _i64, _i64, _f32, _vp]),
    "hfm_mg_galerkin": (_int, [_vp, _vp, _i64, _i64, _i64, _vp]),
    "hfm_mg_convert": (_int, [_vp, _vp, _i64, _int, _vp]),
    "hfm_mg_faces_batched": (_int, [_vp, _vp, _i64, _f32, _vp]),
    "hfm_mg_galerkin_batched": (_int, [_vp, _vp, _i64, _vp]),
    "hfm_mg_convert_batched": (_int, [_vp, _vp, _i64, _int, _vp]),
    "hfm_mg_coarse_inverse": (_int, [_vp, _i64, _vp]),
    "hfm_mg_persistent": (_int, [_vp, C.c_int32, _vp, _vp, _int, _int, C.c_int32, C.c_double,
                                 _vp]),
    "hfm_mg_persistent_workgroups": (_int, [_int, _int, C.POINTER(C.c_int32)]),
    "hfm_mg_persist_struct_sizes": (_int, [C.POINTER(_i64), C.POINTER(_i64)]),
    "hfm_flow2d_steps": (_int, [_vp, _i64, _vp, _vp, _int, _int, _vp]),
    "hfm_flow2d_struct_sizes": (_int, [C.POINTER(_i64), C.POINTER(_i64)]),
}

PHASE_DTYPE = np.dtype([
    ("kind", "<i4"), ("mode", "<i4"), ("dot", "<i4"), ("nitems", "<i4"),
    ("nu1", "<i4"), ("nu2", "<i4"), ("slot0", "<i4"), ("pad", "<i4"),
    ("omega", "<f4"), ("alpha", "<f4"),
    ("tasks", "<u8"), ("blk_task", "<u8"), ("chain", "<u8"), ("dots", "<u8"),
])
PH = {"sweep": 0, "restrict": 1, "smooth_restrict": 2, "smooth2": 3, "coarse_cycle": 4,
      "finalize": 5, "dir": 6, "update": 7, "flush": 8}


class PersistCtl(C.Structure):
    _fields_ = [("cnt", C.c_void_p), ("need", C.c_void_p), ("prev", C.c_void_p),
                ("fail", C.c_void_p), ("nslot", C.c_int32), ("pad", C.c_int32)]


FLOW_TASK_DTYPE = np.dtype([
    ("u", "<u8"), ("v", "<u8"), ("uh", "<u8"), ("vh", "<u8"), ("z", "<u8"), ("p", "<u8"),
    ("phi", "<u8"), ("r", "<u8"), ("mgx", "<u8"), ("mgy", "<u8"), ("mgb", "<u8"),
    ("fu", "<u8"), ("stats", "<u8"), ("team", "<u8"),
])

FLOW_PARAMS_DTYPE = np.dtype([
    ("nu_h2", "<f4"), ("inv_h", "<f4"), ("h", "<f4"), ("dt", "<f4"), ("cfl_coef", "<f4"),
    ("omega", "<f4"), ("alpha", "<f4"),
    ("N", "<i4"), ("nu1", "<i4"), ("nu2", "<i4"), ("nlev", "<i4"), ("maxit", "<i4"),
    ("max_sub", "<i4"), ("team", "<i4"), ("tol", "<f8"),
])

FLOW_STATS = ("iters", "rel", "iters_sum", "status", "steps", "substeps", "rel_max", "nsub")
FLOW_STATUS = {**STATUS, 4: "cfl", 5: "team_timeout"}
FLOW_TEAM_BYTES = 4096      # team scratch per flow (zeroed before each launch)
FLOW_TEAM_MAX = 64
FLOW_PROJECT_ONLY = 1
BC_NOALLOC_SCRATCH_FLOATS = 1 << 20   # scratch of hfm_bc_velocity_inplace_noalloc_f32
ADV_INDICATOR, ADV_BLUFF_SHIFT = 1, 4          # hfm_advect27_post_f32
GRAD_P_BC, GRAD_PRE_BLUFF, GRAD_POST_BLUFF = 1, 2, 4   # hfm_grad3_f32

_lock = threading.Lock()
_lib = None


class HfmError(RuntimeError):
    """An entry point of libhfm.so returned an error code."""


def lib_path() -> str:
    return os.environ.get("HFM_LIB", LIB_PATH)


def load() -> C.CDLL:
    """Load libhfm.so once (raises if it has not been built)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        path = lib_path()
        if not os.path.exists(path):
            raise HfmError(f"libhfm.so not found at {path}: run __graft_entry__.build() "
                           "(the HIP extension is required; there is no CPU fallback)")
        lib = C.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        ts, ss = _i64(), _i64()
        if lib.hfm_struct_sizes(C.byref(ts), C.byref(ss)) != HFM_OK:
            raise HfmError("hfm_struct_sizes failed")
        if ts.value != TASK_DTYPE.itemsize or ss.value != SCALARS_DTYPE.itemsize:
            raise HfmError(f"ABI mismatch: task {ts.value} vs {TASK_DTYPE.itemsize}, "
                           f"scalars {ss.value} vs {SCALARS_DTYPE.itemsize}")
        if lib.hfm_flow2d_struct_sizes(C.byref(ts), C.byref(ss)) != HFM_OK:
            raise HfmError("hfm_flow2d_struct_sizes failed")
        if ts.value != FLOW_TASK_DTYPE.itemsize or ss.value != FLOW_PARAMS_DTYPE.itemsize:
            raise HfmError(f"ABI mismatch: flow task {ts.value} vs {FLOW_TASK_DTYPE.itemsize}, "
                           f"params {ss.value} vs {FLOW_PARAMS_DTYPE.itemsize}")
        if lib.hfm_mg_persist_struct_sizes(C.byref(ts), C.byref(ss)) != HFM_OK:
            raise HfmError("hfm_mg_persist_struct_sizes failed")
        if ts.value != PHASE_DTYPE.itemsize or ss.value != C.sizeof(PersistCtl):
            raise HfmError(f"ABI mismatch: phase {ts.value} vs {PHASE_DTYPE.itemsize}, "
                           f"persist ctl {ss.value} vs {C.sizeof(PersistCtl)}")
        _lib = lib
        return lib


def call(name: str, *args) -> None:
    """Call entry point ``name``; raise HfmError with hfm_last_error() on failure."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != HFM_OK:
        msg = lib.hfm_last_error().decode(errors="replace")
        raise HfmError(f"{name} returned {rc}: {msg}")


def stream_handle(device=None) -> int:
    """hipStream_t of torch's current stream (as an int for ctypes)."""
    return torch.cuda.current_stream(device).cuda_stream


def ptr(t) -> int:
    """Device address of a tensor (None -> NULL)."""
    return 0 if t is None else t.data_ptr()


def ptr_array(ptrs):
    """A ctypes array of void* (host memory) for the pointer-array arguments."""
    arr = (C.c_void_p * len(ptrs))()
    for q, p in enumerate(ptrs):
        arr[q] = p
    return arr


def tile(shape, max_planes: int = 0):
    """(tile_rows, tile_planes, tiles) of the fused pass pairs for a level of
    extents (D, H, W) with at most ``max_planes`` planes per tile (0: the
    library's 32), or None when the level cannot be tiled (hfm_mg_tile)."""
    lib = load()
    R, Z, n = C.c_int32(), C.c_int32(), _i64()
    D, H, W = (int(v) for v in shape)
    if lib.hfm_mg_tile(D, H, W, int(max_planes), C.byref(R), C.byref(Z), C.byref(n)) != HFM_OK:
        return None
    return R.value, Z.value, n.value


def magic(d: int) -> int:
    """floor(2^64 / d) + 1 (0 for d == 1): exact division of n < 2^32 by d."""
    return 0 if d == 1 else ((1 << 64) - 1) // d + 1
