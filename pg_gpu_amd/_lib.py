"""ctypes binding of libpgld.so and the device plumbing around it.

The library is the only compute path: if libpgld.so is missing or no GPU is
visible, every call raises.  Torch is used for device memory and the current
HIP stream only.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpgld.so")

FAMILY_HAP, FAMILY_GENO = 0, 1
EINVAL, EDOMAIN, ECAPACITY, EUNSUPPORTED = -1, -2, -3, -4
STAT_DD, STAT_DZ, STAT_PI2, STAT_PI2_SYM, STAT_SIGMA_D2 = 0, 1, 2, 3, 4
KEY_BIN_ABS, KEY_BIN, KEY_WINDOW = 0, 1, 2
EST_R2, EST_R2_NAIVE, EST_SIGMA_D2, EST_RH_R, EST_DIP_R2, EST_D_NAIVE = 0, 1, 2, 3, 4, 5
SET_MOMENTS, SET_SELF = 0, 1

TERM_DTYPE = np.dtype([("type", "<i4"), ("pops", "<i4", (4,)), ("pad", "<i4"), ("weight", "<f8")])
assert TERM_DTYPE.itemsize == 32

_vp, _i32, _i64, _f64 = C.c_void_p, C.c_int32, C.c_int64, C.c_double
_PP = C.POINTER(C.c_void_p)

_SIGNATURES = {
    "pgld_version": ([], C.c_int),
    "pgld_last_error": ([], C.c_char_p),
    "pgld_scan_rows": ([_vp, _i64, _i64, _vp, _i64, _i32, _vp, _vp], C.c_int),
    "pgld_allele_counts": ([_vp, _i64, _i64, _vp, _i64, _vp, _vp, _vp], C.c_int),
    "pgld_pairs_per_variant": ([_vp, _i64, _f64, _vp, _vp], C.c_int),
    "pgld_pairs_fill": ([_vp, _i64, _i64, _i64, _vp, _vp, _vp], C.c_int),
    "pgld_counts_pairs_hap": ([_vp, _i64, _i64, _vp, _i64, _vp, _vp, _i64, _vp, _vp, _vp], C.c_int),
    "pgld_counts_pairs_geno": ([_vp, _i64, _i64, _vp, _i64, _vp, _vp, _i64, _vp, _vp, _vp], C.c_int),
    "pgld_stats_pairs": ([_i32, _i32, _PP, _PP, _i64, _vp, _vp, _i32, _vp, _vp], C.c_int),
    "pgld_r_family": ([_vp, _vp, _i64, _i32, _vp, _vp], C.c_int),
    "pgld_pair_keys": ([_vp, _vp, _vp, _i64, _i32, _vp, _i32, _f64, _f64, _vp, _vp], C.c_int),
    "pgld_keyed_sums": ([_vp, _vp, _i64, _i32, _i32, _vp, _vp], C.c_int),
    "pgld_pack_planes": ([_i32, _vp, _i64, _i64, _vp, _i64, _i64, _vp, _vp, _vp, _vp], C.c_int),
    "pgld_pack_planes_fp4": ([_i32, _vp, _i64, _i64, _vp, _i64, _i64, _vp, _vp, _vp, _vp], C.c_int),
    "pgld_pack_batch_create": ([_i32, _i32, _i32, _PP, C.POINTER(_i64), C.POINTER(_i64), _PP, C.POINTER(_i64),
                                C.POINTER(_i64), _PP, _PP, _PP, C.POINTER(_vp)], C.c_int),
    "pgld_pack_batch_run": ([_vp, _vp], C.c_int),
    "pgld_pack_batch_destroy": ([_vp], C.c_int),
    "pgld_band_plan_create": ([_vp, _i64, _f64, _i32, _vp, _vp, _i32, _f64, _f64, C.POINTER(_vp)], C.c_int),
    "pgld_band_plan_info": ([_vp, C.POINTER(_i64), C.POINTER(_i64), C.POINTER(_i64)], C.c_int),
    "pgld_band_plan_destroy": ([_vp], C.c_int),
    "pgld_band_plan_set_timing": ([_vp, _i32], C.c_int),
    "pgld_band_ld_pairs": ([_vp, _vp, _vp, _i64, _i32, _i32, _i32, _i32, _f64, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp,
                            _vp, _vp], C.c_int),
    "pgld_band_ld_pairs_fp4": ([_vp, _vp, _vp, _i64, _i32, _i32, _i32, _i32, _f64, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp,
                            _vp, _vp], C.c_int),
    "pgld_pack_bits": ([_vp, _i64, _i64, _i64, _vp, _vp], C.c_int),
    "pgld_band_counts_bits": ([_vp, _i32, _vp, _vp, _i64, _i32, _i32, _vp, _vp, _vp], C.c_int),
    "pgld_band_r2_bits": ([_vp, _vp, _vp, _i64, _i32, _i32, _i32, _f64, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp],
                          C.c_int),
    "pgld_batch_plan_create": ([_i32, _vp, _vp, _f64, _i32, _vp, _vp, _i32, _f64, _f64, C.POINTER(_vp)], C.c_int),
    "pgld_batch_bind": ([_vp, _i32, _i32, _i32, _PP, _PP, C.POINTER(_i64), C.POINTER(_i32)], C.c_int),
    "pgld_batch_bind_fp4": ([_vp, _i32, _i32, _i32, _PP, _PP, C.POINTER(_i64), C.POINTER(_i32)], C.c_int),
    "pgld_batch_sums": ([_vp, _i32, _vp, _vp, _vp], C.c_int),
    "pgld_batch_schedule": ([_vp, _i32, _i32, _i32, _i32], C.c_int),
    "pgld_batch_autotune": ([_vp, _i32, _vp, _vp, _vp, C.POINTER(_i32), C.POINTER(_i32), C.POINTER(_i32),
                             C.POINTER(_i32), C.POINTER(C.c_float)], C.c_int),
    "pgld_graph_begin": ([_vp], C.c_int),
    "pgld_graph_end": ([_vp, C.POINTER(_vp)], C.c_int),
    "pgld_graph_launch": ([_vp, _vp], C.c_int),
    "pgld_graph_destroy": ([_vp], C.c_int),
    "pgld_band_r2": ([_vp, _vp, _vp, _i64, _i32, _i32, _i32, _f64, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp],
                     C.c_int),
    "pgld_band_plan_kernel_ms": ([_vp, C.POINTER(C.c_float)], C.c_int),
    "pgld_band_sums": ([_vp, _i32, _i32, _PP, _PP, C.POINTER(_i64), C.POINTER(_i32), _i32, _i32, _vp, _vp, _vp],
                       C.c_int),
    "pgld_band_sums_fp4": ([_vp, _i32, _i32, _PP, _PP, C.POINTER(_i64), C.POINTER(_i32), _i32, _i32, _vp, _vp,
                            _vp], C.c_int),
    "pgld_band_sums_fp4_noalloc": ([_vp, _i32, _i32, _PP, _PP, C.POINTER(_i64), C.POINTER(_i32), _i32, _i32, _vp,
                                    _vp, _vp], C.c_int),
    "pgld_band_counts": ([_vp, _i32, _vp, _vp, _i64, _i32, _i32, _vp, _vp, _vp], C.c_int),
}

_lib = None


class PgldError(RuntimeError):
    """A libpgld call returned a nonzero status (kept in .status)."""

    def __init__(self, msg, status=None):
        super().__init__(msg)
        self.status = status


def lib():
    """Load libpgld.so (once).  Raises if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is not built; run `python -m pg_gpu_amd.build` (hipcc, gfx950)")
        h = C.CDLL(LIB_PATH)
        for name, (args, res) in _SIGNATURES.items():
            fn = getattr(h, name)
            fn.argtypes = args
            fn.restype = res
        _lib = h
    return _lib


def exported_symbols():
    return list(_SIGNATURES)


def check(rc, name):
    if rc != 0:
        msg = lib().pgld_last_error().decode(errors="replace")
        raise PgldError(f"{name} failed with status {rc}: {msg}", rc)


def call(name, *args):
    rc = getattr(lib(), name)(*args)
    check(rc, name)


# ---------------------------------------------------------------------------
# device plumbing (torch owns memory and streams)
# ---------------------------------------------------------------------------

def torch():
    import torch as _t
    return _t


def require_gpu():
    t = torch()
    if not t.cuda.is_available():
        raise RuntimeError("pg_gpu_amd needs a ROCm GPU: torch.cuda.is_available() is False")
    lib()


def device():
    require_gpu()
    return torch().device("cuda", torch().cuda.current_device())


def stream_handle():
    return _vp(torch().cuda.current_stream().cuda_stream)


def ptr(t):
    """Raw device pointer of a torch tensor (None for None)."""
    if t is None:
        return None
    return _vp(t.data_ptr())


_NP2T = {np.int8: "int8", np.int32: "int32", np.int64: "int64", np.float64: "float64", np.float32: "float32"}


def to_device(x, dtype):
    """numpy / torch / list -> contiguous torch tensor of dtype on the GPU."""
    t = torch()
    tdt = getattr(t, _NP2T[np.dtype(dtype).type])
    if isinstance(x, t.Tensor):
        return x.to(device=device(), dtype=tdt).contiguous()
    return t.as_tensor(np.ascontiguousarray(np.asarray(x), dtype=dtype), device=device())


def empty(shape, dtype):
    t = torch()
    return t.empty(shape, dtype=getattr(t, _NP2T[np.dtype(dtype).type]), device=device())


def zeros(shape, dtype):
    t = torch()
    return t.zeros(shape, dtype=getattr(t, _NP2T[np.dtype(dtype).type]), device=device())


def is_tensor(x):
    try:
        return isinstance(x, torch().Tensor)
    except ImportError:
        return False


def to_numpy(x):
    if is_tensor(x):
        return x.detach().cpu().numpy()
    return np.asarray(x)


def ptr_array(tensors):
    arr = (C.c_void_p * len(tensors))(*[t.data_ptr() if t is not None else None for t in tensors])
    return C.cast(arr, _PP), arr
