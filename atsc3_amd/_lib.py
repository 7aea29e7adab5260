"""ctypes binding of libatsc3fec.so (declarations in include/atsc3_fec.h).

The library is required: there is no CPU fallback anywhere in the package.
If the .so is missing or fails to load, ``lib()`` raises.
"""
from __future__ import annotations

import ctypes
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libatsc3fec.so")

_lock = threading.Lock()
_lib = None

c_i32, c_f32, c_f64, c_vp = ctypes.c_int32, ctypes.c_float, ctypes.c_double, ctypes.c_void_p

# name -> argtypes (all pointers are passed as integers / c_void_p)
_SIGS = {
    "ldpc_kernel_abi": [],
    "ldpc_minsum_f32": [c_vp, c_i32, c_i32, c_i32, c_i32, c_vp, c_vp, c_vp, c_vp,
                        c_i32, c_f32, c_vp, c_vp, c_vp, c_vp, c_vp],
    "atsc3_dropin_engine": [c_i32, c_i32, c_i32, c_vp, c_vp, c_vp, c_vp],
    "atsc3_minsum_csr_f32_dev": [c_vp, c_i32, c_i32, c_i32, c_i32, c_vp, c_vp, c_vp,
                                 c_vp, c_i32, c_f32, c_vp, c_vp, c_vp, c_vp, c_vp,
                                 c_vp, c_vp],
    "atsc3_minsum_qc_dev": [c_vp, c_i32, c_i32, c_i32, c_i32, c_i32, c_i32, c_vp, c_vp,
                            c_vp, c_vp, c_vp, c_vp, c_vp, c_i32, c_vp, c_i32, c_i32,
                            c_i32, c_i32, c_f32, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp,
                            c_vp, c_vp, c_vp],
    "atsc3_minsum_qc_ctl_words": [c_i32],
    "atsc3_minsum_qc_phased_dev": [c_vp, c_i32, c_i32, c_i32, c_i32, c_i32, c_i32, c_vp,
                                   c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_i32, c_vp, c_i32,
                                   c_i32, c_i32, c_i32, c_f32, c_i32, c_vp, c_vp, c_vp,
                                   c_vp, c_vp, c_vp, c_vp, c_vp, c_vp],
    "atsc3_minsum_batch_ctl_words": [c_i32, c_i32],
    "atsc3_minsum_batch_max_grid": [c_i32],
    "atsc3_minsum_batch_status_word": [],
    "atsc3_minsum_batch_dev": [c_vp, c_i32, c_vp, c_vp, c_vp, c_i32, c_i32, c_i32, c_i32,
                               c_i32, c_f32, c_vp, c_i32, c_i32, c_vp, c_i32, c_i32, c_vp,
                               c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp],
    "atsc3_minsum_batch_xcd_ctl_words": [c_i32, c_i32],
    "atsc3_minsum_batch_xcd_max_grid": [c_i32],
    "atsc3_minsum_batch_xcd_dev": [c_vp, c_i32, c_vp, c_vp, c_vp, c_i32, c_i32, c_i32, c_i32,
                                   c_i32, c_f32, c_vp, c_i32, c_i32, c_vp, c_vp, c_vp, c_i32,
                                   c_i32, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                                   c_vp],
    "atsc3_unrolled_check_fwd_f32": [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_i32, c_i32,
                                     c_i32, c_i32, c_vp, c_vp, c_vp, c_vp],
    "atsc3_unrolled_check_fwd_f64": [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_i32, c_i32,
                                     c_i32, c_i32, c_vp, c_vp, c_vp, c_vp],
    "atsc3_unrolled_var_fwd_f32": [c_vp, c_vp, c_vp, c_vp, c_vp, c_i32, c_i32, c_i32, c_vp],
    "atsc3_unrolled_var_fwd_f64": [c_vp, c_vp, c_vp, c_vp, c_vp, c_i32, c_i32, c_i32, c_vp],
    "atsc3_unrolled_check_bwd_f32": [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                                     c_i32, c_i32, c_i32, c_i32, c_vp, c_vp, c_vp],
    "atsc3_unrolled_check_bwd_f64": [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                                     c_i32, c_i32, c_i32, c_i32, c_vp, c_vp, c_vp],
    "atsc3_unrolled_var_bwd_f32": [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_i32, c_i32, c_i32,
                                   c_vp],
    "atsc3_unrolled_var_bwd_f64": [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_i32, c_i32, c_i32,
                                   c_vp],
    "atsc3_soft_demap_fwd_f32": [c_vp, c_vp, c_vp, c_vp, c_i32, c_i32, c_vp, c_i32, c_i32,
                                 c_vp, c_vp],
    "atsc3_soft_demap_fwd_f64": [c_vp, c_vp, c_vp, c_vp, c_i32, c_i32, c_vp, c_i32, c_i32,
                                 c_vp, c_vp],
    "atsc3_soft_demap_bwd_grid": [c_i32, c_i32],
    "atsc3_soft_demap_bwd_f32": [c_vp, c_vp, c_vp, c_vp, c_i32, c_i32, c_vp, c_i32, c_i32,
                                 c_vp, c_vp, c_vp, c_vp, c_vp],
    "atsc3_soft_demap_bwd_f64": [c_vp, c_vp, c_vp, c_vp, c_i32, c_i32, c_vp, c_i32, c_i32,
                                 c_vp, c_vp, c_vp, c_vp, c_vp],
    "atsc3_baseband_dev": [c_vp, c_i32, c_i32, c_i32, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp,
                           c_vp],
    "atsc3_gather_cells_dev": [c_vp, c_vp, c_i32, c_i32, c_vp, c_vp],
    "atsc3_demap_gather_dev": [c_vp, c_vp, c_vp, c_vp, c_i32, c_i32, c_vp, c_i32, c_vp, c_vp,
                               c_vp, c_vp, c_vp],
    "atsc3_demap_weighted_dev": [c_vp, c_vp, c_vp, c_i32, c_i32, c_vp, c_i32, c_vp, c_vp, c_vp,
                                 c_vp, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp],
    "atsc3_cpe_ws_doubles": [c_i32],
    "atsc3_cpe_dev": [c_vp, c_i32, c_vp, c_vp, c_vp, c_vp, c_i32, c_vp, c_vp, c_i32, c_vp, c_i32,
                      c_vp, c_i32, c_vp, c_vp, c_vp, c_vp],
    "atsc3_dummy_snr_dev": [c_vp, c_i32, c_vp, c_vp, c_vp],
    "atsc3_sp_ctl_words": [c_i32],
    "atsc3_sum_product_dev": [c_vp, c_vp, c_vp, c_vp, c_i32, c_i32, c_i32, c_i32, c_vp, c_vp,
                              c_vp, c_i32, c_f64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                              c_vp, c_vp],
    "atsc3_sp_select_dev": [c_vp, c_vp, c_vp, c_i32, c_i32, c_i32, c_i32, c_vp, c_vp, c_vp],
    "atsc3_sp_merge_dev": [c_vp, c_vp, c_i32, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                           c_vp, c_vp],
    "atsc3_ldm_ws_doubles": [c_i32, c_i32],
    "atsc3_ldm_cancel_dev": [c_vp, c_vp, c_vp, c_i32, c_i32, c_i32, c_vp, c_vp, c_i32,
                             c_f32, c_vp, c_vp, c_vp, c_vp],
    "atsc3_minsum_resident_ws_bytes": [c_i32, c_i32, c_i32, c_i32],
    "atsc3_minsum_resident_lds_rows": [c_i32],
    "atsc3_minsum_resident_grid": [],
    "atsc3_minsum_resident_dev": [c_vp, c_i32, c_vp, c_vp, c_vp, c_i32, c_i32, c_i32, c_i32,
                                  c_i32, c_i32, c_i32, c_f32, c_i32, c_vp, c_i32, c_vp,
                                  ctypes.c_int64, c_vp, c_vp, c_vp, c_vp, c_vp],
    "atsc3_minsum_resident_llr_dev": [c_vp, c_i32, c_vp, c_vp, c_vp, c_i32, c_i32, c_i32,
                                      c_i32, c_i32, c_i32, c_i32, c_f32, c_i32, c_vp, c_i32,
                                      c_vp, ctypes.c_int64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp],
    "atsc3_minsum_resident_team_ws_bytes": [c_i32, c_i32, c_i32, c_i32, c_i32],
    "atsc3_minsum_resident_team_status_word": [],
    "atsc3_minsum_resident_team_order_dev": [c_vp, c_i32, c_vp, c_vp, c_vp, c_i32, c_vp, c_vp,
                                             c_vp],
    "atsc3_minsum_resident_team_dev": [c_vp, c_i32, c_vp, c_vp, c_vp, c_i32, c_i32, c_i32,
                                       c_i32, c_i32, c_i32, c_i32, c_f32, c_i32, c_vp, c_vp,
                                       c_vp, c_i32, c_i32, c_i32, c_vp, ctypes.c_int64, c_vp,
                                       c_vp, c_vp, c_vp, c_vp],
    "atsc3_minsum_resident_noalloc_max_bytes": [],
    "atsc3_minsum_resident_noalloc_dev": [c_vp, c_i32, c_vp, c_vp, c_vp, c_i32, c_i32, c_i32,
                                          c_i32, c_i32, c_i32, c_i32, c_f32, c_i32, c_vp,
                                          c_i32, c_vp, c_vp, c_vp, c_vp, c_vp],
    "atsc3_demap_ws_doubles": [c_i32, c_i32],
    "atsc3_demap_groups_dev": [c_vp, c_i32, c_i32, c_i32, c_vp],
    "atsc3_demap_raw_dev": [c_vp, c_vp, c_i32, c_i32, c_vp, c_i32, c_vp, c_vp, c_vp, c_vp,
                            c_vp],
    "atsc3_minsum_utc_ws_bytes": [c_i32, c_i32, c_i32, c_i32],
    "atsc3_minsum_utc_dev": [c_vp, c_i32, c_i32, c_i32, c_i32, c_i32, c_i32, c_vp, c_vp, c_vp,
                             c_vp, c_i32, c_i32, c_f64, c_vp, ctypes.c_int64, c_vp, c_vp, c_vp,
                             c_vp, c_vp],
    "atsc3_demap_dev": [c_vp, c_vp, c_vp, c_i32, c_i32, c_vp, c_i32, c_vp, c_vp, c_vp,
                        c_vp, c_vp],
}

# functions whose C return type is not int32_t
_RESTYPES = {"atsc3_minsum_resident_ws_bytes": ctypes.c_int64,
             "atsc3_minsum_resident_team_ws_bytes": ctypes.c_int64,
             "atsc3_minsum_utc_ws_bytes": ctypes.c_int64,
             "atsc3_minsum_resident_noalloc_max_bytes": ctypes.c_int64}


def restype_of(name):
    return _RESTYPES.get(name, c_i32)


ERRORS = {-1: "check degree exceeds the supported maximum",
          -2: "inconsistent sizes or NULL pointers",
          -3: "HIP runtime error or kernel launch failure",
          -4: "size exceeds a fixed capacity"}


class FecError(RuntimeError):
    def __init__(self, fn, rc):
        super().__init__(f"{fn} returned {rc}: {ERRORS.get(rc, 'unknown error')}")
        self.rc = rc


def lib():
    """Load the library (once).  Raises if it is missing."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    f"{LIB_PATH} is not built: run python -m atsc3_amd.build "
                    "(hipcc, gfx950)")
            L = ctypes.CDLL(LIB_PATH)
            for name, args in _SIGS.items():
                f = getattr(L, name)
                f.restype = restype_of(name)
                f.argtypes = args
            if L.ldpc_kernel_abi() != 1:
                raise RuntimeError("libatsc3fec ABI mismatch")
            _lib = L
        return _lib


SIZE_QUERIES = ("atsc3_demap_ws_doubles", "atsc3_minsum_qc_ctl_words", "atsc3_ldm_ws_doubles",
                "atsc3_minsum_batch_ctl_words", "atsc3_minsum_batch_max_grid",
                "atsc3_minsum_batch_xcd_ctl_words", "atsc3_minsum_batch_xcd_max_grid",
                "atsc3_minsum_batch_status_word", "atsc3_soft_demap_bwd_grid",
                "atsc3_sp_ctl_words", "atsc3_cpe_ws_doubles", "atsc3_minsum_resident_ws_bytes",
                "atsc3_minsum_resident_grid", "atsc3_minsum_resident_lds_rows",
                "atsc3_minsum_utc_ws_bytes", "atsc3_minsum_resident_noalloc_max_bytes",
                "atsc3_minsum_resident_team_ws_bytes", "atsc3_minsum_resident_team_status_word",
                "atsc3_dropin_engine")


def call(name, *args):
    rc = getattr(lib(), name)(*args)
    if rc != 0 and name not in SIZE_QUERIES:
        raise FecError(name, rc)
    return rc


def exported_symbols():
    return list(_SIGS)
