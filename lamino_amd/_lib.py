"""ctypes binding of the C ABI in include/lamino.h (lamino_amd/liblamino_amd.so).

The library is loaded from the package directory only; a missing or stale
build raises immediately -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblamino_amd.so")

c_i64 = ctypes.c_int64
c_i64p = ctypes.POINTER(ctypes.c_int64)
c_vp = ctypes.c_void_p
c_fp = ctypes.c_void_p   # device float* (as integer address)
c_dp = ctypes.POINTER(ctypes.c_double)

OP_FN = ctypes.CFUNCTYPE(ctypes.c_int, c_vp, ctypes.POINTER(c_vp), ctypes.POINTER(c_vp))
LOSS_FN = ctypes.CFUNCTYPE(ctypes.c_int, c_vp, ctypes.POINTER(c_vp), c_dp)


class lam_dataset(ctypes.Structure):
    _fields_ = [("nproj", c_i64), ("nrows", c_i64), ("ncols", c_i64),
                ("proj", c_vp), ("alpha", c_vp), ("gamma", c_vp), ("gamma_const", ctypes.c_double)]


class lam_recon_params(ctypes.Structure):
    _fields_ = [("method", ctypes.c_int), ("max_iters", c_i64), ("inner_iters", c_i64),
                ("lambda_", ctypes.c_double), ("mu", ctypes.c_double), ("sigma", ctypes.c_double),
                ("p", ctypes.c_double), ("tol", ctypes.c_double), ("xtol", ctypes.c_double),
                ("pad_factor", ctypes.c_double), ("eps", ctypes.c_double)]


class lam_recon_stats(ctypes.Structure):
    _fields_ = [("setup_s", ctypes.c_double), ("solve_s", ctypes.c_double), ("normal_ops", c_i64),
                ("cg_iters", c_i64), ("outer_iters", c_i64), ("lipschitz", ctypes.c_double)]


class lam_job_info(ctypes.Structure):
    _fields_ = [("id", c_i64), ("status", ctypes.c_int), ("slot", ctypes.c_int), ("submit_s", ctypes.c_double),
                ("start_s", ctypes.c_double), ("end_s", ctypes.c_double), ("footprint_est", ctypes.c_size_t),
                ("footprint_peak", ctypes.c_size_t), ("stats", lam_recon_stats)]


class lam_server_stats(ctypes.Structure):
    _fields_ = [("submitted", c_i64), ("completed", c_i64), ("failed", c_i64), ("running", c_i64),
                ("queued", c_i64), ("budget", ctypes.c_size_t), ("in_use", ctypes.c_size_t),
                ("peak_in_use", ctypes.c_size_t), ("slots", ctypes.c_int), ("mode", ctypes.c_int)]


LAM_SERVE_PLAIN, LAM_SERVE_ASYNC, LAM_SERVE_GRAPH = 0, 1, 2
LAM_OUT_TIFF, LAM_OUT_VTI, LAM_OUT_TIFF_VECTORS = 1, 2, 4


class lam_output(ctypes.Structure):
    _fields_ = [("path", ctypes.c_char_p), ("formats", ctypes.c_int), ("transpose120", ctypes.c_int)]


# name -> (restype, argtypes)
_SIGS = {
    "lam_last_error": (ctypes.c_char_p, []),
    "lam_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(c_vp)]),
    "lam_ctx_destroy": (ctypes.c_int, [c_vp]),
    "lam_ctx_set_stream": (ctypes.c_int, [c_vp, c_vp]),
    "lam_ctx_stream": (c_vp, [c_vp]),
    "lam_ctx_sync": (ctypes.c_int, [c_vp]),
    "lam_malloc": (ctypes.c_int, [c_vp, ctypes.c_size_t, ctypes.POINTER(c_vp)]),
    "lam_free": (ctypes.c_int, [c_vp, c_vp]),
    "lam_upload": (ctypes.c_int, [c_vp, c_vp, c_vp, ctypes.c_size_t]),
    "lam_download": (ctypes.c_int, [c_vp, c_vp, c_vp, ctypes.c_size_t]),
    "lam_memset": (ctypes.c_int, [c_vp, c_vp, ctypes.c_int, ctypes.c_size_t]),
    "lam_copy_device": (ctypes.c_int, [c_vp, c_vp, c_vp, ctypes.c_size_t]),
    "lam_grid_polar": (ctypes.c_int, [c_vp, c_vp, c_i64, c_i64, c_i64, ctypes.c_double, ctypes.POINTER(c_vp)]),
    "lam_grid_polar_pp": (ctypes.c_int, [c_vp, c_vp, c_vp, c_i64, c_i64, c_i64, ctypes.POINTER(c_vp)]),
    "lam_grid_destroy": (ctypes.c_int, [c_vp]),
    "lam_grid_shape": (ctypes.c_int, [c_vp, c_i64p, c_i64p, c_i64p]),
    "lam_grid_coords_host": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp]),
    "lam_nufft3d1": (ctypes.c_int, [c_vp, c_vp, c_i64p, ctypes.c_double, c_vp, c_vp]),
    "lam_nufft3d2": (ctypes.c_int, [c_vp, c_vp, c_i64p, ctypes.c_double, c_vp, c_vp]),
    "lam_forward": (ctypes.c_int, [c_vp, c_vp, c_i64p, ctypes.c_double, c_fp, c_fp, c_fp, c_fp]),
    "lam_adjoint": (ctypes.c_int, [c_vp, c_vp, c_i64p, ctypes.c_double, c_fp, c_fp, c_fp, c_fp]),
    "lam_adjoint_accumulate": (ctypes.c_int, [c_vp, c_vp, c_i64p, ctypes.c_double, c_fp, c_fp, c_fp, c_fp]),
    "lam_forward_multi": (ctypes.c_int, [c_vp, c_vp, ctypes.c_int, c_i64p, ctypes.c_double, c_fp, c_fp, c_fp, c_vp]),
    "lam_adjoint_multi": (ctypes.c_int, [c_vp, c_vp, ctypes.c_int, c_i64p, ctypes.c_double, c_vp, c_fp, c_fp, c_fp,
                                         ctypes.c_int]),
    "lam_sysmat": (ctypes.c_int, [c_vp, c_vp, c_i64p, ctypes.c_double, ctypes.POINTER(c_vp), ctypes.POINTER(c_vp)]),
    "lam_toep_create": (ctypes.c_int, [c_vp, ctypes.POINTER(c_vp), ctypes.c_int, c_i64p, ctypes.c_int,
                                       ctypes.POINTER(c_vp)]),
    "lam_toep_destroy": (ctypes.c_int, [c_vp]),
    "lam_toep_apply": (ctypes.c_int, [c_vp, ctypes.POINTER(c_vp), ctypes.POINTER(c_vp), ctypes.c_float]),
    "lam_toep_quad": (ctypes.c_int, [c_vp, ctypes.POINTER(c_vp), c_dp]),
    "lam_fft2": (ctypes.c_int, [c_vp, c_i64p, ctypes.c_int, ctypes.c_int, c_vp, c_vp]),
    "lam_precond": (ctypes.c_int, [c_vp, c_i64p, c_fp, c_fp]),
    "lam_grad_u": (ctypes.c_int, [c_vp, c_i64p, c_fp, c_fp, c_fp, c_fp]),
    "lam_divergence": (ctypes.c_int, [c_vp, c_i64p, c_fp, c_fp, c_fp, c_fp]),
    "lam_laplacian": (ctypes.c_int, [c_vp, c_i64p, c_fp, c_fp]),
    "lam_qggmrf": (ctypes.c_int, [c_vp, c_i64p, c_fp, c_fp, ctypes.c_float, ctypes.c_float]),
    "lam_dot": (ctypes.c_int, [c_vp, c_i64, c_fp, c_fp, c_dp]),
    "lam_max": (ctypes.c_int, [c_vp, c_i64, c_fp, ctypes.POINTER(ctypes.c_float)]),
    "lam_dot_noalloc": (ctypes.c_int, [c_vp, c_i64, c_fp, c_fp, c_dp]),
    "lam_cgsolver": (ctypes.c_int, [c_vp, c_i64p, OP_FN, c_vp, ctypes.POINTER(c_vp), ctypes.POINTER(c_vp), c_i64,
                                    ctypes.c_double, ctypes.POINTER(c_vp), c_i64p]),
    "lam_split_bregman": (ctypes.c_int, [c_vp, c_i64p, OP_FN, c_vp, ctypes.POINTER(c_vp), ctypes.POINTER(c_vp),
                                         ctypes.c_double, ctypes.c_double, c_i64, c_i64, ctypes.c_double,
                                         ctypes.c_double, ctypes.POINTER(c_vp)]),
    "lam_nagopt": (ctypes.c_int, [c_vp, c_i64p, OP_FN, LOSS_FN, c_vp, ctypes.POINTER(c_vp), c_i64, ctypes.c_double,
                                  ctypes.c_double, ctypes.c_double, c_i64]),
    "lam_lipschitz": (ctypes.c_int, [c_vp, c_i64p, OP_FN, c_vp, c_i64, ctypes.c_double, ctypes.c_uint64, c_dp]),
    "lam_toep_op": (ctypes.c_int, [c_vp, ctypes.POINTER(c_vp), ctypes.POINTER(c_vp)]),
    "lam_autotune_bins": (ctypes.c_int, [c_vp, c_vp, c_i64p, ctypes.c_double, ctypes.POINTER(ctypes.c_int), c_dp]),
    "lam_timing_enable": (ctypes.c_int, [c_vp, ctypes.c_int]),
    "lam_timing_only": (ctypes.c_int, [c_vp, ctypes.c_char_p]),
    "lam_timing_query": (ctypes.c_int, [c_vp, ctypes.c_char_p, c_dp, c_i64p]),
    "lam_mbir": (ctypes.c_int, [c_vp, ctypes.POINTER(lam_dataset), ctypes.c_int, c_i64p,
                                ctypes.POINTER(lam_recon_params), c_vp, ctypes.POINTER(lam_recon_stats)]),
    "lam_mbir_footprint": (ctypes.c_int, [ctypes.POINTER(lam_dataset), ctypes.c_int, c_i64p,
                                          ctypes.POINTER(lam_recon_params), ctypes.POINTER(ctypes.c_size_t)]),
    "lam_server_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_size_t, ctypes.c_int, ctypes.c_int,
                                         ctypes.POINTER(c_vp)]),
    "lam_server_destroy": (ctypes.c_int, [c_vp]),
    "lam_server_submit": (ctypes.c_int, [c_vp, ctypes.POINTER(lam_dataset), ctypes.c_int, c_i64p,
                                         ctypes.POINTER(lam_recon_params), c_vp, c_i64p]),
    "lam_server_submit_output": (ctypes.c_int, [c_vp, ctypes.POINTER(lam_dataset), ctypes.c_int, c_i64p,
                                                ctypes.POINTER(lam_recon_params), ctypes.POINTER(lam_output), c_vp,
                                                c_i64p]),
    "lam_server_wait": (ctypes.c_int, [c_vp, c_i64, ctypes.POINTER(lam_job_info)]),
    "lam_server_query": (ctypes.c_int, [c_vp, ctypes.POINTER(lam_server_stats)]),
    "lam_tiff_info": (ctypes.c_int, [ctypes.c_char_p, c_i64p]),
    "lam_tiff_read": (ctypes.c_int, [ctypes.c_char_p, c_vp, c_i64p]),
    "lam_tiff_write": (ctypes.c_int, [ctypes.c_char_p, c_vp, c_i64p]),
    "lam_tiff_write_vectors": (ctypes.c_int, [ctypes.c_char_p, c_vp, c_vp, c_vp, c_i64p]),
    "lam_tiff_write3": (ctypes.c_int, [ctypes.c_char_p, c_vp, c_vp, c_vp, c_i64p]),
    "lam_vti_write_vectors": (ctypes.c_int, [ctypes.c_char_p, c_vp, c_vp, c_vp, c_i64p, ctypes.c_char_p]),
    "lam_read_angles": (ctypes.c_int, [ctypes.c_char_p, c_dp, c_i64, c_i64p]),
    "lam_mask_infs_nans": (ctypes.c_int, [c_vp, c_i64, c_fp]),
    "lam_forward_simulate": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, c_i64p, c_dp, c_i64, ctypes.c_double, c_i64,
                                            c_i64, ctypes.c_double, c_vp]),
    "lam_demag": (ctypes.c_int, [c_vp, c_i64p, c_fp, c_fp, c_fp, c_fp, c_fp, c_fp]),
    "lam_transpose120": (ctypes.c_int, [c_vp, c_i64p, c_fp, c_fp]),
}

_lib = None


class LaminoError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[lamino status {code}] {msg}")
        self.code = code


def load():
    """Load liblamino_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build() (hipcc, gfx950) first")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols():
    return sorted(_SIGS)


def check(rc: int) -> None:
    if rc != 0:
        msg = load().lam_last_error()
        raise LaminoError(rc, msg.decode() if msg else "")
