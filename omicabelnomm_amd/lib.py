"""ctypes binding of liboagpu.so (include/oagpu.h).

The library is the product: there is no CPU fallback.  Loading fails loudly
when liboagpu.so has not been built.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboagpu.so")

OK = 0
ERR_ARG, ERR_HIP, ERR_STATE, ERR_HIT_OVERFLOW = -1, -2, -3, -4
ERR_NUMERIC, ERR_UNSUPPORTED, ERR_COVARIATES, ERR_NO_DEVICE = -5, -6, -7, -8

FP64, FP32, BF16X3 = 0, 1, 2
F64, F32, F16 = 0, 1, 2
DOSAGE_NONE, DOSAGE_SUM, DOSAGE_SCALE = 0, 1, 2
HOST, DEVICE = 0, 1

PRECISIONS = {"fp64": FP64, "fp32": FP32, "bf16x3": BF16X3}


class Params(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int), ("l", ctypes.c_int), ("r", ctypes.c_int),
        ("real_n", ctypes.c_int), ("precision", ctypes.c_int),
        ("disp_cov", ctypes.c_int), ("store_all", ctypes.c_int),
        ("min_r2_store", ctypes.c_float), ("min_p_store", ctypes.c_float),
        ("min_p_disp", ctypes.c_float), ("max_traits", ctypes.c_int),
        ("max_snps", ctypes.c_int), ("max_hits", ctypes.c_longlong),
        ("device", ctypes.c_int),
    ]


class Hit(ctypes.Structure):
    _fields_ = [
        ("trait", ctypes.c_int32), ("snp", ctypes.c_int32), ("h", ctypes.c_int32),
        ("n_used", ctypes.c_int32), ("b", ctypes.c_double), ("se", ctypes.c_double),
        ("t", ctypes.c_double), ("p", ctypes.c_double), ("r2", ctypes.c_double),
    ]


HIT_DTYPE = np.dtype([("trait", "<i4"), ("snp", "<i4"), ("h", "<i4"), ("n_used", "<i4"),
                      ("b", "<f8"), ("se", "<f8"), ("t", "<f8"), ("p", "<f8"), ("r2", "<f8")])
assert HIT_DTYPE.itemsize == ctypes.sizeof(Hit)


class Dense(ctypes.Structure):
    _fields_ = [
        ("beta", ctypes.c_void_p), ("se", ctypes.c_void_p), ("tstat", ctypes.c_void_p),
        ("pval", ctypes.c_void_p), ("kept", ctypes.c_void_p), ("r2", ctypes.c_void_p),
        ("rss", ctypes.c_void_p), ("emitted", ctypes.c_void_p), ("n_used", ctypes.c_void_p),
        ("var_x", ctypes.c_void_p),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("n_hits", ctypes.c_longlong), ("n_emitted", ctypes.c_longlong),
        ("n_sig", ctypes.c_longlong), ("n_fallback", ctypes.c_longlong),
        ("n_patterns", ctypes.c_int), ("n_chunks", ctypes.c_int), ("t_star", ctypes.c_double),
    ]


class ServerParams(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("hbm_budget", ctypes.c_longlong),
                ("max_active", ctypes.c_int), ("chunk_snps", ctypes.c_int), ("graphs", ctypes.c_int)]


class Request(ctypes.Structure):
    _fields_ = [
        ("t", ctypes.c_int), ("l", ctypes.c_int),
        ("Y", ctypes.c_void_p), ("ldy", ctypes.c_longlong),
        ("XL", ctypes.c_void_p), ("ldxl", ctypes.c_longlong),
        ("keep", ctypes.c_void_p),
        ("min_p_store", ctypes.c_float), ("min_p_disp", ctypes.c_float), ("min_r2_store", ctypes.c_float),
        ("disp_cov", ctypes.c_int), ("store_all", ctypes.c_int),
        ("max_hits", ctypes.c_longlong),
    ]


EXPORTS = ["oagpu_create", "oagpu_destroy", "oagpu_last_error", "oagpu_set_subset", "oagpu_set_covariates",
           "oagpu_set_interactions", "oagpu_set_dosage_model", "oagpu_load_traits", "oagpu_load_genotypes",
           "oagpu_scan", "oagpu_scan_async", "oagpu_scan_wait", "oagpu_host_alloc", "oagpu_host_free",
           "oagpu_last_timings", "oagpu_gemm_variant", "oagpu_gemm_tunings", "oagpu_tvalue",
           "oagpu_panel_create", "oagpu_panel_destroy", "oagpu_server_create", "oagpu_server_destroy",
           "oagpu_server_last_error", "oagpu_server_graph_chunks", "oagpu_server_chunk_snps", "oagpu_server_submit", "oagpu_server_poll", "oagpu_server_wait",
           "oagpu_scan_noalloc"]

_lib = None


def load():
    """Load liboagpu.so and declare its signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("liboagpu.so is not built (%s); run omicabelnomm_amd.build.build()" % LIB_PATH)
    # PyTorch ships its own HIP runtime under the same soname as /opt/rocm's.
    # Whichever is loaded first serves both; when liboagpu.so comes first,
    # torch's GPU support later finds no device.  Loading torch first (when
    # it is installed) lets a process use torch tensors beside this library.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    vp, ll, i = ctypes.c_void_p, ctypes.c_longlong, ctypes.c_int
    L.oagpu_create.argtypes = [ctypes.POINTER(Params), ctypes.POINTER(vp)]
    L.oagpu_create.restype = i
    L.oagpu_destroy.argtypes = [vp]
    L.oagpu_destroy.restype = None
    L.oagpu_last_error.argtypes = [vp]
    L.oagpu_last_error.restype = ctypes.c_char_p
    L.oagpu_set_subset.argtypes = [vp, vp]
    L.oagpu_set_covariates.argtypes = [vp, vp, ll]
    L.oagpu_set_interactions.argtypes = [vp, vp, ll, i]
    L.oagpu_set_dosage_model.argtypes = [vp, i, i, i, vp]
    L.oagpu_load_traits.argtypes = [vp, vp, i, i, ll, i, i]
    L.oagpu_load_genotypes.argtypes = [vp, vp, i, i, ll, i, i, vp, vp, i]
    L.oagpu_scan.argtypes = [vp, ctypes.POINTER(Dense), vp, ll, ctypes.POINTER(Stats)]
    L.oagpu_scan_noalloc.argtypes = [vp, vp, ll, ctypes.POINTER(Stats)]
    L.oagpu_scan_async.argtypes = [vp, i]
    L.oagpu_scan_wait.argtypes = [vp, ctypes.POINTER(Dense), vp, ll, ctypes.POINTER(Stats)]
    L.oagpu_last_timings.argtypes = [vp, ctypes.POINTER(ctypes.c_double), i]
    L.oagpu_gemm_variant.argtypes = [vp]
    L.oagpu_gemm_variant.restype = ctypes.c_char_p
    L.oagpu_gemm_tunings.argtypes = []
    L.oagpu_gemm_tunings.restype = ctypes.c_int
    L.oagpu_tvalue.argtypes = [ctypes.c_float]
    L.oagpu_tvalue.restype = ctypes.c_double
    L.oagpu_panel_create.argtypes = [i, i, ll, vp, i, i, ll, ctypes.POINTER(vp)]
    L.oagpu_panel_destroy.argtypes = [vp]
    L.oagpu_panel_destroy.restype = None
    L.oagpu_server_create.argtypes = [vp, ctypes.POINTER(ServerParams), ctypes.POINTER(vp)]
    L.oagpu_server_destroy.argtypes = [vp]
    L.oagpu_server_destroy.restype = None
    L.oagpu_server_last_error.argtypes = [vp]
    L.oagpu_server_last_error.restype = ctypes.c_char_p
    L.oagpu_server_graph_chunks.argtypes = [vp]
    L.oagpu_server_graph_chunks.restype = ctypes.c_longlong
    L.oagpu_server_chunk_snps.argtypes = [vp]
    L.oagpu_server_chunk_snps.restype = ctypes.c_int
    L.oagpu_server_submit.argtypes = [vp, ctypes.POINTER(Request), ctypes.POINTER(ll)]
    L.oagpu_server_poll.argtypes = [vp, ll]
    L.oagpu_server_wait.argtypes = [vp, ll, vp, ll, ctypes.POINTER(Stats), ctypes.POINTER(ctypes.c_double),
                                    ctypes.c_char_p, i]
    for f in ["oagpu_set_subset", "oagpu_set_covariates", "oagpu_set_interactions", "oagpu_set_dosage_model",
              "oagpu_load_traits", "oagpu_load_genotypes", "oagpu_scan", "oagpu_scan_noalloc", "oagpu_last_timings", "oagpu_panel_create",
              "oagpu_server_create", "oagpu_server_submit", "oagpu_server_poll", "oagpu_server_wait"]:
        getattr(L, f).restype = i
    _lib = L
    return L


class OagpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("oagpu error %d: %s" % (code, msg))
        self.code = code
