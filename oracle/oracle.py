"""ctypes wrapper of the FP64 CPU oracle (oracle/oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker, and by bench.py's cpu_baseline leg.  The product path
(omicabelnomm_amd) never imports this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "liboaoracle.so")
_lib = None


class _Params(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int), ("real_n", ctypes.c_int), ("l", ctypes.c_int),
        ("r", ctypes.c_int), ("t", ctypes.c_int), ("m", ctypes.c_int),
        ("disp_cov", ctypes.c_int), ("store_all", ctypes.c_int),
        ("min_r2_store", ctypes.c_double), ("t_star", ctypes.c_double),
        ("min_p_disp", ctypes.c_double), ("threads", ctypes.c_int),
    ]


class _Out(ctypes.Structure):
    _fields_ = [
        ("beta", ctypes.c_void_p), ("se", ctypes.c_void_p),
        ("tstat", ctypes.c_void_p), ("pval", ctypes.c_void_p),
        ("kept", ctypes.c_void_p), ("r2", ctypes.c_void_p),
        ("rss", ctypes.c_void_p), ("emitted", ctypes.c_void_p),
        ("n_used", ctypes.c_void_p), ("var_x", ctypes.c_void_p),
        ("n_sig", ctypes.c_int64), ("n_kept", ctypes.c_int64),
        ("n_emitted", ctypes.c_int64),
    ]


def build() -> str:
    """Compile the oracle with its Makefile (g++)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.oa_oracle_scan.argtypes = [ctypes.POINTER(_Params), ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.POINTER(_Out)]
        L.oa_oracle_scan.restype = ctypes.c_int
        L.oa_oracle_tvalue.argtypes = [ctypes.c_float]
        L.oa_oracle_tvalue.restype = ctypes.c_double
        L.oa_oracle_erfinv.argtypes = [ctypes.c_double]
        L.oa_oracle_erfinv.restype = ctypes.c_double
        L.oa_oracle_pvalue.argtypes = [ctypes.c_double]
        L.oa_oracle_pvalue.restype = ctypes.c_double
        L.oa_oracle_min_singular.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.oa_oracle_min_singular.restype = ctypes.c_double
        L.oa_oracle_exclusions.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_void_p]
        L.oa_oracle_exclusions.restype = ctypes.c_int
        _lib = L
    return _lib


def tvalue(pval: float) -> float:
    return lib().oa_oracle_tvalue(float(np.float32(pval)))


def erfinv(x: float) -> float:
    return lib().oa_oracle_erfinv(float(x))


def pvalue(t: float) -> float:
    return lib().oa_oracle_pvalue(float(t))


def min_singular(XL: np.ndarray) -> float:
    XL = np.asfortranarray(XL, dtype=np.float64)
    return lib().oa_oracle_min_singular(XL.ctypes.data, XL.shape[0], XL.shape[1])


def exclusions(text: str, n: int):
    keep = np.zeros(n, dtype=np.uint8)
    rc = lib().oa_oracle_exclusions(text.encode(), n, keep.ctypes.data)
    return rc, keep.astype(bool)


def scan(XL, XR, Y, r, *, min_p_store=0.1, min_p_disp=0.05, min_r2_store=1e-5,
         disp_cov=False, store_all=False, real_n=None, threads=1):
    """Dense FP64 restatement of the scan.

    XL: (n, l), XR: (n, r*m) with NaN = missing genotype, Y: (n, t) with NaN =
    missing phenotype.  Thresholds are the reference Settings values after its
    normalisation (src/Algorithm.cpp:338-342); they are rounded to float like
    the reference's members.  Returns a dict of numpy arrays shaped
    (t, m, p) / (t, m) / (t,).
    """
    XL = np.asfortranarray(XL, dtype=np.float64)
    XR = np.asfortranarray(XR, dtype=np.float64)
    Y = np.asfortranarray(Y, dtype=np.float64)
    n, l = XL.shape
    assert XR.shape[0] == n and Y.shape[0] == n and XR.shape[1] % r == 0
    m = XR.shape[1] // r
    t = Y.shape[1]
    p = l + r
    prm = _Params(n=n, real_n=real_n or n, l=l, r=r, t=t, m=m,
                  disp_cov=int(disp_cov), store_all=int(store_all),
                  min_r2_store=float(np.float32(min_r2_store)),
                  t_star=tvalue(min_p_store),
                  min_p_disp=float(np.float32(min_p_disp)), threads=threads)
    res = {
        "beta": np.zeros((t, m, p)), "se": np.zeros((t, m, p)),
        "tstat": np.zeros((t, m, p)), "pval": np.zeros((t, m, p)),
        "kept": np.zeros((t, m, p), dtype=np.uint8),
        "r2": np.zeros((t, m)), "rss": np.zeros((t, m)),
        "emitted": np.zeros((t, m), dtype=np.uint8),
        "n_used": np.zeros(t, dtype=np.int32), "var_x": np.zeros((m, p)),
    }
    out = _Out(**{k: v.ctypes.data for k, v in res.items()})
    rc = lib().oa_oracle_scan(ctypes.byref(prm), XL.ctypes.data, XR.ctypes.data,
                              Y.ctypes.data, ctypes.byref(out))
    res["status"] = rc
    res["n_sig"] = out.n_sig
    res["n_kept"] = out.n_kept
    res["n_emitted"] = out.n_emitted
    res["t_star"] = prm.t_star
    return res
