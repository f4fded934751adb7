"""ctypes wrapper of the FP64 CPU oracle (oracle/oracle.cpp) and of the FP32
restatement of the reference's blocked algorithm (oracle/ref32.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() to
compare its scan with, and by bench.py's cpu_baseline leg.  The product path
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
        L.oa_ref32_scan.argtypes = [ctypes.POINTER(_Params), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_int, ctypes.c_int, ctypes.POINTER(_Out)]
        L.oa_ref32_scan.restype = ctypes.c_int
        L.oa_ref32_blas.argtypes = []
        L.oa_ref32_blas.restype = ctypes.c_char_p
        _lib = L
    return _lib


def ref32_blas() -> str:
    """The BLAS behind the FP32 restatement's products."""
    return lib().oa_ref32_blas().decode()


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


def scan_ref32(XL, XR, Y, r, *, tb=1000, mb=500, min_p_store=0.1, min_p_disp=0.05, min_r2_store=1e-5,
               disp_cov=False, store_all=False, real_n=None, threads=1):
    """The FP32 restatement of the reference's blocked algorithm (tb traits
    and mb SNPs per block, the reference's defaults), same inputs and
    outputs as scan()."""
    XL = np.asfortranarray(XL, dtype=np.float32)
    XR = np.asfortranarray(XR, dtype=np.float32)
    Y = np.asfortranarray(Y, dtype=np.float32)
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
    out = _Out(**{k: v.ctypes.data for k, v in res.items() if k != "var_x"})
    rc = lib().oa_ref32_scan(ctypes.byref(prm), XL.ctypes.data, XR.ctypes.data, Y.ctypes.data,
                             int(tb), int(mb), ctypes.byref(out))
    res["status"] = rc
    res["n_sig"] = out.n_sig
    res["n_kept"] = out.n_kept
    res["n_emitted"] = out.n_emitted
    return res


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


# ---------------------------------------------------------------------------
# Dosage models (host-side in the reference).

def read_dosage_factors(text: str, expected: int):
    """read_dosages (src/AIOwrapper.cpp:2236-2256): `expected` floats read
    with operator>> while the stream is not at its end.  Where the reference
    accepts a list with every factor read, the result is the same.  A read
    that fails (the end of the file after a trailing blank, or a token that
    is not a number) leaves the reference's factor uninitialised while still
    counting it; here such a list is refused.  Returns (ok, factors as
    float32 values)."""
    toks = text.split()
    vals = []
    for tok in toks[:expected]:
        try:
            vals.append(float(np.float32(float(tok))))
        except ValueError:
            break
    return len(vals) == expected, np.array(vals, dtype=np.float64)


def dosage_model(model: int, file_r: int, factor_text: str = ""):
    """Factor vector of a dosage model (src/AIOwrapper.cpp:256-340):
    model 0 additive (2*AA + AB), 1 dominant (AA + AB), 2 recessive (AA),
    3 linear from a file (one column per factor), 4 additive from a file.
    Returns (status, factors, summed): status 0 ok, -1 wrong column count,
    -2 unreadable factor list.  The reference's recessive model reads an
    uninitialised second factor (file_r = 3) or no column at all
    (file_r = 2); here it is the model its own message states,
    (1*AA, 0*AB, 0*BB)."""
    if model in (0, 1, 2):
        if file_r not in (2, 3):
            return -1, None, True
        f = {0: [2.0, 1.0], 1: [1.0, 1.0], 2: [1.0]}[model]
        return 0, np.array(f), True
    ok, f = read_dosage_factors(factor_text, file_r)
    if not ok:
        return -2, None, model == 4
    return 0, f, model == 4


def dosage_apply(G, file_r: int, factors, summed: bool):
    """Genotype columns of a dosage model (src/AIOwrapper.cpp:1258-1290):
    G holds file_r columns per SNP; summed: one column per SNP,
    sum_i f[i] G[:, s*file_r + i]; otherwise f[i] G[:, s*file_r + i] for
    each factor.  NaN stays NaN (imputed later, as in the reference)."""
    G = np.asarray(G, dtype=np.float64)
    n, cols = G.shape
    assert cols % file_r == 0
    k = len(factors)
    blocks = G.reshape(n, cols // file_r, file_r)[:, :, :k]
    if summed:
        out = np.zeros((n, cols // file_r))
        for i in range(k):
            out = out + factors[i] * blocks[:, :, i]
        return out
    return (blocks * np.asarray(factors)[None, None, :]).reshape(n, -1)
