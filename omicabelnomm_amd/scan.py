"""Host-side mirror of the reference's scan interface over liboagpu.so.

`Settings`, `Outputs`, `Algorithm.applyDefaultParams` and the threshold
normalisation follow src/Definitions.h:168-257 and src/Algorithm.cpp:277-360.
`Scanner` exposes the in-memory block interface of the C ABI
(include/oagpu.h): covariates, a trait block, a genotype block, scan.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import lib as L

P_NEQ_B_OPT_MD = 7  # the only solver id the reference implements (src/Algorithm.h:29)


@dataclass
class Settings:
    """Subset of the reference Settings struct (src/Definitions.h:168-219)
    that the scan reads, plus the precision/device of this implementation."""
    n: int = 0
    t: int = 0
    m: int = 0
    l: int = 0
    r: int = 1
    p: int = 0
    tb: int = 0
    mb: int = 0
    threads: int = 1
    storeBin: bool = False
    minPdisp: float = 0.05
    minPstore: float = 0.1
    minR2disp: float = 0.000001
    minR2store: float = 0.00001
    storePInd: bool = False
    disp_cov: bool = False
    fnameAL: str = ""
    fnameAR: str = ""
    fnameY: str = ""
    fnameOutFiles: str = ""
    fname_excludelist: str = ""
    use_interactions: bool = False
    keep_depVar: bool = False
    use_multiple_interaction_sets: bool = False
    fname_interactions: str = ""
    ini_IDX_interactions: int = -1
    end_IDX_interactions: int = -1
    limit_m: int = 2**31 - 1
    limit_t: int = 2**31 - 1
    limit_n: int = 2**31 - 1
    dosages: bool = False
    model: int = -1
    fname_dosages: str = ""
    precision: str = "fp64"
    device: int = 0


@dataclass
class Outputs:
    """Reference Outputs (src/Definitions.h:222-257): the fields this
    implementation fills."""
    duration: float = 0.0
    total_sig_results: int = 0
    n_hits: int = 0
    n_emitted: int = 0
    n_fallback: int = 0
    timings_ms: dict = field(default_factory=dict)


class _CSettings(ctypes.Structure):
    _fields_ = [
        ("fnameAL", ctypes.c_char_p), ("fnameAR", ctypes.c_char_p), ("fnameY", ctypes.c_char_p),
        ("fnameOutFiles", ctypes.c_char_p), ("fname_excludelist", ctypes.c_char_p),
        ("fname_interactions", ctypes.c_char_p), ("r", ctypes.c_int), ("threads", ctypes.c_int),
        ("storeBin", ctypes.c_int), ("storePInd", ctypes.c_int), ("disp_cov", ctypes.c_int),
        ("minPdisp", ctypes.c_float), ("minPstore", ctypes.c_float), ("minR2disp", ctypes.c_float),
        ("minR2store", ctypes.c_float), ("use_interactions", ctypes.c_int),
        ("keep_depVar", ctypes.c_int), ("use_multiple_interaction_sets", ctypes.c_int),
        ("ini_IDX_interactions", ctypes.c_int), ("end_IDX_interactions", ctypes.c_int),
        ("limit_m", ctypes.c_int), ("limit_t", ctypes.c_int), ("limit_n", ctypes.c_int),
        ("precision", ctypes.c_int), ("device", ctypes.c_int), ("verbose", ctypes.c_int),
        ("dosages", ctypes.c_int), ("model", ctypes.c_int), ("fname_dosages", ctypes.c_char_p),
    ]


class _COutputs(ctypes.Structure):
    _fields_ = [
        ("duration", ctypes.c_double), ("acc_loady", ctypes.c_double),
        ("acc_loadxr", ctypes.c_double), ("acc_gpu", ctypes.c_double),
        ("acc_storeb", ctypes.c_double), ("total_sig_results", ctypes.c_longlong),
        ("n_txt_lines", ctypes.c_longlong), ("n_bin_results", ctypes.c_longlong),
        ("n", ctypes.c_int), ("l", ctypes.c_int), ("r", ctypes.c_int), ("t", ctypes.c_int),
        ("m", ctypes.c_int), ("realN", ctypes.c_int),
    ]


_host = None


def _host_lib():
    global _host
    if _host is None:
        import os
        path = os.path.join(L.HERE, "libomicabelnomm.so")
        if not os.path.exists(path):
            raise RuntimeError("libomicabelnomm.so is not built; run omicabelnomm_amd.build.build()")
        L.load()
        h = ctypes.CDLL(path)
        h.omicabelnomm_solve.argtypes = [ctypes.POINTER(_CSettings), ctypes.POINTER(_COutputs),
                                         ctypes.c_char_p, ctypes.c_int]
        h.omicabelnomm_solve.restype = ctypes.c_int
        _host = h
    return _host


class Algorithm:
    """The reference's entry points (src/Algorithm.h:41-47)."""

    def applyDefaultParams(self, params: Settings) -> None:
        """src/Algorithm.cpp:277-311."""
        d = Settings()
        for k in ("fname_excludelist", "disp_cov", "storePInd", "storeBin", "threads", "r",
                  "use_interactions", "keep_depVar", "use_multiple_interaction_sets",
                  "fname_interactions", "ini_IDX_interactions", "end_IDX_interactions",
                  "minR2store", "minR2disp", "minPstore", "minPdisp", "limit_m", "limit_t", "limit_n"):
            setattr(params, k, getattr(d, k))

    def solve(self, params: Settings, out: Outputs, type: int = P_NEQ_B_OPT_MD, verbose=False) -> None:
        """Drop-in file scan: DatABEL inputs named in params, outputs
        <fnameOutFiles>_dis.txt (+ _sto.ibin/_sto.dbin with storeBin).
        Raises OagpuError where the reference prints and exits."""
        if type != P_NEQ_B_OPT_MD:
            return   # the reference's solve ignores every other id (src/Algorithm.cpp:37-48)
        h = _host_lib()
        enc = lambda x: (x or "").encode()
        cs = _CSettings(
            fnameAL=enc(params.fnameAL), fnameAR=enc(params.fnameAR), fnameY=enc(params.fnameY),
            fnameOutFiles=enc(params.fnameOutFiles), fname_excludelist=enc(params.fname_excludelist),
            fname_interactions=enc(params.fname_interactions), r=params.r, threads=params.threads,
            storeBin=int(params.storeBin), storePInd=int(params.storePInd),
            disp_cov=int(params.disp_cov), minPdisp=params.minPdisp, minPstore=params.minPstore,
            minR2disp=params.minR2disp, minR2store=params.minR2store,
            use_interactions=int(params.use_interactions), keep_depVar=int(params.keep_depVar),
            use_multiple_interaction_sets=int(params.use_multiple_interaction_sets),
            ini_IDX_interactions=params.ini_IDX_interactions,
            end_IDX_interactions=params.end_IDX_interactions, limit_m=params.limit_m,
            limit_t=params.limit_t, limit_n=params.limit_n,
            precision=L.PRECISIONS[params.precision], device=params.device, verbose=int(verbose),
            dosages=int(params.dosages), model=params.model, fname_dosages=enc(params.fname_dosages))
        co = _COutputs()
        err = ctypes.create_string_buffer(4096)
        rc = h.omicabelnomm_solve(ctypes.byref(cs), ctypes.byref(co), err, 4096)
        if rc != 0:
            raise L.OagpuError(rc, err.value.decode(errors="replace"))
        out.duration = co.duration
        out.total_sig_results = co.total_sig_results
        out.n_hits = co.n_bin_results
        for k in ("n", "l", "r", "t", "m"):
            setattr(params, k, getattr(co, k))
        params.p = co.l + co.r


def normalize_thresholds(params: Settings) -> Settings:
    """Threshold normalisation at the top of partialNEQ_Blocked_STL_MD
    (src/Algorithm.cpp:338-342): with -b, or a display threshold looser
    than the storage one, storage uses the display threshold."""
    if params.minR2disp < params.minR2store or params.storeBin:
        params.minR2store = params.minR2disp
    if params.minPdisp > params.minPstore or params.storeBin:
        params.minPstore = params.minPdisp
    return params


def tvalue(pval: float) -> float:
    """getTvalue (src/Utility.cpp:244-250)."""
    return L.load().oagpu_tvalue(float(np.float32(pval)))


def _device_ptr(x):
    """(pointer, where, dtype code, ld, ncols) for a numpy or torch array.

    numpy arrays are (rows, cols) matrices stored column-major (Fortran) or
    made so; torch CUDA tensors are given as (cols, ld) row-major, i.e. one
    row per column of the matrix."""
    try:
        import torch  # noqa: F401
        is_torch = type(x).__module__.startswith("torch")
    except ImportError:  # pragma: no cover
        is_torch = False
    if is_torch:
        if not x.is_cuda:
            x = x.numpy()
        else:
            if x.dim() != 2 or x.stride(1) != 1:
                raise ValueError("device arrays must be (columns, ld) row-major tensors")
            dt = {"torch.float64": L.F64, "torch.float32": L.F32, "torch.float16": L.F16}[str(x.dtype)]
            return x.data_ptr(), L.DEVICE, dt, x.stride(0), x.shape[0], x
    a = np.asarray(x)
    if a.dtype not in (np.float64, np.float32, np.float16):
        a = a.astype(np.float64)
    a = np.asfortranarray(a)
    dt = {np.dtype(np.float64): L.F64, np.dtype(np.float32): L.F32, np.dtype(np.float16): L.F16}[a.dtype]
    return a.ctypes.data, L.HOST, dt, a.shape[0], a.shape[1], a


@dataclass
class ScanResult:
    hits: np.ndarray
    stats: dict
    dense: Optional[dict]
    timings_ms: dict


class Scanner:
    """One liboagpu context: covariates, then trait and genotype blocks."""

    def __init__(self, n, l, r=1, precision="fp64", *, real_n=None, disp_cov=False,
                 store_all=False, min_p_store=0.1, min_p_disp=0.05, min_r2_store=1e-5,
                 max_traits=1024, max_snps=0, max_hits=1 << 22, device=0):
        self._L = L.load()
        self.n, self.l, self.r, self.p = n, l, r, l + r
        self.params = L.Params(n=n, l=l, r=r, real_n=real_n or n,
                               precision=L.PRECISIONS[precision] if isinstance(precision, str) else precision,
                               disp_cov=int(disp_cov), store_all=int(store_all),
                               min_r2_store=min_r2_store, min_p_store=min_p_store,
                               min_p_disp=min_p_disp, max_traits=max_traits, max_snps=max_snps,
                               max_hits=max_hits, device=device)
        self._ctx = ctypes.c_void_p()
        rc = self._L.oagpu_create(ctypes.byref(self.params), ctypes.byref(self._ctx))
        self._check(rc)
        self._keep = {}
        self.t = 0
        self.m = 0

    def _check(self, rc):
        if rc != L.OK:
            msg = self._L.oagpu_last_error(self._ctx).decode() if self._ctx else "create failed"
            raise L.OagpuError(rc, msg)

    def close(self):
        if self._ctx:
            self._L.oagpu_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_subset(self, keep):
        """Analyse only the individuals with keep[k] true (before set_covariates)."""
        if keep is None:
            self._check(self._L.oagpu_set_subset(self._ctx, None))
            return
        k = np.ascontiguousarray(keep, dtype=np.uint8)
        assert k.shape == (self.n,)
        self._keep["subset"] = k
        self._check(self._L.oagpu_set_subset(self._ctx, k.ctypes.data))

    def set_covariates(self, XL):
        XL = np.asfortranarray(XL, dtype=np.float64)
        assert XL.shape == (self.n, self.l)
        self._check(self._L.oagpu_set_covariates(self._ctx, XL.ctypes.data, XL.shape[0]))

    def set_interactions(self, E):
        E = np.asfortranarray(E, dtype=np.float64)
        assert E.shape[0] == self.n
        self._check(self._L.oagpu_set_interactions(self._ctx, E.ctypes.data, E.shape[0], E.shape[1]))

    def set_dosage_model(self, factors, file_r=1, sum=True):
        """Dosage model for the genotype blocks loaded afterwards
        (src/AIOwrapper.cpp:256-340,1208-1297): each file SNP has file_r
        columns, the first len(factors) are read; sum=True makes one column
        per SNP (sum of factor x column), sum=False one column per factor.
        factors=None switches the model off."""
        if factors is None:
            self._check(self._L.oagpu_set_dosage_model(self._ctx, L.DOSAGE_NONE, 1, 1, None))
            self._dos = None
            return
        f = np.ascontiguousarray(factors, dtype=np.float64)
        mode = L.DOSAGE_SUM if sum else L.DOSAGE_SCALE
        self._check(self._L.oagpu_set_dosage_model(self._ctx, mode, int(file_r), int(f.size), f.ctypes.data))
        self._dos = (int(file_r), 1 if sum else int(f.size))

    def load_traits(self, Y, trait_offset=0):
        ptr, where, dt, ld, t, keep = _device_ptr(Y)
        self._keep["Y"] = keep
        self._check(self._L.oagpu_load_traits(self._ctx, ptr, dt, where, ld, t, trait_offset))
        self.t = t

    def load_genotypes(self, G, m=None, col_src=None, col_env=None, snp_offset=0):
        ptr, where, dt, ld, n_src, keep = _device_ptr(G)
        self._keep["G"] = keep
        n_col = n_src
        dos = getattr(self, "_dos", None)
        if dos is not None:
            n_col = (n_src // dos[0]) * dos[1]
        if m is None:
            m = n_col // self.r if col_src is None else len(col_src) // self.r
        src = env = None
        if col_src is not None:
            src = np.ascontiguousarray(col_src, dtype=np.int32)
            assert src.size == m * self.r
        if col_env is not None:
            env = np.ascontiguousarray(col_env, dtype=np.int32)
            assert env.size == m * self.r
        self._keep["desc"] = (src, env)
        self._check(self._L.oagpu_load_genotypes(
            self._ctx, ptr, dt, where, ld, n_src, m,
            src.ctypes.data if src is not None else None,
            env.ctypes.data if env is not None else None, snp_offset))
        self.m = m

    def scan(self, dense=False, hit_cap=None) -> ScanResult:
        t, m, p = self.t, self.m, self.p
        dres = None
        dstruct = L.Dense()
        if dense:
            dres = {
                "beta": np.zeros((t, m, p)), "se": np.zeros((t, m, p)),
                "tstat": np.zeros((t, m, p)), "pval": np.zeros((t, m, p)),
                "kept": np.zeros((t, m, p), dtype=np.uint8),
                "r2": np.zeros((t, m)), "rss": np.zeros((t, m)),
                "emitted": np.zeros((t, m), dtype=np.uint8),
                "n_used": np.zeros(t, dtype=np.int32), "var_x": np.zeros((m, p)),
            }
            for k, v in dres.items():
                setattr(dstruct, k, v.ctypes.data)
        cap = int(hit_cap if hit_cap is not None else self.params.max_hits)
        hits = np.zeros(cap, dtype=L.HIT_DTYPE)
        st = L.Stats()
        rc = self._L.oagpu_scan(self._ctx, ctypes.byref(dstruct) if dense else None,
                                hits.ctypes.data, cap, ctypes.byref(st))
        stats = {k: getattr(st, k) for k, _ in L.Stats._fields_}
        if rc == L.ERR_HIT_OVERFLOW:
            raise L.OagpuError(rc, self._L.oagpu_last_error(self._ctx).decode() + " stats=%s" % stats)
        self._check(rc)
        return ScanResult(hits=hits[: stats["n_hits"]].copy(), stats=stats, dense=dres,
                          timings_ms=self.timings())

    def scan_async(self) -> None:
        """Enqueue the scan of the loaded traits and genotypes and return; the
        next genotype block may be loaded (its upload overlaps this scan)
        before scan_wait collects the hits (oagpu_scan_async)."""
        self._check(self._L.oagpu_scan_async(self._ctx, 0))

    def scan_wait(self, hit_cap=None) -> ScanResult:
        """Hits and statistics of the scan started by scan_async."""
        cap = int(hit_cap if hit_cap is not None else self.params.max_hits)
        hits = np.zeros(cap, dtype=L.HIT_DTYPE)
        st = L.Stats()
        rc = self._L.oagpu_scan_wait(self._ctx, None, hits.ctypes.data, cap, ctypes.byref(st))
        stats = {k: getattr(st, k) for k, _ in L.Stats._fields_}
        if rc == L.ERR_HIT_OVERFLOW:
            raise L.OagpuError(rc, self._L.oagpu_last_error(self._ctx).decode() + " stats=%s" % stats)
        self._check(rc)
        return ScanResult(hits=hits[: stats["n_hits"]].copy(), stats=stats, dense=None,
                          timings_ms=self.timings())

    def gemm_variant(self) -> str:
        return self._L.oagpu_gemm_variant(self._ctx).decode()

    def timings(self) -> dict:
        buf = (ctypes.c_double * 7)()
        k = self._L.oagpu_last_timings(self._ctx, buf, 7)
        names = ["prep", "gemm", "snp_gram", "corrections", "epilogue", "fallback", "total"]
        return {names[i]: buf[i] for i in range(k)}
