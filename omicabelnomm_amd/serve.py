"""Resident genotype panel and request server (include/oagpu_serve.h).

Panel holds n individuals x m SNPs in HBM; Server scans submitted trait sets
against it with continuous batching: requests are admitted at chunk
boundaries within an HBM budget and share one contraction per chunk.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import lib as L
from .scan import _device_ptr


class Panel:
    def __init__(self, G, device=0):
        """G: (m, n) torch tensor on the GPU (used in place) or host array
        (copied), one row per SNP, NaN = missing."""
        self._L = L.load()
        ptr, where, dt, ld, m, keep = _device_ptr(G)
        self._keep = keep
        self.n = int(ld) if where == L.DEVICE else int(np.asarray(keep).shape[0])
        self.m = int(m)
        self._p = ctypes.c_void_p()
        rc = self._L.oagpu_panel_create(device, self.n, self.m, ptr, dt, where, ld, ctypes.byref(self._p))
        if rc != L.OK:
            raise L.OagpuError(rc, "oagpu_panel_create failed")

    def close(self):
        if self._p:
            self._L.oagpu_panel_destroy(self._p)
            self._p = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class Answer:
    hits: np.ndarray
    stats: dict
    latency_ms: float


class Server:
    def __init__(self, panel: Panel, precision="bf16x3", hbm_budget=0, max_active=0, chunk_snps=0, graphs=False):
        self._L = L.load()
        self.panel = panel
        prm = L.ServerParams(precision=L.PRECISIONS[precision], hbm_budget=hbm_budget, max_active=max_active,
                             chunk_snps=chunk_snps, graphs=1 if graphs else 0)
        self._s = ctypes.c_void_p()
        rc = self._L.oagpu_server_create(panel._p, ctypes.byref(prm), ctypes.byref(self._s))
        if rc != L.OK:
            msg = self._L.oagpu_server_last_error(self._s).decode() if self._s else ""
            raise L.OagpuError(rc, msg)
        self._max_hits = {}

    def submit(self, Y, XL, keep=None, *, min_p_store=0.1, min_p_disp=0.05, min_r2_store=1e-5,
               disp_cov=False, store_all=False, max_hits=1 << 20) -> int:
        """Y: (n, t) and XL: (n, l) host arrays (NaN = missing in Y)."""
        Y = np.asfortranarray(Y, dtype=np.float64)
        XL = np.asfortranarray(XL, dtype=np.float64)
        n = self.panel.n
        assert Y.shape[0] == n and XL.shape[0] == n
        k = None if keep is None else np.ascontiguousarray(keep, dtype=np.uint8)
        req = L.Request(t=Y.shape[1], l=XL.shape[1], Y=Y.ctypes.data, ldy=n, XL=XL.ctypes.data, ldxl=n,
                        keep=k.ctypes.data if k is not None else None, min_p_store=min_p_store,
                        min_p_disp=min_p_disp, min_r2_store=min_r2_store, disp_cov=int(disp_cov),
                        store_all=int(store_all), max_hits=max_hits)
        rid = ctypes.c_longlong()
        rc = self._L.oagpu_server_submit(self._s, ctypes.byref(req), ctypes.byref(rid))
        if rc != L.OK:
            raise L.OagpuError(rc, self._L.oagpu_server_last_error(self._s).decode())
        self._max_hits[rid.value] = max_hits
        return rid.value

    def chunk_snps(self) -> int:
        """SNPs per chunk: the configured size, or (chunk_snps=-1) the size
        chosen by timing once the first requests were admitted."""
        return int(self._L.oagpu_server_chunk_snps(self._s))

    def graph_chunks(self) -> int:
        """Chunks replayed from a captured HIP graph so far."""
        return int(self._L.oagpu_server_graph_chunks(self._s))

    def done(self, rid) -> bool:
        return self._L.oagpu_server_poll(self._s, rid) == 1

    def wait(self, rid) -> Answer:
        cap = self._max_hits.pop(rid, 1 << 20)
        hits = np.zeros(cap, dtype=L.HIT_DTYPE)
        st = L.Stats()
        lat = ctypes.c_double()
        err = ctypes.create_string_buffer(512)
        rc = self._L.oagpu_server_wait(self._s, rid, hits.ctypes.data, cap, ctypes.byref(st), ctypes.byref(lat),
                                       err, 512)
        stats = {k: getattr(st, k) for k, _ in L.Stats._fields_}
        if rc != L.OK:
            raise L.OagpuError(rc, err.value.decode())
        return Answer(hits=hits[: stats["n_hits"]].copy(), stats=stats, latency_ms=lat.value)

    def close(self):
        if self._s:
            self._L.oagpu_server_destroy(self._s)
            self._s = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
