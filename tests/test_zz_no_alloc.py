"""oagpu_scan_noalloc (the scan without device allocation) against
oagpu_scan, through the C ABI.

- two calls in a row, the second on a smaller input (the shared scratch
  array keeps the first call's contents), each equal to oagpu_scan on the
  same context with the same contraction kernel (OAGPU_GEMM_VARIANT pins
  oagpu_scan's autotuner to the hand-written kernel this entry point
  launches): hits and statistics identical;
- a workload whose chunks do not fit the array at once runs in several
  chunks and still matches;
- the call leaves the device's free memory unchanged;
- a workload whose single-SNP chunk exceeds the array is refused with a
  clear error, and the next call still works.
"""
import ctypes

import numpy as np
import pytest
import torch

from omicabelnomm_amd import Scanner, synth
from omicabelnomm_amd import lib as L

pytestmark = pytest.mark.gpu

THR = dict(min_p_store=0.2, min_p_disp=0.05, min_r2_store=-10.0)


def scanner(data, precision, **kw):
    sc = Scanner(data.n, data.l, data.r, precision, max_traits=max(64, data.t), max_hits=1 << 22, **THR, **kw)
    sc.set_covariates(data.XL)
    if data.E is not None:
        sc.set_interactions(data.E)
    sc.load_traits(data.Y.to("cuda").contiguous())
    sc.load_genotypes(data.G.to("cuda").contiguous(), m=data.m, col_src=data.col_src, col_env=data.col_env)
    return sc


def noalloc(sc, cap=1 << 22):
    lib = L.load()
    hits = np.zeros(cap, dtype=L.HIT_DTYPE)
    st = L.Stats()
    rc = lib.oagpu_scan_noalloc(sc._ctx, hits.ctypes.data, cap, ctypes.byref(st))
    stats = {k: getattr(st, k) for k, _ in L.Stats._fields_}
    return rc, hits[: stats["n_hits"]].copy(), stats


CASES = [
    # a larger input, then a smaller one
    ("groups_gxe", dict(n=900, l=4, m=300, t=90, interaction=True, miss="groups", pattern_groups=4,
                        pattern_rate=0.03, planted=12, seed=71)),
    ("mcar_small", dict(n=300, l=3, m=40, t=20, miss="mcar", miss_rate=0.05, planted=4, seed=72)),
]


FIXED = {"fp64": "f64_direct_128x128", "bf16x3": "bf16x3_lds_128x128"}


@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
def test_two_calls_match_oagpu_scan(precision, monkeypatch):
    monkeypatch.setenv("OAGPU_GEMM_VARIANT", FIXED[precision])
    for name, kw in CASES:
        data = synth.make(**kw)
        print("case noalloc", precision, name, kw, flush=True)
        with scanner(data, precision) as sc:
            rc, got, st = noalloc(sc)
            assert rc == L.OK, sc._L.oagpu_last_error(sc._ctx)
            want = sc.scan()
            assert got.tobytes() == want.hits.tobytes()
            for k in ("n_hits", "n_emitted", "n_sig", "n_fallback", "n_patterns"):
                assert st[k] == want.stats[k], (k, st[k], want.stats[k])


def test_several_chunks_and_no_allocation(monkeypatch):
    """n = 20,000 with 1,500 traits: about 1,400 SNPs of chunk buffers fit
    the array, so 3,000 SNPs take more than one chunk; the device's free
    memory is the same before and after the call."""
    data = synth.make(n=20000, l=3, m=3000, t=1500, miss="groups", pattern_groups=3, pattern_rate=0.01,
                      planted=10, seed=73)
    print("case noalloc chunks", flush=True)
    monkeypatch.setenv("OAGPU_GEMM_VARIANT", FIXED["bf16x3"])
    with scanner(data, "bf16x3") as sc:
        rc, first, st = noalloc(sc)   # code objects, first launches
        assert rc == L.OK, sc._L.oagpu_last_error(sc._ctx)
        assert st["n_chunks"] >= 2, st
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        rc, again, st2 = noalloc(sc)
        torch.cuda.synchronize()
        free1 = torch.cuda.mem_get_info()[0]
        assert rc == L.OK
        assert free1 == free0, (free0, free1)
        assert again.tobytes() == first.tobytes()
        want = sc.scan()
        torch.cuda.synchronize()
        assert torch.cuda.mem_get_info()[0] < free1   # oagpu_scan allocates its chunk buffers
        assert again.tobytes() == want.hits.tobytes()


def test_too_large_is_refused_then_next_call_works():
    """FP64 with r = 2 at n = 9,000,000: one SNP's operand rows alone take
    144 MB, beyond the 128 MiB array."""
    n = 9_000_000
    rng = np.random.default_rng(3)
    XL = np.column_stack([np.ones(n), rng.standard_normal(n)])
    print("case noalloc too large", flush=True)
    with Scanner(n, 2, 2, "fp64", max_traits=1, max_hits=1 << 10, **THR) as sc:
        sc.set_covariates(XL)
        sc.set_interactions(rng.standard_normal((n, 1)))
        sc.load_traits(torch.randn((1, n), dtype=torch.float64, device="cuda"))
        G = torch.rand((1, n), dtype=torch.float32, device="cuda") * 2
        sc.load_genotypes(G, m=1, col_src=np.array([0, 0], dtype=np.int32), col_env=np.array([0, -1], dtype=np.int32))
        rc, _, _ = noalloc(sc, cap=1 << 10)
        msg = sc._L.oagpu_last_error(sc._ctx).decode()
        assert rc == L.ERR_ARG and "too large for oagpu_scan_noalloc" in msg, (rc, msg)
    data = synth.make(**CASES[1][1])
    with scanner(data, "fp64") as sc:
        rc, got, _ = noalloc(sc)
        assert rc == L.OK
        assert got.tobytes() == sc.scan().hits.tobytes()
