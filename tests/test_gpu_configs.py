"""The benchmark workloads, scanned on the GPU and checked against the FP64
oracle.

The reference's own check program scans its full scenarios
(tests/test.cpp:155-308) and only counts significant results; here each
workload is scanned through the C ABI and its results are held to the
oracle:

- C0 (n=1,000, l=3, m=10,000, t=10, no missing data) at full size, dense;
- C1 (n=5,000, l=6, m=100,000, t=1,000, 5% MCAR per trait) and C2
  (n=10,000, l=5, r=2 [E*G, G], m=200,000, t=2,000, 8 pattern groups) at
  full size -- chunks sized from free device memory, one chunk for the
  FP32 and split-BF16 scans -- with the hits of the
  first SNPs x every trait compared with the oracle run on those SNPs; and
  the same generators at m=256 scanned densely, where the planted effects
  give hits at p < 1e-8;
- C3 (64 uneven jobs with individual subsets) and C4 (request stream) as a
  16,384-SNP float16 panel served by the request server, the first requests
  of each job list checked against the oracle on the request's rows.

Bars: FP64 beta and SE within 1e-10 (relative to max(|beta|, SE) and SE),
identical kept/emitted records and counts; FP32 and split-BF16 within
max |dbeta|/SE <= 1e-3 and max |d -log10 p| <= 1e-3 with the p < 1e-8 hit
set identical.
"""
import numpy as np
import pytest

from omicabelnomm_amd import Scanner, synth
from omicabelnomm_amd.serve import Panel, Server
from parity_util import (MIXED_TOL, STRONG_P, assert_fp64_records, assert_mixed_records, compare_fp64, keys,
                         oracle_records, precision_report, window)

pytestmark = pytest.mark.gpu


def scanner_for(data, precision, **thr):
    sc = Scanner(data.n, data.l, data.r, precision, max_traits=data.t, max_hits=1 << 22, **thr)
    sc.set_covariates(data.XL)
    if data.E is not None:
        sc.set_interactions(data.E)
    return sc


def scan_device(data, precision, dense=False, **thr):
    with scanner_for(data, precision, **thr) as sc:
        sc.load_traits(data.Y.contiguous())
        sc.load_genotypes(data.G.contiguous(), m=data.m, col_src=data.col_src, col_env=data.col_env)
        res = sc.scan(dense=dense)
        res.stats["variant"] = sc.gemm_variant()
    return res


def strong_set(d):
    kept = d["kept"].astype(bool) & d["emitted"].astype(bool)[:, :, None] & (d["pval"] < STRONG_P)
    return set(zip(*np.nonzero(kept)))


# ---------------------------------------------------------------------------
# C0: full size, dense

def test_c0_full_size(oracle):
    data = synth.config(0, device="cuda")
    print("case C0 full", data.n, data.l, data.m, data.t, flush=True)
    o = oracle.scan(data.XL, data.xr_matrix(), data.y_matrix(), data.r, threads=16)
    g = scan_device(data, "fp64", dense=True)
    print("C0 fp64", compare_fp64(g, o), g.stats, flush=True)
    for prec in ("fp32", "bf16x3"):
        gm = scan_device(data, prec, dense=True)
        db, dl = precision_report(gm, o)
        print("C0", prec, "max|db|/SE %.2e max|d-log10p| %.2e" % (db, dl), gm.stats["variant"], flush=True)
        assert db <= MIXED_TOL and dl <= MIXED_TOL
        assert strong_set(gm.dense) == strong_set(o)


# ---------------------------------------------------------------------------
# C1, C2: full-size scans (one chunk) compared on a SNP window, and the same
# generators at m = 256 (every planted effect inside) compared densely

WINDOW = 256
THR_FULL = dict(min_p_store=1e-3, min_p_disp=1e-3, min_r2_store=-10.0)


@pytest.mark.parametrize("cfg", [1, 2])
def test_full_size_scan_window(cfg, oracle):
    data = synth.config(cfg, device="cuda")
    print("case C%d full size" % cfg, data.n, data.l, data.r, data.m, data.t, flush=True)
    o = oracle.scan(data.XL, data.xr_matrix(np.arange(WINDOW)), data.y_matrix(), data.r, threads=16,
                    **THR_FULL)
    rec = oracle_records(o)
    g = scan_device(data, "fp64", **THR_FULL)
    assert_fp64_records(window(g.hits, 0, WINDOW), rec)
    print("C%d fp64: %d hits in the window, %d in all" % (cfg, len(rec["trait"]), len(g.hits)), flush=True)
    for prec in ("fp32", "bf16x3"):
        gm = scan_device(data, prec, **THR_FULL)
        print("C%d %s: %d chunk(s)" % (cfg, prec, gm.stats["n_chunks"]), flush=True)
        assert_mixed_records(window(gm.hits, 0, WINDOW), rec, "C%d %s (%s)" % (cfg, prec, gm.stats["variant"]))
        # the whole scan against the FP64 path: strong hits identical
        strong = lambda h: set(keys(h[h["p"] < STRONG_P]))
        assert strong(gm.hits) == strong(g.hits)


@pytest.mark.parametrize("cfg", [1, 2])
def test_config_generator_dense(cfg, oracle):
    data = synth.config(cfg, device="cuda", m=WINDOW)
    print("case C%d generator at m=%d" % (cfg, WINDOW), flush=True)
    thr = dict(min_r2_store=-10.0)
    o = oracle.scan(data.XL, data.xr_matrix(), data.y_matrix(), data.r, threads=16, **thr)
    g = scan_device(data, "fp64", dense=True, **thr)
    print("C%d fp64" % cfg, compare_fp64(g, o), flush=True)
    assert len(strong_set(o)) > 0
    for prec in ("fp32", "bf16x3"):
        gm = scan_device(data, prec, dense=True, **thr)
        db, dl = precision_report(gm, o)
        print("C%d %s max|db|/SE %.2e max|d-log10p| %.2e strong %d" % (cfg, prec, db, dl, len(strong_set(o))),
              flush=True)
        assert db <= MIXED_TOL and dl <= MIXED_TOL
        assert strong_set(gm.dense) == strong_set(o)


# ---------------------------------------------------------------------------
# C3, C4: reduced float16 panel through the request server

PANEL_SNPS = 16384
CHECK_SNPS = 48
CHECK_TRAITS = 96
THR_SERVE = dict(min_p_store=1e-2, min_p_disp=1e-2, min_r2_store=-10.0)


@pytest.mark.parametrize("cfg", [3, 4])
def test_served_requests_match_oracle(cfg, oracle):
    specs = synth.job_specs(cfg)[:6]
    n = synth.SERVE_CONFIGS[cfg]["n"]
    G = synth.panel(cfg, m=PANEL_SNPS)
    print("case C%d served" % cfg, [(s.t, s.l, None if s.keep is None else int(s.keep.sum())) for s in specs],
          flush=True)
    arrays = [s.arrays(n) for s in specs]
    XRw = G[:CHECK_SNPS].double().cpu().numpy().T            # (n, CHECK_SNPS)
    panel = Panel(G)
    answers = {}
    for prec in ("fp64", "bf16x3"):
        with Server(panel, prec, chunk_snps=4096) as srv:    # 4 chunks
            ids = [srv.submit(Y, XL, s.keep, max_hits=1 << 22, **THR_SERVE) for s, (Y, XL) in zip(specs, arrays)]
            answers[prec] = [srv.wait(i) for i in ids]
    panel.close()
    for j, (s, (Y, XL)) in enumerate(zip(specs, arrays)):
        rows = np.ones(n, dtype=bool) if s.keep is None else s.keep.astype(bool)
        tw = min(s.t, CHECK_TRAITS)
        o = oracle.scan(XL[rows], XRw[rows], Y[rows][:, :tw], 1, threads=16, **THR_SERVE)
        rec = oracle_records(o)
        a64 = answers["fp64"][j]
        assert a64.stats["n_chunks"] == PANEL_SNPS // 4096
        assert_fp64_records(window(a64.hits, 0, CHECK_SNPS, tw), rec)
        assert_mixed_records(window(answers["bf16x3"][j].hits, 0, CHECK_SNPS, tw), rec,
                             "C%d request %d (t=%d, l=%d, rows=%d)" % (cfg, j, s.t, s.l, int(rows.sum())))
