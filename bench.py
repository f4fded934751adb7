#!/usr/bin/env python3
"""Benchmark of the MI355X scan: SNP x trait tests/s on one GPU.

Workload (default): the n=10,000 interaction configuration -- l = 4
covariates + intercept, r = 2 columns per SNP [E*G, G], m = 200,000 SNPs,
t = 2,000 traits with 8 structured missingness groups (2% rows blanked for
each block of 250 traits), seeded synthetic data generated on the GPU.
A step is one full scan of that workload: trait block preparation
(missing-row patterns, XL'y), every SNP chunk (operand prep, contraction,
Gram pieces, missing-row corrections, per-pair solve and statistics) and
the compacted hit list copied back and sorted on the host.  Inputs are
resident in HBM when the timed region starts.

One JSON line is printed: the mixed-precision (split-BF16 x3) value, the
FP64 path beside it, the roofline of the dominant kernel (the contraction),
the FP32 restatement of the reference's blocked CPU algorithm timed on a
bounded sample of the same workload (`cpu_baseline`), the timed runs' hits
on a sample of SNPs checked against the FP64 oracle (`oracle_check`; the
bench exits with status 3 when they disagree), and
the latency half of the metric: p50/p99 of a bounded request stream
(config-4 requests on a 100,000-SNP float16 panel at 20 requests/s).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SNPxtrait tests/sec, mixed vs FP64 at max |Δ-log10p|≤1e-3; p50/p99 latency"
PEAK = {"fp64": 78.6, "fp32": 157.3, "bf16x3": 2500.0}   # TFLOP/s, MI355X dense MFMA
# MFMA products per algorithmic multiply-add of the contraction
MFMA_PRODUCTS = {"bf16x3": 3, "fp32": 1, "fp64": 1}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=2,
                    help="0-2: one scan per step; 3: 64 batched uneven jobs; 4: Poisson request stream")
    ap.add_argument("--jobs", type=int, default=0, help="requests of configs 3/4 (0 = the config's count)")
    ap.add_argument("--chunk", type=int, default=-1, help="server chunk (SNPs), configs 3/4 (-1: chosen by timing)")
    ap.add_argument("--budget-gib", type=float, default=160.0, help="server HBM budget, configs 3/4")
    ap.add_argument("--graphs", action="store_true",
                    help="configs 3/4: replay each chunk's launches from a captured HIP graph")
    ap.add_argument("--fp64-jobs", type=int, default=-1,
                    help="requests of configs 3/4 rerun one at a time in FP64 (-1 = all)")
    ap.add_argument("--precision", default="bf16x3", choices=["fp64", "fp32", "bf16x3"])
    ap.add_argument("--fp64-steps", type=int, default=1, help="timed FP64 steps for the ratio (0 = skip)")
    ap.add_argument("--m", type=int, default=0, help="override SNP count (testing only)")
    ap.add_argument("--t", type=int, default=0, help="override trait count (testing only)")
    ap.add_argument("--p-store", type=float, default=1e-6)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--serve-leg-m", type=int, default=100_000, metavar="SNPS",
                    help="configs 0-2: panel SNPs of the bounded request-stream leg (0 = skip it)")
    ap.add_argument("--serve-leg-jobs", type=int, default=32)
    ap.add_argument("--serve-leg-rate", type=float, default=20.0, help="offered requests/s of that leg")
    ap.add_argument("--serve-leg-chunk", type=int, default=-1, help="server chunk of that leg (-1: chosen by timing)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--oracle-seconds", type=float, default=5.0,
                    help="FP64 oracle work behind the hit check (planted SNPs, then the first ones)")
    ap.add_argument("--host-fed", type=int, default=0, metavar="BLOCKS",
                    help="also time the scan with genotypes in page-locked host memory, uploaded in BLOCKS "
                         "blocks, each upload overlapping the previous block's scan (configs 0-2)")
    ap.add_argument("--dosage", type=int, default=0, metavar="SNPS",
                    help="also time the device-side dosage combine (additive model, float16 probability "
                         "columns P(AA), P(AB), P(BB) of SNPS SNPs at the workload's n)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run one 1 GiB device copy first (calibrates FETCH_SIZE/WRITE_SIZE counters)")
    return ap.parse_args()


def make_scanner(data, precision, p_store, max_hits=1 << 22):
    from omicabelnomm_amd import Scanner
    # minR2store = -10 (the CLI's documented lower end): with missing
    # phenotypes the reference's R2 is negative and its default keeps nothing.
    sc = Scanner(data.n, data.l, data.r, precision, min_p_store=p_store, min_p_disp=p_store,
                 min_r2_store=-10.0, max_traits=data.t, max_hits=max_hits)
    sc.set_covariates(data.XL)
    if data.E is not None:
        sc.set_interactions(data.E)
    return sc


def one_step(sc, data):
    sc.load_traits(data.Y)
    sc.load_genotypes(data.G, m=data.m, col_src=data.col_src, col_env=data.col_env)
    return sc.scan()


def timed(sc, data, steps, warmup):
    import torch
    for _ in range(warmup):
        one_step(sc, data)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = None
    gemm_ms = 0.0
    chunks = 0
    for _ in range(steps):
        res = one_step(sc, data)
        gemm_ms += res.timings_ms["gemm"]
        chunks += res.stats["n_chunks"]
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    return dt, res, gemm_ms / max(chunks, 1), chunks / steps


def dosage_leg(n, snps, reps=5):
    """The dosage combine alone: float16 probability columns (3 per SNP)
    already on the device, additive model 2*P(AA) + P(AB), one FP64 column
    per SNP.  Algorithmic bytes per element of the output: two float16
    reads and one FP64 write (12 B)."""
    import torch
    from omicabelnomm_amd import Scanner
    gen = torch.Generator(device="cuda")
    gen.manual_seed(0x0D05)
    P = torch.rand((snps * 3, n), generator=gen, device="cuda", dtype=torch.float32).to(torch.float16)
    XL = np.column_stack([np.ones(n), np.random.default_rng(5).standard_normal((n, 2))])
    with Scanner(n, 3, 1, "bf16x3", max_traits=8) as sc:
        sc.set_covariates(XL)
        sc.set_dosage_model([2.0, 1.0], 3, sum=True)
        sc.load_genotypes(P)          # allocates the combined block
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            sc.load_genotypes(P)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / reps
    nbytes = 12.0 * n * snps
    return {"snps": snps, "n": n, "ms_per_block": dt * 1e3, "algorithmic_gb": nbytes / 1e9,
            "gb_per_s": nbytes / dt / 1e9, "frac_of_hbm_peak": nbytes / dt / 8e12}


def host_fed(data, precision, p_store, blocks, steps):
    """The scan with the genotypes in page-locked host memory, as the drop-in
    driver streams DatABEL blocks: each block is uploaded while the previous
    block's scan runs (oagpu_scan_async), so the rate is bounded by PCIe or
    by the scan, whichever is slower."""
    import torch
    m, r = data.m, data.r
    mb = (m + blocks - 1) // blocks
    parts = []
    for s0 in range(0, m, mb):
        s1 = min(m, s0 + mb)
        if data.col_src is None:
            lo, hi, src, env = s0 * r, s1 * r, None, None
        else:
            src = np.asarray(data.col_src[s0 * r:s1 * r])
            lo, hi = int(src.min()), int(src.max()) + 1
            assert hi - lo <= (s1 - s0) * r
            src = src - lo
            env = None if data.col_env is None else np.asarray(data.col_env[s0 * r:s1 * r])
        hb = torch.empty((hi - lo, data.n), dtype=data.G.dtype, pin_memory=True)
        hb.copy_(data.G[lo:hi])
        parts.append((hb, s1 - s0, src, env, s0))
    # PCIe rate of one block, for the record
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dev = parts[0][0].to("cuda", non_blocking=True)
    torch.cuda.synchronize()
    h2d = parts[0][0].numel() * parts[0][0].element_size() / (time.perf_counter() - t0) / 1e9
    del dev
    sc = make_scanner(data, precision, p_store)

    def step():
        sc.load_traits(data.Y)
        hb, mcnt, src, env, s0 = parts[0]
        sc.load_genotypes(hb.numpy().T, m=mcnt, col_src=src, col_env=env, snp_offset=s0)
        sc.scan_async()
        for hb, mcnt, src, env, s0 in parts[1:]:
            sc.load_genotypes(hb.numpy().T, m=mcnt, col_src=src, col_env=env, snp_offset=s0)
            sc.scan_wait()
            sc.scan_async()
        sc.scan_wait()

    step()   # warm-up (autotuner, buffers)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    sc.close()
    return {"value": m * data.t / dt, "ms_per_step": dt * 1e3, "blocks": len(parts),
            "block_gb": round(parts[0][0].numel() * parts[0][0].element_size() / 1e9, 3),
            "h2d_gb_per_s": round(h2d, 1)}


def measured_traffic(variant, config):
    """HBM bytes per contraction from the committed rocprofv3 counter
    summary of this kernel and workload, or None."""
    path = os.path.join(ROOT, "profiles", "c%d_contraction_traffic.json" % config)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f)
    if rec.get("kernel") != variant:
        return None
    per_launch = 1024.0 * (rec["fetch_kib_per_launch"] * rec["fetch_scale"] +
                           rec["write_kib_per_launch"] * rec["write_scale"])
    return per_launch * rec["launches_per_contraction"]


def cpu_model():
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(data, budget_s, p_store):
    """The FP32 restatement of the reference's blocked CPU algorithm
    (oracle/ref32.cpp: per block of 1,000 traits x 500 SNPs the sgemm/ssyrk
    products, per pair build_S, the missing-row corrections and sposv, the
    explicit RSS over the n rows, hpc_statistics) on the first m' SNPs x all
    traits, m' sized to about budget_s of work."""
    from oracle import oracle as O
    threads = min(16, os.cpu_count() or 1)   # the host's CPU share beside one GPU
    XL = data.XL
    Y = data.y_matrix()
    thr = dict(min_p_store=p_store, min_p_disp=p_store, min_r2_store=-10.0)

    def run(ms):
        XR = data.xr_matrix(np.arange(ms))
        t0 = time.perf_counter()
        O.scan_ref32(XL, XR, Y, data.r, threads=threads, **thr)
        return time.perf_counter() - t0

    probe = max(1, min(2 * threads, data.m))
    dt = run(probe)
    ms = int(min(data.m, max(probe, probe * budget_s / max(dt, 1e-3))))
    dt = run(ms)
    if dt < 0.6 * budget_s and ms < data.m:
        # the probe's fixed costs (covariate products, thread start-up) made
        # the first estimate short: size once more from the longer run
        ms = int(min(data.m, ms * budget_s / max(dt, 1e-3)))
        dt = run(ms)
    return {"value": ms * data.t / dt, "unit": "SNPxtrait tests/s", "cores": threads, "kind": "port",
            "sample": "first %d SNPs x %d traits of the same workload" % (ms, data.t),
            "seconds": round(dt, 3), "precision": "fp32",
            "structure": "the reference's blocked algorithm (tb=1000, mb=500): XL'XL, XL'Y, XL'XR, XR'Y "
                         "products, per-pair build_S + missing-row corrections + Cholesky solve, explicit "
                         "RSS over the n rows, statistics",
            "blas": O.ref32_blas(),
            "nproc": os.cpu_count(), "cpu_model": cpu_model()}


def oracle_sample(data, budget_s, p_store):
    """The FP64 restatement (oracle) on the SNPs with planted effects, then
    the first SNPs, x all traits, about budget_s of work: (SNP indices,
    dense result)."""
    from oracle import oracle as O
    threads = min(16, os.cpu_count() or 1)
    XL = data.XL
    Y = data.y_matrix()
    planted = np.unique(data.planted[:, 0]) if data.planted is not None else np.zeros(0, dtype=np.int64)
    thr = dict(min_p_store=p_store, min_p_disp=p_store, min_r2_store=-10.0)

    def run(snps):
        t0 = time.perf_counter()
        o = O.scan(XL, data.xr_matrix(snps), Y, data.r, threads=threads, **thr)
        return time.perf_counter() - t0, o

    probe = max(1, min(threads, data.m))
    dt, _ = run(np.arange(probe))
    ms = int(min(data.m, max(probe, probe * budget_s / max(dt, 1e-3))))
    rest = np.setdiff1d(np.arange(min(data.m, ms + len(planted))), planted)
    snps = np.sort(np.concatenate([planted, rest])[:max(ms, len(planted))])
    _, o = run(snps)
    return snps, o


def oracle_check(o, snps, hits, fp64_hits):
    """The timed runs' hits on the oracle's SNPs against the oracle: the FP64
    path record for record (beta and SE within 1e-10 relative), the mixed
    path on the records both keep (max |dbeta|/SE, max |d -log10 p|, both
    <= 1e-3) and its p < 1e-8 set identical."""
    kept = o["kept"].astype(bool) & o["emitted"].astype(bool)[:, :, None]
    y, s, h = np.nonzero(kept)
    rec = {(int(a), int(snps[b]), int(c)): (o["beta"][a, b, c], o["se"][a, b, c], o["pval"][a, b, c])
           for a, b, c in zip(y, s, h)}

    def window(hh):
        return hh[np.isin(hh["snp"], snps)]

    def mlog10(p):
        return -np.log10(max(float(p), 1e-300))

    out = {"vs": "FP64 oracle (CPU) on the %d sampled SNPs x all traits, hits stored at the bench thresholds"
                 % len(snps),
           "n_oracle_records": len(rec)}
    ok = True
    if fp64_hits is not None:
        w = window(fp64_hits)
        keys = {(int(x["trait"]), int(x["snp"]), int(x["h"])) for x in w}
        same = keys == set(rec)
        rb = max((abs(x["b"] - rec[k][0]) / max(abs(rec[k][0]), rec[k][1])
                  for x in w for k in [(int(x["trait"]), int(x["snp"]), int(x["h"]))] if k in rec), default=0.0)
        out["fp64_records_identical"] = same
        out["fp64_max_rel_beta"] = float(rb)
        ok = ok and same and rb <= 1e-10
    w = window(hits)
    db = dl = 0.0
    for x in w:
        k = (int(x["trait"]), int(x["snp"]), int(x["h"]))
        if k in rec:
            b, se, pv = rec[k]
            db = max(db, abs(x["b"] - b) / se)
            dl = max(dl, abs(mlog10(x["p"]) - mlog10(pv)))
    strong_g = {(int(x["trait"]), int(x["snp"]), int(x["h"])) for x in w if x["p"] < 1e-8}
    strong_o = {k for k, v in rec.items() if v[2] < 1e-8}
    out.update({"max_dbeta_over_se": float(db), "max_abs_dlog10p": float(dl),
                "hits_p_lt_1e-8_identical": strong_g == strong_o, "n_hits_p_lt_1e-8": len(strong_o)})
    ok = ok and db <= 1e-3 and dl <= 1e-3 and strong_g == strong_o
    out["ok"] = bool(ok)
    return out


def percentile(xs, q):
    return float(np.percentile(np.asarray(xs, dtype=np.float64), q)) if xs else None


def serve_leg(m, jobs, rate, p_store, precision, chunk=-1):
    """The latency half of the metric beside the throughput headline: the
    request workload of config 4 (float16 panel of 20,000 individuals,
    requests with 1-8,192 traits, 2-16 covariates, MCAR 0-5 %) on a panel of
    m SNPs, its first `jobs` requests replayed as a Poisson stream at `rate`
    requests/s, latency = submission to answer."""
    import threading
    import torch
    from omicabelnomm_amd import synth
    from omicabelnomm_amd.serve import Panel, Server
    n = synth.SERVE_CONFIGS[4]["n"]
    specs = synth.job_specs(4, jobs=jobs)
    G = synth.panel(4, m=m)
    torch.cuda.synchronize()
    arrays = [sp.arrays(n) for sp in specs]
    thr = dict(min_p_store=p_store, min_p_disp=p_store, min_r2_store=-10.0)
    panel = Panel(G)
    # the specs' arrivals are a unit-rate Poisson process: scale to `rate`
    base = synth.SERVE_CONFIGS[4]["rate"]
    arrivals = [sp.arrival * base / rate for sp in specs]
    with Server(panel, precision, chunk_snps=chunk) as srv:
        warm = synth.JobSpec(t=64, l=3, keep=None, miss_rate=0.02, arrival=0.0, seed=1)
        srv.wait(srv.submit(*warm.arrays(n), **thr))   # autotuner, code objects
        ids = [None] * len(specs)
        start = time.perf_counter()

        def submitter():
            for i, sp in enumerate(specs):
                delay = start + arrivals[i] - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                ids[i] = srv.submit(arrays[i][0], arrays[i][1], sp.keep, max_hits=1 << 22, **thr)

        th = threading.Thread(target=submitter)
        th.start()
        th.join()
        answers = [srv.wait(i) for i in ids]
        wall = time.perf_counter() - start
        chunk_used = srv.chunk_snps()
    panel.close()
    del G
    lat = [a.latency_ms for a in answers]
    tests = sum(sp.t for sp in specs) * m
    return {"workload": "config-4 requests on a float16 panel of %d individuals x %d SNPs resident in HBM"
                        % (n, m),
            "requests": len(specs), "traits": sum(sp.t for sp in specs), "offered_rate_per_s": rate,
            "seconds": round(wall, 3), "tests_per_s": tests / wall,
            "p50_ms": percentile(lat, 50), "p99_ms": percentile(lat, 99), "max_ms": max(lat),
            "precision": precision, "chunk_snps": chunk_used, "chunk_timed": chunk < 0}


def bench_serve(args):
    """Configs 3 and 4: a float16 panel resident in HBM and requests answered
    by the server (continuous batching).  value = tests/s with every request
    submitted at once (the server's capacity); for config 4 the same
    requests are then replayed as a Poisson stream for the latency figures;
    fp64 = the same requests run one at a time as plain FP64 scans."""
    import threading
    import torch
    from omicabelnomm_amd import Scanner, synth
    from omicabelnomm_amd.serve import Panel, Server
    cfg = synth.SERVE_CONFIGS[args.config]
    n = cfg["n"]
    m = args.m or cfg["m"]
    thr = dict(min_p_store=args.p_store, min_p_disp=args.p_store, min_r2_store=-10.0)
    specs = synth.job_specs(args.config, jobs=args.jobs or None)
    t0 = time.perf_counter()
    G = synth.panel(args.config, m=m)
    torch.cuda.synchronize()
    gen_s = time.perf_counter() - t0
    arrays = []
    for i, sp in enumerate(specs):
        arrays.append(sp.arrays(n))
        if i % 16 == 15:
            print("generated %d/%d requests" % (i + 1, len(specs)), file=sys.stderr, flush=True)
    tests = sum(sp.t for sp in specs) * m
    panel = Panel(G)

    chunk_used = [args.chunk]

    def serve(arrivals):
        srv = Server(panel, args.precision, hbm_budget=int(args.budget_gib * 2**30), max_active=64,
                     chunk_snps=args.chunk, graphs=args.graphs)
        warm = synth.JobSpec(t=64, l=3, keep=None, miss_rate=0.02, arrival=0.0, seed=1)
        srv.wait(srv.submit(*warm.arrays(n), **thr))   # autotuner, code objects
        ids = [None] * len(specs)
        start = time.perf_counter()

        def submitter():
            for i, sp in enumerate(specs):
                delay = start + arrivals[i] - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                ids[i] = srv.submit(arrays[i][0], arrays[i][1], sp.keep, max_hits=1 << 22, **thr)

        th = threading.Thread(target=submitter)
        th.start()
        th.join()
        answers = [srv.wait(i) for i in ids]
        wall = time.perf_counter() - start
        replayed = srv.graph_chunks()
        chunk_used[0] = srv.chunk_snps()
        srv.close()
        print("chunks replayed from graphs: %d" % replayed, file=sys.stderr, flush=True)
        return answers, wall

    answers, wall = serve([0.0] * len(specs))
    value = tests / wall
    lat = [a.latency_ms for a in answers]
    print("served %d requests at once in %.2f s" % (len(specs), wall), file=sys.stderr, flush=True)
    stream = None
    if cfg["rate"]:
        # the same requests as a Poisson stream, at the configured rate or 80 %
        # of the measured capacity, whichever is lower
        rate = min(cfg["rate"], 0.8 * len(specs) / wall)
        scale = cfg["rate"] / rate
        s_answers, s_wall = serve([sp.arrival * scale for sp in specs])
        s_lat = [a.latency_ms for a in s_answers]
        stream = {"rate_per_s": rate, "seconds": s_wall, "tests_per_s": tests / s_wall,
                  "p50_ms": percentile(s_lat, 50), "p99_ms": percentile(s_lat, 99), "max_ms": max(s_lat)}
        print("stream at %.2f/s: p50 %.0f ms p99 %.0f ms" % (rate, stream["p50_ms"], stream["p99_ms"]),
              file=sys.stderr, flush=True)

    # the same requests one at a time, plain FP64 scans of the panel
    fp64 = prec = None
    k = len(specs) if args.fp64_jobs < 0 else min(args.fp64_jobs, len(specs))
    if k > 0:
        dt64, tests64, dmax, same, nh = 0.0, 0, 0.0, True, 0
        for sp, (Y, XL), a in zip(specs[:k], arrays[:k], answers[:k]):
            with Scanner(n, sp.l, 1, "fp64", max_traits=sp.t, max_hits=1 << 22, **thr) as sc:
                if sp.keep is not None:
                    sc.set_subset(sp.keep)
                sc.set_covariates(XL)
                sc.load_traits(Y)
                sc.load_genotypes(G)
                torch.cuda.synchronize()
                t1 = time.perf_counter()
                res = sc.scan()
                dt64 += time.perf_counter() - t1
            tests64 += sp.t * m
            print("fp64 request t=%d, %.2f s so far" % (sp.t, dt64), file=sys.stderr, flush=True)
            keys = lambda h: list(zip(h["trait"].tolist(), h["snp"].tolist(), h["h"].tolist()))
            strong = lambda h: set(keys(h[h["p"] < 1e-8]))
            same = same and strong(a.hits) == strong(res.hits)
            where = {kk: i for i, kk in enumerate(keys(res.hits))}
            for j, kk in enumerate(keys(a.hits)):
                i = where.get(kk)
                if i is not None and a.hits["p"][j] > 0 and res.hits["p"][i] > 0:
                    dmax = max(dmax, abs(np.log10(a.hits["p"][j]) - np.log10(res.hits["p"][i])))
            nh += int((res.hits["p"] < 1e-8).sum())
        print("fp64 one at a time: %d requests in %.2f s" % (k, dt64), file=sys.stderr, flush=True)
        fp64 = {"value": tests64 / dt64, "mode": "one request at a time, plain FP64 scan of the resident panel",
                "requests": k, "seconds": round(dt64, 3)}
        prec = {"vs": "fp64 one-at-a-time answers of %d requests (hits stored at p<%g)" % (k, args.p_store),
                "hits_p_lt_1e-8_identical": same, "n_hits_p_lt_1e-8": nh, "max_abs_dlog10p": dmax}
    panel.close()
    widths = [sp.t for sp in specs]
    line = {
        "metric": METRIC,
        "value": value, "unit": "SNPxtrait tests/s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": wall * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"bf16x3": "bf16x3(fp32 acc)+f64", "fp32": "f32+f64", "fp64": "f64"}[args.precision],
        "data": "synthetic (seeded; float16 dosage panel generated on the GPU, host traits)",
        "config": {"workload": "C%d: panel n=%d x m=%d float16 resident; %d requests, traits %d..%d (sum %d), "
                               "%s, MCAR 0-5%%, covariates 2-16 + intercept"
                               % (args.config, n, m, len(specs), min(widths), max(widths), sum(widths),
                                  "individual subsets %d-%d" % cfg["subset"] if cfg["subset"] else
                                  "all individuals"),
                   "thresholds": {"minPstore": args.p_store, "minPdisp": args.p_store, "minR2store": -10.0},
                   "precision": args.precision, "chunk_snps": chunk_used[0], "chunk_timed": args.chunk < 0, "gen_seconds": round(gen_s, 2),
                   "graphs": args.graphs,
                   "value_is": "all requests submitted at once: total tests / time to the last answer"},
        "latency_ms": {"p50": percentile(lat, 50), "p99": percentile(lat, 99), "max": max(lat),
                       "note": "submitted together"},
        "stream": stream,
        "roofline": None,
        "cpu_baseline": None,
        "fp64": fp64,
        "speedup_vs_fp64": (value / fp64["value"]) if fp64 else None,
        "precision_check": prec,
    }
    print(json.dumps(line))


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the scan's multi-GPU split (SNP-range "
              "shards, no collectives) is described, not implemented.",
              file=sys.stderr)
        sys.exit(2)
    import torch
    from omicabelnomm_amd import synth
    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.config in (3, 4):
        bench_serve(args)
        return
    if args.copy_probe:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    over = {}
    if args.m:
        over["m"] = args.m
    if args.t:
        over["t"] = args.t
    t0 = time.perf_counter()
    data = synth.config(args.config, device="cuda", **over)
    torch.cuda.synchronize()
    gen_s = time.perf_counter() - t0
    n, l, r, m, t = data.n, data.l, data.r, data.m, data.t
    tests = m * t
    flops = 2.0 * n * r * m * (t + l)   # algorithmic flops of the contraction X'[Y | XL]

    sc = make_scanner(data, args.precision, args.p_store)
    dt, res, gemm_avg_ms, chunks = timed(sc, data, args.steps, args.warmup)
    value = tests / dt
    gemm_tflops = (flops / chunks) / (gemm_avg_ms * 1e-3) / 1e12
    hits_mixed = res.hits
    timings = res.timings_ms
    variant = sc.gemm_variant()
    sc.close()

    fp64 = None
    prec = None
    hits64 = None
    if args.fp64_steps > 0 and args.precision != "fp64":
        sc = make_scanner(data, "fp64", args.p_store)
        dt64, res64, gemm64_ms, chunks64 = timed(sc, data, args.fp64_steps, 1)
        variant64 = sc.gemm_variant()
        sc.close()
        fp64 = {"value": tests / dt64, "ms_per_step": dt64 * 1e3,
                "gemm_tflops": (flops / chunks64) / (gemm64_ms * 1e-3) / 1e12, "gemm_variant": variant64,
                "timings_ms": res64.timings_ms}
        hits64 = res64.hits
        # hit set at p < 1e-8 and the -log10 p / beta agreement on shared hits
        key = lambda h: set(zip(h["trait"][h["p"] < 1e-8], h["snp"][h["p"] < 1e-8], h["h"][h["p"] < 1e-8]))
        a, b = key(hits_mixed), key(res64.hits)
        idx64 = {(x["trait"], x["snp"], x["h"]): x for x in res64.hits}
        dl, db = 0.0, 0.0
        for x in hits_mixed:
            y = idx64.get((x["trait"], x["snp"], x["h"]))
            if y is not None:
                dl = max(dl, abs(np.log10(max(x["p"], 1e-300)) - np.log10(max(y["p"], 1e-300))))
                db = max(db, abs(x["b"] - y["b"]) / y["se"])
        prec = {"vs": "fp64 GPU path (hits stored at p<%g)" % args.p_store,
                "hits_p_lt_1e-8_identical": a == b, "n_hits_p_lt_1e-8": len(b),
                "max_abs_dlog10p": dl, "max_dbeta_over_se": db}

    hostfed = host_fed(data, args.precision, args.p_store, args.host_fed, args.steps) if args.host_fed else None
    serving = None
    if args.serve_leg_m > 0:
        serving = serve_leg(args.serve_leg_m, args.serve_leg_jobs, args.serve_leg_rate, args.p_store,
                            args.precision, args.serve_leg_chunk)
        print("request stream: p50 %.0f ms p99 %.0f ms" % (serving["p50_ms"], serving["p99_ms"]),
              file=sys.stderr, flush=True)
    dosage = dosage_leg(data.n, args.dosage) if args.dosage else None

    cpu = None
    ocheck = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(data, args.cpu_seconds, args.p_store)
        snps_o, o = oracle_sample(data, args.oracle_seconds, args.p_store)
        if args.precision == "fp64":
            hits64, hits_mixed = hits_mixed, hits_mixed
        ocheck = oracle_check(o, snps_o, hits_mixed, hits64)

    peak = PEAK[args.precision]
    line = {
        "metric": METRIC, "value": value, "unit": "SNPxtrait tests/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"bf16x3": "bf16x3(fp32 acc)+f64", "fp32": "f32+f64", "fp64": "f64"}[args.precision],
        "data": "synthetic (seeded, generated on the GPU)",
        "config": {"workload": "C%d: n=%d, l=%d, r=%d%s, m=%d, t=%d, %s" % (
            args.config, n, l, r, " ([E*G, G] interaction)" if data.E is not None else "", m, t,
            synth.CONFIGS[args.config].get("miss", "")),
            "thresholds": {"minPstore": args.p_store, "minPdisp": args.p_store, "minR2store": -10.0},
            "precision": args.precision, "chunks_per_step": chunks,
            "gen_seconds": round(gen_s, 2)},
        "roofline": {"bound": "mfma", "achieved": gemm_tflops, "peak": peak, "unit": "TFLOP/s",
                     "frac": gemm_tflops / peak, "traffic": None if over else measured_traffic(variant, args.config),
                     "kernel": variant,
                     "mfma_products": MFMA_PRODUCTS[args.precision],
                     "hardware_frac": gemm_tflops * MFMA_PRODUCTS[args.precision] / peak,
                     "note": "algorithmic 2*n*r*m*(t+l) flops / HIP-event time of the contraction; "
                             "hardware_frac counts the MFMA products issued per algorithmic multiply-add "
                             "(3 for split-BF16: lo*hi, hi*lo, hi*hi); "
                             "traffic = HBM bytes per contraction (rocprofv3 FETCH_SIZE, WRITE_SIZE, "
                             "copy-calibrated) when a profile of this kernel is committed"},
        "cpu_baseline": cpu,
        "fp64": fp64,
        "speedup_vs_fp64": (value / fp64["value"]) if fp64 else None,
        "precision_check": prec,
        "oracle_check": ocheck,
        "latency": serving,
        "timings_ms": timings,
        "stats": {k: int(v) if isinstance(v, (int, np.integer)) else v for k, v in res.stats.items()},
    }
    if hostfed:
        line["host_fed"] = hostfed
    if dosage:
        line["dosage_combine"] = dosage
    print(json.dumps(line))
    if ocheck is not None and not ocheck["ok"]:
        print("oracle check failed: %s" % json.dumps(ocheck), file=sys.stderr)
        sys.exit(3)


if __name__ == "__main__":
    main()
