"""Individual subsets and the request server against the oracle and the
plain scan.

- a subset scan (oagpu_set_subset) equals the oracle run on the subset's
  rows alone;
- every request the server answers equals the same scan run on its own
  context, for requests with shared and independent missing patterns and
  with subsets, over ragged chunks;
- a request's answer is the same alone and in a mix of concurrent requests
  (bit-identical in FP64; split-BF16 within 1e-4 SE on beta, as the mixed
  path is held to against FP64);
- a budget that admits one request at a time still answers every request;
- a panel with missing genotypes (mean-imputed, compact rows from the
  operand pass) answers as the plain scan does.
"""
import numpy as np
import pytest
import torch

from omicabelnomm_amd import Scanner, synth
from omicabelnomm_amd.serve import Panel, Server
from parity_util import compare_fp64, run_oracle

pytestmark = pytest.mark.gpu

THR = dict(min_p_store=0.2, min_p_disp=0.05, min_r2_store=-10.0)


def hit_key(h):
    return np.stack([h["trait"], h["snp"], h["h"], h["n_used"]], axis=1)


def assert_hits_equal(a, b, tol):
    """Same hit set; |db| <= tol*SE, SE and t within tol relative, R2 within
    tol, -log10 p within 1000*tol (1e-9 at least)."""
    assert len(a) == len(b), (len(a), len(b))
    np.testing.assert_array_equal(hit_key(a), hit_key(b))
    assert np.all(np.abs(a["b"] - b["b"]) <= tol * b["se"]), np.max(np.abs(a["b"] - b["b"]) / b["se"])
    np.testing.assert_allclose(a["se"], b["se"], rtol=tol, atol=0)
    np.testing.assert_allclose(a["t"], b["t"], rtol=tol, atol=tol)
    np.testing.assert_allclose(a["r2"], b["r2"], rtol=0, atol=tol)
    np.testing.assert_allclose(-np.log10(a["p"]), -np.log10(b["p"]), rtol=0, atol=max(tol * 1e3, 1e-9))


def panel_data(n=520, m=230, seed=41):
    d = synth.make(n=n, l=2, m=m, t=2, miss="none", seed=seed)
    return d.G.contiguous()   # (m, n) float32


def request_data(n, t, l, miss, seed, rate=0.06):
    rng = np.random.default_rng(1000 + seed)
    XL = np.column_stack([np.ones(n), rng.standard_normal((n, l - 1))])
    Y = rng.standard_normal((n, t))
    if miss == "mcar":
        Y[rng.random((n, t)) < rate] = np.nan
    elif miss == "groups":
        for q in range(3):
            rows = rng.choice(n, size=int(rate * n), replace=False)
            Y[np.ix_(rows, np.arange(q, t, 3))] = np.nan
    return Y, XL


def scan_alone(G, Y, XL, keep, precision, **thr):
    """The same analysis on a context of its own (the plain scan path)."""
    n, t = Y.shape
    with Scanner(n, XL.shape[1], 1, precision, max_traits=t, max_hits=1 << 20, store_all=True, **thr) as sc:
        if keep is not None:
            sc.set_subset(keep)
        sc.set_covariates(XL)
        sc.load_traits(Y)
        sc.load_genotypes(G.to("cuda"))
        return sc.scan().hits


def test_subset_scan_matches_oracle_on_the_subset(oracle):
    data = synth.make(n=420, l=4, m=35, t=28, miss="mcar", miss_rate=0.08, planted=6, seed=12)
    rng = np.random.default_rng(5)
    keep = rng.random(data.n) < 0.7
    print("case subset scan", int(keep.sum()), flush=True)
    with Scanner(data.n, data.l, 1, "fp64", max_traits=data.t, store_all=True, disp_cov=True) as sc:
        sc.set_subset(keep)
        sc.set_covariates(data.XL)
        sc.load_traits(data.Y.cpu().numpy().T)
        sc.load_genotypes(data.G.cpu().numpy().T)
        g = sc.scan(dense=True)
    XR = data.xr_matrix()[keep]
    o = oracle.scan(data.XL[keep], XR, data.y_matrix()[keep], 1, store_all=True, disp_cov=True, threads=4)
    compare_fp64(g, o)


@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
def test_server_answers_equal_plain_scans(precision):
    G = panel_data()
    n = G.shape[1]
    rng = np.random.default_rng(7)
    keep = rng.random(n) < 0.8
    reqs = [request_data(n, 30, 3, "groups", 1), request_data(n, 9, 5, "mcar", 2),
            request_data(n, 17, 2, "none", 3), request_data(n, 4, 6, "mcar", 4)]
    keeps = [None, None, keep, keep]
    print("case server", precision, flush=True)
    panel = Panel(G.to("cuda"))
    with Server(panel, precision, chunk_snps=64) as srv:   # 4 chunks, the last ragged
        ids = [srv.submit(Y, XL, k, store_all=True, **THR) for (Y, XL), k in zip(reqs, keeps)]
        answers = [srv.wait(i) for i in ids]
    for (Y, XL), k, a in zip(reqs, keeps, answers):
        want = scan_alone(G, Y, XL, k, precision, **THR)
        assert_hits_equal(a.hits, want, 1e-12 if precision == "fp64" else 1e-4)
        assert a.stats["n_chunks"] == 4 and a.latency_ms > 0
    panel.close()


def test_chunk_size_chosen_by_timing():
    """chunk_snps=-1: the server times one chunk of each candidate size with
    the first requests and keeps the fastest per SNP; the timing runs leave
    nothing in the answers, which equal plain scans of the whole panel."""
    G = panel_data(n=520, m=40_000, seed=43)
    n = G.shape[1]
    reqs = [request_data(n, 12, 3, "groups", 11), request_data(n, 5, 4, "mcar", 12),
            request_data(n, 20, 2, "none", 13)]
    print("case chunk tuning", flush=True)
    panel = Panel(G.to("cuda"))
    with Server(panel, "fp64", chunk_snps=-1) as srv:
        assert srv.chunk_snps() == 32768   # buffers for the largest candidate until the first requests arrive
        ids = [srv.submit(Y, XL, store_all=True, **THR) for Y, XL in reqs]
        answers = [srv.wait(i) for i in ids]
        chosen = srv.chunk_snps()
    panel.close()
    print("chosen chunk", chosen, flush=True)
    assert chosen in (4096, 8192, 16384, 32768)
    for (Y, XL), a in zip(reqs, answers):
        want = scan_alone(G, Y, XL, None, "fp64", **THR)
        assert_hits_equal(a.hits, want, 1e-12)
        assert a.stats["n_chunks"] == -(-40_000 // chosen)


# split-BF16: the batch changes the contraction's shape, so the autotuner may
# pick another kernel (another summation order); the bound is the one the
# mixed path is held to against FP64
@pytest.mark.parametrize("precision,rtol", [("fp64", 0.0), ("bf16x3", 1e-4)])
def test_request_alone_and_in_a_mix(precision, rtol):
    G = panel_data(seed=43)
    n = G.shape[1]
    target = request_data(n, 21, 4, "mcar", 11)
    others = [request_data(n, 40, 3, "groups", 12), request_data(n, 5, 2, "none", 13),
              request_data(n, 64, 7, "mcar", 14)]
    print("case alone/mix", precision, flush=True)
    panel = Panel(G.to("cuda"))
    with Server(panel, precision, chunk_snps=48) as srv:
        alone = srv.wait(srv.submit(*target, store_all=True, **THR))
        ids = [srv.submit(*others[0], store_all=True, **THR), srv.submit(*others[1], store_all=True, **THR)]
        tid = srv.submit(*target, store_all=True, **THR)
        ids.append(srv.submit(*others[2], store_all=True, **THR))
        mixed = srv.wait(tid)
        for i in ids:
            srv.wait(i)
    if rtol == 0.0:
        assert alone.hits.tobytes() == mixed.hits.tobytes()
    else:
        assert_hits_equal(alone.hits, mixed.hits, rtol)
    panel.close()


def test_budget_admits_one_at_a_time():
    G = panel_data(n=300, m=100, seed=44)
    n = G.shape[1]
    reqs = [request_data(n, 6 + 3 * i, 2 + i, "mcar", 20 + i) for i in range(3)]
    print("case budget", flush=True)
    panel = Panel(G.to("cuda"))
    with Server(panel, "fp64", hbm_budget=1, chunk_snps=32) as srv:
        ids = [srv.submit(Y, XL, store_all=True, **THR) for Y, XL in reqs]
        answers = [srv.wait(i) for i in ids]
    for (Y, XL), a in zip(reqs, answers):
        assert_hits_equal(a.hits, scan_alone(G, Y, XL, None, "fp64", **THR), 1e-12)
    # one at a time: each finished before the next started, so latencies grow
    assert answers[0].latency_ms < answers[2].latency_ms
    panel.close()


def test_small_request_passes_one_over_budget():
    """Admission is first fit: while a long request runs, one that needs more
    than the budget beside it waits, and a small one submitted after it is
    admitted at once (it finishes first); the large one then runs alone."""
    G = panel_data(n=300, m=20_000, seed=45)
    n = G.shape[1]
    long_req = request_data(n, 8, 3, "mcar", 30)
    big_req = request_data(n, 6, 2, "mcar", 31)
    small_req = request_data(n, 5, 2, "mcar", 32)
    print("case first fit", flush=True)
    panel = Panel(G.to("cuda"))
    # the big request's hit arena (2^24 records of 56 bytes) exceeds the
    # budget by itself; the others need a few MB
    with Server(panel, "fp64", hbm_budget=600 << 20, chunk_snps=512) as srv:
        a = srv.submit(*long_req, store_all=True, max_hits=1 << 18, **THR)
        b = srv.submit(*big_req, store_all=True, max_hits=1 << 24, **THR)
        c = srv.submit(*small_req, store_all=True, max_hits=1 << 18, **THR)
        ans = {k: srv.wait(i) for k, i in (("a", a), ("b", b), ("c", c))}
    for k, (Y, XL) in (("a", long_req), ("b", big_req), ("c", small_req)):
        assert_hits_equal(ans[k].hits, scan_alone(G, Y, XL, None, "fp64", **THR), 1e-12)
    print("latency ms", {k: round(v.latency_ms, 1) for k, v in ans.items()}, flush=True)
    assert ans["c"].latency_ms < ans["b"].latency_ms
    panel.close()


@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
def test_panel_with_missing_genotypes(precision):
    """A panel holding missing genotypes (imputed with the panel-wide mean;
    the compact missing-row values are then written, imputed, by the operand
    pass in FP64): every answer equals the plain scan of the same panel."""
    d = synth.make(n=400, l=2, m=150, t=2, miss="none", geno_nan=0.02, seed=45)
    G = d.G.contiguous()
    n = G.shape[1]
    assert torch.isnan(G).any()
    reqs = [request_data(n, 12, 3, "mcar", 31), request_data(n, 20, 2, "groups", 32)]
    print("case nan panel", precision, flush=True)
    panel = Panel(G.to("cuda"))
    with Server(panel, precision, chunk_snps=64) as srv:
        ids = [srv.submit(Y, XL, store_all=True, **THR) for Y, XL in reqs]
        answers = [srv.wait(i) for i in ids]
    for (Y, XL), a in zip(reqs, answers):
        assert_hits_equal(a.hits, scan_alone(G, Y, XL, None, precision, **THR), 1e-12 if precision == "fp64" else 1e-4)
    panel.close()


def test_subset_refused_on_a_panel_with_missing_genotypes():
    """The shared operand pass imputes a missing genotype with the
    panel-wide mean, where a scan of a subset uses the subset's mean: a
    request with an individual subset on such a panel is refused
    (OAGPU_ERR_UNSUPPORTED), one with every flag set is served."""
    from omicabelnomm_amd import lib as L
    d = synth.make(n=300, l=2, m=40, t=2, miss="none", geno_nan=0.02, seed=46)
    G = d.G.contiguous()
    n = G.shape[1]
    Y, XL = request_data(n, 6, 3, "mcar", 33)
    keep = np.ones(n, dtype=np.uint8)
    keep[::7] = 0
    print("case nan panel subset", flush=True)
    panel = Panel(G.to("cuda"))
    with Server(panel, "fp64", chunk_snps=64) as srv:
        with pytest.raises(L.OagpuError) as e:
            srv.submit(Y, XL, keep, **THR)
        assert e.value.code == L.ERR_UNSUPPORTED
        assert "subset" in str(e.value)
        a = srv.wait(srv.submit(Y, XL, np.ones(n, dtype=np.uint8), store_all=True, **THR))
    assert_hits_equal(a.hits, scan_alone(G, Y, XL, None, "fp64", **THR), 1e-12)
    panel.close()


@pytest.mark.skip(reason="graph replay is switched off: the default benchmark run ended in a host "
                         "segmentation fault with replay on in its request-stream leg, cause not found")
@pytest.mark.parametrize("precision,tol", [("fp64", 0.0), ("bf16x3", 1e-4)])
def test_graph_replay_equals_launches(precision, tol):
    """A chunk's launches replayed from a captured HIP graph (the chunk start
    written to device memory before each replay) give the same answers as
    launching them one by one: two requests over 9 chunks (the first full
    chunk of an active set runs eagerly, the second is captured, the rest
    replay it; the short last chunk has its own graph), then a third request
    alone (a new active set, captured again).  FP64 answers are equal bit for
    bit; split-BF16 ones within the tolerance of a request served alone or
    in a mix (when the two requests are admitted in different iterations the
    active sets differ, and with them the contraction's column count and
    kernel)."""
    G = panel_data(m=270, seed=47)
    n = G.shape[1]
    reqs = [request_data(n, 24, 3, "mcar", 34), request_data(n, 40, 2, "groups", 35),
            request_data(n, 9, 4, "mcar", 36)]
    panel = Panel(G.to("cuda"))
    out = {}
    for graphs in (False, True):
        print("case graphs", graphs, flush=True)
        with Server(panel, precision, chunk_snps=32, graphs=graphs) as srv:
            ids = [srv.submit(Y, XL, store_all=True, **THR) for Y, XL in reqs[:2]]
            answers = [srv.wait(i) for i in ids]
            answers.append(srv.wait(srv.submit(*reqs[2], store_all=True, **THR)))
            replayed = srv.graph_chunks()
        out[graphs] = answers
        print("graphs", graphs, "replayed chunks", replayed, flush=True)
        if graphs:
            assert replayed >= 7          # the third request alone: 8 full chunks, the first eager
        else:
            assert replayed == 0
    panel.close()
    for a, b in zip(out[False], out[True]):
        assert a.stats["n_hits"] == b.stats["n_hits"] > 0
        if tol == 0.0:
            for f in ("trait", "snp", "h", "n_used", "b", "se", "t", "p", "r2"):
                np.testing.assert_array_equal(a.hits[f], b.hits[f])
        else:
            assert_hits_equal(a.hits, b.hits, tol)
