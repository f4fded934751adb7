"""Shared helpers of the parity tests: run the HIP path and the FP64 oracle
on the same inputs and compare."""
import numpy as np

LOG10 = np.log(10.0)


def run_gpu(data, precision="fp64", *, dense=True, device_inputs=False, max_snps=0,
            min_p_store=0.1, min_p_disp=0.05, min_r2_store=1e-5, disp_cov=False,
            store_all=False, max_hits=1 << 20):
    from omicabelnomm_amd import Scanner
    sc = Scanner(data.n, data.l, data.r, precision, disp_cov=disp_cov, store_all=store_all,
                 min_p_store=min_p_store, min_p_disp=min_p_disp, min_r2_store=min_r2_store,
                 max_traits=max(64, data.t), max_snps=max_snps, max_hits=max_hits)
    try:
        sc.set_covariates(data.XL)
        if data.E is not None:
            sc.set_interactions(data.E)
        if device_inputs:
            import torch
            Y = data.Y.to("cuda").contiguous()
            G = data.G.to("cuda").contiguous()
        else:
            Y = data.Y.cpu().numpy().T
            G = data.G.cpu().numpy().T
        sc.load_traits(Y)
        sc.load_genotypes(G, m=data.m, col_src=data.col_src, col_env=data.col_env)
        res = sc.scan(dense=dense)
    finally:
        sc.close()
    return res


def run_oracle(oracle, data, *, min_p_store=0.1, min_p_disp=0.05, min_r2_store=1e-5,
               disp_cov=False, store_all=False):
    return oracle.scan(data.XL, data.xr_matrix(), data.y_matrix(), data.r,
                       min_p_store=min_p_store, min_p_disp=min_p_disp,
                       min_r2_store=min_r2_store, disp_cov=disp_cov, store_all=store_all,
                       threads=4)


def mlog10(p):
    with np.errstate(divide="ignore", invalid="ignore"):
        return -np.log10(p)


def mlog10p_t(t):
    """-log10 of the p-value rule of hpc_statistics (src/Algorithm.cpp:
    1159-1167) computed from |t| without forming p: 0.5 - 0.1 (4.4 - t) t
    below t = 1.6, erfc(t / sqrt 2) = 2 Phi(-t) above, whose logarithm
    log 2 + log Phi(-t) stays exact where p itself is a subnormal double
    (|t| > ~37.5) or underflows."""
    from scipy.special import log_ndtr
    t = np.abs(np.asarray(t, dtype=np.float64))
    lo = t < 1.6
    with np.errstate(divide="ignore", invalid="ignore"):
        small = -np.log10(0.5 - 0.1 * (4.4 - t) * t)
    big = -(np.log(2.0) + log_ndtr(-t)) / LOG10
    return np.where(lo, small, big)


T_SWITCH = 1.6


def dlog10p_t(ta, tb):
    """|d -log10 p| between two t statistics under the reference's p-value
    rule, and the pairs it leaves out.  The rule is discontinuous at
    t = 1.6 (0.5 - 0.1 (4.4 - t) t = 0.052 just below, erfc(1.6 / sqrt 2)
    = 0.110 at 1.6), so two t values 1e-6 apart on either side of it differ
    by 0.32 in -log10 p; such straddling pairs are measured by |dt| instead.
    Returns (dl over the other pairs, |dt| over the straddling pairs)."""
    ta = np.abs(np.asarray(ta, dtype=np.float64))
    tb = np.abs(np.asarray(tb, dtype=np.float64))
    straddle = (ta < T_SWITCH) != (tb < T_SWITCH)
    dl = np.abs(mlog10p_t(ta[~straddle]) - mlog10p_t(tb[~straddle]))
    return dl, np.abs(ta[straddle] - tb[straddle])


def dlog10p(pa, pb):
    """|-log10 pa + log10 pb|, 0 where both underflowed to 0."""
    d = np.abs(mlog10(pa) - mlog10(pb))
    d[(pa == 0) & (pb == 0)] = 0.0
    return d


def compare_fp64(g, o, *, rtol=1e-10):
    """FP64 parity: beta and SE within rtol (relative to max(|beta|, SE)
    for beta, to SE for SE), masks and the kept/emitted sets identical."""
    d = g.dense
    np.testing.assert_array_equal(d["n_used"], o["n_used"])
    np.testing.assert_array_equal(d["kept"], o["kept"])
    np.testing.assert_array_equal(d["emitted"], o["emitted"])
    ok = ~np.isnan(o["se"])
    np.testing.assert_array_equal(ok, ~np.isnan(d["se"]))
    se = o["se"][ok]
    scale = np.maximum(np.abs(o["beta"][ok]), se)
    err_b = np.abs(d["beta"][ok] - o["beta"][ok]) / scale
    err_se = np.abs(d["se"][ok] - se) / se
    assert err_b.max(initial=0) <= rtol, ("beta", err_b.max())
    assert err_se.max(initial=0) <= rtol, ("se", err_se.max())
    kept = o["kept"].astype(bool)
    # p as stored, where it is a normal double; -log10 p from t everywhere
    normal = kept & (o["pval"] > 1e-300)
    dl = dlog10p(d["pval"][normal], o["pval"][normal])
    assert dl.max(initial=0) <= 1e-8, ("-log10 p", dl.max())
    # -log10 p ~ t^2 / (2 ln 10): a relative error e in t moves it by about
    # 2 e (-log10 p), which is large for the near-perfect fits' t ~ 1e7
    tk = d["tstat"][kept]
    dlt, dt_sw = dlog10p_t(tk, o["tstat"][kept])
    lim = 1e-8 + 4.0 * rtol * mlog10p_t(o["tstat"][kept])[(np.abs(tk) < T_SWITCH) == (np.abs(o["tstat"][kept]) < T_SWITCH)]
    assert np.all(dlt <= lim), ("-log10 p from t", dlt.max())
    assert dt_sw.max(initial=0) <= 1e-9, ("t at the p-rule switch", dt_sw.max())
    r2ok = ~np.isnan(o["r2"])
    np.testing.assert_allclose(d["r2"][r2ok], o["r2"][r2ok], rtol=1e-9, atol=1e-12)
    assert g.stats["n_sig"] == o["n_sig"]
    assert g.stats["n_emitted"] == o["n_emitted"]
    assert g.stats["n_hits"] == o["n_kept"]
    return dict(max_rel_beta=float(err_b.max(initial=0)), max_rel_se=float(err_se.max(initial=0)))


def hits_from_dense(o, m):
    """(trait, snp, h) of every kept coefficient inside an emitted record,
    in the reference's order (trait, snp, h descending)."""
    kept = o["kept"].astype(bool) & o["emitted"].astype(bool)[:, :, None]
    y, s, h = np.nonzero(kept)
    order = np.lexsort((-h, s, y))
    return np.stack([y[order], s[order], h[order]], axis=1)


def precision_report(g, o):
    """max |dbeta|/SE and max |d -log10 p| over the coefficients the oracle
    evaluated."""
    d = g.dense
    ok = ~np.isnan(o["se"]) & ~np.isnan(d["se"])
    db = np.abs(d["beta"][ok] - o["beta"][ok]) / o["se"][ok]
    dl, dt_sw = dlog10p_t(d["tstat"][ok], o["tstat"][ok])
    # t straddling the reference's p-rule switch: held to |dt| instead
    assert dt_sw.max(initial=0) <= 1e-3, ("t at the p-rule switch", dt_sw.max())
    return float(db.max(initial=0)), float(dl.max(initial=0))


# ---------------------------------------------------------------------------
# hit records (non-dense scans) against the oracle

MIXED_TOL = 1e-3
STRONG_P = 1e-8


def oracle_records(o, snp0=0):
    """Kept coefficients of emitted (trait, SNP) records of a dense oracle
    result, in the reference's order (trait, SNP, h descending)."""
    kept = o["kept"].astype(bool) & o["emitted"].astype(bool)[:, :, None]
    y, s, h = np.nonzero(kept)
    order = np.lexsort((-h, s, y))
    y, s, h = y[order], s[order], h[order]
    return {"trait": y, "snp": s + snp0, "h": h, "b": o["beta"][y, s, h], "se": o["se"][y, s, h],
            "t": o["tstat"][y, s, h], "p": o["pval"][y, s, h], "n_used": o["n_used"][y]}


def window(hits, snp_lo, snp_hi, traits=None):
    sel = (hits["snp"] >= snp_lo) & (hits["snp"] < snp_hi)
    if traits is not None:
        sel &= hits["trait"] < traits
    return hits[sel]


def keys(rec):
    return list(zip(np.asarray(rec["trait"]).tolist(), np.asarray(rec["snp"]).tolist(),
                    np.asarray(rec["h"]).tolist()))


def assert_fp64_records(hits, rec, tol=1e-10):
    """GPU hits (FP64 path) equal the oracle's kept records."""
    assert len(hits) == len(rec["trait"]), (len(hits), len(rec["trait"]))
    assert keys(hits) == keys(rec)
    np.testing.assert_array_equal(hits["n_used"], rec["n_used"])
    scale = np.maximum(np.abs(rec["b"]), rec["se"])
    assert np.all(np.abs(hits["b"] - rec["b"]) <= tol * scale), np.max(np.abs(hits["b"] - rec["b"]) / scale)
    assert np.all(np.abs(hits["se"] - rec["se"]) <= tol * rec["se"])
    normal = rec["p"] > 1e-300
    assert dlog10p(hits["p"][normal], rec["p"][normal]).max(initial=0) <= 1e-8
    # relative to -log10 p, as in compare_fp64 (near-perfect fits: t ~ 1e7)
    dlt, dt_sw = dlog10p_t(hits["t"], rec["t"])
    same_side = (np.abs(hits["t"]) < T_SWITCH) == (np.abs(rec["t"]) < T_SWITCH)
    lim = 1e-8 + 4.0 * tol * mlog10p_t(rec["t"])[same_side]
    assert np.all(dlt <= lim), ("-log10 p from t", dlt.max(initial=0))
    assert dt_sw.max(initial=0) <= 1e-9, ("t at the p-rule switch", dt_sw.max())


def mixed_report(hits, rec):
    """max |dbeta|/SE and max |d -log10 p| over the records both sides kept,
    and whether the p < 1e-8 sets agree."""
    idx = {k: i for i, k in enumerate(keys(rec))}
    pairs = [(j, idx[k]) for j, k in enumerate(keys(hits)) if k in idx]
    db = dl = 0.0
    if pairs:
        jj, ii = (np.array(v) for v in zip(*pairs))
        db = float(np.max(np.abs(hits["b"][jj] - rec["b"][ii]) / rec["se"][ii]))
        dls, dt_sw = dlog10p_t(hits["t"][jj], rec["t"][ii])
        assert dt_sw.max(initial=0) <= MIXED_TOL, ("t at the p-rule switch", dt_sw.max())
        dl = float(dls.max(initial=0))
    strong_g = {k for k, p in zip(keys(hits), hits["p"]) if p < STRONG_P}
    strong_o = {k for k, p in zip(keys(rec), rec["p"]) if p < STRONG_P}
    return db, dl, strong_g == strong_o, len(strong_o)


def assert_mixed_records(hits, rec, label):
    db, dl, same, ns = mixed_report(hits, rec)
    print(label, "max|db|/SE %.2e max|d-log10p| %.2e strong hits %d same %s" % (db, dl, ns, same), flush=True)
    assert db <= MIXED_TOL and dl <= MIXED_TOL, (db, dl)
    assert same


def near_singular(n, sv, seed):
    """n x len(sv) matrix with the given singular values (random orthonormal
    left and right factors)."""
    rng = np.random.default_rng(seed)
    U, _ = np.linalg.qr(rng.standard_normal((n, len(sv))))
    V, _ = np.linalg.qr(rng.standard_normal((len(sv), len(sv))))
    return np.asfortranarray(U @ np.diag(sv) @ V.T)
