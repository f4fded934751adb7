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
