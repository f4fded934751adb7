"""Pins of the FP64 oracle against independent references (numpy / scipy
least squares, closed forms) -- CPU only.

Each check is chosen so that a plausible restatement mistake (a dropped
correction term, a wrong index, a misplaced column, a wrong degree of
freedom) fails it.
"""
import numpy as np
import pytest
import scipy.special as sp

from omicabelnomm_amd import synth
from parity_util import near_singular


def _data(n=300, l=4, m=8, t=6, r=1, miss_rate=0.0, seed=0):
    rng = np.random.default_rng(seed)
    XL = np.column_stack([np.ones(n), rng.standard_normal((n, l - 1))])
    XR = rng.binomial(2, 0.3, size=(n, r * m)).astype(float) + rng.normal(0, 0.05, (n, r * m))
    Y = rng.standard_normal((n, t))
    if miss_rate:
        Y[rng.random((n, t)) < miss_rate] = np.nan
    return XL, XR, Y


def _lstsq(XL, XR, Y, r, y, s):
    obs = ~np.isnan(Y[:, y])
    X = np.column_stack([XL, XR[:, s * r:(s + 1) * r]])[obs]
    b, *_ = np.linalg.lstsq(X, Y[obs, y], rcond=None)
    res = X @ b - Y[obs, y]
    return b, float(res @ res), X


@pytest.mark.parametrize("miss", [0.0, 0.1])
def test_r1_beta_rss_equal_least_squares_on_observed_rows(oracle, miss):
    XL, XR, Y = _data(miss_rate=miss, seed=1)
    o = oracle.scan(XL, XR, Y, 1, store_all=True)
    for y in range(Y.shape[1]):
        for s in range(XR.shape[1]):
            b, rss, _ = _lstsq(XL, XR, Y, 1, y, s)
            np.testing.assert_allclose(o["beta"][y, s], b, rtol=1e-9, atol=1e-12)
            np.testing.assert_allclose(o["rss"][y, s], rss, rtol=1e-9)
    assert o["status"] == 0


def test_no_missing_se_is_textbook_ols(oracle):
    XL, XR, Y = _data(n=250, l=3, m=5, t=4, r=2, seed=2)
    o = oracle.scan(XL, XR, Y, 2, store_all=True, disp_cov=True)
    n, p = XL.shape[0], XL.shape[1] + 2
    for y in range(4):
        for s in range(5):
            b, rss, X = _lstsq(XL, XR, Y, 2, y, s)
            cov = rss / (n - p) * np.linalg.inv(X.T @ X)
            se = np.sqrt(np.diag(cov))
            np.testing.assert_allclose(o["beta"][y, s], b, rtol=1e-9, atol=1e-12)
            np.testing.assert_allclose(o["se"][y, s, 1:], se[1:], rtol=1e-9)
            np.testing.assert_allclose(o["tstat"][y, s, 1:], np.abs(b[1:] / se[1:]), rtol=1e-9)


def test_var_x_is_unmasked_inverse_with_mean_imputation(oracle):
    XL, XR, Y = _data(n=200, l=3, m=4, t=3, miss_rate=0.1, seed=3)
    XR[5, 1] = np.nan
    XR[17, 1] = np.nan
    o = oracle.scan(XL, XR, Y, 1, store_all=True)
    x = XR[:, 1].copy()
    x[np.isnan(x)] = np.nanmean(x)
    X = np.column_stack([XL, x])
    np.testing.assert_allclose(o["var_x"][1], np.diag(np.linalg.inv(X.T @ X)), rtol=1e-10)
    # SE uses the unmasked var_x scaled by (n - Ymiss)/n
    y = 0
    n_obs = int((~np.isnan(Y[:, y])).sum())
    n, p = 200, 4
    sig = np.sqrt(o["rss"][y, 1] / (n_obs - p))
    se = sig * np.sqrt(o["var_x"][1, 3] * n_obs / n)
    np.testing.assert_allclose(o["se"][y, 1, 3], se, rtol=1e-12)


def test_r2_misplaced_correction(oracle):
    """With r = 2 and missing phenotypes the reference subtracts every SNP
    column's missing-row correction from the first SNP column
    (src/Algorithm.cpp:1390-1396).  When the second column is zero on the
    missing rows, that correction is zero and b is the least-squares
    solution on the observed rows; otherwise it is not."""
    n, l = 240, 3
    XL, XR, Y = _data(n=n, l=l, m=3, t=2, r=2, miss_rate=0.0, seed=4)
    miss = np.zeros(n, bool)
    miss[::9] = True
    Y[miss, :] = np.nan
    XR[miss, 1] = 0.0      # SNP 0, second column: zero on the missing rows
    o = oracle.scan(XL, XR, Y, 2, store_all=True)
    b, _, _ = _lstsq(XL, XR, Y, 2, 0, 0)
    np.testing.assert_allclose(o["beta"][0, 0], b, rtol=1e-9, atol=1e-12)
    b1, _, _ = _lstsq(XL, XR, Y, 2, 0, 1)
    assert np.max(np.abs(o["beta"][0, 1] - b1)) > 1e-6
    # RSS is still the explicit residual of that b over observed rows
    obs = ~miss
    X = np.column_stack([XL, XR[:, 2:4]])[obs]
    res = X @ o["beta"][0, 1] - Y[obs, 0]
    np.testing.assert_allclose(o["rss"][0, 1], res @ res, rtol=1e-10)


def test_rss_skips_observed_zero(oracle):
    XL, XR, Y = _data(n=150, l=2, m=2, t=1, seed=5)
    Y[:10, 0] = 0.0
    o = oracle.scan(XL, XR, Y, 1, store_all=True)
    X = np.column_stack([XL, XR[:, 0]])
    b = o["beta"][0, 0]
    keep = Y[:, 0] != 0
    res = X[keep] @ b - Y[keep, 0]
    np.testing.assert_allclose(o["rss"][0, 0], res @ res, rtol=1e-10)
    # b itself uses every observed row
    b_ls, _, _ = _lstsq(XL, XR, Y, 1, 0, 0)
    np.testing.assert_allclose(b, b_ls, rtol=1e-9)


def test_singular_snp_keeps_rhs(oracle):
    """sposv fails on an all-zero SNP column and the reference ignores info:
    b stays the right-hand side [XL'y; 0] and var_x is undefined (NaN here)."""
    XL, XR, Y = _data(n=120, l=3, m=2, t=2, seed=6)
    XR[:, 1] = 0.0
    o = oracle.scan(XL, XR, Y, 1, store_all=True)
    rhs = np.concatenate([XL.T @ Y[:, 0], [0.0]])
    np.testing.assert_allclose(o["beta"][0, 1], rhs, rtol=1e-12)
    assert np.all(np.isnan(o["var_x"][1]))


def test_sst_and_nused(oracle):
    XL, XR, Y = _data(n=180, l=2, m=1, t=3, miss_rate=0.2, seed=7)
    o = oracle.scan(XL, XR, Y, 1, store_all=True)
    for y in range(3):
        obs = ~np.isnan(Y[:, y])
        assert o["n_used"][y] == obs.sum()
    # R2 with SST over the observed rows and the reference's dof
    y, n, p = 0, 180, 3
    obs = ~np.isnan(Y[:, y])
    sst = np.sum((Y[obs, y] - Y[obs, y].mean()) ** 2)
    nc = obs.sum() - p
    np.testing.assert_allclose(o["r2"][y, 0], 1 - (o["rss"][y, 0] / (nc - 1)) / (sst / (n - 1)), rtol=1e-10)


def test_pvalue_rule(oracle):
    for t in [1.6, 2.0, 3.7, 8.0, 12.5, 30.0]:
        np.testing.assert_allclose(oracle.pvalue(t), sp.erfc(t / np.sqrt(2)), rtol=1e-12)
    assert oracle.pvalue(0.0) == pytest.approx(0.5)
    assert oracle.pvalue(1.0) == pytest.approx(0.5 - 0.1 * 3.4)


@pytest.mark.parametrize("p", [0.1, 0.05, 1e-3, 5e-5, 1e-8, 0.5])
def test_tvalue_and_erfinv(oracle, p):
    want = np.float32(np.sqrt(2.0) * sp.erfinv(1.0 - np.float64(np.float32(p))))
    assert oracle.tvalue(p) == float(want)


@pytest.mark.parametrize("x", [0.3, -0.7, 1e-12, 0.49999, 0.5, 0.9, 1.0 - 1e-10, 1.0 - 2.0 ** -52])
def test_erfinv_against_scipy(oracle, x):
    np.testing.assert_allclose(oracle.erfinv(x), sp.erfinv(x), rtol=2e-15)


def test_min_singular(oracle):
    rng = np.random.default_rng(8)
    A = rng.standard_normal((50, 4))
    np.testing.assert_allclose(oracle.min_singular(A), np.linalg.svd(A, compute_uv=False).min(), rtol=1e-10)
    A[:, 3] = A[:, 2]
    assert oracle.min_singular(A) < 1e-6


@pytest.mark.parametrize("smin", [0.8e-6, 1.25e-6, 1e-9])
def test_min_singular_near_the_threshold(oracle, smin):
    """At n = 10,000 with sigma_max = 100, sigma_min near the reference's
    1e-6 threshold: Jacobi on XL'XL could not resolve it (its eigenvalues
    near 1e-12 sit below the rounding of 1e4); the one-sided Jacobi on XL
    agrees with numpy's SVD to 1e-6 relative and puts the matrix on the
    right side of the threshold."""
    A = near_singular(10_000, [100.0, 30.0, 5.0, 1.0, smin], seed=9)
    want = np.linalg.svd(A, compute_uv=False).min()
    got = oracle.min_singular(A)
    np.testing.assert_allclose(got, want, rtol=1e-6)
    np.testing.assert_allclose(got, smin, rtol=1e-6)
    assert (got < 1e-6) == (smin < 1e-6)


def test_exclusion_list_example(oracle):
    import os
    text = open(os.path.join(os.path.dirname(__file__), "golden", "exclude_individuals.txt")).read()
    rc, keep = oracle.exclusions(text, 40)
    assert rc == 8
    assert sorted(np.nonzero(~keep)[0] + 1) == [1, 2, 5, 7, 9, 10, 11, 13]
    # a first line with a single number is rejected ("Please provide ordered pairs!")
    assert oracle.exclusions("4\n7 9\n", 40)[0] == -1
    # unordered ranges are rejected
    assert oracle.exclusions("5 9\n2 3\n", 40)[0] == -2
    # ranges beyond n are clipped
    rc, keep = oracle.exclusions("38 45\n", 40)
    assert rc == 3


def test_synth_generator_shapes():
    d = synth.make(n=100, l=3, m=7, t=5, interaction=True, miss="groups", pattern_groups=2, seed=1)
    assert d.r == 2 and d.m == 7 and d.t == 5
    XR = d.xr_matrix()
    G = d.G.double().numpy()
    np.testing.assert_allclose(XR[:, 0], G[0] * d.E[:, 0])
    np.testing.assert_allclose(XR[:, 1], G[0])
    Y = d.y_matrix()
    miss = np.isnan(Y)
    # 2 groups, each with its own row pattern
    assert (miss[:, 0] == miss[:, 1]).all() and not (miss[:, 0] == miss[:, 4]).all()


def test_mlog10p_from_t_matches_the_pvalue_rule(oracle):
    """The tests' -log10 p taken from t (no subnormal p) agrees with the
    oracle's p-value rule wherever that p is a normal double, on both sides
    of the t = 1.6 switch."""
    from parity_util import mlog10p_t
    ts = np.concatenate([np.linspace(0.0, 1.59, 40), np.linspace(1.6, 37.0, 200)])
    want = np.array([-np.log10(oracle.pvalue(t)) for t in ts])
    np.testing.assert_allclose(mlog10p_t(ts), want, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(mlog10p_t(-ts), want, rtol=1e-12, atol=1e-12)
    # beyond the normal range it keeps growing smoothly (p would underflow)
    big = mlog10p_t(np.array([40.0, 60.0]))
    assert np.all(np.isfinite(big)) and big[1] > big[0] > 340
