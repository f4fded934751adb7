"""Dosage models through the GPU path (src/AIOwrapper.cpp:256-340 set-up,
1208-1297 per genotype block).  The genotype block holds file_r
probability columns per SNP and the library forms the model's columns on
the device; the oracle scans the columns its own dosage_apply forms.

FP64: beta and SE within 1e-10 relative, masks and hit sets identical.
Split-BF16: max |dbeta|/SE and max |d -log10 p| below 1e-3.
The reference check program's dosage runs on its interaction fixtures
(tests/test.cpp:201-345) keep its known-answer lower bounds."""
import os

import numpy as np
import pytest

from parity_util import compare_fp64, precision_report
from refmodel import CHECK_SETTINGS, GOLDEN, load_fixture, prepare, scenarios

pytestmark = pytest.mark.gpu

# name: (factors, file_r, summed) -- the reference's models (additive
# 2*AA + AB, dominant AA + AB, recessive AA) and two factor-file models.
MODELS = {
    "additive3": ([2.0, 1.0], 3, True),
    "additive2": ([2.0, 1.0], 2, True),
    "dominant3": ([1.0, 1.0], 3, True),
    "recessive2": ([1.0], 2, True),
    "linear3": ([0.5, 1.5, -2.0], 3, False),
    "myadd3": ([1.0, 0.5, 0.25], 3, True),
    "linear2of3": ([0.5, 1.5], 3, False),
}


def probability_panel(n, m, file_r, seed, nan_rate=0.03):
    """Genotype probabilities (P(AA), P(AB), P(BB))[:file_r] per SNP, each
    0.8 x the true genotype's indicator plus 0.2 x a random split; a few
    missing values.  Returns (probabilities n x m*file_r, allele counts)."""
    rng = np.random.default_rng(seed)
    maf = rng.uniform(0.05, 0.5, m)
    g = rng.binomial(2, maf, size=(n, m))
    onehot = np.zeros((n, m, 3))
    np.put_along_axis(onehot, (2 - g)[:, :, None], 1.0, axis=2)
    noise = rng.dirichlet(np.ones(3), size=(n, m))
    P = (0.8 * onehot + 0.2 * noise)[:, :, :file_r].reshape(n, m * file_r)
    P[rng.random(P.shape) < nan_rate] = np.nan
    return P, g


def make_inputs(n, l, m, t, file_r, seed):
    rng = np.random.default_rng(seed + 1)
    XL = np.column_stack([np.ones(n), rng.standard_normal((n, l - 1))])
    P, g = probability_panel(n, m, file_r, seed)
    Y = rng.standard_normal((n, t))
    for j in range(min(t, m)):       # trait j carries SNP j's allele count
        Y[:, j] += 0.4 * g[:, j]
    Y[rng.random(Y.shape) < 0.05] = np.nan
    return XL, P, Y


def gpu_dosage_scan(XL, G, Y, r, dos, precision="fp64", g_dtype=np.float32):
    from omicabelnomm_amd import Scanner
    factors, file_r, summed = dos
    with Scanner(XL.shape[0], XL.shape[1], r, precision, min_p_store=0.1, min_p_disp=0.05,
                 min_r2_store=-10, disp_cov=True, max_traits=64, max_hits=1 << 20) as sc:
        sc.set_covariates(XL)
        sc.set_dosage_model(factors, file_r, sum=summed)
        sc.load_traits(Y)
        sc.load_genotypes(np.asfortranarray(G.astype(g_dtype)))
        return sc.scan(dense=True)


def oracle_dosage_scan(oracle, XL, G, Y, r, dos):
    factors, file_r, summed = dos
    XR = oracle.dosage_apply(G, file_r, factors, summed)
    return oracle.scan(XL, XR, Y, r, min_p_store=0.1, min_p_disp=0.05, min_r2_store=-10,
                       disp_cov=True, threads=4)


@pytest.mark.parametrize("name", list(MODELS))
def test_dosage_model_fp64(name, oracle):
    dos = MODELS[name]
    r = 1 if dos[2] else len(dos[0])
    XL, G, Y = make_inputs(n=611, l=3, m=37, t=24, file_r=dos[1], seed=81)
    print("case dosage", name, flush=True)
    g = gpu_dosage_scan(XL, G, Y, r, dos)
    Gf = G.astype(np.float32).astype(np.float64)
    o = oracle_dosage_scan(oracle, XL, Gf, Y, r, dos)
    compare_fp64(g, o)
    assert g.stats["n_sig"] > 0


@pytest.mark.parametrize("g_dtype", [np.float16, np.float64])
def test_dosage_genotype_types(g_dtype, oracle):
    dos = MODELS["additive3"]
    XL, G, Y = make_inputs(n=402, l=4, m=29, t=16, file_r=3, seed=83)
    print("case dosage dtype", np.dtype(g_dtype).name, flush=True)
    g = gpu_dosage_scan(XL, G, Y, 1, dos, g_dtype=g_dtype)
    o = oracle_dosage_scan(oracle, XL, G.astype(g_dtype).astype(np.float64), Y, 1, dos)
    compare_fp64(g, o)


@pytest.mark.parametrize("name", ["additive3", "linear3"])
def test_dosage_device_input_with_padded_columns(name, oracle):
    """Genotype probabilities already on the device (float16, leading
    dimension n + 5): the combine reads the caller's buffer in place."""
    import torch
    from omicabelnomm_amd import Scanner
    dos = MODELS[name]
    r = 1 if dos[2] else len(dos[0])
    n = 517
    XL, G, Y = make_inputs(n=n, l=3, m=23, t=14, file_r=dos[1], seed=89)
    print("case dosage device", name, flush=True)
    G16 = G.astype(np.float16)
    full = torch.zeros((G.shape[1], n + 5), dtype=torch.float16, device="cuda")
    full[:, :n] = torch.from_numpy(np.ascontiguousarray(G16.T)).to("cuda")
    Gd = full[:, :n]
    assert Gd.stride(0) == n + 5
    with Scanner(n, 3, r, "fp64", min_p_store=0.1, min_p_disp=0.05, min_r2_store=-10, disp_cov=True,
                 max_traits=64) as sc:
        sc.set_covariates(XL)
        sc.set_dosage_model(dos[0], dos[1], sum=dos[2])
        sc.load_traits(Y)
        sc.load_genotypes(Gd)
        g = sc.scan(dense=True)
    torch.cuda.synchronize()
    o = oracle_dosage_scan(oracle, XL, G16.astype(np.float64), Y, r, dos)
    compare_fp64(g, o)


# All three probability columns of a SNP sum to 1, the intercept: "linear3"
# is singular up to rounding, which FP64 parity covers but no precision
# bound can; the mixed-precision check uses two of the three columns.
@pytest.mark.parametrize("name", ["additive3", "linear2of3"])
def test_dosage_model_bf16x3(name, oracle):
    dos = MODELS[name]
    r = 1 if dos[2] else len(dos[0])
    XL, G, Y = make_inputs(n=1500, l=3, m=48, t=40, file_r=dos[1], seed=85)
    print("case dosage bf16x3", name, flush=True)
    g = gpu_dosage_scan(XL, G, Y, r, dos, precision="bf16x3")
    o = oracle_dosage_scan(oracle, XL, G.astype(np.float32).astype(np.float64), Y, r, dos)
    db, dl = precision_report(g, o)
    print(name, "max|db|/SE", db, "max|d-log10p|", dl, flush=True)
    assert db < 1e-3 and dl < 1e-3
    d = g.dense
    hit_o = (o["pval"] < 1e-8) & o["kept"].astype(bool)
    hit_g = (d["pval"] < 1e-8) & d["kept"].astype(bool)
    np.testing.assert_array_equal(hit_g, hit_o)


def test_dosage_blocks_and_switch_off(oracle):
    """Two genotype blocks in a row (the second smaller) under one model,
    then the model switched off: each scan equals its own oracle scan, and
    a block whose width is not a whole number of file SNPs is refused."""
    from omicabelnomm_amd import Scanner
    from omicabelnomm_amd.lib import OagpuError
    dos = MODELS["dominant3"]
    XL, G, Y = make_inputs(n=350, l=3, m=40, t=12, file_r=3, seed=87)
    print("case dosage blocks", flush=True)
    G32 = G.astype(np.float32)
    with Scanner(350, 3, 1, "fp64", min_p_store=0.1, min_p_disp=0.05, min_r2_store=-10,
                 disp_cov=True, max_traits=64) as sc:
        sc.set_covariates(XL)
        sc.load_traits(Y)
        sc.set_dosage_model(dos[0], 3, sum=True)
        for c0, c1 in [(0, 75), (75, 105)]:
            sc.load_genotypes(np.asfortranarray(G32[:, c0:c1]))
            g = sc.scan(dense=True)
            o = oracle_dosage_scan(oracle, XL, G32[:, c0:c1].astype(np.float64), Y, 1, dos)
            compare_fp64(g, o)
        with pytest.raises(OagpuError):
            sc.load_genotypes(np.asfortranarray(G32[:, :74]))
        sc.set_dosage_model(None)
        sc.load_genotypes(np.asfortranarray(G32[:, :30]))
        g = sc.scan(dense=True)
        o = oracle.scan(XL, G32[:, :30].astype(np.float64), Y, 1, min_p_store=0.1, min_p_disp=0.05,
                        min_r2_store=-10, disp_cov=True, threads=4)
        compare_fp64(g, o)


# The check program's dosage runs (tests/test.cpp:207-241,315-345): model
# d-1 with --ngpred dosage_r[d] and the factor file dosages_<r>.txt; the
# recessive model (d = 3) is the one its message states (the reference's
# reads no column for r = 2) and its count is not asserted, as there.
CHECK_MODELS = {1: ([2.0, 1.0], 2, True), 2: ([1.0, 1.0], 2, True), 3: ([1.0], 2, True),
                4: ("dosages_1.txt", 1, False), 5: ("dosages_2.txt", 2, True)}


def check_model(d):
    from oracle import oracle as O
    f, file_r, summed = CHECK_MODELS[d]
    if isinstance(f, str):
        ok, f = O.read_dosage_factors(open(os.path.join(GOLDEN, f)).read(), file_r)
        assert ok
    return list(f), file_r, summed


def check_cases():
    out = []
    for d in CHECK_MODELS:
        for multi, keep, inter, expected in scenarios():
            out.append((d, multi, keep, inter, True, expected))
        # "dosages alone": the program leaves its last interaction settings
        # (INTRR, several sets, keep) in place and toggles the exclusion list
        for use_excl in (False, True):
            out.append((d, True, True, "INTRR", use_excl, 50))
    return out


@pytest.fixture(scope="module")
def fx():
    return load_fixture()


@pytest.mark.parametrize("d,multi,keep,inter,use_excl,expected", check_cases())
def test_check_program_dosage_runs(d, multi, keep, inter, use_excl, expected, fx, oracle):
    from omicabelnomm_amd import Scanner
    print("case check dosage", d, inter, multi, keep, use_excl, flush=True)
    excl = open(os.path.join(GOLDEN, "exclude_individuals.txt")).read() if use_excl else None
    dos = check_model(d)
    dd = prepare(fx, inter, multi=multi, keep_dep=keep, excl_text=excl, dosage=dos)
    with Scanner(dd["n"], dd["XL"].shape[1], dd["r"], "fp64", real_n=dd["realN"],
                 min_p_store=CHECK_SETTINGS["min_p_store"], min_p_disp=CHECK_SETTINGS["min_p_disp"],
                 min_r2_store=CHECK_SETTINGS["min_r2_store"], disp_cov=True, max_traits=64) as sc:
        sc.set_covariates(dd["XL"])
        if dd["E"] is not None:
            sc.set_interactions(dd["E"])
        sc.set_dosage_model(dos[0], dos[1], sum=dos[2])
        sc.load_traits(dd["Y"])
        sc.load_genotypes(dd["G"], m=dd["m"], col_src=dd["col_src"], col_env=dd["col_env"])
        g = sc.scan(dense=True)
    o = oracle.scan(dd["XL"], dd["XR"], dd["Y"], dd["r"], real_n=dd["realN"], threads=4, **CHECK_SETTINGS)
    compare_fp64(g, o, rtol=1e-8)
    if d != 3:
        assert g.stats["n_sig"] >= expected
