"""The reference's interaction fixtures through the GPU path: per-scenario
parity with the FP64 oracle (in-memory C ABI) and the reference's own
known-answer lower bounds; then the drop-in file path (DatABEL in,
_dis.txt / _sto.ibin / _sto.dbin out) through Algorithm.solve and the
omicabelnomm CLI."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from omicabelnomm_amd import databel
from omicout import (expected_dis_rows, expected_records, pind_order, read_dbin, read_dbin_pind, read_dis,
                     read_ibin)
from parity_util import compare_fp64
from refmodel import CHECK_SETTINGS, GOLDEN, load_fixture, prepare, scenarios

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fx():
    return load_fixture()


@pytest.fixture(scope="module")
def excl():
    return open(os.path.join(GOLDEN, "exclude_individuals.txt")).read()


def gpu_scan(d, precision="fp64"):
    from omicabelnomm_amd import Scanner
    with Scanner(d["n"], d["XL"].shape[1], d["r"], precision, real_n=d["realN"],
                 min_p_store=CHECK_SETTINGS["min_p_store"], min_p_disp=CHECK_SETTINGS["min_p_disp"],
                 min_r2_store=CHECK_SETTINGS["min_r2_store"], disp_cov=True, max_traits=64) as sc:
        sc.set_covariates(d["XL"])
        if d["E"] is not None:
            sc.set_interactions(d["E"])
        sc.load_traits(d["Y"])
        sc.load_genotypes(d["G"], m=d["m"], col_src=d["col_src"], col_env=d["col_env"])
        return sc.scan(dense=True)


@pytest.mark.parametrize("multi,keep,inter,expected", scenarios())
def test_fixture_scenario(oracle, fx, excl, multi, keep, inter, expected):
    print("scenario", inter, "multi" if multi else "single", "keep" if keep else "", flush=True)
    d = prepare(fx, inter, multi=multi, keep_dep=keep, excl_text=excl)
    g = gpu_scan(d)
    o = oracle.scan(d["XL"], d["XR"], d["Y"], d["r"], real_n=d["realN"], threads=4, **CHECK_SETTINGS)
    # near-perfect fits (the fixtures plant XR = Y/2 - 0.01*sum(XL)) amplify
    # rounding differences of the two solve orders in SE
    compare_fp64(g, o, rtol=1e-8)
    assert g.stats["n_sig"] >= expected


def _write_inputs(fx, tmp, inter_names=()):
    obs = list(fx["obs_names"])
    databel.write(os.path.join(tmp, "XL"), fx["XL"], obs, list(fx["xl_names"]))
    databel.write(os.path.join(tmp, "XR"), fx["XR"], obs, list(fx["xr_names"]))
    databel.write(os.path.join(tmp, "Y"), fx["Y"], obs, list(fx["y_names"]))
    for nm in inter_names:
        databel.write(os.path.join(tmp, nm), fx["E_" + nm], obs, list(fx["Enames_" + nm]))
    shutil.copy(os.path.join(GOLDEN, "exclude_individuals.txt"), os.path.join(tmp, "excl.txt"))


def _settings(tmp, out):
    from omicabelnomm_amd import Algorithm, Settings
    s = Settings()
    Algorithm().applyDefaultParams(s)
    s.fnameAL, s.fnameAR, s.fnameY = (os.path.join(tmp, x) for x in ("XL", "XR", "Y"))
    s.fnameOutFiles = os.path.join(tmp, out)
    s.fname_excludelist = os.path.join(tmp, "excl.txt")
    s.minPstore, s.minPdisp, s.minR2store, s.minR2disp = 0.1, 0.00005, 0.00001, 0.000001
    s.disp_cov = True
    s.storeBin = True
    return s


def test_dropin_files_basic_model(oracle, fx, tmp_path):
    from omicabelnomm_amd import Algorithm, Outputs
    tmp = str(tmp_path)
    _write_inputs(fx, tmp)
    s = _settings(tmp, "basic")
    out = Outputs()
    Algorithm().solve(s, out)
    # oracle view of the same run (storeBin normalises: minPstore = minPdisp, minR2store = minR2disp)
    excl = open(os.path.join(GOLDEN, "exclude_individuals.txt")).read()
    d = prepare(fx, None, excl_text=excl, limit_t=50, limit_m=100)
    o = oracle.scan(d["XL"], d["XR"], d["Y"], 1, real_n=d["realN"], min_p_store=0.00005,
                    min_p_disp=0.00005, min_r2_store=0.000001, disp_cov=True, threads=4)
    assert out.total_sig_results == o["n_sig"]
    rows = read_dis(os.path.join(tmp, "basic_dis.txt"))
    names = dict(y=list(fx["y_names"]), al=list(fx["xl_names"]), ar=list(fx["xr_names"]))
    want = expected_dis_rows(o, names, 4, 1, 0.00005, 0.000001, d["realN"])
    assert [(r["phe"], r["snp"]) for r in rows] == [(w["phe"], w["snp"]) for w in want]
    for r, w in zip(rows, want):
        assert r["n"] == w["n"]
        np.testing.assert_allclose([r["B"], r["SE"], r["T"]], np.float32([w["B"], w["SE"], w["T"]]),
                                   rtol=2e-6, atol=2e-7)
        np.testing.assert_allclose(r["P"], w["P"], rtol=1e-5)
        assert abs(r["npct"] - w["npct"]) <= 0.006
    ib = read_ibin(os.path.join(tmp, "basic_sto.ibin"))
    assert (ib["n"], ib["l"], ib["r"], ib["t"], ib["m"]) == (d["n"], 4, 1, 50, 100)
    assert ib["al_names"] == list(fx["xl_names"][1:]) and ib["y_names"] == list(fx["y_names"])
    recs = read_dbin(os.path.join(tmp, "basic_sto.dbin"), disp_cov=True, store_all=False, l=4, r=1)
    want_recs = expected_records(o, l=4, r=1)
    assert len(recs) == len(want_recs) == o["n_emitted"]
    assert sum(len(rc["results"]) for rc in recs) == o["n_kept"]
    for rc, (blk_y, ar_off, n_used, res) in zip(recs, want_recs):
        assert (rc["y"], rc["ar_offset"], rc["n_used"]) == (blk_y, ar_off, n_used)
        assert [(a, b) for a, b, _, _ in rc["results"]] == [(a, b) for a, b, _, _ in res]
        for (_, _, B, SE), (_, _, wB, wSE) in zip(rc["results"], res):
            assert B == pytest.approx(np.float32(wB), rel=1e-6, abs=1e-7)
            assert SE == pytest.approx(np.float32(wSE), rel=1e-6, abs=1e-7)
    # the CLI writes the same _dis.txt
    cli = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "omicabelnomm_amd", "omicabelnomm")
    cmd = [cli, "-c", os.path.join(tmp, "XL.fvi"), "-g", os.path.join(tmp, "XR.fvi"), "-p",
           os.path.join(tmp, "Y.fvi"), "-o", os.path.join(tmp, "cli.txt"), "-x", os.path.join(tmp, "excl.txt"),
           "-d", "0.00005", "-s", "0.1", "-e", "0.00001", "-r", "0.000001", "-i", "-b"]
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert p.returncode == 0, p.stdout
    assert "Number of Significant Results: %d" % o["n_sig"] in p.stdout
    assert open(os.path.join(tmp, "cli_dis.txt")).read() == open(os.path.join(tmp, "basic_dis.txt")).read()
    assert open(os.path.join(tmp, "cli_sto.dbin"), "rb").read() == open(os.path.join(tmp, "basic_sto.dbin"), "rb").read()


def test_dropin_files_store_pind(oracle, fx, tmp_path):
    """-f (storePInd) with -b and -i: every coefficient of every (trait, SNP)
    pair is stored, in the positional layout (src/AIOwrapper.cpp:1454-1531:
    no block header, no size_p, no AL/AR index), and every one of them is a
    _dis.txt row.  The records are read back as OmicOUT reads them
    (OmicOUT/main.cpp:347-420) and compared with the oracle's store-all scan;
    the CLI's -f files are byte-equal to Algorithm.solve's."""
    from omicabelnomm_amd import Algorithm, Outputs
    tmp = str(tmp_path)
    _write_inputs(fx, tmp)
    s = _settings(tmp, "pind")
    s.storePInd = True
    out = Outputs()
    Algorithm().solve(s, out)
    excl = open(os.path.join(GOLDEN, "exclude_individuals.txt")).read()
    d = prepare(fx, None, excl_text=excl, limit_t=50, limit_m=100)
    o = oracle.scan(d["XL"], d["XR"], d["Y"], 1, real_n=d["realN"], min_p_store=0.00005,
                    min_p_disp=0.00005, min_r2_store=0.000001, disp_cov=True, store_all=True, threads=4)
    t, m, l, r = 50, 100, 4, 1
    assert o["emitted"].all()                       # every pair is a record
    ib = read_ibin(os.path.join(tmp, "pind_sto.ibin"))
    assert ib["store_all"] and ib["disp_cov"]
    h_max = r + l - 1                                # the intercept is not stored
    recs = read_dbin_pind(os.path.join(tmp, "pind_sto.dbin"), h_max)
    order = pind_order(t, m)
    assert len(recs) == len(order) == t * m
    for rc, (y, snp) in zip(recs, order):
        assert rc["n_used"] == o["n_used"][y]
        assert rc["r2"] == pytest.approx(float(np.float32(o["r2"][y, snp])), rel=2 ** -7, abs=1e-6)
        want = [(o["beta"][y, snp, h], o["se"][y, snp, h]) for h in range(l + r - 1, 0, -1)]
        for (B, SE), (wB, wSE) in zip(rc["results"], want):
            assert B == pytest.approx(np.float32(wB), rel=1e-6, abs=1e-7)
            assert SE == pytest.approx(np.float32(wSE), rel=1e-6, abs=1e-7)
    rows = read_dis(os.path.join(tmp, "pind_dis.txt"))
    names = dict(y=list(fx["y_names"]), al=list(fx["xl_names"]), ar=list(fx["xr_names"]))
    want_rows = expected_dis_rows(o, names, l, r, 0.00005, 0.000001, d["realN"], store_all=True)
    assert len(rows) == len(want_rows) == t * m * h_max
    assert [(x["phe"], x["snp"]) for x in rows] == [(w["phe"], w["snp"]) for w in want_rows]
    cli = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "omicabelnomm_amd", "omicabelnomm")
    cmd = [cli, "-c", os.path.join(tmp, "XL.fvi"), "-g", os.path.join(tmp, "XR.fvi"), "-p",
           os.path.join(tmp, "Y.fvi"), "-o", os.path.join(tmp, "cli.txt"), "-x", os.path.join(tmp, "excl.txt"),
           "-d", "0.00005", "-s", "0.1", "-e", "0.00001", "-r", "0.000001", "-i", "-b", "-f"]
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert p.returncode == 0, p.stdout
    for suffix in ("_dis.txt", "_sto.dbin", "_sto.ibin"):
        assert open(os.path.join(tmp, "cli" + suffix), "rb").read() == \
            open(os.path.join(tmp, "pind" + suffix), "rb").read(), suffix


def test_dropin_files_additive_dosage(oracle, fx, tmp_path):
    """-j (additive dosage model, 2*AA + AB) with --ngpred 2 through
    Algorithm.solve and the CLI: the file's 100 columns are 50 SNPs of two
    probability columns; every output names a SNP by its first column
    (src/AIOwrapper.cpp:563-588,709-718)."""
    from omicabelnomm_amd import Algorithm, Outputs
    tmp = str(tmp_path)
    _write_inputs(fx, tmp)
    s = _settings(tmp, "dos")
    s.r = 2
    s.dosages = True
    s.model = 0
    out = Outputs()
    Algorithm().solve(s, out)
    excl = open(os.path.join(GOLDEN, "exclude_individuals.txt")).read()
    d = prepare(fx, None, excl_text=excl, limit_t=50, limit_m=100, dosage=([2.0, 1.0], 2, True))
    assert d["m"] == 50
    o = oracle.scan(d["XL"], d["XR"], d["Y"], 1, real_n=d["realN"], min_p_store=0.00005,
                    min_p_disp=0.00005, min_r2_store=0.000001, disp_cov=True, threads=4)
    assert out.total_sig_results == o["n_sig"] and o["n_sig"] > 0
    rows = read_dis(os.path.join(tmp, "dos_dis.txt"))
    names = dict(y=list(fx["y_names"]), al=list(fx["xl_names"]), ar=list(fx["xr_names"][::2]))
    want = expected_dis_rows(o, names, 4, 1, 0.00005, 0.000001, d["realN"])
    assert len(rows) > 0
    assert [(r["phe"], r["snp"]) for r in rows] == [(w["phe"], w["snp"]) for w in want]
    for r, w in zip(rows, want):
        np.testing.assert_allclose([r["B"], r["SE"], r["T"]], np.float32([w["B"], w["SE"], w["T"]]),
                                   rtol=2e-6, atol=2e-7)
    ib = read_ibin(os.path.join(tmp, "dos_sto.ibin"))
    assert (ib["r"], ib["m"]) == (1, 50)
    assert ib["ar_names"] == list(fx["xr_names"][::2])
    cli = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "omicabelnomm_amd", "omicabelnomm")
    cmd = [cli, "-c", os.path.join(tmp, "XL.fvi"), "-g", os.path.join(tmp, "XR.fvi"), "-p",
           os.path.join(tmp, "Y.fvi"), "-o", os.path.join(tmp, "cli.txt"), "-x", os.path.join(tmp, "excl.txt"),
           "-d", "0.00005", "-s", "0.1", "-e", "0.00001", "-r", "0.000001", "-i", "-b", "-n", "2", "-j"]
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert p.returncode == 0, p.stdout
    assert open(os.path.join(tmp, "cli_dis.txt")).read() == open(os.path.join(tmp, "dos_dis.txt")).read()


@pytest.mark.parametrize("multi", [False, True])
def test_dropin_files_interactions(oracle, fx, tmp_path, multi):
    from omicabelnomm_amd import Algorithm, Outputs
    tmp = str(tmp_path)
    _write_inputs(fx, tmp, ["INTR"])
    s = _settings(tmp, "inter")
    s.storeBin = False
    s.use_interactions = True
    s.use_multiple_interaction_sets = multi
    s.keep_depVar = True
    s.fname_interactions = os.path.join(tmp, "INTR")
    out = Outputs()
    Algorithm().solve(s, out)
    excl = open(os.path.join(GOLDEN, "exclude_individuals.txt")).read()
    d = prepare(fx, "INTR", multi=multi, keep_dep=True, excl_text=excl)
    o = oracle.scan(d["XL"], d["XR"], d["Y"], d["r"], real_n=d["realN"], threads=4, **CHECK_SETTINGS)
    assert out.total_sig_results == o["n_sig"]
    assert (s.r, s.m) == (d["r"], d["m"])
    rows = read_dis(os.path.join(tmp, "inter_dis.txt"))
    assert len(rows) > 0


def test_dropin_streams_several_genotype_blocks(oracle, tmp_path, monkeypatch):
    """1200 SNPs in genotype blocks of 500 (three uploads, the next one
    queued while the current one is scanned): the outputs equal a single
    oracle pass, in the reference's (trait block, SNP block) order."""
    from omicabelnomm_amd import Algorithm, Outputs, Settings, synth
    tmp = str(tmp_path)
    data = synth.make(n=300, l=3, m=1200, t=12, miss="mcar", miss_rate=0.05, planted=20, beta=0.5, seed=17)
    XL, XR, Y = data.XL, data.xr_matrix(), data.y_matrix()
    obs = ["id%d" % i for i in range(data.n)]
    databel.write(os.path.join(tmp, "XL"), XL.astype(np.float32), obs, ["c%d" % i for i in range(3)])
    databel.write(os.path.join(tmp, "XR"), XR.astype(np.float32), obs, ["snp%d" % i for i in range(1200)])
    databel.write(os.path.join(tmp, "Y"), Y.astype(np.float32), obs, ["y%d" % i for i in range(12)])
    monkeypatch.setenv("OMICABELNOMM_BLOCK_SNPS", "500")
    s = Settings()
    Algorithm().applyDefaultParams(s)
    s.fnameAL, s.fnameAR, s.fnameY = (os.path.join(tmp, x) for x in ("XL", "XR", "Y"))
    s.fnameOutFiles = os.path.join(tmp, "blk")
    s.minPstore, s.minPdisp, s.minR2store, s.minR2disp = 0.01, 0.01, -10.0, -10.0
    s.storeBin = True
    out = Outputs()
    Algorithm().solve(s, out)
    f32 = lambda a: a.astype(np.float32).astype(np.float64)
    o = oracle.scan(f32(XL), f32(XR), f32(Y), 1, min_p_store=0.01, min_p_disp=0.01, min_r2_store=-10.0,
                    threads=4)
    assert out.total_sig_results == o["n_sig"]
    rows = read_dis(os.path.join(tmp, "blk_dis.txt"))
    names = dict(y=["y%d" % i for i in range(12)], al=["c%d" % i for i in range(3)],
                 ar=["snp%d" % i for i in range(1200)])
    want = expected_dis_rows(o, names, 3, 1, 0.01, -10.0, data.n)
    assert [(r["phe"], r["snp"]) for r in rows] == [(w["phe"], w["snp"]) for w in want]
    recs = read_dbin(os.path.join(tmp, "blk_sto.dbin"), disp_cov=False, store_all=False, l=3, r=1)
    want_recs = expected_records(o, l=3, r=1)
    assert [(r["y"], r["ar_offset"], len(r["results"])) for r in recs] == \
        [(w[0], w[1], len(w[3])) for w in want_recs]
