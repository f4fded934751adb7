"""Error behaviour of the C ABI on the GPU.  Where the reference prints a
message and exits (src/Algorithm.cpp:497-527 singular covariates; its size
and order checks in src/AIOwrapper.cpp:87-145), the library returns a status
code with a message, and the context stays usable: a correct call after the
error gives the oracle's answer.  Every case here is refused by host-side
checks before any kernel runs, except the hit overflow, which runs the scan
and reports the needed capacity."""
import ctypes

import numpy as np
import pytest

from omicabelnomm_amd import OagpuError, Scanner, synth
from omicabelnomm_amd import lib as L
from parity_util import compare_fp64, near_singular, run_oracle

pytestmark = pytest.mark.gpu


def _code(fn, *a, **k):
    with pytest.raises(OagpuError) as e:
        fn(*a, **k)
    return e.value.code


def test_create_refuses_bad_shapes():
    print("case create errors", flush=True)
    assert _code(Scanner, 100, 40, 1, max_traits=8) == L.ERR_UNSUPPORTED        # l beyond the compiled buckets
    assert _code(Scanner, 100, 3, 5, max_traits=8) == L.ERR_UNSUPPORTED         # r beyond 4
    assert _code(Scanner, 100, 3, 1, precision=7, max_traits=8) == L.ERR_ARG
    assert _code(Scanner, 100, 3, 1, max_traits=0) == L.ERR_ARG
    assert _code(Scanner, 100, 3, 1, max_traits=8, device=99) == L.ERR_ARG


def test_call_order_and_bad_blocks_then_a_correct_scan(oracle):
    data = synth.make(n=300, l=3, m=25, t=20, miss="mcar", miss_rate=0.05, planted=5, beta=0.4, seed=71)
    Y = data.Y.cpu().numpy().T
    G = data.G.cpu().numpy().T
    print("case call order and bad blocks", flush=True)
    with Scanner(data.n, data.l, data.r, "fp64", min_p_store=0.1, min_p_disp=0.05, min_r2_store=-10,
                 max_traits=32) as sc:
        assert _code(sc.load_traits, Y) == L.ERR_STATE                          # covariates first
        assert _code(sc.scan) == L.ERR_STATE                                    # nothing loaded
        # linearly dependent covariates: the reference's rank check exits
        XLdep = np.column_stack([data.XL, data.XL[:, 1] * 2.0])[:, [0, 1, 3]]
        assert _code(sc.set_covariates, XLdep) == L.ERR_COVARIATES
        XLnan = data.XL.copy()
        XLnan[5, 1] = np.nan
        assert _code(sc.set_covariates, XLnan) == L.ERR_ARG
        sc.set_covariates(data.XL)
        assert _code(sc.load_traits, np.zeros((data.n, 33))) == L.ERR_ARG      # more than max_traits
        sc.load_traits(Y)
        assert _code(sc.load_genotypes, G, m=2, col_src=[0, 99]) == L.ERR_ARG   # descriptor out of range
        assert _code(sc.load_genotypes, G[:, :5], m=6) == L.ERR_ARG             # more SNPs than columns
        assert _code(sc.set_dosage_model, [1.0, 1.0, 1.0], 2) == L.ERR_ARG      # more factors than columns
        L_ = L.load()
        st = L.Stats()
        rc = L_.oagpu_scan_wait(sc._ctx, None, None, 0, ctypes.byref(st))
        assert rc == L.ERR_STATE                                                # no scan in flight
        sc.load_genotypes(G, m=data.m)
        g = sc.scan(dense=True)
    compare_fp64(g, run_oracle(oracle, data, min_r2_store=-10))


def test_hit_overflow_reports_the_needed_capacity(oracle):
    """A host buffer smaller than the hits: the call fails with the count it
    needed, and the same loaded blocks scan again into a large enough one."""
    data = synth.make(n=300, l=3, m=30, t=24, miss="none", planted=10, beta=0.5, seed=73)
    print("case hit overflow", flush=True)
    with Scanner(data.n, data.l, data.r, "fp64", min_p_store=0.5, min_p_disp=0.05, min_r2_store=-10,
                 max_traits=32) as sc:
        sc.set_covariates(data.XL)
        sc.load_traits(data.Y.cpu().numpy().T)
        sc.load_genotypes(data.G.cpu().numpy().T, m=data.m)
        with pytest.raises(OagpuError) as e:
            sc.scan(hit_cap=5)
        assert e.value.code == L.ERR_HIT_OVERFLOW
        g = sc.scan(dense=True)
    o = run_oracle(oracle, data, min_p_store=0.5, min_r2_store=-10)
    assert g.stats["n_hits"] == o["n_kept"] > 5
    assert "'n_hits': %d" % o["n_kept"] in str(e.value)
    compare_fp64(g, o)


@pytest.mark.parametrize("smin", [0.8e-6, 1.25e-6])
def test_rank_check_resolves_the_threshold(smin):
    """The covariate rank check (the reference's sgesvd test: a singular
    value below 1e-6 means linearly dependent columns) at n = 10,000, with
    sigma_min 25 % either side of the threshold:
    - sigma_max = 100: the rank check must refuse 0.8e-6 and pass 1.25e-6,
      which the Gram matrix's FP64 Cholesky then refuses on its own
      (condition number 8e7, squared 6e15); an eigenvalue test on XL'XL
      could not tell the two apart (eigenvalues near 1e-12 under the
      rounding of 1e4), so the messages show which test refused;
    - sigma_max = 3: 1.25e-6 is accepted and the context scans."""
    XL = near_singular(10_000, [100.0, 30.0, 5.0, 1.0, smin], seed=9)
    print("case rank check", smin, flush=True)
    with Scanner(10_000, 5, 1, "fp64", max_traits=4) as sc:
        with pytest.raises(OagpuError) as e:
            sc.set_covariates(XL)
        assert e.value.code == L.ERR_COVARIATES
        assert ("linearly dependent" in str(e.value)) == (smin < 1e-6), str(e.value)
    XL = near_singular(10_000, [3.0, 1.0, 0.3, 0.05, smin], seed=10)
    with Scanner(10_000, 5, 1, "fp64", max_traits=4) as sc:
        if smin < 1e-6:
            assert _code(sc.set_covariates, XL) == L.ERR_COVARIATES
        else:
            sc.set_covariates(XL)
