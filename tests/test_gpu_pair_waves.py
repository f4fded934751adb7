"""The shared-pattern pair kernel over its trait waves, against the FP64
oracle.

When missing-phenotype patterns are few (4 G <= t) the pairs are finished
from a per-(SNP, pattern) table.  The traits of a pattern with at least 32
traits fill waves of their own (the table rows are wave-uniform), the others
share mixed waves; each wave runs its traits over a run of SNPs, first
solving every pair for the SNP terms and the RSS only, then finishing in full
the SNPs at which some trait needs more (a record, the explicit residual).
These tests scan without dense output, so that first pass decides what is
reported, and compare the hit records with the oracle's kept records:

- pattern sizes that give only own waves, only mixed waves, and both, with
  the traits of each pattern contiguous or interleaved;
- run lengths 1, 7 and 64 (OAGPU_PST_RUN) over 150 SNPs: full, partial and
  single-SNP runs;
- a trait with observed zeros and one with a near-perfect fit, which take
  the explicit residual from the first pass;
- one and two columns per SNP; FP64 and FP32 contraction outputs (the
  near-perfect fit is held to the FP64 bars only).
"""
import numpy as np
import pytest
import torch

from omicabelnomm_amd import synth
from parity_util import assert_fp64_records, assert_mixed_records, oracle_records, run_gpu, run_oracle

pytestmark = pytest.mark.gpu

THR = dict(min_p_store=0.05, min_p_disp=0.05, min_r2_store=-10.0)

LAYOUTS = {
    # name: (pattern sizes, interleaved traits)
    "own_waves": ([64, 128, 37, 45], False),
    "mixed_waves": ([20] * 12, False),
    "both_interleaved": ([70, 40, 31, 33, 5, 3, 2, 100], True),
}


def patterned(sizes, interleave, r, seed):
    """Traits in len(sizes) missing-row patterns of the given sizes (each
    pattern blanks its own 3% of rows), plus an observed-zero trait and a
    near-perfect fit on SNP 3 inside the first pattern."""
    t = int(sum(sizes))
    n, m = 420, 150
    data = synth.make(n=n, l=4, m=m, t=t, interaction=(r == 2), miss="none", planted=12, seed=seed)
    rng = np.random.default_rng(seed)
    order = rng.permutation(t) if interleave else np.arange(t)
    Y = data.Y.clone()
    src3 = int(data.col_src[3 * r + r - 1]) if data.col_src is not None else 3 * r
    Y[order[1]] = 0.5 * data.G[src3].double() + 1e-6 * torch.randn(n, dtype=torch.float64)
    Y[order[0], :30] = 0.0
    pos = 0
    for sz in sizes:
        rows = torch.as_tensor(rng.choice(n, size=int(0.03 * n), replace=False))
        ys = torch.as_tensor(order[pos:pos + sz])
        Y[ys[:, None], rows[None, :]] = float("nan")
        pos += sz
    data.Y = Y
    return data, int(order[1])


@pytest.mark.parametrize("run", ["1", "7", "64"])
@pytest.mark.parametrize("r", [1, 2])
@pytest.mark.parametrize("layout", list(LAYOUTS))
def test_trait_waves_match_oracle(layout, r, run, oracle_mod, monkeypatch):
    sizes, inter = LAYOUTS[layout]
    data, perfect = patterned(sizes, inter, r, seed=17 + r)
    assert 4 * len(sizes) <= data.t             # the table path
    print("case waves", layout, "r", r, "run", run, "t", data.t, flush=True)
    monkeypatch.setenv("OAGPU_PST_RUN", run)
    o = run_oracle(oracle_mod, data, **THR)
    rec = oracle_records(o)
    assert len(rec["trait"]) > 100
    g = run_gpu(data, "fp64", dense=False, **THR)
    print(layout, r, run, "records", len(rec["trait"]), g.stats, flush=True)
    assert g.stats["n_fallback"] >= data.m       # the observed-zero trait at every SNP
    assert_fp64_records(g.hits, rec)
    # FP32 contraction outputs: the near-perfect fit's SE (~1e-8) is below
    # their rounding, so that trait is left out of the mixed-precision bars
    gm = run_gpu(data, "fp32", dense=False, **THR)
    keep = rec["trait"] != perfect
    assert_mixed_records(gm.hits[gm.hits["trait"] != perfect], {k: v[keep] for k, v in rec.items()},
                         "%s r%d run %s fp32" % (layout, r, run))


@pytest.fixture(scope="module")
def oracle_mod(oracle):
    return oracle
