"""Genotype blocks loaded one after another into one scan context.

The context alternates two upload slots for the genotype block and its
column descriptors (src/AIOwrapper.cpp:767-840 builds the interaction
columns of every block).  Every scan of the sequence must equal a fresh
context's scan of the same block: a slot loaded again with its previous
descriptors, with other descriptors of the same length, with a pair layout
([E*G, G], the fused contraction's) and with a non-pair one.
"""
import numpy as np
import pytest
import torch

from omicabelnomm_amd import Scanner, synth

pytestmark = pytest.mark.gpu

M_BLOCK = 48


def _layouts():
    a = np.arange(M_BLOCK, dtype=np.int32)
    pair_env = np.tile(np.array([0, -1], np.int32), M_BLOCK)
    return {
        "A": (np.repeat(a, 2), pair_env),                      # [E*G, G] of columns 0..47
        "B": (np.repeat(a + M_BLOCK, 2), pair_env),            # columns 48..95
        "C": (np.repeat(a[::-1].copy(), 2), pair_env),         # 47..0: A's length, other values
        "D": (np.repeat(a, 2), np.tile(np.array([-1, 0], np.int32), M_BLOCK)),   # [G, E*G]
    }


def _scanner(data, precision):
    sc = Scanner(data.n, data.l, data.r, precision, min_p_store=0.5, min_p_disp=0.5, min_r2_store=-10.0,
                 max_traits=64, max_hits=1 << 16)
    sc.set_covariates(data.XL)
    sc.set_interactions(data.E)
    sc.load_traits(data.Y.to("cuda").contiguous())
    return sc


def _scan(sc, G, src, env):
    sc.load_genotypes(G, m=src.size // 2, col_src=src, col_env=env)
    return sc.scan(dense=True)


def _same(a, b, what):
    for k in ("beta", "se", "pval", "r2", "kept", "emitted"):
        np.testing.assert_array_equal(a.dense[k], b.dense[k], err_msg=f"{what}: {k}")
    np.testing.assert_array_equal(a.hits, b.hits, err_msg=f"{what}: hits")
    assert a.stats["n_hits"] == b.stats["n_hits"], what


@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
def test_descriptor_reuse_across_slots(precision):
    data = synth.make(n=1500, l=3, m=2 * M_BLOCK, t=40, interaction=True, miss="groups", pattern_groups=4,
                      pattern_rate=0.03, planted=12, beta=0.3, seed=31)
    G = data.G.to("cuda").contiguous()
    lay = _layouts()
    seq = ["A", "B", "A", "C", "D", "C"]   # slots 0, 1, 0, 1, 0, 1
    expected = {}
    for name in set(seq):
        with _scanner(data, precision) as fresh:
            expected[name] = _scan(fresh, G, *lay[name])
    with _scanner(data, precision) as sc:
        for i, name in enumerate(seq):
            print(f"descriptor reuse: {precision} block {i} layout {name}", flush=True)
            _same(_scan(sc, G, *lay[name]), expected[name], f"block {i} ({name})")
    # the layouts differ in their results, so a kept stale copy would show
    assert not np.array_equal(expected["A"].dense["beta"], expected["C"].dense["beta"])
    assert not np.array_equal(expected["A"].dense["beta"], expected["D"].dense["beta"])
