"""The reference's own known answers on its shipped interaction fixtures
(tests/test.cpp:201-308): lower bounds on the number of significant
results per scenario, checked on the FP64 oracle (CPU)."""
import os

import pytest

from refmodel import CHECK_SETTINGS, GOLDEN, load_fixture, prepare, scenarios


@pytest.fixture(scope="module")
def fx():
    return load_fixture()


@pytest.fixture(scope="module")
def excl():
    return open(os.path.join(GOLDEN, "exclude_individuals.txt")).read()


@pytest.mark.parametrize("multi,keep,inter,expected", scenarios())
def test_known_answer_counts(oracle, fx, excl, multi, keep, inter, expected):
    d = prepare(fx, inter, multi=multi, keep_dep=keep, excl_text=excl)
    o = oracle.scan(d["XL"], d["XR"], d["Y"], d["r"], real_n=d["realN"], threads=4, **CHECK_SETTINGS)
    assert o["status"] == 0
    assert o["n_sig"] >= expected, (inter, multi, keep, o["n_sig"], expected)
