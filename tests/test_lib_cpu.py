"""CPU checks of the C ABI library: it loads, exports every symbol that
include/oagpu.h declares, and reports the absence of a GPU as an error
instead of falling back to CPU code."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "oagpu.h")).read()
    return sorted(set(re.findall(r"\b(oagpu_[a-z_]+)\s*\(", txt)))


def test_header_declares_expected_entry_points():
    syms = declared_symbols()
    for s in ["oagpu_create", "oagpu_destroy", "oagpu_set_covariates", "oagpu_load_traits",
              "oagpu_load_genotypes", "oagpu_scan", "oagpu_last_error"]:
        assert s in syms


def test_library_exports_every_declared_symbol():
    from omicabelnomm_amd import lib
    L = lib.load()
    for s in declared_symbols():
        assert hasattr(L, s), s


def test_tvalue_matches_oracle(oracle):
    from omicabelnomm_amd import tvalue
    for p in [0.1, 0.05, 5e-5, 1e-8]:
        assert tvalue(p) == oracle.tvalue(p)


def test_no_gpu_is_an_error_not_a_fallback():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from omicabelnomm_amd import Scanner, OagpuError
    with pytest.raises(OagpuError):
        Scanner(100, 3, 1, max_traits=4)
