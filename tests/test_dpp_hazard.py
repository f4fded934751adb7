"""The missing-row corrections broadcast covariates with DPP row_newbcast
multiply-adds written as inline asm (omicabelnomm_amd/csrc/
oagpu_kernels_corr.hip).  On CDNA a DPP read of a VGPR needs two wait
states after a VALU instruction wrote it, and the compiler does not see
inside the asm.  This test compiles the kernels for gfx950 to assembly and
checks that no DPP source was written by a VALU instruction within the two
instructions before it."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc"


def _vregs(op):
    m = re.match(r"v\[(\d+):(\d+)\]", op)
    if m:
        return set(range(int(m.group(1)), int(m.group(2)) + 1))
    m = re.match(r"v(\d+)$", op)
    return {int(m.group(1))} if m else set()


def dpp_hazards(asm_lines):
    """(number of DPP multiply-adds, list of hazards) in gfx9 assembly."""
    ins = [l.strip() for l in asm_lines]
    ins = [l for l in ins if l and not l.startswith((";", ".")) and not l.endswith(":")]
    n, bad = 0, []
    for i, l in enumerate(ins):
        if not l.startswith("v_fmac_f64_dpp"):
            continue
        n += 1
        src0 = _vregs(l.split(None, 1)[1].split(",")[1].strip().split()[0])
        waits, k = 0, i - 1
        while k >= 0 and waits < 2:
            p = ins[k]
            name = p.split()[0]
            if name == "s_nop":
                waits += int(p.split()[1]) + 1
            else:
                if name.startswith("v_") and not name.startswith(("v_readlane", "v_readfirstlane")):
                    dst = p.split(None, 1)[1].split(",")[0].strip() if " " in p else ""
                    if _vregs(dst) & src0:
                        bad.append((p, l))
                        break
                waits += 1
            k -= 1
    return n, bad


def test_hazard_scan_finds_a_planted_hazard():
    ok = ["ds_read2_b64 v[38:41], v34 offset1:6", "s_waitcnt lgkmcnt(0)",
          "v_fmac_f64_dpp v[2:3], v[38:39], v[20:21] row_newbcast:1 row_mask:0xf bank_mask:0xf"]
    assert dpp_hazards(ok) == (1, [])
    bad = ["v_mov_b64_e32 v[38:39], v[50:51]",
           "v_fmac_f64_dpp v[2:3], v[38:39], v[20:21] row_newbcast:1 row_mask:0xf bank_mask:0xf"]
    assert len(dpp_hazards(bad)[1]) == 1
    padded = ["v_mov_b64_e32 v[38:39], v[50:51]", "s_nop 1",
              "v_fmac_f64_dpp v[2:3], v[38:39], v[20:21] row_newbcast:1 row_mask:0xf bank_mask:0xf"]
    assert dpp_hazards(padded)[1] == []


def test_corrections_dpp_sources_are_not_fresh_valu_writes(tmp_path):
    if not os.path.exists(HIPCC):
        pytest.skip("no hipcc")
    out = tmp_path / "corr.s"
    cmd = [HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", "-I" + os.path.join(ROOT, "include"), "--cuda-device-only", "-S", "-o", str(out),
           os.path.join(ROOT, "omicabelnomm_amd", "csrc", "oagpu_kernels_corr.hip")]
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, timeout=600)
    n, bad = dpp_hazards(open(out).read().splitlines())
    assert n > 1000, n
    assert bad == [], bad[:3]
