"""Build liboagpu.so (HIP, gfx950) and the omicabelnomm drop-in CLI.

Run as ``python -m omicabelnomm_amd.build`` or through
``__graft_entry__.build()``.  Objects go to omicabelnomm_amd/build/, the
shared library and the CLI binary next to this file.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "build")
LIB = os.path.join(HERE, "liboagpu.so")
HOSTLIB = os.path.join(HERE, "libomicabelnomm.so")
CLI = os.path.join(HERE, "omicabelnomm")

HIP_SOURCES = ["oagpu_api.hip", "oagpu_kernels_trait.hip", "oagpu_kernels_prep.hip", "oagpu_kernels_snp.hip",
               "oagpu_kernels_corr.hip", "oagpu_kernels_pst.hip", "oagpu_kernels_epi.hip", "oagpu_gemm.hip",
               "oagpu_gemm_lib.hip", "oagpu_gemm_fused.hip", "oagpu_serve.hip", "oagpu_noalloc.hip"]
# library GEMM candidates of the contraction autotuner
HIP_LIBS = ["-L/opt/rocm/lib", "-lhipblaslt", "-Wl,-rpath,/opt/rocm/lib"]
HOST_SOURCES = ["host/databel.cpp", "host/writers.cpp", "host/driver.cpp"]
CLI_SOURCES = ["host/main.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-I" + os.path.join(ROOT, "include"),
]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(ln for ln in p.stdout.splitlines()
                    if "packed-fp32-ops' is not a recognized feature" not in ln)
    if p.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), out))
    return out


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src).replace(".hip", ".o").replace(".cpp", ".o"))
    _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
    return obj


def build(verbose: bool = True, cli: bool = True) -> str:
    os.makedirs(OBJ, exist_ok=True)
    jobs = min(8, os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, HIP_SOURCES))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + HIP_LIBS)
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    if cli:
        inc = "-I" + os.path.join(ROOT, "include")

        def host_obj(src):
            o = os.path.join(OBJ, "host_" + os.path.basename(src).replace(".cpp", ".o"))
            _run(["g++", "-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter", inc,
                  "-c", os.path.join(CSRC, src), "-o", o])
            return o

        with ThreadPoolExecutor(max_workers=jobs) as ex:
            hobjs = list(ex.map(host_obj, HOST_SOURCES + CLI_SOURCES))
        tmp = HOSTLIB + ".tmp"
        _run(["g++", "-shared", "-o", tmp] + hobjs[:len(HOST_SOURCES)] +
             ["-L" + HERE, "-loagpu", "-Wl,-rpath,$ORIGIN", "-lpthread"])
        os.replace(tmp, HOSTLIB)
        tmpc = CLI + ".tmp"
        _run(["g++", "-o", tmpc] + hobjs[len(HOST_SOURCES):] +
             ["-L" + HERE, "-lomicabelnomm", "-loagpu", "-Wl,-rpath,$ORIGIN", "-lpthread"])
        os.replace(tmpc, CLI)
        if verbose:
            print("built", HOSTLIB, "and", CLI)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
