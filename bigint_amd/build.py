"""Compiles the HIP engine for gfx950 into bigint_amd/libbigint_hip.so (in-tree).

Every HIP source is built with -Xclang -target-feature -Xclang -packed-fp32-ops so that no
packed FP32 instructions are emitted.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libbigint_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++20", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + os.path.join(ROOT, "include"), "-Wall"]
SOURCES = ["ntt_kernels.hip", "add_kernels.hip", "planner_kernels.hip", "engine.cpp"]
HOST_SOURCES = ["host/mul.cpp", "host/mul_backend_hip.cpp", "host/bigint.cpp", "host/device_int.cpp"]
HOST_LIB = os.path.join(HERE, "libbigint.so")
RATES_LIB = os.path.join(HERE, "libbigint_rates.so")
CXX = os.environ.get("BIGINT_CXX", "/opt/rocm/llvm/bin/clang++")
HOST_FLAGS = ["-O2", "-std=gnu++20", "-fPIC", "-Wall", "-Wextra", "-I" + os.path.join(ROOT, "include"),
              "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd))


def build(verbose: bool = False) -> str:
    """Compile every HIP source and link the shared library; returns its path."""
    objs = []
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    for src in SOURCES:
        obj = os.path.join(bdir, src + ".o")
        lang = ["-x", "hip"] if src.endswith(".cpp") else []
        cmd = [HIPCC] + HIP_FLAGS + lang + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        _run(cmd)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    # issue-rate microbenchmark of the integer VALU forms (measurement support, bench.py --valu-rates)
    tmp = RATES_LIB + ".tmp"
    _run([HIPCC] + HIP_FLAGS + ["-shared", os.path.join(CSRC, "valu_rates.hip"), "-o", tmp])
    os.replace(tmp, RATES_LIB)
    # host C++ library (BigInt/BigFloat API) on top of the C ABI
    hobjs = []
    for src in HOST_SOURCES:
        obj = os.path.join(bdir, os.path.basename(src) + ".o")
        cmd = [CXX] + HOST_FLAGS + ["-c", os.path.join(HERE, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        _run(cmd)
        hobjs.append(obj)
    tmp = HOST_LIB + ".tmp"
    _run([CXX, "-shared", "-fPIC", "-o", tmp] + hobjs + ["-L" + HERE, "-lbigint_hip",
                                                        "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, HOST_LIB)
    # test driver of the host API
    calc = os.path.join(bdir, "bigint_calc")
    _run([CXX] + HOST_FLAGS + [os.path.join(ROOT, "tests", "cpp", "bigint_calc.cpp"), "-o", calc,
                               "-L" + HERE, "-lbigint", "-lbigint_hip",
                               "-Wl,-rpath," + HERE, "-Wl,-rpath,$ORIGIN/.."])
    # host schoolbook against the engine, end to end (the host route's size limit)
    _run([CXX] + HOST_FLAGS + [os.path.join(HERE, "examples", "crossover.cpp"), "-o",
                               os.path.join(bdir, "crossover"), "-L" + HERE, "-lbigint", "-lbigint_hip",
                               "-Wl,-rpath," + HERE, "-Wl,-rpath,$ORIGIN/.."])
    # π example on the engine
    _run([CXX] + HOST_FLAGS + [os.path.join(HERE, "examples", "pi.cpp"), "-o", os.path.join(bdir, "pi"),
                               "-L" + HERE, "-lbigint", "-lbigint_hip", "-Wl,-rpath," + HERE,
                               "-Wl,-rpath,$ORIGIN/.."])
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
