"""Builds libopenferric_amd.so (HIP kernels + C ABI) for gfx950 with hipcc.

Run `python -m openferric_amd.build` or call build(). The library is written
in-tree next to this file so it travels with the repository snapshot.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libopenferric_amd.so")
SOURCES = ["ofg_kernels.hip", "ofg_adjoint.hip", "ofg_mcgreeks.hip", "ofg_pathwise.hip",
           "ofg_lsm.hip", "ofg_qmc.hip", "ofg_multi.hip", "ofg_dsl.hip", "ofg_api.cpp",
           "ofg_lsm_api.cpp", "ofg_qmc_api.cpp", "ofg_multi_api.cpp", "ofg_basket_api.cpp",
           "ofg_dsl_front.cpp", "ofg_dsl_api.cpp", "ofg_exact.hip", "ofg_exact_api.cpp",
           "ofg_spread.hip", "ofg_spread_api.cpp", "ofg_plan.cpp", "ofg_normals.hip", "ofg_device_io.hip",
           "ofg_dsl_jit.cpp"]
HEADERS = ["ofg_math.hpp", "ofg_internal.hpp", "ofg_adjoint.hpp", "ofg_paths.hpp",
           "ofg_mcgreeks.hpp", "ofg_pathwise.hpp", "ofg_lsm.hpp", "ofg_qmc.hpp", "ofg_multi.hpp",
           "ofg_host.hpp", "ofg_jump.hpp", "ofg_corr.hpp", "ofg_dsl.hpp", "ofg_dsl_front.hpp",
           "ofg_fast_tables.inc", "ofg_exact.hpp", "ofg_spread.hpp", "ofg_plan.hpp", "ofg_normals.hpp", "ofg_device_io.hpp",
           "ofg_dsl_jit.hpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: the reference never fuses a*b+c implicitly, so neither do
# we (explicit fma() calls only). -packed-fp32-ops off: packed FP32 forms are
# not used. -fmacf64-inst off: every FP64 fma is the three-address v_fma_f64;
# the two-address v_fmac_f64 form makes the register allocator copy a
# loop-invariant addend (a polynomial coefficient, a tangent input) into the
# destination first (a v_mov_b64 per use): measured 5 % faster on the
# reference-draw pricing loop and 3 % on the Heston forward tangent, the
# fast-draw loops unchanged. O3 + gfx950 only.
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Xclang", "-target-feature", "-Xclang", "-fmacf64-inst",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]


def _compile(src: str, obj: str, extra=()):
    cmd = [HIPCC, "-x", "hip", *FLAGS, *extra, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed for {src}")
    return obj


def _stale(src: str, obj: str) -> bool:
    if not os.path.exists(obj):
        return True
    deps = [os.path.join(CSRC, src)] + [os.path.join(CSRC, h) for h in HEADERS] + \
        [os.path.join(os.path.dirname(HERE), "include", "openferric_amd.h")]
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(d) > t for d in deps if os.path.exists(d))


def build(verbose: bool = False, incremental: bool = False) -> str:
    """Compiles every HIP/C++ source and links the shared library
    (incremental: only sources older than their object or any header)."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = [os.path.join(objdir, src.rsplit(".", 1)[0] + ".o") for src in SOURCES]
    todo = [(s, o) for s, o in zip(SOURCES, objs) if not incremental or _stale(s, o)]

    def one(job):
        if verbose:
            print("hipcc", job[0], flush=True)
        return _compile(*job)

    # one hipcc per source, as many at once as the host has cores (at most 16)
    jobs = max(1, min(16, os.cpu_count() or 1, len(todo) or 1))
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        list(pool.map(one, todo))
    tmp = LIB + ".tmp"
    # hipRTC (the structured-product engine's run-time compiler) is opened on
    # first use, not linked: the library loads where it is absent
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-ldl"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, LIB)
    return LIB


def build_cpp_tests(verbose: bool = False) -> str:
    """Builds tests/cpp/test_engine (the reference's engine tests on the C++
    mirror of the pricing contract), linked against the library."""
    root = os.path.dirname(HERE)
    src = os.path.join(root, "tests", "cpp", "test_engine.cpp")
    exe = os.path.join(root, "tests", "cpp", "test_engine")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(root, "include"), src,
           "-L", HERE, "-lopenferric_amd", "-Wl,-rpath,$ORIGIN/../../openferric_amd", "-o", exe]
    if verbose:
        print("g++ test_engine.cpp", flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("C++ mirror test build failed")
    return exe


def build_math_shim(verbose: bool = False) -> str:
    """Builds tests/cpu_shim/libofg_math_shim.so: a host-only build of the
    device math header, used by the tests to check the kernels' own math on
    the CPU and to recompute fast-mode paths for the GPU parity tests."""
    root = os.path.dirname(HERE)
    src = os.path.join(root, "tests", "cpu_shim", "math_shim.hip")
    out = os.path.join(root, "tests", "cpu_shim", "libofg_math_shim.so")
    cmd = [HIPCC, "-x", "hip", "--offload-host-only", "-O2", "-ffp-contract=off", "-fPIC",
           "-shared", "-o", out + ".tmp", src]
    if verbose:
        print("hipcc (host only) math_shim.hip", flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("host build of ofg_math.hpp failed")
    os.replace(out + ".tmp", out)
    return out


def dump_asm(src: str = "ofg_kernels.hip", out: str | None = None) -> str:
    """Writes the gfx950 assembly of one source (for instruction counts)."""
    out = out or os.path.join(HERE, "_build", src + ".s")
    cmd = [HIPCC, "-x", "hip", *FLAGS, "--offload-device-only", "-S",
           os.path.join(CSRC, src), "-o", out]
    subprocess.run(cmd, check=True)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_cpp_tests(verbose=True))
    print(build_math_shim(verbose=True))
