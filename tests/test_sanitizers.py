"""The library's host code under AddressSanitizer and UndefinedBehaviorSanitizer
(CPU only). tests/sanitize/host_check.cpp is compiled with g++
-fsanitize=address,undefined together with the library's own host sources
(the structured-product front end, the launch-plan builder; the correlation
repair is header-only) and driven with:

* every product of the test corpus plus seeded mutations of each
  (truncations, byte substitutions, deleted and duplicated lines) -- the
  front end must report either a compiled product or a DslError, never a
  sanitizer finding, and agree line for line with the shipped library's
  ofg_dsl_compile on the same text;
* seeded correlation matrices (PSD, indefinite, non-finite, asymmetric);
* seeded option batches through the pricing plan builder, valid and invalid,
  with the work-queue invariants checked.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "tests", "sanitize")
EXE = os.path.join(SAN, "host_check")
CSRC = os.path.join(ROOT, "openferric_amd", "csrc")
SOURCES = [os.path.join(SAN, "host_check.cpp"), os.path.join(CSRC, "ofg_dsl_front.cpp"),
           os.path.join(CSRC, "ofg_plan.cpp")]
HEADERS = ["ofg_dsl_front.hpp", "ofg_dsl.hpp", "ofg_plan.hpp", "ofg_corr.hpp", "ofg_host.hpp",
           "ofg_internal.hpp", "ofg_math.hpp"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:exitcode=86",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1:exitcode=87")


def _build():
    deps = SOURCES + [os.path.join(CSRC, h) for h in HEADERS]
    if os.path.exists(EXE) and all(os.path.getmtime(d) <= os.path.getmtime(EXE) for d in deps):
        return EXE
    # per-process output name: parallel test workers may build at once
    tmp = f"{EXE}.tmp{os.getpid()}"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", *SOURCES, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    os.replace(tmp, EXE)
    return EXE


@pytest.fixture(scope="module")
def exe():
    return _build()


def _run(exe, *args):
    r = subprocess.run([exe, *args], capture_output=True, text=True, env=ENV, timeout=600)
    assert r.returncode == 0, (r.returncode, r.stderr[-6000:])
    assert "runtime error" not in r.stderr and "Sanitizer" not in r.stderr, r.stderr[-6000:]
    return r.stdout


def _corpus():
    from tests import dsl_products as P
    import bench
    base = [v for k, v in vars(P).items() if isinstance(v, str) and v.startswith("product")]
    base += [P.standard_autocall(3, 1.0, 0.08, 1.5), P.phoenix_memory(4, 1.0, 0.7, 0.06, 3.0),
             bench.DSL_PRODUCT]
    rng = np.random.default_rng(20261204)
    subs = list("():\n \t#\"x1.-_=<>*/") + ["é", "if", "then", "else", "0.0.1", "1e999"]
    out = list(base)
    for src in base:
        n = len(src)
        for cut in np.linspace(1, n - 1, 24).astype(int):
            out.append(src[:cut])
        for _ in range(40):
            i = int(rng.integers(0, n))
            out.append(src[:i] + subs[int(rng.integers(0, len(subs)))] + src[i + 1:])
        lines = src.split("\n")
        for _ in range(12):
            j = int(rng.integers(0, len(lines)))
            out.append("\n".join(lines[:j] + lines[j + 1:]))
            out.append("\n".join(lines[:j + 1] + [lines[j]] + lines[j + 1:]))
    return out


def _library_outcome(L, src):
    h = C.c_void_p()
    rc = L.ofg_dsl_compile(src.encode(), C.byref(h))
    if rc == 0:
        L.ofg_dsl_free(h)
        return "ok"
    return "err " + L.ofg_last_error().decode()


def test_front_end_under_sanitizers_matches_library(exe, tmp_path):
    from openferric_amd import _lib
    L = _lib.load()
    corpus = _corpus()
    files = []
    for i, src in enumerate(corpus):
        p = tmp_path / f"p{i:05d}.dsl"
        p.write_bytes(src.encode())
        files.append(str(p))
    lines = _run(exe, "dsl", *files).splitlines()
    assert len(lines) == len(corpus)
    n_ok = 0
    for src, line in zip(corpus, lines):
        lib = _library_outcome(L, src)
        if line.startswith("ok"):
            n_ok += 1
            assert lib == "ok", (src, lib)
        else:
            assert line == lib, (src, line, lib)
    # the corpus reaches both outcomes in numbers
    assert n_ok >= 20 and len(lines) - n_ok >= 200, n_ok


def test_correlation_repair_under_sanitizers(exe):
    out = _run(exe, "corr").split()
    repaired, refused, factored = int(out[2]), int(out[4]), int(out[6])
    assert repaired > 0 and refused > 0 and factored > 0


def test_plan_builder_under_sanitizers(exe):
    out = _run(exe, "plan").split()
    built, refused = int(out[2]), int(out[4])
    assert built > 50 and refused > 0
