"""The reference's per-sample normal streams drawn on the GPU
(ofg_fill_normals) against the oracle's restatement of the same streams
(or_sample_normals: resolve_stream_seed, SplitMix64 seeding, xoshiro256++
or Pcg64, uniform_open01, Acklam's inverse CDF).

The draw code is the pricing kernels' (acklam_block_wave: central rational
per lane, tail values compacted per wave and inverted 64 at a time, the
division as the compiler's Newton-Markstein sequence). Central-region
values (p in [0.02425, 0.97575]) use only +, *, fma and that division, so
they must equal the oracle bit for bit. Tail values also take a log and a
sqrt, where the device library and the host's libm may round differently:
they are held to 4 ulp and the number of bitwise-equal ones is reported.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [(42, 0, 3000, 253), (2**64 - 5, 49_990, 130, 64), (7, 123, 1, 400)]


def _oracle_rows(oracle, seed, first, n_streams, n):
    return np.stack([oracle.sample_normals(seed, first + i, n, 1)[0] for i in range(n_streams)])


def _compare(got, ref):
    # tail region p < 0.02425 or p > 0.97575 is |z| > 1.9728; the margin keeps
    # every tail value on this side (a few central ones join them)
    tail = np.abs(ref) > 1.97
    central = ~tail
    assert np.array_equal(got[central], ref[central]), \
        f"{np.count_nonzero(got[central] != ref[central])} central values differ"
    ulp = np.abs(got[tail] - ref[tail]) / np.spacing(np.abs(ref[tail]))
    assert ulp.max(initial=0) <= 4, ulp.max()
    return int(np.count_nonzero(tail)), int(np.count_nonzero(got[tail] == ref[tail]))


@pytest.mark.parametrize("seed,first,n_streams,n", CASES)
def test_reference_streams_match_oracle(oracle, gpu, seed, first, n_streams, n):
    from openferric_amd import fill_normals
    got = fill_normals(seed, n_streams, n, first_stream=first)
    ref = _oracle_rows(oracle, seed, first, n_streams, n)
    n_tail, n_equal = _compare(got, ref)
    print(f"seed {seed}: {n_tail} tail values, {n_equal} bitwise equal")


def test_pcg64_streams_match_oracle(oracle, gpu):
    from openferric_amd import RngMode, fill_normals
    got = fill_normals(99, 700, 41, first_stream=3, rng_mode=RngMode.Pcg64)
    with oracle.rng_kind(oracle.RNG_PCG64):
        ref = _oracle_rows(oracle, 99, 3, 700, 41)
    _compare(got, ref)


def test_layouts_and_offsets(gpu):
    from openferric_amd import fill_normals
    a = fill_normals(5, 200, 37)
    b = fill_normals(5, 200, 37, step_major=True)
    assert np.array_equal(a.T, b)
    c = fill_normals(5, 150, 37, first_stream=50)
    assert np.array_equal(a[50:], c)
    assert fill_normals(5, 0, 10).shape == (0, 10)


def test_argument_errors(gpu):
    import ctypes as C
    from openferric_amd import InvalidInput, RngMode, _lib, fill_normals
    with pytest.raises(InvalidInput):
        fill_normals(1, 4, 4, rng_mode=RngMode.Fast)
    # more than 2^40 values: refused before anything is allocated or written
    L = _lib.load()
    buf = (C.c_double * 1)()
    assert L.ofg_fill_normals(_lib.context(0), _lib.RNG_REFERENCE, 1, 0, 1 << 21, 1 << 20, 0,
                              buf, None) == 1
    assert "2^40" in _lib.last_error()
