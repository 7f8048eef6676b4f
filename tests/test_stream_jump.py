"""Jump-ahead of the xoshiro256++ stream (ofg_stream_jump, host-only code of
the C ABI): the least-squares Monte Carlo path's reference draws start path p
at draw p * d of the engine's single stream. Pinned against brute-force
stepping of the oracle's generator, across the polynomial reduction (k >= 256)
and for large k by composition."""
import numpy as np
import pytest

from oracle import oracle as O


STATES = [
    [0x123456789ABCDEF, 42, 7, 0xDEADBEEFCAFEBABE],
    [1, 0, 0, 0],
]


@pytest.mark.parametrize("state", STATES)
@pytest.mark.parametrize("k", [0, 1, 2, 63, 64, 255, 256, 257, 511, 1000, 123_457, 3_000_017])
def test_jump_equals_stepping(state, k):
    from openferric_amd import stream_jump
    assert stream_jump(state, k) == O.xoshiro_advance(state, k)


def test_jump_composes_for_large_distances():
    from openferric_amd import stream_jump
    st = STATES[0]
    rng = np.random.default_rng(5)
    for _ in range(6):
        a, b = (int(x) for x in rng.integers(1 << 40, 1 << 62, size=2))
        assert stream_jump(stream_jump(st, a), b) == stream_jump(st, a + b)


def test_seeded_stream_positions_match_path_major_draws():
    """Path p of a 5-step engine starts where the reference's loop leaves off
    after p paths: the p*5-th state of seed_from_u64(seed)."""
    from openferric_amd import stream_jump
    s = (O.C.c_uint64 * 4)()
    O.lib().or_xoshiro_seed(42, s)
    base = list(s)
    for p in (0, 1, 17, 999):
        assert stream_jump(base, 5 * p) == O.xoshiro_advance(base, 5 * p)
