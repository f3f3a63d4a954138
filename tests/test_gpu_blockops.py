"""Device permute / fuse_legs / split_leg against the oracle, bit-exact
(pure data movement), on the reference fixture and random U(1) and
U(1)xU(1) tensors with mixed directions and nonzero flux."""
import itertools

import numpy as np
import pytest

import oracle
import tk_hip
from test_oracle_blockops import CASES, _filled_rank3, _random

pytestmark = pytest.mark.gpu


def test_fixture_fuse_split_gpu(ctx):
    print("case blockops fixture", flush=True)
    legs, dirs, x = _filled_rank3()
    for start, end in [(1, 3), (0, 2), (0, 3), (1, 2)]:
        y = tk_hip.block_fuse(ctx, 1, legs, dirs, (0, 0), start, end, x)
        assert np.array_equal(y, oracle.block_fuse(1, legs, dirs, (0, 0), start, end, x))
        assert np.array_equal(tk_hip.block_split(ctx, 1, legs, dirs, (0, 0), start, end, y), x)


@pytest.mark.parametrize("case", range(len(CASES)))
def test_permute_gpu_matches_oracle(ctx, case):
    sym, legs, dirs, flux = CASES[case]
    print(f"case blockops permute {case}", flush=True)
    x = _random(sym, legs, dirs, flux, 20 + case)
    for perm in itertools.permutations(range(len(legs))):
        y = tk_hip.block_permute(ctx, sym, legs, dirs, flux, list(perm), x)
        assert np.array_equal(y, oracle.block_permute(sym, legs, dirs, flux, list(perm), x))


@pytest.mark.parametrize("case", range(len(CASES)))
def test_fuse_split_gpu_matches_oracle(ctx, case):
    sym, legs, dirs, flux = CASES[case]
    print(f"case blockops fuse {case}", flush=True)
    x = _random(sym, legs, dirs, flux, 30 + case)
    r = len(legs)
    for start in range(r):
        for end in range(start + 1, r + 1):
            y = tk_hip.block_fuse(ctx, sym, legs, dirs, flux, start, end, x)
            assert np.array_equal(y, oracle.block_fuse(sym, legs, dirs, flux, start, end, x))
            z = tk_hip.block_split(ctx, sym, legs, dirs, flux, start, end, y)
            assert np.array_equal(z, x)


def test_permute_large_block_gpu(ctx):
    """Blocks larger than one work chunk (4096 elements)."""
    print("case blockops permute large", flush=True)
    legs = [[(0, 40), (1, 30)], [(0, 50), (1, 20)], [(0, 10), (1, 12), (2, 8)]]
    dirs, flux = [1, 1, 0], (0, 0)
    x = _random(1, legs, dirs, flux, 7)
    for perm in ([2, 0, 1], [1, 2, 0], [0, 2, 1]):
        y = tk_hip.block_permute(ctx, 1, legs, dirs, flux, perm, x)
        assert np.array_equal(y, oracle.block_permute(1, legs, dirs, flux, perm, x))
