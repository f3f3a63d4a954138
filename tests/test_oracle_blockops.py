"""The oracle's block-sparse permute / fuse_legs / split_leg
(tk-symmetry/src/block_sparse.rs:400-793), pinned by the reference's own
fixture tests (block_sparse.rs:966-1115: nnz preserved, single-leg identity,
fuse -> split round trips, mixed directions) and by dense equivalence: the
dense form of the result equals the transposed / regrouped dense form of the
input, computed from global sector offsets without the per-block loop."""
import itertools

import numpy as np
import pytest

import oracle

IN, OUT = 0, 1
U1, U1U1 = 1, 2


def _filled_rank3():
    """make_filled_rank3 (block_sparse.rs:966-985): legs [(-1,2),(0,3),(1,2)] x 3,
    directions (In, In, Out), flux 0, block value q0*100 + q1*10 + q2."""
    idx = [(-1, 2), (0, 3), (1, 2)]
    legs, dirs = [idx, idx, idx], [IN, IN, OUT]
    keys, offs, tot = oracle.block_table(U1, legs, dirs)
    x = np.zeros(tot)
    dims = dict(((q, 0), d) for q, d in idx)
    for k, off in zip(keys, offs):
        qs = oracle.unpack_key(U1, k, 3)
        n = int(np.prod([dims[q] for q in qs]))
        x[off:off + n] = qs[0][0] * 100 + qs[1][0] * 10 + qs[2][0]
    return legs, dirs, x


def _random(sym, legs, dirs, flux, seed):
    _, _, tot = oracle.block_table(sym, legs, dirs, flux)
    return np.random.default_rng(seed).standard_normal(tot)


CASES = [
    (U1, [[(-1, 2), (0, 3), (1, 2)], [(0, 1), (1, 2)], [(-1, 1), (0, 2), (1, 3), (2, 1)]], [IN, OUT, OUT], (0, 0)),
    (U1, [[(0, 2), (1, 1)], [(-1, 2), (0, 1), (1, 2)], [(0, 3), (1, 2)], [(-1, 1), (0, 2), (1, 2)]],
     [OUT, OUT, IN, IN], (1, 0)),
    (U1U1, [[((0, 0), 1), ((1, 0), 1), ((0, 1), 1), ((1, 1), 1)], [((0, 0), 2), ((1, 0), 3), ((0, 1), 2)],
            [((0, 0), 1), ((1, 0), 2), ((0, 1), 2), ((1, 1), 3)]], [OUT, OUT, IN], (0, 0)),
]


def test_reference_fixture_nnz_and_rank():
    legs, dirs, x = _filled_rank3()
    for start, end, rank in [(1, 3, 2), (0, 2, 2), (0, 3, 1)]:
        nl, nd = oracle.fused_legs(U1, legs, dirs, start, end)
        assert len(nl) == rank
        y = oracle.block_fuse(U1, legs, dirs, (0, 0), start, end, x)
        assert len(y) == len(x)                    # nnz preserved
        assert np.isclose(y.sum(), x.sum())         # every element placed once
    # single leg: unchanged
    nl, nd = oracle.fused_legs(U1, legs, dirs, 1, 2)
    assert nl == legs and nd == dirs
    assert np.array_equal(oracle.block_fuse(U1, legs, dirs, (0, 0), 1, 2, x), x)


def test_reference_fixture_round_trips():
    legs, dirs, x = _filled_rank3()
    for start, end in [(1, 3), (0, 2), (0, 3)]:
        y = oracle.block_fuse(U1, legs, dirs, (0, 0), start, end, x)
        assert np.array_equal(oracle.block_split(U1, legs, dirs, (0, 0), start, end, y), x)


def test_mixed_directions_nnz():
    idx = [(0, 2), (1, 3)]
    legs, dirs = [idx, idx], [IN, OUT]
    _, _, tot = oracle.block_table(U1, legs, dirs)
    x = np.arange(tot, dtype=float)
    y = oracle.block_fuse(U1, legs, dirs, (0, 0), 0, 2, x)
    nl, nd = oracle.fused_legs(U1, legs, dirs, 0, 2)
    assert len(nl) == 1 and nd == [IN] and len(y) == tot


@pytest.mark.parametrize("case", range(len(CASES)))
def test_permute_dense_equivalence(case):
    sym, legs, dirs, flux = CASES[case]
    x = _random(sym, legs, dirs, flux, case)
    d = oracle.to_dense(sym, legs, dirs, flux, x)
    for perm in itertools.permutations(range(len(legs))):
        y = oracle.block_permute(sym, legs, dirs, flux, list(perm), x)
        dp = oracle.to_dense(sym, [legs[i] for i in perm], [dirs[i] for i in perm], flux, y)
        assert np.array_equal(dp, d.transpose(perm))
        inv = list(np.argsort(perm))
        z = oracle.block_permute(sym, [legs[i] for i in perm], [dirs[i] for i in perm], flux, inv, y)
        assert np.array_equal(z, x)


def _fused_dense_expected(sym, legs, dirs, flux, start, end, d):
    """Dense regrouping: the fused axis index of sub-sector combination c,
    local index t is offset(fused sector) + combo offset + t."""
    fused, cs, co = oracle.fuse_map(sym, legs[start:end], dirs[start:end])
    subs = [oracle._sorted_leg(l) for l in legs[start:end]]
    foff, acc = [], 0
    for _, dim in fused:
        foff.append(acc)
        acc += dim
    shape = list(d.shape)
    out = np.zeros(shape[:start] + [acc] + shape[end:])
    starts = []
    for sl in subs:
        o, s0 = [], 0
        for _, dim in sl:
            o.append((s0, dim))
            s0 += dim
        starts.append(o)
    for c, combo in enumerate(itertools.product(*[range(len(sl)) for sl in subs])):
        sl = [starts[i][j] for i, j in enumerate(combo)]
        sub = d[(slice(None),) * start + tuple(slice(a, a + b) for a, b in sl)]
        n = int(np.prod([b for _, b in sl]))
        f0 = foff[int(cs[c])] + int(co[c])
        sub = sub.reshape(shape[:start] + [n] + shape[end:])
        out[(slice(None),) * start + (slice(f0, f0 + n),)] = sub
    return out, fused


@pytest.mark.parametrize("case", range(len(CASES)))
def test_fuse_dense_equivalence_and_split(case):
    sym, legs, dirs, flux = CASES[case]
    x = _random(sym, legs, dirs, flux, 10 + case)
    d = oracle.to_dense(sym, legs, dirs, flux, x)
    r = len(legs)
    for start in range(r):
        for end in range(start + 2, r + 1):
            y = oracle.block_fuse(sym, legs, dirs, flux, start, end, x)
            nl, nd = oracle.fused_legs(sym, legs, dirs, start, end)
            exp, fused = _fused_dense_expected(sym, legs, dirs, flux, start, end, d)
            got = oracle.to_dense(sym, nl, nd, flux, y)
            assert np.array_equal(got, exp)
            assert np.array_equal(oracle.block_split(sym, legs, dirs, flux, start, end, y), x)
