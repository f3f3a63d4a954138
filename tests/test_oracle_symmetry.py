"""Pins the oracle's sector bookkeeping on the reference's own known answers,
then checks the product's host tables against it bit for bit (host-only calls,
no GPU needed)."""
import itertools

import numpy as np
import pytest

import oracle
import tk_hip

U1, UU = 1, 2
IN, OUT = 0, 1


# ---- oracle pinned by reference tests ------------------------------------------------------
def test_u1_pack_round_trip():
    # tk-symmetry/src/builtins.rs: u1_pack_round_trip over -128..=127
    for n in range(-128, 128):
        k = oracle.pack_key(U1, [n])
        assert k == (n & 0xFF)
        assert oracle.unpack_key(U1, k, 1) == [(n, 0)]


def test_u1u1_pack_round_trip():
    for a in (-128, -5, -1, 0, 1, 7, 127):
        for b in (-128, -2, 0, 3, 127):
            k = oracle.pack_key(UU, [(a, b)])
            assert k == (a & 0xFF) | ((b & 0xFF) << 8)
            assert oracle.unpack_key(UU, k, 1) == [(a, b)]


def test_packed_key_round_trip():
    # sector_key.rs: packed_key_round_trip [3, -2, 0]
    k = oracle.pack_key(U1, [3, -2, 0])
    assert oracle.unpack_key(U1, k, 3) == [(3, 0), (-2, 0), (0, 0)]
    # leg i at bits [8i, 8i+8)
    assert k == 3 | (0xFE << 8) | (0 << 16)


def test_packed_key_sort_order():
    # sector_key.rs:210-226: -1 packs to 0xFF, so (0,0) < (1,0) < (-1,0) as raw u64
    k1 = oracle.pack_key(U1, [-1, 0])
    k2 = oracle.pack_key(U1, [0, 0])
    k3 = oracle.pack_key(U1, [1, 0])
    assert sorted([k1, k2, k3]) == [k2, k3, k1]


def test_pack_rejects_overflow():
    # rank 9 x 8 bits > 64 (sector_key.rs packed_key_overflow_debug_panic)
    with pytest.raises(ValueError):
        oracle.pack_key(U1, list(range(9)))
    with pytest.raises(ValueError):
        oracle.pack_key(U1, [200])


def test_qindex_offset_of():
    # sector_key.rs:246-254: offsets 0, 2, 5; total 9
    secs, offs, tot = oracle.qindex_layout([(1, 4), (-1, 2), (0, 3)])
    assert [q for q, _ in secs] == [(-1, 0), (0, 0), (1, 0)]
    assert list(offs) == [0, 2, 5]
    assert tot == 9


def test_enumerate_sector_counts():
    # flux.rs:259-329: 7 (rank 3, in,in,out), 141 (rank 6 all in), 51 (rank 5 all in)
    idx = [(-1, 1), (0, 1), (1, 1)]
    keys, _, _ = oracle.block_table(U1, [idx] * 3, [IN, IN, OUT])
    assert len(keys) == 7
    keys, _, _ = oracle.block_table(U1, [idx] * 6, [IN] * 6)
    assert len(keys) == 141
    keys, _, _ = oracle.block_table(U1, [idx] * 5, [IN] * 5)
    assert len(keys) == 51


def test_block_sparse_get_block():
    # block_sparse.rs:886-930: (0,0,0) present with dims [3,3,3]; (0,0,1) absent
    idx = [(-1, 2), (0, 3), (1, 2)]
    keys, offs, tot = oracle.block_table(U1, [idx] * 3, [IN, IN, OUT])
    ks = set(int(k) for k in keys)
    assert oracle.pack_key(U1, [0, 0, 0]) in ks
    assert oracle.pack_key(U1, [0, 0, 1]) not in ks
    i = list(keys).index(oracle.pack_key(U1, [0, 0, 0]))
    size = (offs[i + 1] if i + 1 < len(offs) else tot) - offs[i]
    assert size == 27
    assert list(keys) == sorted(int(k) for k in keys)


def test_flatten_offsets():
    # flat_storage.rs:201-252: blocks (0,0) 2x2 then (1,1) 3x3 -> offsets 0, 4; total 13
    idx = [(0, 2), (1, 3)]
    keys, offs, tot = oracle.block_table(U1, [idx, idx], [IN, OUT])
    assert list(keys) == [oracle.pack_key(U1, [0, 0]), oracle.pack_key(U1, [1, 1])]
    assert list(offs) == [0, 4]
    assert tot == 13


def test_fuse_map_hand_checked():
    # fuse_legs (block_sparse.rs:474-518) on two incoming legs {-1:2, 0:3, 1:2}:
    # combos in leg order, later leg fastest; groups by fused q ascending
    idx = [(-1, 2), (0, 3), (1, 2)]
    fused, cs, co = oracle.fuse_map(U1, [idx, idx], [IN, IN])
    assert fused == [((-2, 0), 4), ((-1, 0), 12), ((0, 0), 17), ((1, 0), 12), ((2, 0), 4)]
    # combos (-1,-1)(-1,0)(-1,1)(0,-1)(0,0)(0,1)(1,-1)(1,0)(1,1)
    assert list(cs) == [0, 1, 2, 1, 2, 3, 2, 3, 4]
    assert list(co) == [0, 0, 0, 6, 4, 0, 13, 6, 0]


def test_fuse_map_outgoing_uses_dual():
    # Outgoing legs contribute dual(q) (block_sparse.rs:486-490)
    idx = [(0, 2), (1, 3)]
    fused, cs, co = oracle.fuse_map(U1, [idx, idx], [IN, OUT])
    assert fused == [((-1, 0), 6), ((0, 0), 13), ((1, 0), 6)]


def _brute_table(sym, legs, dirs, flux):
    srt = [sorted(((q, 0) if np.isscalar(q) else tuple(q), d) for q, d in l) for l in legs]
    out = []
    for combo in itertools.product(*srt):
        tot = [0, 0]
        for (q, _), dr in zip(combo, dirs):
            s = 1 if dr == IN else -1
            tot[0] += s * q[0]
            tot[1] += s * q[1]
        if tuple(tot) == tuple(flux):
            out.append((oracle.pack_key(sym, [q for q, _ in combo]), int(np.prod([d for _, d in combo]))))
    out.sort()
    return out


def test_oracle_table_matches_python_brute_force():
    rng = np.random.default_rng(1)
    for trial in range(20):
        rank = int(rng.integers(2, 5))
        legs = []
        for _ in range(rank):
            qs = sorted(set(int(q) for q in rng.integers(-3, 4, size=int(rng.integers(1, 5)))))
            legs.append([(q, int(rng.integers(1, 4))) for q in qs])
        dirs = [int(d) for d in rng.integers(0, 2, size=rank)]
        flux = (int(rng.integers(-1, 2)), 0)
        ref = _brute_table(U1, legs, dirs, flux)
        keys, offs, tot = oracle.block_table(U1, legs, dirs, flux)
        assert [int(k) for k in keys] == [k for k, _ in ref]
        sizes = np.diff(np.append(offs, tot))
        assert list(sizes) == [n for _, n in ref]


# ---- product host tables vs the oracle, bit for bit ---------------------------------------
def test_product_pack_and_qindex_match_oracle():
    for n in range(-128, 128):
        assert tk_hip.pack_key(U1, [n]) == oracle.pack_key(U1, [n])
    assert tk_hip.pack_key(UU, [(3, -2), (0, 1), (-1, -1), (5, 0)]) == \
        oracle.pack_key(UU, [(3, -2), (0, 1), (-1, -1), (5, 0)])
    secs = [(1, 4), (-1, 2), (0, 3)]
    a = tk_hip.qindex_layout(U1, secs)
    b = oracle.qindex_layout(secs)
    assert a[0] == b[0] and list(a[1]) == list(b[1]) and a[2] == b[2]


def test_product_rejects_out_of_range_qn():
    with pytest.raises(tk_hip.TkError):
        tk_hip.pack_key(U1, [128])
    with pytest.raises(tk_hip.TkError):
        tk_hip.pack_key(U1, [0] * 9)


@pytest.mark.parametrize("seed", range(12))
def test_product_block_table_matches_oracle(seed):
    rng = np.random.default_rng(100 + seed)
    sym = U1 if seed % 2 == 0 else UU
    rank = int(rng.integers(2, 5))
    legs = []
    for _ in range(rank):
        n = int(rng.integers(1, 6))
        if sym == U1:
            qs = sorted(set(int(q) for q in rng.integers(-4, 5, size=n)))
        else:
            qs = sorted(set((int(a), int(b)) for a, b in rng.integers(-2, 3, size=(n, 2))))
        legs.append([(q, int(rng.integers(1, 5))) for q in qs])
        rng.shuffle(legs[-1])
    dirs = [int(d) for d in rng.integers(0, 2, size=rank)]
    flux = (0, 0)
    pk, po, _, pt = tk_hip.block_table(sym, legs, dirs, flux)
    ok, oo, ot = oracle.block_table(sym, legs, dirs, flux)
    assert np.array_equal(pk, ok)
    assert np.array_equal(po, oo)
    assert pt == ot


def test_product_fuse_map_matches_oracle():
    rng = np.random.default_rng(7)
    for _ in range(10):
        legs = [[(q, int(rng.integers(1, 4))) for q in sorted(set(int(x) for x in rng.integers(-2, 3, 3)))]
                for _ in range(int(rng.integers(2, 4)))]
        dirs = [int(d) for d in rng.integers(0, 2, size=len(legs))]
        a = tk_hip.fuse_map(U1, legs, dirs)
        b = oracle.fuse_map(U1, legs, dirs)
        assert a[0] == b[0]
        assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])


def test_c2_tables_match_oracle():
    """Sector tables of the single-matvec workload (145-sector U(1)xU(1) bond)."""
    bl = tk_hip.synthetic_uu_bond(1024)
    assert len(bl) == 145 and sum(d for _, d in bl) == 1024 and max(d for _, d in bl) <= 30
    br = [((a + 1, b + 1), d) for (a, b), d in bl]
    W = tk_hip.hubbard_w()
    legs = [W.phys, W.phys, bl, br]
    pk, po, _, pt = tk_hip.block_table(UU, legs, [OUT, OUT, OUT, IN])
    ok, oo, ot = oracle.block_table(UU, legs, [OUT, OUT, OUT, IN])
    assert np.array_equal(pk, ok) and np.array_equal(po, oo) and pt == ot
    for bond, wq in ((bl, W.wl), (br, W.wr)):
        pk, po, _, pt = tk_hip.env_table(UU, bond, wq)
        ok, oo, ot = oracle.block_table(UU, [bond, wq, bond], [IN, IN, OUT])
        assert np.array_equal(pk, ok) and np.array_equal(po, oo) and pt == ot
