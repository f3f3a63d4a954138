"""The truncation rule and the bond-dimension schedule, host side.

Both the library's rule (tk_truncation_select, the one the two-site split and
the dense SVD entry points apply) and the oracle's restatement
(oracle.truncation_select, the one its DMRG split applies) are checked against
the known answers of tk-dmrg/src/truncation/mod.rs:185-236, against answers
the reference's loop fixes by hand (relative, not absolute, cutoff; sigma, not
sigma^2), against the degeneracy and numerical-zero rules of this build, and
against each other on random spectra with ties.  The schedule is checked
against truncation/mod.rs:238-259 and proptest_dmrg.rs:53-71.
"""
import numpy as np
import pytest

import oracle
import tk_hip

RULES = {"product": tk_hip.truncation_select, "oracle": oracle.truncation_select}


@pytest.fixture(params=sorted(RULES))
def select(request):
    return RULES[request.param]


def test_truncate_svd_basic(select):
    # truncation/mod.rs:196-212: [[3,0],[0,2],[0,0]] has sigma (3, 2); the
    # default TruncationConfig (max 200, cutoff 1e-12, min 1) keeps both
    keep, nk, nd, err = select([[3.0, 2.0]], 200, 1e-12, 1)
    assert (keep, nk, nd) == ([2], 2, 0)
    assert err < 1e-14


def test_truncate_svd_with_truncation(select):
    # truncation/mod.rs:214-227: diag(3, 2, 1) cut to bond 2: one value
    # discarded, error 1^2 / (3^2 + 2^2 + 1^2) = 1/14
    keep, nk, nd, err = select([[3.0, 2.0, 1.0]], 2, 1e-12, 1)
    assert (keep, nk, nd) == ([2], 2, 1)
    assert abs(err - 1.0 / 14.0) < 1e-10


def test_truncate_svd_min_bond_dim(select):
    # truncation/mod.rs:229-236: diag(10, 1e-15), cutoff 1e-10 would drop the
    # second value, min_bond_dim 2 keeps it
    keep, nk, nd, err = select([[10.0, 1e-15]], 200, 1e-10, 2)
    assert nk >= 2 and keep == [2] and nd == 0


def test_cutoff_is_relative_and_on_sigma(select):
    # the reference's loop: break when s < cutoff * sigma_max (truncation/mod.rs:72-80)
    # relative: 1e-9 < 1e-10 * 100 = 1e-8 is cut (an absolute 1e-10 would keep it)
    assert select([[100.0, 1e-9]], 200, 1e-10, 1)[1] == 1
    # on sigma, not sigma^2: 1e-6 >= 1e-10 * 1 is kept (sigma^2 = 1e-12 would not pass 1e-10)
    assert select([[1.0, 1e-6]], 200, 1e-10, 1)[1] == 2
    # exactly at the threshold is kept (s < threshold breaks, s == threshold does not)
    assert select([[1.0, 0.5]], 200, 0.5, 1)[1] == 2
    # cutoff 0 keeps zeros too, up to max_bond
    assert select([[1.0, 0.0, 0.0]], 2, 0.0, 0)[1] == 2


def test_blocks_are_one_spectrum(select):
    # the split's blocks (one per middle charge) share one global cut: sigma
    # (3, 1) and (2): bond 2 keeps 3 and 2, error 1/14
    keep, nk, nd, err = select([[3.0, 1.0], [2.0]], 2, 1e-12, 1)
    assert (keep, nk, nd) == ([1, 1], 2, 1)
    assert abs(err - 1.0 / 14.0) < 1e-15
    # equal values across blocks: the earlier block first
    assert select([[2.0], [2.0]], 1, 0.0, 0)[0] == [1, 0]


def test_degeneracy_rule(select):
    # never cut between sigma_k and sigma_k+1 when they differ by <= tol * sigma_max
    s = [[1.0, 0.5, 0.5 - 1e-14, 0.1]]
    assert select(s, 2, 0.0, 0)[1] == 2                         # rule off: the pair is split
    assert select(s, 2, 0.0, 0, degeneracy_tol=1e-12)[1] == 1   # rule on: cut before the pair
    assert select(s, 3, 0.0, 0, degeneracy_tol=1e-12)[1] == 3   # a cut after the pair stands
    # a pair across blocks is one group too
    assert select([[1.0, 0.5], [0.5 - 1e-14]], 2, 0.0, 0, degeneracy_tol=1e-12)[0] == [1, 0]
    # a cut that would move to zero kept values stays where it was
    assert select([[2.0], [2.0]], 1, 0.0, 0, degeneracy_tol=1e-12)[0] == [1, 0]


def test_numerical_zero_floor(select):
    # sigma^2 < floor * sigma_max^2 is never kept by the cutoff
    assert select([[1.0, 1e-14]], 200, 0.0, 1, sigma2_floor=1e-26)[1] == 1
    assert select([[1.0, 1e-12]], 200, 0.0, 1)[1] == 2


def _call(select, sig, mb, cut, mn, deg, floor):
    return select(sig, mb, cut, mn, degeneracy_tol=deg, sigma2_floor=floor)


def test_product_rule_equals_oracle_rule():
    rng = np.random.default_rng(5)
    for case in range(300):
        nb = int(rng.integers(1, 6))
        sig = []
        for _ in range(nb):
            v = np.sort(10.0 ** rng.uniform(-15, 1, int(rng.integers(1, 9))))[::-1]
            if rng.random() < 0.4 and len(v) > 1:  # exact and near ties
                j = int(rng.integers(1, len(v)))
                v[j] = v[j - 1] * (1 - (0 if rng.random() < 0.5 else 1e-15))
            sig.append(v.tolist())
        if rng.random() < 0.3 and nb > 1:  # ties across blocks
            sig[1][0] = sig[0][0]
            sig[1] = sorted(sig[1], reverse=True)
        mb = int(rng.integers(0, 12))
        cut = float(rng.choice([0.0, 1e-12, 1e-6, 1e-2]))
        mn = int(rng.integers(0, 4))
        deg = float(rng.choice([0.0, 1e-12, 1e-3]))
        floor = float(rng.choice([0.0, 1e-26, 1e-10]))
        a = _call(tk_hip.truncation_select, sig, mb, cut, mn, deg, floor)
        b = _call(oracle.truncation_select, sig, mb, cut, mn, deg, floor)
        assert a[:3] == b[:3], (case, sig, mb, cut, mn, deg, floor, a, b)
        assert abs(a[3] - b[3]) <= 1e-15, (case, a, b)


def test_select_rejects_bad_input():
    with pytest.raises(tk_hip.TkError):
        tk_hip.truncation_select([[1.0, 2.0]], 5, 0.0)  # not descending
    with pytest.raises(tk_hip.TkError):
        tk_hip.truncation_select([[1.0, -1.0]], 5, 0.0)
    with pytest.raises(tk_hip.TkError):
        tk_hip.truncation_select([[1.0]], 5, -1.0)


def test_schedule_warmup_known_values():
    # truncation/mod.rs:238-245: warmup(10, 200, 5) non-decreasing and ending at 200
    d = tk_hip.BondDimensionSchedule.warmup(10, 200, 5).dims
    assert all(b >= a for a, b in zip(d, d[1:])) and d[-1] == 200 and len(d) == 6
    # the ramp of the D = 2048 Hubbard workload
    assert tk_hip.BondDimensionSchedule.warmup(256, 2048, 7).dims == [256, 345, 464, 624, 840, 1131, 1522, 2048]
    # round(d_init * r^i), half away from zero, r = (d_max / d_init)^(1/n)
    d = tk_hip.BondDimensionSchedule.warmup(10, 200, 5).dims
    r = (200 / 10) ** (1 / 5)
    assert d == [min(200, max(1, int(np.floor(10 * r ** i + 0.5)))) for i in range(6)]
    assert tk_hip.BondDimensionSchedule.warmup(7, 100, 0).dims == [100]


def test_schedule_fixed_and_custom():
    # truncation/mod.rs:247-259
    s = tk_hip.BondDimensionSchedule.fixed(100)
    assert s.bond_dim_at_sweep(0) == 100 and s.bond_dim_at_sweep(999) == 100
    s = tk_hip.BondDimensionSchedule.custom([10, 20, 30])
    assert [s.bond_dim_at_sweep(i) for i in (0, 2, 100)] == [10, 30, 30]


def test_schedule_warmup_property():
    # proptest_dmrg.rs:53-71 over its whole input box
    for d_init in (1, 2, 7, 13, 50):
        for d_max in (50, 51, 99, 256, 500):
            for n in (1, 2, 5, 11, 20):
                d = tk_hip.BondDimensionSchedule.warmup(d_init, d_max, n).dims
                assert all(b >= a for a, b in zip(d, d[1:])), (d_init, d_max, n, d)
                assert d[n] == d_max and d[0] == d_init
