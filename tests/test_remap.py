"""The value-encoding restatement (normalEncode / encodeQuery / findKeyByValue,
remap.cpp:110-291) and the library's host-side lookups, without a GPU.

remap.cpp includes bindex.h, which includes cuda_runtime.h, so the reference's encoder does
not build here; the oracle is pinned instead by a worked example and by closed forms that
count rows (numpy searchsorted / unique) rather than sort them."""
import numpy as np
import pytest

import oracle
import rtscan_amd as R

MAXU = 0xFFFFFFFF


def test_worked_example():
    # rows 5,3,5,9 with codes in [0, 4): interval 1, so codes are sorted ranks
    codes, pos, sc, mv, ms = oracle.normal_encode([5, 3, 5, 9], 0, 4)
    assert codes.tolist() == [1, 0, 2, 3]
    assert pos.tolist() == [1, 0, 2, 3] and sc.tolist() == [0, 1, 2, 3]
    assert mv.tolist() == [3, 5, 9] and ms.tolist() == [0, 1, 3]
    q = lambda v, op: oracle.encode_query(mv, ms, v, op, 4)
    assert [q(5, R.LT), q(5, R.LE), q(5, R.GT), q(5, R.GE)] == [1, 2, 3, 1]
    assert q(10, R.LT) == 4 and q(9, R.GT) == 4  # no such value: the row count
    assert q(2, R.LE) == MAXU  # start 0 minus one wraps, as CODE arithmetic does
    assert q(5, R.EQ) is None and q(5, R.BT) is None
    k = lambda c: oracle.find_key_by_value(mv, ms, c)
    assert [k(0), k(1), k(2), k(3), k(100)] == [3, 5, 5, 0, 0]


def random_column(rng, n, distinct):
    return rng.choice(rng.integers(0, 2 ** 32, distinct, dtype=np.uint64), n).astype(np.uint32)


@pytest.mark.parametrize("n,distinct,lo,hi", [
    (1000, 37, 0, 1000), (4096, 4096, 0, 4096), (5000, 3, 7, 2 ** 32 - 1),
    (3001, 500, 2, 4294967294), (777, 1, 0, 777), (2000, 300, 100, 50)])
def test_rank_encoding_closed_forms(n, distinct, lo, hi):
    rng = np.random.default_rng(n + distinct)
    x = random_column(rng, n, distinct)
    codes, pos, sc, mv, ms = oracle.normal_encode(x, lo, hi)
    I = int((float(hi) - float(lo)) / n) & MAXU
    xs = np.sort(x)
    # the map: distinct values, each starting at lo + I * (rows below it)
    assert np.array_equal(mv, np.unique(x))
    below = np.searchsorted(xs, mv, "left").astype(np.uint64)
    assert np.array_equal(ms, ((lo + I * below) & MAXU).astype(np.uint32))
    # sorted outputs: positions sort the column, codes are lo + I * i
    assert np.array_equal(x[pos], xs)
    assert np.array_equal(sc, ((lo + I * np.arange(n, dtype=np.uint64)) & MAXU).astype(np.uint32))
    assert np.array_equal(codes[pos], sc)
    # each value's rows hold its consecutive codes, in row order
    for v in mv[:20]:
        rows = np.nonzero(x == v)[0]
        start = (lo + I * int(np.searchsorted(xs, v, "left"))) & MAXU
        want = (start + I * np.arange(rows.size, dtype=np.uint64)) & MAXU
        assert np.array_equal(codes[rows], want.astype(np.uint32))
    # encodeQuery: counts of rows below / at-or-below the query
    qs = np.concatenate([rng.choice(x, 20), rng.integers(0, 2 ** 32, 20, dtype=np.uint64)])
    for qv in [int(v) for v in qs] + [0, MAXU]:
        lt = int(np.searchsorted(xs, qv, "left"))
        le = int(np.searchsorted(xs, qv, "right"))
        exp_lt = (lo + I * lt) & MAXU if lt < n else n
        exp_le = (lo + I * le - 1) & MAXU if le < n else n
        exp_gt = (lo + I * le) & MAXU if le < n else n
        assert oracle.encode_query(mv, ms, qv, R.LT, n) == exp_lt
        assert oracle.encode_query(mv, ms, qv, R.GE, n) == exp_lt
        assert oracle.encode_query(mv, ms, qv, R.LE, n) == exp_le
        assert oracle.encode_query(mv, ms, qv, R.GT, n) == exp_gt


def test_find_key_by_value_on_ranks():
    """With codes = sorted ranks, the key of code c is the value at sorted index c, for every
    c below the last value's start; 0 from there on."""
    rng = np.random.default_rng(3)
    x = random_column(rng, 3000, 400)
    n = x.size
    codes, pos, sc, mv, ms = oracle.normal_encode(x, 0, n)
    xs = np.sort(x)
    for c in list(range(0, int(ms[-1]), 7)) + [int(ms[-1]) - 1]:
        assert oracle.find_key_by_value(mv, ms, c) == int(xs[c])
    for c in (int(ms[-1]), n - 1, n, MAXU):
        assert oracle.find_key_by_value(mv, ms, c) == 0


@pytest.mark.parametrize("n,distinct,lo,hi", [(5000, 40, 2, 4294967294), (3000, 3000, 0, 3000),
                                              (4000, 7, 0, 4000), (999, 80, 10, 2 ** 31)])
def test_dispersed_map_closed_form(n, distinct, lo, hi):
    rng = np.random.default_rng(distinct)
    x = random_column(rng, n, distinct)
    mv, ms, st = oracle.dispersed_map(x, lo, hi)
    vals, counts = np.unique(x, return_counts=True)
    D = (float(hi) - float(lo)) / n
    steps = np.array([int(D * int(c)) & MAXU for c in counts], np.uint64)
    starts = (np.uint64(lo) + np.concatenate([np.zeros(1, np.uint64), np.cumsum(steps)[:-1]])) & np.uint64(MAXU)
    assert np.array_equal(mv, vals)
    assert np.array_equal(st, steps.astype(np.uint32))
    assert np.array_equal(ms, starts.astype(np.uint32))


def test_mistakes_are_caught():
    """Plausible slips fail the closed forms above: lower_bound for le, or codes that restart
    per value instead of continuing the sorted index."""
    x = np.array([5, 3, 5, 9, 5], np.uint32)
    codes, pos, sc, mv, ms = oracle.normal_encode(x, 0, 5)
    # le 5 must skip every 5 (3 rows at or below it -> code 4 - 1); lower_bound would give 0
    assert oracle.encode_query(mv, ms, 5, R.LE, 5) == 3
    assert sorted(codes.tolist()) == [0, 1, 2, 3, 4]


# ------------------------------------------------------------- library host lookups

@pytest.mark.parametrize("seed", [0, 1, 2])
def test_library_lookups_equal_oracle(seed):
    rng = np.random.default_rng(seed)
    x = random_column(rng, 2000, 150)
    codes, pos, sc, mv, ms = oracle.normal_encode(x, 3, 2 ** 32 - 3)
    for qv in [int(v) for v in rng.integers(0, 2 ** 32, 50, dtype=np.uint64)] + [int(v) for v in mv[:20]]:
        for op in (R.LT, R.LE, R.GT, R.GE):
            assert R.encode_lookup(mv, ms, qv, op, 2000) == oracle.encode_query(mv, ms, qv, op, 2000)
    for c in [int(v) for v in rng.integers(0, 2 ** 32, 100, dtype=np.uint64)] + [int(v) for v in ms]:
        assert R.encode_key_of(mv, ms, c) == oracle.find_key_by_value(mv, ms, c)


def test_library_lookup_rejects_eq_and_bt():
    for op in (R.EQ, R.BT):
        with pytest.raises(R.HipError):
            R.encode_lookup([1, 2], [0, 1], 1, op, 2)
    assert R.encode_lookup([], [], 5, R.LT, 77) == 77  # empty map: the sentinel
