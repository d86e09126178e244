"""Pins the CPU oracle (oracle/) before it is trusted as the checker.

raw_scan family: closed-form numpy predicates, hand-written bit-order fixtures and the
compiled reference raw_scan.  init_bindex: numpy closed forms and the compiled reference
(oracle/_ref).  bindex_scan: the compiled reference, bit for bit, including the inputs
where the reference disagrees with its own raw_scan (documented below).
"""
import os

import numpy as np
import pytest

import oracle as O
from _util import MAXU, closed_form, pack_rows, unpack_rows

REF_DIR = os.path.join(os.path.dirname(O.__file__), "_ref")


def ref_lib(n):
    try:
        return O.RefBindexLib(n)
    except FileNotFoundError:
        pytest.skip("reference checker library oracle/_ref/libref_bindex_%d.so not built" % n)


# ------------------------------------------------------------------------- bit order

def test_bit_order_fixture():
    # rows 0, 63, 64 set (bindex.cpp:1217 refine): words 0x80000000, 0x00000001, 0x80000000;
    # rows 0 and 31 share word 0 (0x80000001)
    raw = np.full(96, 7, np.uint32)
    raw[[0, 63, 64]] = 3
    raw[[5, 31]] = 4
    assert list(O.raw_scan(raw, O.EQ, 3)) == [0x80000000, 0x00000001, 0x80000000]
    assert list(pack_rows(raw == 3)) == [0x80000000, 0x00000001, 0x80000000]
    assert list(O.raw_scan(raw, O.EQ, 4)) == [0x04000001, 0, 0]


def test_bits_num_needed():
    # bindex.h:140-143: ((n - 1) / 32) + 1, so one word for n = 0 as well
    assert [O.words_for(n) for n in (1, 31, 32, 33, 64, 65)] == [1, 1, 1, 2, 2, 3]
    assert O.lib().orc_bits_num_needed(0) == 1


# ----------------------------------------------------------------------------- raw_scan

EDGE_VALUES = [0, 1, 2, 0x7FFFFFFF, 0x80000000, MAXU - 1, MAXU]


@pytest.mark.parametrize("op", [O.EQ, O.LT, O.GT, O.LE, O.GE, O.BT])
def test_raw_scan_matches_closed_form(op):
    rng = np.random.default_rng(op)
    for n in (1, 31, 32, 33, 1000, 4099):
        x = rng.integers(0, 2 ** 32, n, dtype=np.uint32)
        x[: min(n, len(EDGE_VALUES))] = EDGE_VALUES[: min(n, len(EDGE_VALUES))]
        targets = EDGE_VALUES + [int(v) for v in rng.choice(x, 3)]
        for t1 in targets:
            t2 = int(rng.integers(0, 2 ** 32)) if op == O.BT else 0
            want = pack_rows(closed_form(x, op, t1, t2))
            assert np.array_equal(O.raw_scan(x, op, t1, t2), want)
            assert np.array_equal(O.raw_scan_st(x, op, t1, t2), want)


def test_raw_scan_xors_into_bitmap():
    # refine() toggles (bindex.cpp:1217): scanning twice restores the input bitmap
    x = np.arange(100, dtype=np.uint32)
    bm = np.full(4, 0xDEADBEEF, np.uint32)
    O.raw_scan(x, O.LT, 50, bitmap=bm)
    assert np.array_equal(bm ^ np.uint32(0xDEADBEEF), pack_rows(x < 50))
    O.raw_scan(x, O.LT, 50, bitmap=bm)
    assert (bm == 0xDEADBEEF).all()


def test_raw_scan_entry_targets_and_merge():
    # bindex.cpp:1894-1957: comma-separated targets XOR together; bt bounds are swapped;
    # a column's bitmap is ANDed into the merge bitmap
    rng = np.random.default_rng(5)
    x = rng.integers(0, 1000, 777, dtype=np.uint32)
    y = rng.integers(0, 1000, 777, dtype=np.uint32)
    got = O.raw_scan_entry(x, O.LT, [300, 600])
    assert np.array_equal(got, pack_rows((x < 300) ^ (x < 600)))
    got = O.raw_scan_entry(x, O.BT, [800], [100])
    assert np.array_equal(got, pack_rows((x > 100) & (x < 800)))
    got = O.conjunction([x, y], [(O.GE, [250], None), (O.BT, [10], [900])])
    assert np.array_equal(got, pack_rows((x >= 250) & (y > 10) & (y < 900)))


def test_compare_bitmap_counts():
    # compare_bitmap (bindex.cpp:1959-1978): counts truth hits and matches up to the first
    # mismatching row, then stops
    n = 70
    truth = pack_rows(np.arange(n) % 3 == 0)
    assert O.compare_bitmap(truth, truth, n) == (24, 24, -1)
    bad = truth.copy()
    bad[1] ^= np.uint32(1 << (31 - (40 - 32)))  # flip row 40 (not a multiple of 3)
    total, good, first = O.compare_bitmap(truth, bad, n)
    assert first == 40 and total == 14 and good == 14  # rows 0,3,...,39 counted


def test_popcount_ignores_tail():
    bm = np.array([0xFFFFFFFF, 0xFFFFFFFF], np.uint32)
    assert O.popcount_rows(bm, 40) == 40
    assert O.popcount_rows(bm, 64) == 64


def test_raw_scan_matches_compiled_reference():
    R = ref_lib(524288)
    rng = np.random.default_rng(11)
    x = rng.integers(0, 2 ** 32, 5000, dtype=np.uint32)
    for op in range(6):
        for t1 in [0, 5, int(x[7]), MAXU]:
            t2 = int(x[9]) if op == O.BT else 0
            assert np.array_equal(R.raw_scan(x, op, t1, t2), O.raw_scan(x, op, t1, t2))


# ------------------------------------------------------------------------- init_bindex

def _datasets(n):
    rng = np.random.default_rng(n)
    z = rng.zipf(1.5, n)
    keys = rng.permutation(2 ** 20).astype(np.uint32) * np.uint32(4093)
    return {
        "uniform": rng.integers(0, 2 ** 32, n, dtype=np.uint32),
        "small_range": rng.integers(0, 700, n, dtype=np.uint32),
        "zipf15": keys[np.minimum(z, 2 ** 20) - 1],
        "normal_narrow": np.clip(np.round(rng.normal(2 ** 31, 2 ** 10, n)), 0, MAXU).astype(
            np.uint32),
    }


@pytest.mark.parametrize("name", ["uniform", "small_range", "zipf15"])
def test_init_bindex_closed_form(name):
    n, K = 70000, 128
    x = _datasets(n)[name]
    b = O.Bindex(x, K)
    s = np.sort(x, kind="stable")
    avg = n // K
    sv = b.start_values
    assert sv[0] == s[0] and all(sv[i] == s[i * avg] for i in range(K))
    # areaStartIdx: mid-run split if equal to the previous start, else first occurrence
    idx = b.area_start
    for i in range(1, K):
        if sv[i] == sv[i - 1]:
            assert idx[i] == i * avg
        else:
            assert idx[i] == np.searchsorted(s, sv[i], side="left")
    counts = np.diff(np.append(idx, n))
    assert np.array_equal(b.area_counts, np.cumsum(counts))
    fv = b.fvs()
    for k in range(K - 1):
        assert np.array_equal(fv[k], pack_rows(x < sv[k + 1])), k
    assert np.array_equal(b.sorted_val, s)
    assert np.array_equal(x[b.sorted_pos], s)


@pytest.mark.parametrize("n", [524288, 1000000])
def test_init_bindex_matches_compiled_reference(n):
    R = ref_lib(n)
    for name, x in _datasets(n).items():
        rb = R.bindex(x)
        sv, cnt, bn, fv = rb.read()
        ob = O.Bindex(x)
        assert np.array_equal(sv, ob.start_values), name
        assert np.array_equal(cnt, ob.area_counts), name
        assert np.array_equal(bn, ob.block_nums()), name
        assert np.array_equal(fv, ob.fvs()), name
        for c in [int(v) for v in x[:5]] + [0, MAXU, int(sv[3]), int(sv[3]) + 1]:
            a = ob.in_which_area(c)
            assert a == rb.in_which_area(c)
            if a >= 0:
                blk = ob.in_which_block(a, c)
                assert blk == rb.in_which_block(a, c)
                if c >= sv[a]:
                    assert ob.on_which_pos(a, blk, c) == rb.on_which_pos(a, blk, c)


# --------------------------------------------------------------------------- bindex_scan

def _query_values(x, sv, rng):
    u = np.unique(x)
    vals = [int(v) for v in u[[0, 1, len(u) // 2, -2, -1]]]
    vals += [int(v) for v in sv[[1, 2, 40, 64, 126, 127]]]
    vals += [int(v) for v in rng.integers(0, 2 ** 32, 12)]
    out = set()
    for v in vals:
        for d in (-1, 0, 1):
            out.add((v + d) & MAXU)
    return sorted(out | {0, MAXU})


def same_up_to_ties(x, a, b):
    """Bitmaps a and b select the same multiset of values.  The reference's argsort is an
    unstable parallel sort (bindex.cpp:249-257), so where its BinDex toggles only part of a
    run of equal values (the D1 defect below) WHICH rows of the run flip depends on the
    sort's tie order; everything else must match bit for bit."""
    n = x.size
    ua, ub = unpack_rows(a, n), unpack_rows(b, n)
    if np.array_equal(ua, ub):
        return True
    return np.array_equal(np.sort(x[ua]), np.sort(x[ub]))


@pytest.mark.parametrize("n", [524288, 1000000])
def test_bindex_scan_matches_compiled_reference(n):
    """The oracle's bindex_scan restatement equals the compiled reference on every query it
    does not assert on -- also where both disagree with raw_scan -- bit for bit except for
    the tie order inside runs of one value."""
    R = ref_lib(n)
    rng = np.random.default_rng(3)
    compared = 0
    for name, x in _datasets(n).items():
        rb = R.bindex(x)
        ob = O.Bindex(x)
        vals = _query_values(x, ob.start_values, rng)
        for op in (O.LT, O.LE, O.GT, O.GE, O.EQ):
            for v in vals[:: 2 if op == O.EQ else 1]:
                o = ob.scan(op, v)
                if o is None:
                    continue
                r = rb.scan(op, v)
                assert same_up_to_ties(x, O.mask_tail(o, n), O.mask_tail(r, n)), (name, op, v)
                compared += 1
        for _ in range(12):
            a, b = sorted(int(v) for v in rng.choice(vals, 2, replace=False))
            o = ob.scan(O.BT, a, b)
            if o is None:
                continue
            r = rb.scan(O.BT, a, b)
            assert same_up_to_ties(x, O.mask_tail(o, n), O.mask_tail(r, n)), (name, "bt", a, b)
            compared += 1
    assert compared > 200


def _ref_vs_truth(ob, x, op, t1, t2=0):
    n = x.size
    got = ob.scan(op, t1, t2)
    truth = O.mask_tail(pack_rows(closed_form(x, op, t1, t2)), n)
    return got, truth


def test_reference_bindex_defect_wraparound():
    """D3: le MAX and gt MAX wrap (bindex.cpp:1419-1436): le 4294967295 runs lt 0 (all
    zero) and gt 4294967295 runs gt-path with c+1 = 0 (all ones), while raw_scan gives all
    ones and all zeros.  The oracle reproduces the reference; the engine follows raw_scan."""
    n = 524288
    x = np.random.default_rng(9).integers(0, 2 ** 32, n, dtype=np.uint32)
    ob = O.Bindex(x)
    got, truth = _ref_vs_truth(ob, x, O.LE, MAXU)
    assert O.popcount_rows(got, n) == 0 and O.popcount_rows(truth, n) == n
    got, truth = _ref_vs_truth(ob, x, O.GT, MAXU)
    assert O.popcount_rows(got, n) == n and O.popcount_rows(truth, n) == 0
    R = ref_lib(n)
    rb = R.bindex(x)
    assert np.array_equal(O.mask_tail(rb.scan(O.LE, MAXU), n), O.mask_tail(ob.scan(O.LE, MAXU), n))


def test_reference_bindex_defect_bt_below_min():
    """D2: bt whose lower bound + 1 is below the data minimum returns all ones regardless
    of the upper bound (bindex.cpp:1493-1501)."""
    n = 524288
    x = np.random.default_rng(10).integers(1000, 2 ** 32, n, dtype=np.uint32)
    ob = O.Bindex(x)
    got, truth = _ref_vs_truth(ob, x, O.BT, 5, 2000000)
    assert O.popcount_rows(got, n) == n
    assert O.popcount_rows(truth, n) == int(((x > 5) & (x < 2000000)).sum()) < n


def test_reference_bindex_defect_split_heavy_value():
    """D1: a value with more than N/K duplicates is split mid-run across areas
    (bindex.cpp:618-620 keeps the split when two area starts are equal).  A query bound just
    above it that is not itself an area start lands in the later area (in_which_area,
    bindex.cpp:1138-1151), whose refine covers only that area's share of the run, so
    lt/le miss (and gt/ge keep) the run's rows in the earlier area.  Which rows of the run
    those are depends on the reference's unstable sort; the count does not."""
    n = 524288
    rng = np.random.default_rng(12)
    x = rng.integers(0, 2 ** 32, n, dtype=np.uint32)
    c = np.uint32(2 ** 31 + 12345)
    heavy = rng.choice(n, 3 * (n // 128) + 17, replace=False)
    x[heavy] = c
    x[x == c + 1] = c + 2
    ob = O.Bindex(x)
    sv = ob.start_values
    assert (sv == c).sum() >= 2  # the run is split across areas
    got, truth = _ref_vs_truth(ob, x, O.LT, int(c) + 1)
    assert O.popcount_rows(got, n) < O.popcount_rows(truth, n)
    missed = unpack_rows(truth, n) & ~unpack_rows(got, n)
    assert (x[missed] == c).all()
    got_gt, truth_gt = _ref_vs_truth(ob, x, O.GE, int(c) + 1)
    assert O.popcount_rows(got_gt, n) > O.popcount_rows(truth_gt, n)
    R = ref_lib(n)
    rb = R.bindex(x)
    for op in (O.LT, O.LE, O.GE, O.GT):
        t = int(c) + 1 if op in (O.LT, O.GE) else int(c)
        assert same_up_to_ties(x, O.mask_tail(rb.scan(op, t), n), O.mask_tail(ob.scan(op, t), n))
        assert not np.array_equal(O.mask_tail(rb.scan(op, t), n),
                                  O.mask_tail(pack_rows(closed_form(x, op, t)), n))


def test_oracle_bindex_equals_raw_scan_on_distinct_uniform():
    """Where the reference is right (distinct-ish uniform data, in-range bounds) its BinDex
    result equals raw_scan for every operator."""
    n = 524288
    rng = np.random.default_rng(13)
    x = rng.integers(0, 2 ** 32, n, dtype=np.uint32)
    ob = O.Bindex(x)
    for _ in range(10):
        a, b = sorted(int(v) for v in rng.integers(2 ** 20, 2 ** 32 - 2 ** 20, 2))
        for op, t1, t2 in ((O.LT, a, 0), (O.LE, a, 0), (O.GT, a, 0), (O.GE, a, 0),
                           (O.BT, a, b), (O.EQ, int(x[rng.integers(n)]), 0)):
            got, truth = _ref_vs_truth(ob, x, op, t1, t2)
            assert np.array_equal(O.mask_tail(got, n), truth), (op, t1, t2)


def test_checks_catch_a_planted_mistake():
    """The closed-form check is sensitive: an oracle answer with an off-by-one bound
    (lt computed as le) or a flipped bit order differs from it."""
    rng = np.random.default_rng(14)
    x = rng.integers(0, 50, 4096, dtype=np.uint32)
    good = O.raw_scan(x, O.LT, 25)
    off_by_one = O.raw_scan(x, O.LE, 25)
    assert not np.array_equal(off_by_one, pack_rows(closed_form(x, O.LT, 25)))
    lsb_first = np.packbits(np.pad(x < 25, (0, 0)), bitorder="little").view("<u4")
    assert not np.array_equal(lsb_first, good)
    assert np.array_equal(unpack_rows(good, x.size), x < 25)


# ------------------------------------------------------------- refine-area surface

def test_refine_area_closes_the_sieve():
    """Pins the restatement of GPURefineAreaWithCuda (cuda_and.cu:510-640) by what the
    reference driver relies on (bindex_cuda.cpp:903-987 with refine_with_Cuda,
    bindex_cuda.cpp:1506-1660): for `lt c` with c inside area k, the filter vector
    fv[k-1] plus a non-inverse refine with exclusive bounds (start[k] - 1, c) is exactly
    x < c; fv[k] plus an inverse refine with (c - 1, start[k+1]) is exactly x < c too."""
    n = 524288
    rng = np.random.default_rng(21)
    x = rng.integers(0, 2 ** 32, n, dtype=np.uint32)
    ob = O.Bindex(x)
    sv, fv = ob.start_values, ob.fvs()
    for _ in range(20):
        k = int(rng.integers(1, 126))
        c = int(rng.integers(int(sv[k]) + 1, int(sv[k + 1])))
        want = pack_rows(x < c)
        bm = fv[k - 1].copy()
        assert O.refine_area([ob], [x], [int(sv[k]) - 1, c], 0, False, bm)
        assert np.array_equal(bm, want), (k, c)
        bm = fv[k].copy()
        assert O.refine_area([ob], [x], [c - 1, int(sv[k + 1])], 0, True, bm)
        assert np.array_equal(bm, want), (k, c)
    # second column bounds restrict the rows that change
    y = rng.integers(0, 2 ** 32, n, dtype=np.uint32)
    oy = O.Bindex(y)
    k = 50
    c = int(sv[k]) + (int(sv[k + 1]) - int(sv[k])) // 3
    bm = np.zeros(ob.words, np.uint32)
    assert O.refine_area([ob, oy], [x, y], [int(sv[k]) - 1, c, 5, 2 ** 31], 0, False, bm)
    assert np.array_equal(bm, pack_rows((x >= sv[k]) & (x < c) & (y > 5) & (y < 2 ** 31)))


def test_refine_eq_area_covers_the_value_range():
    """GPURefineEqAreaWithCuda (cuda_and.cu:642-812) refines every block from the smaller
    bound's block to the larger bound's: on a zero bitmap with exclusive bounds (a, b)
    spanning several areas it sets exactly a < x < b."""
    n = 524288
    rng = np.random.default_rng(22)
    x = rng.integers(0, 2 ** 32, n, dtype=np.uint32)
    ob = O.Bindex(x)
    for a, b in [(int(ob.start_values[3]) + 7, int(ob.start_values[90]) + 99),
                 (1000, 2 ** 31), (int(ob.start_values[10]) + 1, int(ob.start_values[10]) + 900000)]:
        bm = np.zeros(ob.words, np.uint32)
        assert O.refine_area([ob], [x], [a, b], 0, False, bm, eq=True)
        assert np.array_equal(bm, pack_rows((x > a) & (x < b))), (a, b)
        # reversed bounds: the block range is taken in order (cuda_and.cu:682-688) but the
        # row test keeps the given (lo, hi) (cuda_and.cu:83-86), so no row passes
        bm = np.zeros(ob.words, np.uint32)
        assert O.refine_area([ob], [x], [b, a], 0, False, bm, eq=True)
        assert not bm.any()


def test_bitmap_digest_properties():
    """The engine's bitmap checksum (not a reference function): it ignores bits past row n,
    sees a single flipped bit and a swap of two different words, and is a sum of per-word
    terms (the digest of a bitmap is the sum of the digests of its words kept alone, the
    other words zeroed, minus the zero terms)."""
    rng = np.random.default_rng(12)
    n = 32 * 40 + 13
    bm = rng.integers(0, 2 ** 32, (n + 31) // 32, dtype=np.uint64).astype(np.uint32)
    d = O.bitmap_digest(bm, n)
    tail = bm.copy()
    tail[-1] ^= np.uint32(1)  # a bit past row n
    assert O.bitmap_digest(tail, n) == d
    for row in (0, 31, 32, 640, n - 1):
        f = bm.copy()
        f[row // 32] ^= np.uint32(1 << (31 - row % 32))
        assert O.bitmap_digest(f, n) != d
    sw = bm.copy()
    sw[[3, 7]] = sw[[7, 3]]
    assert O.bitmap_digest(sw, n) != d
    zero = np.zeros_like(bm)
    dz = O.bitmap_digest(zero, n)
    parts = 0
    for i in range(bm.size):
        one = zero.copy()
        one[i] = bm[i]
        parts += O.bitmap_digest(one, n) - dz
    assert (parts + dz) % 2 ** 64 == d


@pytest.mark.parametrize("n", [1, 31, 33, 1000, 70001])
def test_conjunction_digest_matches_raw_scan_entry(n):
    """orc_conjunction_digest (no bitmap) equals the reference checker's path: raw_scan_entry
    per column into zeroed bitmaps with the AND merge (bindex.cpp:2265-2289), closed ranges
    written in the query-file grammar (bt lo-1 hi+1, ge lo, le hi), then orc_bitmap_digest
    and the popcount; with empty ranges, full ranges and 1..4 columns."""
    rng = np.random.default_rng(n + 3)
    cols = [rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32),
            rng.integers(0, 50, n, dtype=np.uint32),
            (rng.normal(2.0 ** 31, 2.0 ** 28, n).clip(0, MAXU)).astype(np.uint32),
            rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)]
    cases = [[(0, MAXU)], [(10, 20)], [(5, 4)], [(0, 0)], [(MAXU, MAXU)]]
    for _ in range(12):
        k = int(rng.integers(1, 5))
        rs = []
        for c in range(k):
            a, b = sorted(int(v) for v in rng.integers(0, 2 ** 32, 2))
            if c == 1:
                a, b = sorted(int(v) for v in rng.integers(0, 50, 2))
            rs.append((a, b))
        cases.append(rs)
    for rs in cases:
        use = cols[:len(rs)]
        W = O.words_for(n)
        bms = [np.zeros(W, np.uint32) for _ in use]
        for c, (col, (lo, hi)) in enumerate(zip(use, rs)):
            if lo > hi:
                pass  # no row: the zeroed bitmap stays empty
            elif lo == 0 and hi == MAXU:
                O.raw_scan_entry(col, O.GE, [0], bitmap=bms[c], merge=bms[0])
                continue
            elif lo == 0:
                O.raw_scan_entry(col, O.LE, [hi], bitmap=bms[c], merge=bms[0])
                continue
            elif hi == MAXU:
                O.raw_scan_entry(col, O.GE, [lo], bitmap=bms[c], merge=bms[0])
                continue
            else:
                O.raw_scan_entry(col, O.BT, [lo - 1], [hi + 1], bitmap=bms[c], merge=bms[0])
                continue
            if c > 0:
                bms[0] &= bms[c]
        want = (O.bitmap_digest(bms[0], n), O.popcount_rows(bms[0], n))
        assert O.conjunction_digest(use, rs, threads=3) == want, rs
        # and the closed form
        m = np.ones(n, bool)
        for col, (lo, hi) in zip(use, rs):
            m &= (col >= lo) & (col <= hi)
        assert want[1] == int(m.sum())


@pytest.mark.parametrize("n", [1, 31, 32, 33, 1000, 131073, 300001])
def test_conjunction_scan_port_matches_raw_scan_entry(n):
    """The vectorised port timed as the CPU baseline (AVX2 when the host has it) writes the
    same bitmap as raw_scan_entry per column + AND merge (bindex.cpp:2265-2289), with
    ragged tails, empty and full ranges and boundary values that hit lo and hi exactly."""
    rng = np.random.default_rng(n + 11)
    cols = [rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32),
            rng.integers(0, 40, n, dtype=np.uint32),
            rng.integers(2 ** 32 - 40, 2 ** 32, n, dtype=np.uint64).astype(np.uint32),
            rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)]
    cases = [[(0, MAXU)], [(3, 17)], [(MAXU - 5, MAXU)], [(9, 8)], [(0, 0), (0, 39)],
             [(0, MAXU), (5, 5), (MAXU - 39, MAXU - 2)]]
    for _ in range(10):
        k = int(rng.integers(1, 5))
        rs = []
        for c in range(k):
            if c == 1:
                rs.append(tuple(sorted(int(v) for v in rng.integers(0, 40, 2))))
            elif c == 2:
                rs.append(tuple(sorted(int(v) for v in rng.integers(2 ** 32 - 40, 2 ** 32, 2))))
            else:
                rs.append(tuple(sorted(int(v) for v in rng.integers(0, 2 ** 32, 2))))
        cases.append(rs)
    for rs in cases:
        use = cols[:len(rs)]
        m = np.ones(n, bool)
        for col, (lo, hi) in zip(use, rs):
            m &= (col >= lo) & (col <= hi)
        want = pack_rows(m)
        for threads in (1, 5):
            got = O.conjunction_scan(use, rs, threads=threads)
            assert np.array_equal(got, want), (rs, threads)
        # and the reference checker's own path, for the ranges it can express
        if all(lo <= hi for lo, hi in rs):
            W = O.words_for(n)
            bms = [np.zeros(W, np.uint32) for _ in use]
            for c, (col, (lo, hi)) in enumerate(zip(use, rs)):
                if lo == 0 and hi == MAXU:
                    O.raw_scan_entry(col, O.GE, [0], bitmap=bms[c], merge=bms[0])
                elif lo == 0:
                    O.raw_scan_entry(col, O.LE, [hi], bitmap=bms[c], merge=bms[0])
                elif hi == MAXU:
                    O.raw_scan_entry(col, O.GE, [lo], bitmap=bms[c], merge=bms[0])
                else:
                    O.raw_scan_entry(col, O.BT, [lo - 1], [hi + 1], bitmap=bms[c], merge=bms[0])
            assert np.array_equal(O.mask_tail(bms[0], n), want), rs
    assert O.simd_level() in (0, 2)
