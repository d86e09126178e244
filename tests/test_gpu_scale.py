"""Parity at the full serving size (1e9 rows of a Zipf column), through size-independent
properties: the GPU scan's bitmap has the oracle raw_scan's digest and popcount, and the
BinDex query and the batched launch over the same column produce bitmaps with the scan's
digest (so all three equal the oracle at 1e9 rows without holding three host bitmaps)."""
import types

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

N = 1_000_000_000
W = (N + 31) // 32
MAXU = 0xFFFFFFFF


@pytest.fixture(scope="module")
def big(eng):
    x = eng.generate_column(eng.DIST_ZIPF, 21, 0, N, 1.1, 2.0 ** 20)
    print(f"[scale] generated {N} rows", flush=True)
    d = eng.DeviceBuffer.from_numpy(x)
    xs = np.sort(x[::1000])
    # wide, leading, single-value (inside one bin) and wide-high ranges
    preds = [(int(xs[int(a * (xs.size - 1))]), int(xs[int(b * (xs.size - 1))]))
             for a, b in ((0.1, 0.6), (0.0, 0.2), (0.5, 0.5001), (0.3, 0.95))]
    yield x, d, preds
    d.free()


def oracle_bitmap(x, lo, hi):
    """raw_scan's exclusive bt (bindex.cpp:1857-1874) for the closed range [lo, hi]."""
    assert hi < MAXU
    return oracle.raw_scan(x, oracle.BT, lo - 1, hi + 1, threads=16) if lo > 0 else \
        oracle.raw_scan(x, oracle.LE, hi, threads=16)


def test_scan_matches_oracle_digest_and_count(eng, big):
    x, d, preds = big
    out = eng.DeviceBuffer(4 * W)
    cnt = eng.DeviceBuffer(8)
    for lo, hi in preds:
        print(f"[scale] scan [{lo}, {hi}]", flush=True)
        eng.scan_range(d, N, lo, hi, out, cnt)
        truth = oracle_bitmap(x, lo, hi)
        assert eng.bitmap_digest(out, N) == oracle.bitmap_digest(truth, N), (lo, hi)
        assert int(cnt.to_numpy(np.uint64, 1)[0]) == oracle.popcount_rows(truth, N), (lo, hi)
    out.free()
    cnt.free()


def test_scan_ranges_equal_scan(eng, big):
    """rts_scan_ranges over the 1e9-row column: every bitmap and count equal the one-range
    scan's (tiles of 2048 rows, one word per lane, plus the ragged tail rows)."""
    x, d, preds = big
    k = len(preds)
    stride = W + 3
    out = eng.DeviceBuffer(4 * stride * k)
    counts = eng.DeviceBuffer(8 * k)
    print("[scale] rts_scan_ranges of %d predicates" % k, flush=True)
    eng.scan_ranges(d, N, [p[0] for p in preds], [p[1] for p in preds], out, stride, counts)
    got = [int(c) for c in counts.to_numpy(np.uint64, k)]
    one = eng.DeviceBuffer(4 * W)
    cnt = eng.DeviceBuffer(8)
    for i, (lo, hi) in enumerate(preds):
        eng.scan_range(d, N, lo, hi, one, cnt)
        view = types.SimpleNamespace(ptr=out.ptr + 4 * stride * i)
        assert eng.bitmap_digest(view, N) == eng.bitmap_digest(one, N), (lo, hi)
        assert got[i] == int(cnt.to_numpy(np.uint64, 1)[0]), (lo, hi)
    for b in (out, counts, one, cnt):
        b.free()


def test_bindex_query_and_batch_equal_scan(eng, big):
    x, d, preds = big
    print("[scale] BinDex build", flush=True)
    ix = eng.BinDexIndex(d, N)
    # a range strictly inside the widest bin: both bounds in one bin, no filter vector read
    st = ix.start_values().astype(np.int64)
    k = int(np.argmax(st[1:] - st[:-1]))
    assert st[k + 1] - st[k] > 2
    narrow = (int(st[k] + 1), int((st[k] + st[k + 1]) // 2))
    truth = oracle_bitmap(x, *narrow)
    narrow_digest = oracle.bitmap_digest(truth, N)
    del truth
    preds = preds + [narrow]
    scan_out = eng.DeviceBuffer(4 * W)
    b_out = eng.DeviceBuffer(4 * W)
    cnt = eng.DeviceBuffer(8)
    want, want_count = [], []
    for lo, hi in preds:
        print(f"[scale] BinDex query [{lo}, {hi}]", flush=True)
        eng.scan_range(d, N, lo, hi, scan_out, cnt)
        want.append(eng.bitmap_digest(scan_out, N))
        want_count.append(int(cnt.to_numpy(np.uint64, 1)[0]))
        eng.bindex_query([ix], [lo], [hi], b_out, cnt)
        assert eng.bitmap_digest(b_out, N) == want[-1], (lo, hi)
        assert int(cnt.to_numpy(np.uint64, 1)[0]) == want_count[-1], (lo, hi)
    assert want[-1] == narrow_digest
    cnt.free()
    scan_out.free()
    b_out.free()

    print("[scale] batched launch", flush=True)
    plan = eng.BatchPlan([d], [ix], N, len(preds))
    plan.set_queries([[(0, lo, hi, 0)] for lo, hi in preds])
    assert plan.stats()["narrow_cols"] >= 1
    stride = (W + 3) // 4 * 4
    out = eng.DeviceBuffer(4 * stride * len(preds))
    counts = eng.DeviceBuffer(8 * len(preds))
    plan.launch(out, stride, counts)
    assert [int(c) for c in counts.to_numpy(np.uint64, len(preds))] == want_count
    for i in range(len(preds)):
        view = types.SimpleNamespace(ptr=out.ptr + 4 * stride * i)
        assert eng.bitmap_digest(view, N) == want[i], (i, preds[i])
    plan.free()
    out.free()
    counts.free()
    ix.free()
