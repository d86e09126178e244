"""GPU parity of the value encoders (include/rts_remap.h) with the oracle's restatement of
normalEncode (remap.cpp:110-241)."""
import numpy as np
import pytest

import oracle
import rtscan_amd as R

pytestmark = pytest.mark.gpu

MAXU = 0xFFFFFFFF


def columns():
    n = 1_000_003
    yield "uniform", R.generate_column(R.DIST_UNIFORM, 4, 0, n)
    yield "zipf1.1", R.generate_column(R.DIST_ZIPF, 4, 1, n, 1.1, 2.0 ** 16)
    yield "few", (R.generate_column(R.DIST_UNIFORM, 4, 2, 5000) % 7).astype(np.uint32)
    yield "one", np.full(777, 123456, np.uint32)


@pytest.mark.parametrize("name,col", list(columns()), ids=lambda v: v if isinstance(v, str) else "")
@pytest.mark.parametrize("lo,hi", [(0, None), (2, 4294967294)])
def test_rank_encoding_matches_oracle(eng, name, col, lo, hi):
    n = col.size
    hi = n if hi is None else hi
    d = R.DeviceBuffer.from_numpy(col)
    codes, mv, ms, sp, sc = R.encode_column(d, n, lo, hi, want_sorted=True)
    o_codes, o_pos, o_sc, o_mv, o_ms = oracle.normal_encode(col, lo, hi)
    assert np.array_equal(mv, o_mv) and np.array_equal(ms, o_ms)
    assert np.array_equal(sp.to_numpy(np.int32, n), o_pos)
    assert np.array_equal(sc.to_numpy(np.uint32, n), o_sc)
    assert np.array_equal(codes.to_numpy(np.uint32, n), o_codes)
    # in place: codes written over the column
    R.encode_column(d, n, lo, hi, codes=d)
    assert np.array_equal(d.to_numpy(np.uint32, n), o_codes)


@pytest.mark.parametrize("name,col", list(columns()), ids=lambda v: v if isinstance(v, str) else "")
def test_dispersed_encoding_map_and_ranges(eng, name, col):
    n = col.size
    lo, hi = 2, 4294967294
    d = R.DeviceBuffer.from_numpy(col)
    codes, mv, ms = R.encode_column_dispersed(d, n, lo, hi, seed=9)
    o_mv, o_ms, o_st = oracle.dispersed_map(col, lo, hi)
    assert np.array_equal(mv, o_mv) and np.array_equal(ms, o_ms)
    c = codes.to_numpy(np.uint32, n).astype(np.int64)
    r = np.searchsorted(o_mv, col)
    start = o_ms[r].astype(np.int64)
    step = o_st[r].astype(np.int64)
    off = (c - start) & MAXU
    width = np.where(step >= 2, step - 1, 1)
    assert (off < width).all()  # each row's code lies in its value's range
    again, _, _ = R.encode_column_dispersed(d, n, lo, hi, seed=9)
    assert np.array_equal(again.to_numpy(np.uint32, n), c.astype(np.uint32))  # deterministic


def test_encoded_scan_equals_value_scan(eng):
    """Rank codes keep value order: a closed interval of values is the closed interval of
    codes between the lookups, and the GPU scan over codes selects the same rows."""
    col = R.generate_column(R.DIST_ZIPF, 8, 3, 1 << 20, 1.5, 2.0 ** 12)
    n = col.size
    d = R.DeviceBuffer.from_numpy(col)
    codes, mv, ms = R.encode_column(d, n, 0, n)
    rng = np.random.default_rng(1)
    out_v = R.DeviceBuffer(4 * (n // 32))
    out_c = R.DeviceBuffer(4 * (n // 32))
    for _ in range(10):
        a, b = sorted(int(v) for v in rng.choice(col, 2))
        ca = R.encode_lookup(mv, ms, a, R.GE, n)  # first code of the first value >= a
        cb = R.encode_lookup(mv, ms, b, R.LE, n)  # last code of the last value <= b
        cb = n - 1 if cb == n else cb
        R.scan_range(d, n, a, b, out_v)
        R.scan_range(codes, n, ca, cb, out_c)
        assert np.array_equal(out_v.to_numpy(np.uint32), out_c.to_numpy(np.uint32))
