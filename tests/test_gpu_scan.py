"""GPU parity of the direct scan (kernel a), the fused conjunction / AND+popcount (kernel c)
and the cuda_and.cu surface, through the C ABI, against the CPU oracle and numpy closed
forms.  Bit-exact: this is integer predicate work."""
import ctypes

import numpy as np
import pytest

import oracle as O
from _util import MAXU, closed_form, config1_predicates, interval_mask, pack_rows

pytestmark = pytest.mark.gpu


class View:
    """A device pointer offset into a DeviceBuffer (to test unaligned inputs)."""

    def __init__(self, buf, byte_offset):
        self.buf = buf
        self.ptr = buf.ptr + byte_offset


def upload(eng, arr):
    return eng.DeviceBuffer.from_numpy(np.ascontiguousarray(arr))


def read_bitmap(buf, n):
    return buf.to_numpy(np.uint32, (n + 31) // 32)


def read_count(eng, cbuf):
    return int(cbuf.to_numpy(np.uint64, 1)[0])


def ref_predicate(lo, hi):
    """The reference query line for the closed range [lo, hi] (bt is open)."""
    if lo == 0 and hi == MAXU:
        return O.GE, 0, 0
    if lo == 0:
        return O.LE, hi, 0
    if hi == MAXU:
        return O.GE, lo, 0
    return O.BT, lo - 1, hi + 1


def test_config1_plain_scan(eng):
    """2^20 uniform rows, 64 closed ranges with selectivity cycling through
    {0.001, 0.01, 0.1, 0.5, 0.9}: bitmap and popcount equal the oracle's raw_scan."""
    n = 1 << 20
    x = eng.generate_column(eng.DIST_UNIFORM, 1, 0, n)
    dx = upload(eng, x)
    out = eng.DeviceBuffer(4 * (n // 32))
    cnt = eng.DeviceBuffer(8)
    for lo, hi in config1_predicates(64):
        eng.scan_range(dx, n, lo, hi, out, cnt)
        got = read_bitmap(out, n)
        op, t1, t2 = ref_predicate(lo, hi)
        truth = O.raw_scan(x, op, t1, t2)
        assert np.array_equal(got, truth), (lo, hi)
        assert read_count(eng, cnt) == O.popcount_rows(truth, n)


def test_config1_ranges_in_one_pass(eng):
    """The same 64 ranges through rts_scan_ranges (the column read once for all of them):
    every bitmap and count equal the oracle's raw_scan."""
    n = 1 << 20
    x = eng.generate_column(eng.DIST_UNIFORM, 1, 0, n)
    dx = upload(eng, x)
    W = n // 32
    preds = config1_predicates(64)
    out = eng.DeviceBuffer(4 * W * len(preds))
    cnt = eng.DeviceBuffer(8 * len(preds))
    eng.scan_ranges(dx, n, [p[0] for p in preds], [p[1] for p in preds], out, W, cnt)
    got = out.to_numpy(np.uint32, W * len(preds)).reshape(len(preds), W)
    counts = cnt.to_numpy(np.uint64, len(preds))
    for i, (lo, hi) in enumerate(preds):
        op, t1, t2 = ref_predicate(lo, hi)
        truth = O.raw_scan(x, op, t1, t2)
        assert np.array_equal(got[i], truth), (lo, hi)
        assert int(counts[i]) == O.popcount_rows(truth, n)


SIZES = [1, 2, 31, 32, 33, 63, 64, 65, 255, 256, 257, 1023, 1024, 1025, 2047, 2048, 4097,
         65536 + 77, 300001]


@pytest.mark.parametrize("n", [1, 33, 1023, 1024, 1025, 4097, 65536 + 77, 300001])
@pytest.mark.parametrize("npred", [1, 7, 64, 150])
def test_scan_ranges_sizes_passes_and_padding(eng, n, npred):
    """Whole tiles and ragged tails, one pass (<= 64 predicates) and several (150), empty
    ranges (lo > hi) and the full range; words between bitmaps (stride padding) untouched."""
    rng = np.random.default_rng(n * 1000 + npred)
    x = rng.integers(0, 2 ** 32, n, dtype=np.uint32)
    x[: min(n, 3)] = [0, MAXU, 7][: min(n, 3)]
    dx = upload(eng, x)
    W = (n + 31) // 32
    stride = W + 5
    fixed = [(0, MAXU), (0, 0), (MAXU, MAXU), (9, 3), (int(x[n // 2]), int(x[n // 2]))]
    preds = []
    for i in range(npred):
        if i < len(fixed):
            preds.append(fixed[i])
        else:
            a, b = sorted(int(v) for v in rng.integers(0, 2 ** 32, 2))
            preds.append((a, b))
    out = eng.DeviceBuffer(4 * stride * npred)
    out.fill(0x5A)
    cnt = eng.DeviceBuffer(8 * npred)
    eng.scan_ranges(dx, n, [p[0] for p in preds], [p[1] for p in preds], out, stride, cnt)
    got = out.to_numpy(np.uint32, stride * npred).reshape(npred, stride)
    counts = cnt.to_numpy(np.uint64, npred)
    for i, (lo, hi) in enumerate(preds):
        m = interval_mask(x, lo, hi) if lo <= hi else np.zeros(n, bool)
        assert np.array_equal(got[i, :W], pack_rows(m)), (n, i, lo, hi)
        assert np.all(got[i, W:] == 0x5A5A5A5A), (n, i)
        assert int(counts[i]) == int(m.sum()), (n, i, lo, hi)


@pytest.mark.parametrize("offset_rows", [1, 3])
def test_scan_ranges_unaligned_column_without_counts(eng, offset_rows):
    n = 70000 + offset_rows
    rng = np.random.default_rng(offset_rows + 40)
    x = rng.integers(0, 1000, n + 8, dtype=np.uint32)
    dx = upload(eng, x)
    view = View(dx, 4 * offset_rows)
    xs = x[offset_rows: offset_rows + n]
    W = (n + 31) // 32
    preds = [(100, 499), (0, 0), (999, 999), (1, 998)]
    out = eng.DeviceBuffer(4 * W * len(preds))
    eng.scan_ranges(view, n, [p[0] for p in preds], [p[1] for p in preds], out, W)
    got = out.to_numpy(np.uint32, W * len(preds)).reshape(len(preds), W)
    for i, (lo, hi) in enumerate(preds):
        assert np.array_equal(got[i], pack_rows(interval_mask(xs, lo, hi))), (i, lo, hi)


@pytest.mark.parametrize("n", SIZES)
def test_scan_sizes_and_tails(eng, n):
    rng = np.random.default_rng(n)
    x = rng.integers(0, 2 ** 32, n, dtype=np.uint32)
    x[: min(n, 4)] = [0, MAXU, 5, 6][: min(n, 4)]
    dx = upload(eng, x)
    W = (n + 31) // 32
    out = eng.DeviceBuffer(4 * W)
    cnt = eng.DeviceBuffer(8)
    for lo, hi in [(0, MAXU), (0, 0), (MAXU, MAXU), (5, 6), (1 << 30, 3 << 30),
                   (int(x[n // 2]), int(x[n // 2]))]:
        out.fill(0xAB)
        eng.scan_range(dx, n, lo, hi, out, cnt)
        want = pack_rows(interval_mask(x, lo, hi))
        assert np.array_equal(read_bitmap(out, n), want), (n, lo, hi)
        assert read_count(eng, cnt) == int(interval_mask(x, lo, hi).sum())


@pytest.mark.parametrize("offset_rows", [1, 2, 3, 5])
def test_scan_unaligned_column(eng, offset_rows):
    n = 10000 + offset_rows
    rng = np.random.default_rng(offset_rows)
    x = rng.integers(0, 1000, n + 8, dtype=np.uint32)
    dx = upload(eng, x)
    view = View(dx, 4 * offset_rows)
    out = eng.DeviceBuffer(4 * ((n + 31) // 32))
    cnt = eng.DeviceBuffer(8)
    eng.scan_range(view, n, 100, 499, out, cnt)
    xs = x[offset_rows: offset_rows + n]
    assert np.array_equal(read_bitmap(out, n), pack_rows(interval_mask(xs, 100, 499)))
    assert read_count(eng, cnt) == int(interval_mask(xs, 100, 499).sum())


def test_scan_empty_predicate_clears(eng):
    n = 5000
    x = np.arange(n, dtype=np.uint32)
    dx = upload(eng, x)
    out = eng.DeviceBuffer(4 * ((n + 31) // 32)).fill(0xFF)
    cnt = eng.DeviceBuffer(8).fill(0x11)
    eng.scan_range(dx, n, 10, 9, out, cnt)  # lo > hi: empty
    assert not read_bitmap(out, n).any()
    assert read_count(eng, cnt) == 0


def test_scan_xor_accumulates_like_raw_scan(eng):
    """rts_scan_range_xor toggles (raw_scan's refine, bindex.cpp:1217): comma-separated
    targets of one query line XOR together (bindex.cpp:1897-1930)."""
    n = 100000
    x = eng.generate_column(eng.DIST_UNIFORM, 7, 0, n)
    dx = upload(eng, x)
    out = eng.DeviceBuffer(4 * ((n + 31) // 32)).fill(0)
    targets = [1 << 30, 1 << 31, 3 << 30]
    for t in targets:
        lo, hi = eng.predicate_interval(eng.LT, t)
        eng.scan_range_xor(dx, n, lo, hi, out)
    want = O.raw_scan_entry(x, O.LT, targets)
    assert np.array_equal(read_bitmap(out, n), want)


@pytest.mark.parametrize("ncols", [1, 2, 3, 4, 5, 7])
def test_conjunction_scan(eng, ncols):
    n = 200003
    rng = np.random.default_rng(ncols)
    cols = [eng.generate_column(eng.DIST_UNIFORM, 11, c, n) for c in range(ncols)]
    dcols = [upload(eng, c) for c in cols]
    out = eng.DeviceBuffer(4 * ((n + 31) // 32))
    cnt = eng.DeviceBuffer(8)
    for trial in range(4):
        los, his, preds = [], [], []
        for c in range(ncols):
            w = int(2 ** 32 * (0.5 + 0.45 * rng.random()) ** (1.0 / ncols))
            lo = int(rng.integers(0, 2 ** 32 - w))
            los.append(lo)
            his.append(lo + w - 1)
            op, t1, t2 = ref_predicate(lo, lo + w - 1)
            preds.append((op, [t1], [t2] if op == O.BT else None))
        eng.conjunction_scan(dcols, los, his, n, out, cnt)
        truth = O.conjunction(cols, preds)
        assert np.array_equal(read_bitmap(out, n), truth)
        assert read_count(eng, cnt) == O.popcount_rows(truth, n)


@pytest.mark.parametrize("nb", [1, 2, 3, 4, 8, 9, 12])
def test_bitmap_and_popcount(eng, nb):
    n = 100000 + nb
    W = (n + 31) // 32
    rng = np.random.default_rng(nb)
    bms = [rng.integers(0, 2 ** 32, W, dtype=np.uint32) | np.uint32(rng.integers(0, 2 ** 32))
           for _ in range(nb)]
    dbs = [upload(eng, b) for b in bms]
    out = eng.DeviceBuffer(4 * W)
    cnt = eng.DeviceBuffer(8)
    eng.bitmap_and_popcount(dbs, n, out, cnt)
    want = np.bitwise_and.reduce(np.stack(bms), axis=0)
    want = O.mask_tail(want, n)
    assert np.array_equal(read_bitmap(out, n), want)
    assert read_count(eng, cnt) == O.popcount_rows(want, n)
    cnt2 = eng.DeviceBuffer(8)
    eng.bitmap_popcount(dbs[0], n, cnt2)
    assert read_count(eng, cnt2) == O.popcount_rows(bms[0], n)
    assert np.array_equal(dbs[0].to_numpy(np.uint32, W), bms[0])  # popcount does not write


@pytest.mark.parametrize("words", [1, 3, 4, 5, 1000, 1027, 3125000 // 100])
def test_cuda_and_surface(eng, words):
    """GPUbit*WithCuda: units and operand order of cuda_and.cu (and/or/memset count rows,
    copy/negation/SIMD count words; SIMD is result = b & ~a)."""
    L = eng.lib()
    rng = np.random.default_rng(words)
    a = rng.integers(0, 2 ** 32, words, dtype=np.uint32)
    b = rng.integers(0, 2 ** 32, words, dtype=np.uint32)
    rows = words * 32 - 5  # bits_num_needed(rows) == words
    for name, args, want in [
        ("GPUbitAndWithCuda", "ab_rows", a & b),
        ("GPUbitOrWithCuda", "ab_rows", a | b),
        ("GPUbitCopyWithCuda", "ab_words", b),
        ("GPUbitCopyNegationWithCuda", "ab_words", ~b),
    ]:
        da, db = upload(eng, a), upload(eng, b)
        n_arg = rows if args == "ab_rows" else words
        assert getattr(L, name)(da.ptr, db.ptr, n_arg) == 0
        assert np.array_equal(da.to_numpy(np.uint32, words), want), name
    da, db, dr = upload(eng, a), upload(eng, b), eng.DeviceBuffer(4 * words)
    assert L.GPUbitCopySIMDWithCuda(dr.ptr, da.ptr, db.ptr, words) == 0
    assert np.array_equal(dr.to_numpy(np.uint32, words), b & ~a)
    assert L.GPUbitMemsetWithCuda(da.ptr, 0x12345678, rows) == 0
    assert (da.to_numpy(np.uint32, words) == 0x12345678).all()
    # unaligned operands
    if words > 8:
        da, db = upload(eng, a), upload(eng, b)
        assert L.GPUbitAndWithCuda(da.ptr + 4, db.ptr + 8, (words - 3) * 32) == 0
        got = da.to_numpy(np.uint32, words)
        assert np.array_equal(got[1:words - 2], a[1:words - 2] & b[2:words - 1])
        assert got[0] == a[0] and got[words - 1] == a[words - 1]


def test_host_bitmap_wrappers(eng):
    L = eng.lib()
    rng = np.random.default_rng(3)
    rows = 1000
    W = (rows - 1) // 32 + 1
    a = rng.integers(0, 2 ** 32, W, dtype=np.uint32)
    b = rng.integers(0, 2 ** 32, W, dtype=np.uint32)
    aa = a.copy()
    assert L.CPUbitAndWithCuda(aa.ctypes.data, b.ctypes.data, rows) == 0
    assert np.array_equal(aa, a & b)
    # refineWithCuda flips rows 0..size-1 of a host bitmap of `size` words
    size = 45
    h = rng.integers(0, 2 ** 32, size, dtype=np.uint32)
    hh = h.copy()
    assert L.refineWithCuda(hh.ctypes.data, size) == 0
    flip = pack_rows(np.arange(size * 32) < size)
    assert np.array_equal(hh, h ^ flip)


def test_raw_scan_capi(eng):
    """raw_scan (bindex.h:146) through the C ABI: host column, host bitmap, XOR semantics."""
    import rtscan_amd as R
    n = 77777
    x = R.generate_column(R.DIST_UNIFORM, 5, 0, n)
    # the BinDex struct: raw_scan reads only `length`; build one via the descriptor layout
    from rtscan_amd.bindex_struct import new_bindex_struct
    bx = new_bindex_struct(length=n)
    for op, t1, t2 in [(O.LT, 1 << 31, 0), (O.BT, 5, 1 << 30), (O.EQ, int(x[3]), 0),
                       (O.GE, MAXU, 0), (O.BT, 1 << 31, 7)]:
        bm = np.full((n - 1) // 32 + 1, 0x5A5A5A5A, np.uint32)
        want = O.raw_scan(x, op, t1, t2, bitmap=bm.copy())
        eng.lib().raw_scan(ctypes.byref(bx), bm.ctypes.data, t1, t2, op, x.ctypes.data, None)
        # raw_scan itself does not swap bt bounds (only raw_scan_entry does): bt 2^31 7 is empty
        assert np.array_equal(bm, want), (op, t1, t2)


def test_config2_scale_plain_scan(eng):
    """1e8 uniform rows: full bitmaps and popcounts of a few predicates equal the oracle."""
    n = 100_000_000
    x = eng.generate_column(eng.DIST_UNIFORM, 2, 0, n)
    dx = upload(eng, x)
    out = eng.DeviceBuffer(4 * (n // 32))
    cnt = eng.DeviceBuffer(8)
    for lo, hi in [(123456789, 123456789 + 429496729), (0, 2 ** 31), (17, MAXU - 17)]:
        eng.scan_range(dx, n, lo, hi, out, cnt)
        got = read_bitmap(out, n)
        op, t1, t2 = ref_predicate(lo, hi)
        truth = O.raw_scan(x, op, t1, t2, threads=16)
        assert np.array_equal(got, truth)
        assert read_count(eng, cnt) == O.popcount_rows(truth, n)


def test_counted_scans_from_several_threads_on_one_stream(eng):
    """Host threads that count on the same (null) stream at the same time each get their
    own count scratch: every count is right (the reference calls its wrappers from one host
    thread per column, bindex_cuda.cpp:1901-1914)."""
    import threading
    n = 1 << 20
    rng = np.random.default_rng(8)
    cols = [rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32) for _ in range(4)]
    errors = []

    def worker(k):
        try:
            d = eng.DeviceBuffer.from_numpy(cols[k])
            out = eng.DeviceBuffer(4 * (n // 32))
            cnt = eng.DeviceBuffer(8)
            r = np.random.default_rng(k)
            for _ in range(60):
                lo, hi = sorted(int(v) for v in r.integers(0, 2 ** 32, 2))
                eng.scan_range(d, n, lo, hi, out, cnt)  # null stream, synchronous per call
                eng.synchronize()
                got = int(cnt.to_numpy(np.uint64, 1)[0])
                want = int(((cols[k] >= lo) & (cols[k] <= hi)).sum())
                if got != want:
                    errors.append((k, lo, hi, got, want))
        except Exception as e:  # noqa: BLE001
            errors.append(repr(e))

    ts = [threading.Thread(target=worker, args=(k,)) for k in range(4)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errors, errors[:3]


def test_reference_wrappers_from_several_threads(eng):
    """The reference-named wrappers run on each calling thread's default stream and return
    after their work: several host threads calling them at once (as the reference's driver
    does, one thread per column, bindex_cuda.cpp:1901-1914) each get their own right
    result."""
    import threading
    L = eng.lib()
    n = 1 << 20
    W = n // 32
    errors = []

    def worker(k):
        try:
            r = np.random.default_rng(100 + k)
            for _ in range(30):
                a = r.integers(0, 2 ** 32, W, dtype=np.uint64).astype(np.uint32)
                b = r.integers(0, 2 ** 32, W, dtype=np.uint64).astype(np.uint32)
                da, db = eng.DeviceBuffer.from_numpy(a), eng.DeviceBuffer.from_numpy(b)
                eng.check(L.GPUbitAndWithCuda(da.ptr, db.ptr, n), "and")
                if not np.array_equal(da.to_numpy(np.uint32, W), a & b):
                    errors.append(("and", k))
                eng.check(L.GPUbitCopyNegationWithCuda(db.ptr, da.ptr, W), "not")
                if not np.array_equal(db.to_numpy(np.uint32, W), ~(a & b)):
                    errors.append(("not", k))
        except Exception as e:  # noqa: BLE001
            errors.append(repr(e))

    ts = [threading.Thread(target=worker, args=(k,)) for k in range(4)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errors, errors[:3]


def test_scratch_is_bounded_per_thread_and_freed_with_the_thread(eng):
    """Counted scans on many streams keep at most 16 scratch buffers in the calling thread
    (the oldest is freed first) with every count right; threads that exit release theirs
    (the device's free memory comes back after 40 short-lived counting threads)."""
    import threading
    L = eng.lib()
    n = 1 << 16
    x = np.random.default_rng(9).integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    d = eng.DeviceBuffer.from_numpy(x)
    out = eng.DeviceBuffer(4 * (n // 32))
    cnt = eng.DeviceBuffer(8)
    errors = []

    def counted_scans(nstreams):
        streams = [eng.Stream() for _ in range(nstreams)]
        for i, s in enumerate(streams):
            lo, hi = 1000 * i, (1 << 31) + 77 * i
            eng.scan_range(d, n, lo, hi, out, cnt, stream=s)
            s.synchronize()
            if int(cnt.to_numpy(np.uint64, 1)[0]) != int(((x >= lo) & (x <= hi)).sum()):
                errors.append(i)
        held = L.rts_thread_scratch_buffers()
        for s in streams:
            s.destroy()
        return held

    def worker(res):
        try:
            res.append(counted_scans(3))
        except Exception as e:  # noqa: BLE001
            errors.append(repr(e))

    held = []
    t = threading.Thread(target=worker, args=(held,))
    t.start()
    t.join()
    assert held == [3]
    eng.synchronize()
    free0 = eng.device_free_bytes()
    for _ in range(40):
        t = threading.Thread(target=worker, args=(held,))
        t.start()
        t.join()
    eng.synchronize()
    # 40 threads x 3 buffers x 256 KB would be 30 MB if nothing were freed
    assert eng.device_free_bytes() >= free0 - (8 << 20)
    res = []
    t = threading.Thread(target=lambda: res.append(counted_scans(24)))
    t.start()
    t.join()
    assert res == [16]
    assert not errors, errors[:3]
