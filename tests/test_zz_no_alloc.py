"""The allocation-free entry point (rts_conjunction_scan_noalloc, include/rtscan_hip.h): its
bitmaps and counts equal rts_conjunction_scan's on the same inputs, over two calls in a row
(the second on a smaller input, so leftovers of the first in the shared scratch would
show), on aligned and unaligned columns, one to six columns and an empty range; it uses no
per-thread scratch; too many rows are refused before any device work."""
import ctypes
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def call(eng, fn, cols, ranges, n, out, cnt):
    k = len(cols)
    ptrs = (ctypes.c_void_p * k)(*[c.ptr for c in cols])
    lo = np.array([r[0] for r in ranges], np.uint32)
    hi = np.array([r[1] for r in ranges], np.uint32)
    eng.check(fn(ptrs, lo.ctypes.data, hi.ctypes.data, k, n, out.ptr, cnt.ptr if cnt else None, None), "scan")
    eng.synchronize()


class Offset:
    def __init__(self, buf, off):
        self.ptr = buf.ptr + off


def test_noalloc_equals_the_allocating_entry_point(eng):
    L = eng.lib()
    rng = np.random.default_rng(77)
    cases = [  # (rows, columns, unaligned): the second call is the smaller one
        (10_000_019, 3, False), (100_003, 1, False),
        (2_000_001, 6, False), (70_001, 2, True),
        (1 << 20, 4, False), (33, 1, False),
    ]
    held = []

    def run():
        for n, k, unaligned in cases:
            print("noalloc case n=%d k=%d unaligned=%s" % (n, k, unaligned), flush=True)
            xs = [rng.integers(0, 2 ** 32, n + 1, dtype=np.uint64).astype(np.uint32) for _ in range(k)]
            bufs = [eng.DeviceBuffer.from_numpy(x) for x in xs]
            cols = [Offset(b, 4) if unaligned else b for b in bufs]
            ranges = []
            for _ in range(k):
                a, b = sorted(int(v) for v in rng.integers(0, 2 ** 32, 2))
                ranges.append((a, max(b, a + (1 << 31))) if len(ranges) % 2 else (a, b))
            ranges = [(min(a, 0xFFFFFFFF), min(b, 0xFFFFFFFF)) for a, b in ranges]
            W = (n + 31) // 32
            o1, o2 = eng.DeviceBuffer(4 * W), eng.DeviceBuffer(4 * W)
            c1, c2 = eng.DeviceBuffer(8), eng.DeviceBuffer(8)
            o1.fill(0x11)
            o2.fill(0x22)
            call(eng, L.rts_conjunction_scan_noalloc, cols, ranges, n, o1, c1)
            call(eng, L.rts_conjunction_scan, cols, ranges, n, o2, c2)
            assert np.array_equal(o1.to_numpy(np.uint32, W), o2.to_numpy(np.uint32, W)), (n, k)
            got, want = int(c1.to_numpy(np.uint64, 1)[0]), int(c2.to_numpy(np.uint64, 1)[0])
            assert got == want, (n, k, got, want)
            # and against the closed form
            m = np.ones(n, bool)
            off = 1 if unaligned else 0
            for x, (a, b) in zip(xs, ranges):
                m &= (x[off:off + n] >= a) & (x[off:off + n] <= b)
            assert got == int(m.sum())
            # an empty range: zero bitmap and count, no scan
            call(eng, L.rts_conjunction_scan_noalloc, cols[:1], [(9, 8)], n, o1, c1)
            assert int(c1.to_numpy(np.uint64, 1)[0]) == 0 and not o1.to_numpy(np.uint32, W).any()
            for b in bufs + [o1, o2, c1, c2]:
                b.free()
        held.append(L.rts_thread_scratch_buffers())

    # a fresh thread holds no scratch: the no-alloc calls must not create any
    t = threading.Thread(target=lambda: (run_noalloc_only(eng, held)))
    t.start()
    t.join()
    assert held == [0]
    run()


def run_noalloc_only(eng, held):
    L = eng.lib()
    n = 1_000_003
    x = np.random.default_rng(5).integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    d = eng.DeviceBuffer.from_numpy(x)
    o = eng.DeviceBuffer(4 * ((n + 31) // 32))
    c = eng.DeviceBuffer(8)
    for lo, hi in ((1 << 30, 3 << 30), (5, 1 << 20)):
        call(eng, L.rts_conjunction_scan_noalloc, [d], [(lo, hi)], n, o, c)
        assert int(c.to_numpy(np.uint64, 1)[0]) == int(((x >= lo) & (x <= hi)).sum())
    held.append(L.rts_thread_scratch_buffers())


def test_noalloc_refuses_too_many_rows(eng):
    L = eng.lib()
    ptrs = (ctypes.c_void_p * 1)(None)
    lo = np.zeros(1, np.uint32)
    hi = np.ones(1, np.uint32)
    rc = L.rts_conjunction_scan_noalloc(ptrs, lo.ctypes.data, hi.ctypes.data, 1, (1 << 40) + 1, None, None, None)
    assert rc != 0
    assert b"RTS_NOALLOC_MAX_ROWS" in L.rts_last_error()
