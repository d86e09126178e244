"""GPU parity of the device-side checker (include/rts_check.h): compare_bitmap semantics
(bindex.cpp:1959-1978) and the bitmap digest equal the oracle's; pinned-memory egress."""
import ctypes

import numpy as np
import pytest

import oracle
import rtscan_amd as R

pytestmark = pytest.mark.gpu


def rand_bitmap(rng, n, density=0.3):
    return np.packbits(rng.random(n) < density, bitorder="big").view(np.uint8)


def words(rng, n, density=0.3):
    m = rng.random(((n + 31) // 32) * 32) < density
    return np.packbits(m, bitorder="big").view(">u4").astype(np.uint32)


@pytest.mark.parametrize("n", [1, 31, 32, 33, 1000, 262144 + 7, 3_000_001, 33_554_432])
def test_compare_matches_oracle(eng, n):
    rng = np.random.default_rng(n % 1000)
    truth = words(rng, n)
    dt = R.DeviceBuffer.from_numpy(truth)
    r = R.bitmap_compare(dt, dt, n)
    want = oracle.compare_bitmap(truth, truth, n)
    assert (r["total_hit"], r["true_hit"], r["first_mismatch"]) == want
    assert r["truth_count"] == r["got_count"] == want[0]
    # planted mismatches: first row, a chunk boundary, the middle, the last row
    for row in sorted({0, min(n - 1, 8192 * 32), n // 2, n - 1}):
        got = truth.copy()
        got[row // 32] ^= np.uint32(1 << (31 - row % 32))
        if row + 5 < n:  # a later mismatch must not matter
            got[(row + 5) // 32] ^= np.uint32(1 << (31 - (row + 5) % 32))
        dg = R.DeviceBuffer.from_numpy(got)
        r = R.bitmap_compare(dt, dg, n)
        assert (r["total_hit"], r["true_hit"], r["first_mismatch"]) == oracle.compare_bitmap(truth, got, n), row
        assert r["first_mismatch"] == row


def test_compare_ignores_bits_past_n(eng):
    n = 1000
    a = np.zeros((n + 31) // 32, np.uint32)
    b = a.copy()
    b[-1] = np.uint32(0xFF)  # rows 1000.. of the last word
    r = R.bitmap_compare(R.DeviceBuffer.from_numpy(a), R.DeviceBuffer.from_numpy(b), n)
    assert r["first_mismatch"] == -1 and r["got_count"] == 0


@pytest.mark.parametrize("n", [1, 32, 77, 65536, 1_000_003, 100_000_000])
def test_digest_matches_oracle(eng, n):
    rng = np.random.default_rng(n % 977)
    bm = words(rng, n, 0.5)
    d = R.DeviceBuffer.from_numpy(bm)
    assert R.bitmap_digest(d, n) == oracle.bitmap_digest(bm, n)


def test_scan_digest_equals_oracle_scan_digest(eng):
    """At-scale parity through the digest: a GPU scan of 1e8 rows and the oracle's raw_scan
    of the same rows give the same digest and count."""
    n = 100_000_000
    x = R.generate_column(R.DIST_UNIFORM, 5, 0, n)
    dx = R.DeviceBuffer.from_numpy(x)
    out = R.DeviceBuffer(4 * ((n + 31) // 32))
    lo, hi = 1 << 30, 3 << 30
    R.scan_range(dx, n, lo, hi, out)
    want = oracle.raw_scan(x, R.BT, lo - 1, hi + 1)
    assert R.bitmap_digest(out, n) == oracle.bitmap_digest(want, n)


def test_reference_named_compare_bitmap_prints(eng, capfd):
    n = 100
    rng = np.random.default_rng(3)
    truth = words(rng, n)
    got = truth.copy()
    got[1] ^= np.uint32(1 << 31)  # row 32
    cols = [np.arange(n, dtype=np.uint32) * (c + 1) for c in range(3)]
    ptrs = (ctypes.c_void_p * 3)(*[c.ctypes.data for c in cols])
    R.lib().compare_bitmap(truth.ctypes.data, got.ctypes.data, n, ptrs, 3)
    out = capfd.readouterr().out
    total, good, first = oracle.compare_bitmap(truth, got, n)
    assert "[ERROR] check error in raw_data[32]= 32 / 64 / 96 \n" in out
    assert "[CHECK]hit %d/%d" % (good, total) in out


def test_pinned_async_egress(eng):
    n_words = 1 << 20
    src = np.arange(n_words, dtype=np.uint32)
    d = R.DeviceBuffer.from_numpy(src)
    hp = ctypes.c_void_p()
    R.check(R.lib().rts_host_alloc(ctypes.byref(hp), 4 * n_words), "rts_host_alloc")
    s = R.Stream()
    R.check(R.lib().rts_memcpy_dtoh_async(hp, d.ptr, 4 * n_words, s.handle), "dtoh_async")
    s.synchronize()
    got = np.ctypeslib.as_array((ctypes.c_uint32 * n_words).from_address(hp.value))
    assert np.array_equal(got, src)
    R.check(R.lib().rts_host_free(hp), "rts_host_free")
