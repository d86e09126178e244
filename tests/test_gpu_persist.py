"""GPU tests of BinDex persistence (rts_bindex_save / rts_bindex_load): a loaded snapshot
is the same index (start values, counts, blocks, filter vectors, sorted arrays) and answers
queries and the batched launch identically; a damaged snapshot is refused."""
import ctypes

import numpy as np
import pytest

import rtscan_amd as R
from _util import interval_mask, pack_rows

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dist,p0,p1", [(R.DIST_UNIFORM, 0.0, 0.0), (R.DIST_ZIPF, 1.5, 2.0 ** 12)])
def test_snapshot_round_trip(eng, tmp_path, dist, p0, p1):
    n = 1_000_003
    x = R.generate_column(dist, 17, 0, n, p0, p1)
    d = R.DeviceBuffer.from_numpy(x)
    built = R.BinDexIndex(d, n)
    path = str(tmp_path / "col.rtsbdx")
    built.save(path)
    loaded = R.BinDexIndex(d, n, snapshot=path)
    assert np.array_equal(built.start_values(), loaded.start_values())
    assert np.array_equal(built.area_counts(), loaded.area_counts())
    assert built.refine_entries() == loaded.refine_entries()
    for k in (0, 63, 126):
        assert np.array_equal(built.filter_vector(k), loaded.filter_vector(k))
    bv, bp = built.sorted()
    lv, lp = loaded.sorted()
    assert np.array_equal(bv, lv) and np.array_equal(bp, lp)
    L = R.lib()
    for v in [int(v) for v in x[:50]]:
        a = L.in_which_area(built.ptr, v)
        assert a == L.in_which_area(loaded.ptr, v)
        if a >= 0:  # in_which_block takes the area (bindex.cpp:1153)
            pa = ctypes.cast(built.s.areas[a], ctypes.c_void_p).value
            pb = ctypes.cast(loaded.s.areas[a], ctypes.c_void_p).value
            assert L.in_which_block(pa, v) == L.in_which_block(pb, v)
    rng = np.random.default_rng(4)
    W = (n + 31) // 32
    out = R.DeviceBuffer(4 * W)
    cnt = R.DeviceBuffer(8)
    for _ in range(8):
        lo, hi = sorted(int(v) for v in rng.choice(x, 2))
        R.bindex_query([loaded], [lo], [hi], out, cnt)
        m = interval_mask(x, lo, hi)
        assert np.array_equal(out.to_numpy(np.uint32, W), pack_rows(m))
        assert int(cnt.to_numpy(np.uint64, 1)[0]) == int(m.sum())
    # the batched launch over a loaded index
    plan = R.BatchPlan([d], [loaded], n, 4)
    qs = [[(0, int(a), int(b), 0)] for a, b in (sorted(rng.choice(x, 2)) for _ in range(4))]
    plan.set_queries(qs)
    stride = (W + 3) // 4 * 4
    bo = R.DeviceBuffer(4 * stride * 4)
    bc = R.DeviceBuffer(8 * 4)
    plan.launch(bo, stride, bc)
    got = bo.to_numpy(np.uint32, stride * 4).reshape(4, stride)
    for i, q in enumerate(qs):
        assert np.array_equal(got[i, :W], pack_rows(interval_mask(x, q[0][1], q[0][2])))


def test_damaged_snapshot_is_refused(eng, tmp_path):
    n = 300_000
    x = R.generate_column(R.DIST_UNIFORM, 3, 0, n)
    d = R.DeviceBuffer.from_numpy(x)
    ix = R.BinDexIndex(d, n)
    path = tmp_path / "a.rtsbdx"
    ix.save(str(path))
    raw = bytearray(path.read_bytes())
    bad = tmp_path / "bad.rtsbdx"
    bad.write_bytes(bytes(raw[:4]) + b"X" + bytes(raw[5:]))  # magic
    with pytest.raises(R.HipError):
        R.BinDexIndex(d, n, snapshot=str(bad))
    short = tmp_path / "short.rtsbdx"
    short.write_bytes(bytes(raw[: len(raw) // 2]))  # truncated
    with pytest.raises(R.HipError):
        R.BinDexIndex(d, n, snapshot=str(short))


def test_snapshot_refuses_another_column_of_the_same_length(eng, tmp_path):
    n = 200_003
    x = R.generate_column(R.DIST_UNIFORM, 8, 0, n)
    y = x.copy()
    y[n // 2] ^= 1  # one row differs
    dx, dy = R.DeviceBuffer.from_numpy(x), R.DeviceBuffer.from_numpy(y)
    ix = R.BinDexIndex(dx, n)
    path = str(tmp_path / "c.rtsbdx")
    ix.save(path)
    R.BinDexIndex(dx, n, snapshot=path).free()  # its own column: accepted
    with pytest.raises(R.HipError):
        R.BinDexIndex(dy, n, snapshot=path)
