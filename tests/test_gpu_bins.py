"""GPU parity of the engine's bin count (rts_bindex_build_bins: 256, 512, 1024 bins as 2, 4,
8 equal-count bins per reference area), of batch plans over several index sets, the
planner's direct / BinDex cost model and the launch-shape bookkeeping (group counts after a
config change and after an autotuner cache hit)."""
import ctypes

import numpy as np
import pytest

import oracle as O
import snapshot_format as S
from _util import MAXU, interval_mask, pack_rows
from test_gpu_batch import check_results, make_table, random_queries
from test_gpu_bindex import engine_bin_starts

pytestmark = pytest.mark.gpu


def datasets(eng, n):
    rng = np.random.default_rng(n + 7)
    heavy = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    heavy[rng.choice(n, 3 * (n // 128) + 11, replace=False)] = 0x9000_0000
    return {
        "uniform": eng.generate_column(eng.DIST_UNIFORM, 5, 0, n),
        "normal": eng.generate_column(eng.DIST_NORMAL, 5, 1, n, 2.0 ** 31, 2.0 ** 28),
        "zipf15": eng.generate_column(eng.DIST_ZIPF, 5, 3, n, 1.5, 2.0 ** 20),
        "small_range": rng.integers(0, 300, n, dtype=np.uint32),
        "split_heavy": heavy,
    }


def query_ranges(x, rng, count=24):
    xs = np.sort(x)
    n = xs.size
    out = [(0, MAXU), (0, 0), (MAXU, MAXU), (int(xs[0]), int(xs[-1])), (int(xs[n // 2]), int(xs[n // 2]))]
    for _ in range(count):
        sel = float(np.exp(rng.uniform(np.log(1e-5), np.log(0.9))))
        q0 = float(rng.random()) * (1 - sel)
        out.append((int(xs[int(q0 * (n - 1))]), int(xs[min(n - 1, int((q0 + sel) * (n - 1)))])))
    return out


@pytest.mark.parametrize("bins", [256, 512, 1024])
def test_more_bins_answer_exactly_and_keep_the_reference_fields(eng, bins):
    n = 1_000_003
    W = (n + 31) // 32
    rng = np.random.default_rng(bins)
    out = eng.DeviceBuffer(4 * W)
    cnt = eng.DeviceBuffer(8)
    for name, x in datasets(eng, n).items():
        d = eng.DeviceBuffer.from_numpy(x)
        ix = eng.BinDexIndex(d, n, bins=bins)
        ref = eng.BinDexIndex(d, n)
        assert ix.bins == bins and ref.bins == 128
        # the reference-facing descriptor is the 128-area one, whatever the bin count
        assert np.array_equal(ix.start_values(), ref.start_values()), name
        assert np.array_equal(ix.area_counts(), ref.area_counts()), name
        for k in (0, 1, 63, 126):
            assert np.array_equal(ix.filter_vector(k), ref.filter_vector(k)), (name, k)
        # fewer refine entries per bin: the engine bins nest inside the areas
        assert ix.refine_entries() <= ref.refine_entries()
        for lo, hi in query_ranges(x, rng):
            out.fill(0x3C)
            eng.bindex_query([ix], [lo], [hi], out, cnt)
            m = interval_mask(x, lo, hi)
            assert np.array_equal(out.to_numpy(np.uint32, W), pack_rows(m)), (name, bins, lo, hi)
            assert int(cnt.to_numpy(np.uint64, 1)[0]) == int(m.sum())
        ix.free()
        ref.free()
        d.free()


def test_more_bins_multi_column_and_mixed_bin_counts(eng):
    """A conjunction over indexes of different bin counts (each column its own) is exact."""
    n = 600_001
    W = (n + 31) // 32
    rng = np.random.default_rng(4)
    xs = list(datasets(eng, n).values())[:4]
    ds = [eng.DeviceBuffer.from_numpy(x) for x in xs]
    ixs = [eng.BinDexIndex(d, n, bins=b) for d, b in zip(ds, (128, 256, 512, 1024))]
    out = eng.DeviceBuffer(4 * W)
    cnt = eng.DeviceBuffer(8)
    for _ in range(12):
        k = int(rng.integers(2, 5))
        cs = sorted(int(c) for c in rng.choice(4, k, replace=False))
        rs = [query_ranges(xs[c], rng, 1)[-1] for c in cs]
        eng.bindex_query([ixs[c] for c in cs], [r[0] for r in rs], [r[1] for r in rs], out, cnt)
        d, c = O.conjunction_digest([xs[c] for c in cs], rs, threads=8)
        assert eng.bitmap_digest(out, n) == d and int(cnt.to_numpy(np.uint64, 1)[0]) == c, (cs, rs)


def test_snapshot_of_a_512_bin_index_round_trips_and_refuses_damage(eng, tmp_path):
    n = 300_007
    W = (n + 31) // 32
    x = eng.generate_column(eng.DIST_ZIPF, 8, 2, n, 1.1, 2.0 ** 20)
    d = eng.DeviceBuffer.from_numpy(x)
    ix = eng.BinDexIndex(d, n, bins=512)
    path = tmp_path / "b512.rtsbdx"
    ix.save(str(path))
    back = eng.BinDexIndex(d, n, snapshot=str(path))
    assert back.bins == 512
    o1, o2 = eng.DeviceBuffer(4 * W), eng.DeviceBuffer(4 * W)
    for lo, hi in query_ranges(x, np.random.default_rng(1), 10):
        eng.bindex_query([ix], [lo], [hi], o1)
        eng.bindex_query([back], [lo], [hi], o2)
        assert np.array_equal(o1.to_numpy(np.uint32, W), o2.to_numpy(np.uint32, W))
        assert np.array_equal(o2.to_numpy(np.uint32, W), pack_rows(interval_mask(x, lo, hi)))
    raw = bytearray(path.read_bytes())
    f, arrays, off = S.parse(bytes(raw))
    assert f["bins"] == 512 and f["nb"] > 0
    # payload layout: sorted_val, sorted_pos, filter vectors, then the (row, value) entries
    fv_off = off + 8 * n
    bent_off = fv_off + 4 * (f["bins"] - 1) * f["fv_stride"]
    assert bent_off + 8 * f["nb"] == len(raw)
    boff = arrays["boff"].astype(np.int64)
    k = int(np.argmax(boff[1:] - boff[:-1]))
    assert boff[k + 1] - boff[k] >= 10
    damages = {
        "entry row past n": (bent_off + 8 * (f["nb"] // 2), (n + 5).to_bytes(4, "little")),
        # the 6th entry of the longest list gets row 0, below the 5th entry's row
        "entries out of row order": (bent_off + 8 * int(boff[k] + 5), (0).to_bytes(4, "little")),
        "filter-vector bit": (fv_off + 4 * 1000, bytes([raw[fv_off + 4000] ^ 0x10])),
        "sorted position": (off + 4 * n + 40, bytes([raw[off + 4 * n + 40] ^ 1])),
    }
    for name, (pos, patch) in damages.items():
        bad = bytearray(raw)
        bad[pos:pos + len(patch)] = patch
        p = tmp_path / "bad.rtsbdx"
        p.write_bytes(bytes(bad))
        with pytest.raises(eng.HipError):
            eng.BinDexIndex(d, n, snapshot=str(p))
    # a snapshot of another column length is refused by the front end
    with pytest.raises(ValueError):
        eng.BinDexIndex(d, n - 1, snapshot=str(path))


def test_batch_over_index_sets_and_the_cost_model(eng):
    n = 1 << 20
    cols, dcols, idx, sc = make_table(eng, n, seed=17)
    sets = {b: [eng.BinDexIndex(d, n, bins=b) for d in dcols] for b in (256, 512)}
    rng = np.random.default_rng(6)
    nq = 120
    queries = random_queries(rng, sc, nq)
    plan = eng.BatchPlan(dcols, idx, n, nq)
    assert plan.add_indexes(sets[256]) == 256
    assert plan.add_indexes(sets[512]) == 512
    plan.set_queries(queries)
    stride = n // 32
    out = eng.DeviceBuffer(4 * stride * nq)
    counts = eng.DeviceBuffer(8 * nq)
    seen_direct = 0
    for bins in (0, 256, 512, 128):
        for weight in (0, 40, 1000):
            with pytest.raises(eng.HipError):
                plan.set_config(eng.batch_config(bins=1024))  # no such index set
            plan.set_config(eng.batch_config(bins=bins, direct_weight=weight))
            st = plan.stats()
            if weight == 0:
                assert st["auto_direct_cols"] == 0
            seen_direct += st["auto_direct_cols"]
            out.fill(0x77)
            plan.launch(out, stride, counts)
            check_results(eng, cols, queries, out, counts, stride)
    assert seen_direct > 0
    # at weight 1000 a direct column costs its full 4 bytes per row: never cheaper
    plan.set_config(eng.batch_config(direct_weight=1000))
    assert plan.stats()["auto_direct_cols"] == 0


def test_group_size_change_regroups_and_cache_hits_apply_the_cached_shape(eng, tmp_path):
    n = 1 << 18
    cols, dcols, idx, sc = make_table(eng, n, seed=19)
    nq = 96
    queries = random_queries(np.random.default_rng(8), sc, nq)
    plan = eng.BatchPlan(dcols, idx, n, nq, eng.batch_config(512, 64, 4, 1, 512))
    plan.set_queries(queries)
    stride = (n // 32 + 3) // 4 * 4
    out = eng.DeviceBuffer(4 * stride * nq)
    counts = eng.DeviceBuffer(8 * nq)
    g64 = plan.stats()["groups"]
    plan.set_config(eng.batch_config(512, 2, 4, 1, 512))
    g2 = plan.stats()["groups"]
    assert g2 >= nq // 2 and g2 > g64  # two queries per group at most
    plan.launch(out, stride, counts)
    check_results(eng, cols, queries, out, counts, stride)
    cache = str(tmp_path / "tune.txt")
    cands = [eng.batch_config(512, g, 4, 1, 512) for g in (1, 3)]
    chosen, _ = plan.autotune(out, stride, cands, reps=1, cache_path=cache)
    expect = plan.stats()["groups"]
    plan.set_config(eng.batch_config(512, 64, 4, 1, 512))
    again, _ = plan.autotune(out, stride, cands, reps=1, cache_path=cache)  # from the cache
    assert again.astuple() == chosen.astuple()
    assert plan.config().astuple() == chosen.astuple()
    assert plan.stats()["groups"] == expect
    plan.launch(out, stride, counts)
    check_results(eng, cols, queries, out, counts, stride)


class _Offset:
    """A device pointer `off` bytes into a DeviceBuffer."""

    def __init__(self, buf, off):
        self.ptr = buf.ptr + off


@pytest.mark.parametrize("bins", [128, 1024])
def test_build_over_a_column_that_is_not_16_byte_aligned(eng, bins):
    """The filter-vector build loads a lane's 32 rows with 16-byte loads only when the
    column is 16-byte aligned; a column 4 bytes into its buffer takes the row-by-row loads.
    Both builds give the same index: filter vectors, refine entries and query results
    (2048-row tiles with a ragged last one: n is not a multiple of 2048)."""
    n = 300_017
    rng = np.random.default_rng(bins + 3)
    x = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    x[rng.choice(n, 40, replace=False)] = 0x1234_5678  # one duplicated value
    a = eng.DeviceBuffer.from_numpy(x)
    padded = eng.DeviceBuffer.from_numpy(np.concatenate([np.zeros(1, np.uint32), x]))
    ia = eng.BinDexIndex(a, n, bins=bins)
    iu = eng.BinDexIndex(_Offset(padded, 4), n, bins=bins)
    assert ia.refine_entries() == iu.refine_entries()
    for k in (0, 1, 63, 126):
        assert np.array_equal(ia.filter_vector(k), iu.filter_vector(k)), k
    W = (n + 31) // 32
    out = eng.DeviceBuffer(4 * W)
    cnt = eng.DeviceBuffer(8)
    xs = np.sort(x)
    for lo, hi in [(0, MAXU), (int(xs[100]), int(xs[200_000])), (0x1234_5678, 0x1234_5678),
                   (int(xs[n // 2]), int(xs[n // 2 + 1500]))]:
        m = interval_mask(x, lo, hi)
        for ix in (ia, iu):
            eng.bindex_query([ix], [lo], [hi], out, cnt)
            assert np.array_equal(out.to_numpy(np.uint32, W), pack_rows(m)), (lo, hi)
            assert int(cnt.to_numpy(np.uint64, 1)[0]) == int(m.sum()), (lo, hi)


@pytest.mark.parametrize("threads,order", [(256, 1), (256, 0), (512, 1)])
def test_groups_are_packed_to_the_fewest_work_items(eng, threads, order):
    """Queries of 1..4 columns, shuffled, none empty: first fit by decreasing column count
    packs them into ceil(columns / slots) groups when the 1- and 3-column queries pair up
    (4 | 3 + 1 | 2 + 2 ... for slots = 4 at 2048-word tiles and 256 threads, 8 at 512), and
    every query still gets its own exact bitmap and count."""
    n = (1 << 18) + 77
    cols, dcols, idx, sc = make_table(eng, n, seed=29)
    rng = np.random.default_rng(threads + order)
    shape = [4] * 9 + [3] * 14 + [2] * 11 + [1] * 14
    rng.shuffle(shape)
    queries = []
    for k in shape:
        q = random_queries(rng, sc, 1, with_edges=False)[0]
        while len(q) != k:
            q = random_queries(rng, sc, 1, with_edges=False)[0]
        queries.append(q)
    nq = len(queries)
    slots = 32 * threads // 2048
    plan = eng.BatchPlan(dcols, idx, n, nq, eng.batch_config(2048, 64, 4, order, threads))
    plan.set_queries(queries)
    assert plan.stats()["groups"] == -(-sum(shape) // slots)
    stride = ((n + 31) // 32 + 3) // 4 * 4
    out = eng.DeviceBuffer(4 * stride * nq)
    counts = eng.DeviceBuffer(8 * nq)
    out.fill(0x5A)
    plan.launch(out, stride, counts)
    check_results(eng, cols, queries, out, counts, stride)


@pytest.mark.parametrize("name", ["uniform", "zipf15", "small_range", "split_heavy"])
def test_list_form_ranges_at_bin_edges_are_exact(eng, tmp_path, name):
    """A range over one or two bins whose refine lists hold every row of the bin (a bin
    start of at most 64 rows is listed with its bin) is read from those lists alone, with
    no filter vector.  Ranges at and around the engine bin starts of a 256-bin index: one
    bin, two bins, bounds one inside a bin, three bins (filter vectors again), and the
    same through a snapshot of the index and through the one-query kernel (alone and in a
    two-column query).  A heavy bin start (small_range: every value is
    heavy) leaves its bin incomplete, so only a range from inside one bin to that bin's
    end keeps the list form there.  Every bitmap and count equals numpy's."""
    n = 1_000_003
    x = datasets(eng, n)[name]
    d = eng.DeviceBuffer.from_numpy(x)
    ix = eng.BinDexIndex(d, n, bins=256)
    sv = np.sort(x)
    st = engine_bin_starts(ix, sv, 256) + [MAXU + 1]
    qs = []
    for b in (0, 1, 2, 17, 63, 100, 127, 128, 200, 253, 254, 255):
        lo, nxt = st[b], st[min(b + 1, 256)]
        for hi_bin in (b + 1, b + 2, b + 3):
            hi = st[min(hi_bin, 256)] - 1
            if lo > MAXU or hi < lo:
                continue
            qs.append([(0, lo, min(hi, MAXU), 0)])
            if lo + 1 <= min(hi, MAXU):
                qs.append([(0, lo + 1, min(hi, MAXU), 0)])
            if nxt <= MAXU and lo < nxt - 1:
                qs.append([(0, lo + 1, nxt, 0)])
    plan = eng.BatchPlan([d], [ix], n, len(qs), eng.batch_config(2048, 64, 4, 1, 256))
    plan.set_queries(qs)
    stride = ((n + 31) // 32 + 3) // 4 * 4
    out = eng.DeviceBuffer(4 * stride * len(qs))
    counts = eng.DeviceBuffer(8 * len(qs))
    out.fill(0x3C)
    plan.launch(out, stride, counts)
    check_results(eng, [x], qs, out, counts, stride)
    listed = plan.stats()["list_cols"]
    if name == "uniform":
        assert listed >= len(qs) // 3
    # the one-query kernel takes the same list form: each range alone, counted, and as the
    # second column of a two-column query with a full-range first column
    W = (n + 31) // 32
    one = eng.DeviceBuffer(4 * W)
    c1 = eng.DeviceBuffer(8)
    for q in qs[::3]:
        (_c, lo, hi, _m) = q[0]
        m = interval_mask(x, lo, hi)
        eng.bindex_query([ix], [lo], [hi], one, c1)
        assert np.array_equal(one.to_numpy(np.uint32, W), pack_rows(m)), (name, lo, hi)
        assert int(c1.to_numpy(np.uint64, 1)[0]) == int(m.sum()), (name, lo, hi)
        eng.bindex_query([ix, ix], [0, lo], [MAXU, hi], one, c1)
        assert np.array_equal(one.to_numpy(np.uint32, W), pack_rows(m)), (name, lo, hi)
        assert int(c1.to_numpy(np.uint64, 1)[0]) == int(m.sum()), (name, lo, hi)
    path = str(tmp_path / "ix.rtsbdx")
    ix.save(path)
    loaded = eng.BinDexIndex(d, n, snapshot=path)
    plan2 = eng.BatchPlan([d], [loaded], n, len(qs), eng.batch_config(2048, 64, 4, 1, 256))
    plan2.set_queries(qs)
    assert plan2.stats()["list_cols"] == listed
    out.fill(0x3C)
    plan2.launch(out, stride, counts)
    check_results(eng, [x], qs, out, counts, stride)


def test_mix_digest_tracks_the_query_distribution(eng):
    n = 1 << 18
    cols, dcols, idx, sc = make_table(eng, n, seed=23)
    plan = eng.BatchPlan(dcols, idx, n, 2048)
    plan.set_queries(random_queries(np.random.default_rng(1), sc, 2048, with_edges=False))
    a = plan.mix_digest()
    plan.set_queries(random_queries(np.random.default_rng(2), sc, 2048, with_edges=False))
    b = plan.mix_digest()
    # one-column wide queries only: another mix
    wide = [[(0, int(sc[0][int(0.1 * n)]), int(sc[0][int(0.9 * n)]), 0)] for _ in range(2048)]
    plan.set_queries(wide)
    c = plan.mix_digest()
    assert a == b and a != c
