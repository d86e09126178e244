"""GPU parity of the batched persistent launch (include/rts_batch.h): every query's bitmap
and popcount equal the numpy closed form, for BinDex, direct and mixed evaluation, both
tile sizes, several group sizes, graph replays and re-planned queries."""
import numpy as np
import pytest

from _util import MAXU, interval_mask, pack_rows

pytestmark = pytest.mark.gpu

KINDS = [(0, 0.0, 0.0), (1, 2.0 ** 31, 2.0 ** 28), (2, 1.1, 2.0 ** 20), (2, 1.5, 2.0 ** 20)]


def make_table(eng, n, seed=31):
    cols = [eng.generate_column(k, seed, c, n, p0, p1) for c, (k, p0, p1) in enumerate(KINDS)]
    dcols = [eng.DeviceBuffer.from_numpy(c) for c in cols]
    idx = [eng.BinDexIndex(d, n) for d in dcols]
    return cols, dcols, idx, [np.sort(c) for c in cols]


def random_queries(rng, sorted_cols, nq, direct_share=0.0, with_edges=True):
    """1..4 columns per query; per-column selectivity log-uniform in [1e-5, 0.9]; endpoints
    from the column's quantiles; some edge ranges."""
    n = sorted_cols[0].size
    out = []
    for i in range(nq):
        k = int(rng.integers(1, 5))
        cs = rng.choice(4, k, replace=False)
        q = []
        for c in cs:
            sel = float(np.exp(rng.uniform(np.log(1e-5), np.log(0.9))))
            q0 = float(rng.random()) * (1 - sel)
            lo = int(sorted_cols[c][int(q0 * (n - 1))])
            hi = int(sorted_cols[c][min(n - 1, int((q0 + sel) * (n - 1)))])
            if with_edges and i % 37 == 5:
                lo, hi = 0, MAXU
            if with_edges and i % 41 == 7:
                lo, hi = hi, lo  # empty when lo > hi
            mode = 2 if rng.random() < direct_share else 0
            q.append((int(c), lo, hi, mode))
        out.append(q)
    return out


def check_results(eng, cols, queries, out, counts, stride):
    n = cols[0].size
    W = (n + 31) // 32
    got = out.to_numpy(np.uint32, stride * len(queries)).reshape(len(queries), stride)
    cnt = counts.to_numpy(np.uint64, len(queries))
    for i, q in enumerate(queries):
        m = np.ones(n, bool)
        for (c, lo, hi, _mode) in q:
            m &= interval_mask(cols[c], lo, hi)
        assert np.array_equal(got[i, :W], pack_rows(m)), (i, q)
        assert int(cnt[i]) == int(m.sum()), (i, q)


@pytest.mark.parametrize("tile_words,qgroup,direct_share,order,threads", [
    (1024, 16, 0.0, 1, 512), (1024, 16, 0.0, 0, 512), (1024, 4, 0.3, 1, 512), (2048, 8, 0.0, 1, 512),
    (2048, 64, 1.0, 0, 512), (1024, 1, 0.5, 1, 512), (512, 32, 0.0, 1, 512), (512, 64, 0.4, 0, 512),
    (512, 3, 1.0, 1, 512), (2048, 16, 0.1, 1, 256), (1024, 64, 0.0, 0, 256), (512, 8, 1.0, 1, 256)])
def test_batch_matches_closed_form(eng, tile_words, qgroup, direct_share, order, threads):
    n = (1 << 20) + 12345  # partial last tile, partial last word
    cols, dcols, idx, sc = make_table(eng, n)
    rng = np.random.default_rng(tile_words + qgroup)
    queries = random_queries(rng, sc, 150, direct_share)
    cfg = eng.batch_config(tile_words, qgroup, 4, order, threads)
    plan = eng.BatchPlan(dcols, idx, n, len(queries), cfg)
    plan.set_queries(queries)
    st = plan.stats()
    assert st["bindex_cols"] + st["direct_cols"] == sum(len(q) for q in queries)
    if direct_share < 1.0:  # selective ranges put both bounds in one bin: one-pass lists
        assert st["narrow_cols"] > 0
    stride = ((n + 31) // 32 + 3) // 4 * 4
    out = eng.DeviceBuffer(4 * stride * len(queries)).fill(0xEE)
    counts = eng.DeviceBuffer(8 * len(queries))
    plan.launch(out, stride, counts)
    check_results(eng, cols, queries, out, counts, stride)
    plan.free()


@pytest.mark.parametrize("nq", [1, 2, 3, 5])
def test_few_groups_split_items_exactly(eng, nq):
    """One query per group, so the launch has 1, 2, 3 or 5 groups: a work item's index
    splits into (tile, group) through the group count's reciprocal, or none for one group;
    3-tile and 49-tile columns put the items of an XCD on either side of a group boundary."""
    for n in (3 * 65536 - 7, 49 * 65536 + 100):
        cols, dcols, idx, sc = make_table(eng, n, seed=nq)
        queries = random_queries(np.random.default_rng(nq), sc, nq, with_edges=False)
        plan = eng.BatchPlan(dcols, idx, n, nq, eng.batch_config(2048, 1, 4, 1, 256))
        plan.set_queries(queries)
        assert plan.stats()["groups"] == nq
        stride = ((n + 31) // 32 + 3) // 4 * 4
        out = eng.DeviceBuffer(4 * stride * nq).fill(0xEE)
        counts = eng.DeviceBuffer(8 * nq)
        plan.launch(out, stride, counts)
        check_results(eng, cols, queries, out, counts, stride)
        plan.free()


def test_batch_without_indexes_is_direct(eng):
    n = 300001
    cols, dcols, idx, sc = make_table(eng, n, seed=5)
    queries = random_queries(np.random.default_rng(1), sc, 40)
    plan = eng.BatchPlan(dcols, [None] * 4, n, 40)
    plan.set_queries(queries)
    assert plan.stats()["bindex_cols"] == 0
    stride = ((n + 31) // 32 + 3) // 4 * 4
    out = eng.DeviceBuffer(4 * stride * 40)
    counts = eng.DeviceBuffer(8 * 40)
    plan.launch(out, stride, counts)
    check_results(eng, cols, queries, out, counts, stride)


def test_batch_graph_replays_and_requeries(eng):
    """The captured graph replays correctly several times with eager work and
    synchronisation in between, and evaluates new queries set after capture."""
    n = 1 << 20
    cols, dcols, idx, sc = make_table(eng, n, seed=9)
    rng = np.random.default_rng(2)
    nq = 64
    plan = eng.BatchPlan(dcols, idx, n, nq)
    stride = n // 32
    out = eng.DeviceBuffer(4 * stride * nq)
    counts = eng.DeviceBuffer(8 * nq)
    stream = eng.Stream()
    scratch = eng.DeviceBuffer(4 * stride)
    for rep in range(3):
        queries = random_queries(rng, sc, nq, direct_share=0.1)
        plan.set_queries(queries, stream)
        out.fill(0x5A)
        plan.graph_launch(out, stride, counts, stream)
        stream.synchronize()
        check_results(eng, cols, queries, out, counts, stride)
        # eager work between replays
        eng.scan_range(dcols[0], n, 5, 1 << 31, scratch, None, stream)
        stream.synchronize()
        plan.graph_launch(out, stride, counts, stream)  # replay, same queries
        stream.synchronize()
        check_results(eng, cols, queries, out, counts, stride)


def test_batch_autotune_and_cache(eng, tmp_path):
    n = 1 << 20
    cols, dcols, idx, sc = make_table(eng, n, seed=11)
    queries = random_queries(np.random.default_rng(3), sc, 48)
    plan = eng.BatchPlan(dcols, idx, n, 48)
    plan.set_queries(queries)
    stride = n // 32
    out = eng.DeviceBuffer(4 * stride * 48)
    counts = eng.DeviceBuffer(8 * 48)
    cands = [eng.batch_config(t, g, b) for t in (512, 1024, 2048) for g in (4, 16) for b in (2, 4)]
    cache = str(tmp_path / "tune.txt")
    chosen, ms = plan.autotune(out, stride, cands, reps=2, cache_path=cache)
    assert chosen.astuple() in [c.astuple() for c in cands] and ms > 0
    again, ms2 = plan.autotune(out, stride, cands, reps=2, cache_path=cache)
    assert again.astuple() == chosen.astuple()  # served from the cache
    plan.launch(out, stride, counts)
    check_results(eng, cols, queries, out, counts, stride)


@pytest.mark.parametrize("n", [1, 31, 5000, 70001])
def test_batch_small_tables_and_back_to_back_launches(eng, n):
    """Tables of a few tiles leave most XCDs (and most workgroups) without work items; four
    launches in a row, alternating launch shapes, each start from the work counters the
    previous launch left (reset by its last workgroup) and must give the same results."""
    if n >= 128:
        cols, dcols, idx, sc = make_table(eng, n, seed=n % 97)
    else:  # a BinDex needs at least 128 rows: direct evaluation only
        cols = [eng.generate_column(k, 3, c, n, p0, p1) for c, (k, p0, p1) in enumerate(KINDS)]
        dcols = [eng.DeviceBuffer.from_numpy(c) for c in cols]
        idx, sc = [None] * 4, [np.sort(c) for c in cols]
    rng = np.random.default_rng(n)
    stride = ((n + 31) // 32 + 3) // 4 * 4
    for nq in (1, 70):
        queries = random_queries(rng, sc, nq, direct_share=0.2)
        plan = eng.BatchPlan(dcols, idx, n, nq)
        plan.set_queries(queries)
        out = eng.DeviceBuffer(4 * stride * nq)
        counts = eng.DeviceBuffer(8 * nq)
        for rep, shape in enumerate([(2048, 32, 4, 1, 256), (512, 8, 2, 0, 512),
                                     (2048, 32, 4, 1, 256), (1024, 64, 1, 1, 256)]):
            plan.set_config(eng.batch_config(*shape))
            out.fill(0xA5)
            counts.fill(0xFF)
            plan.launch(out, stride, counts)
            check_results(eng, cols, queries, out, counts, stride)
        plan.free()
