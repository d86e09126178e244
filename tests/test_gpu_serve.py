"""GPU tests of the request loop (include/rts_serve.h): results equal the closed form, each
result is the same alone and in a mix of concurrent requests, admission holds requests
back while the result pool is full, and the Poisson replay reports every request."""
import numpy as np
import pytest

from _util import interval_mask, pack_rows

pytestmark = pytest.mark.gpu


def table(eng, n, ncols, seed):
    kinds = [(0, 0.0, 0.0), (1, 2.0 ** 31, 2.0 ** 28), (2, 1.1, 2.0 ** 20), (2, 1.5, 2.0 ** 20)]
    cols = [eng.generate_column(*([kinds[c][0], seed, c, n] + list(kinds[c][1:]))) for c in range(ncols)]
    dcols = [eng.DeviceBuffer.from_numpy(c) for c in cols]
    idx = [eng.BinDexIndex(d, n) for d in dcols]
    return cols, dcols, idx


def random_query(rng, cols):
    k = int(rng.integers(1, len(cols) + 1))
    q = []
    for c in rng.choice(len(cols), k, replace=False):
        a, b = sorted(int(v) for v in rng.choice(cols[c], 2))
        q.append((int(c), a, b, 2 if rng.random() < 0.1 else 0))
    return q


def truth(cols, q):
    m = np.ones(cols[0].size, bool)
    for c, lo, hi, _ in q:
        m &= interval_mask(cols[c], lo, hi)
    return m


@pytest.fixture(scope="module")
def tables(eng):
    t1 = table(eng, (1 << 20) + 5, 2, 1)
    t2 = table(eng, 3_000_017, 4, 2)
    return t1, t2


def test_results_alone_and_in_the_mix(eng, tables):
    (c1, d1, i1), (c2, d2, i2) = tables
    srv = eng.Server(eng.server_config(max_batch=16, streams=3, pool_bytes=64 << 20))
    t1 = srv.add_table(d1, i1, c1[0].size)
    t2 = srv.add_table(d2, i2, c2[0].size)
    srv.start()
    rng = np.random.default_rng(5)
    probe = (t2, random_query(rng, c2))
    # alone
    tk = srv.submit(*probe)
    st, ptr, cnt, _ = srv.wait(tk)
    assert st == eng.REQ_DONE
    alone = srv.read_bitmap(ptr, c2[0].size)
    srv.release(tk)
    assert np.array_equal(alone, pack_rows(truth(c2, probe[1])))
    # in a burst of concurrent requests on both tables
    reqs = []
    for i in range(120):
        tb = t1 if i % 2 else t2
        q = random_query(rng, c1 if tb == t1 else c2)
        reqs.append((tb, q, srv.submit(tb, q)))
        if i == 60:
            reqs.append((t2, probe[1], srv.submit(*probe)))
    for tb, q, tk in reqs:
        st, ptr, cnt, lat = srv.wait(tk)
        assert st == eng.REQ_DONE and lat > 0
        cols = c1 if tb == t1 else c2
        got = srv.read_bitmap(ptr, cols[0].size)
        m = truth(cols, q)
        assert np.array_equal(got, pack_rows(m)), (tb, q)
        assert cnt == int(m.sum())
        if q is probe[1]:
            assert np.array_equal(got, alone)
        srv.release(tk)
    s = srv.stats()
    assert s["completed"] == s["submitted"] == 122
    assert s["max_batch_seen"] > 1  # requests were batched on arrival
    assert s["pool_in_use"] == 0 and 0 < s["pool_peak"] <= 64 << 20
    srv.destroy()


def test_admission_waits_for_pool_memory(eng, tables):
    (c1, d1, i1), (c2, d2, i2) = tables
    n = c2[0].size
    bitmap_bytes = ((n + 31) // 32 * 4 + 255) // 256 * 256
    srv = eng.Server(eng.server_config(max_batch=8, streams=2, pool_bytes=2 * bitmap_bytes))
    t2 = srv.add_table(d2, i2, n)
    srv.start()
    rng = np.random.default_rng(6)
    qs = [random_query(rng, c2) for _ in range(6)]
    tks = [srv.submit(t2, q) for q in qs]
    done = [srv.wait(t, timeout_ms=2000)[0] for t in tks[:2]]
    assert done == [eng.REQ_DONE] * 2
    # the pool holds two results: the third stays queued until one is released
    assert srv.wait(tks[2], timeout_ms=300)[0] == eng.REQ_QUEUED
    for i, tk in enumerate(tks):
        st, ptr, cnt, _ = srv.wait(tk, timeout_ms=5000)
        assert st == eng.REQ_DONE
        assert cnt == int(truth(c2, qs[i]).sum())
        srv.release(tk)
    assert srv.stats()["pool_peak"] <= 2 * bitmap_bytes
    srv.destroy()


def test_replay_reports_every_request(eng, tables):
    (c1, d1, i1), (c2, d2, i2) = tables
    srv = eng.Server(eng.server_config(max_batch=64, streams=3, pool_bytes=256 << 20))
    t1 = srv.add_table(d1, i1, c1[0].size)
    t2 = srv.add_table(d2, i2, c2[0].size)
    srv.start()
    rng = np.random.default_rng(7)
    nq = 400
    tids = [t1 if rng.random() < 0.5 else t2 for _ in range(nq)]
    qs = [random_query(rng, c1 if t == t1 else c2) for t in tids]
    arrivals = np.cumsum(rng.exponential(1e9 / 20000.0, nq)).astype(np.uint64)  # 20k q/s
    lat, cnt, elapsed = srv.replay(tids, qs, arrivals)
    assert (lat > 0).all() and elapsed > 0
    for i in range(0, nq, 7):
        cols = c1 if tids[i] == t1 else c2
        assert int(cnt[i]) == int(truth(cols, qs[i]).sum()), i
    s = srv.stats()
    assert s["completed"] == nq and s["pool_in_use"] == 0
    srv.destroy()


class _Offset:
    def __init__(self, buf, off):
        self.ptr = buf.ptr + off


def test_failed_batches_end_the_replay_and_later_replays_run(eng, tables):
    """A batch that fails (forced from the host: a table whose column is not 16-byte aligned
    and has no index, so planning its direct form is refused) fails its requests: the replay
    returns an error once every one of them is counted, none is left queued or holding pool
    memory, and a later replay on a good table is not cut short by that earlier error."""
    (c1, d1, i1), _ = tables
    n = c1[0].size - 1
    srv = eng.Server(eng.server_config(max_batch=8, streams=2, pool_bytes=64 << 20))
    bad = srv.add_table([_Offset(d1[0], 4)], [None], n)
    good = srv.add_table(d1, i1, c1[0].size)
    srv.start()
    rng = np.random.default_rng(12)
    qs = [[(0, 5, 1 << 31, 0)] for _ in range(40)]
    arrivals = np.cumsum(rng.exponential(2e5, len(qs))).astype(np.uint64)
    with pytest.raises(eng.HipError):
        srv.replay([bad] * len(qs), qs, arrivals)
    st = srv.stats()
    assert st["submitted"] == st["completed"] == len(qs)
    assert st["pool_in_use"] == 0
    qs2 = [random_query(rng, c1) for _ in range(60)]
    lat, cnt, _ = srv.replay([good] * len(qs2), qs2, np.cumsum(rng.exponential(1e5, len(qs2))).astype(np.uint64))
    for q, c in zip(qs2, cnt):
        assert int(c) == int(truth(c1, q).sum())
    st = srv.stats()
    assert st["submitted"] == st["completed"] == len(qs) + len(qs2) and st["pool_in_use"] == 0
    srv.destroy()
