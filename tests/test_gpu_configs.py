"""Parity at the configurations' own sizes, through the C ABI, against the oracle:

* config 3: conjunctions of 2, 3 and 4 columns over the 1e8-row x 4-column table (uniform,
  normal(2^31, 2^28), Zipf 1.1, Zipf 1.5), per-column selectivity in {0.01, 0.1, 0.5}, by the
  fused direct kernel and the fused BinDex kernel: full bitmaps and counts equal the oracle's
  conjunction (raw_scan_entry per column, AND merge, bindex.cpp:2265-2289);
* config 4: the 4096-query batch of the benchmark on that table in one launch, with the
  planner sending some query-columns to the direct form: every query's count and digest
  equal the one-query-per-launch direct path's, and 64 sampled queries equal the oracle's
  digest and count;
* config 5: the two served tables (2^26 x 2 uniform, 1e9 x 4 as config 3): a short Poisson
  replay through the request loop, then 16 sampled requests per table whose counts (from
  the replay) and bitmap digests and counts (from a burst of submitted requests) equal the
  oracle's."""
import types

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

N3 = 100_000_000
KINDS = [(0, 0.0, 0.0), (1, 2.0 ** 31, 2.0 ** 28), (2, 1.1, 2.0 ** 20), (2, 1.5, 2.0 ** 20)]
MAXU = 0xFFFFFFFF


def log(msg):
    print("[configs] " + msg, flush=True)


def ranges_for(samples, cols, sels, rng):
    out = []
    for c, sel in zip(cols, sels):
        s = samples[c]
        q0 = float(rng.random()) * (1.0 - sel)
        out.append((int(s[int(q0 * (s.size - 1))]), int(s[min(s.size - 1, int((q0 + sel) * (s.size - 1)))])))
    return out


def oracle_bitmap(columns, ranges):
    """The reference checker's conjunction, in the query-file grammar: bt lo-1 hi+1 (le / ge
    at the ends of the value range), one zeroed bitmap per column, AND-merged."""
    preds = []
    for lo, hi in ranges:
        if lo == 0 and hi == MAXU:
            preds.append((O.GE, [0], None))
        elif lo == 0:
            preds.append((O.LE, [hi], None))
        elif hi == MAXU:
            preds.append((O.GE, [lo], None))
        else:
            preds.append((O.BT, [lo - 1], [hi + 1]))
    return O.conjunction(columns, preds, threads=16)


@pytest.fixture(scope="module")
def table(eng):
    log("generating the 1e8 x 4 table")
    cols = [eng.generate_column(k, 3, c, N3, p0, p1) for c, (k, p0, p1) in enumerate(KINDS)]
    dcols = [eng.DeviceBuffer.from_numpy(c) for c in cols]
    idx = [eng.BinDexIndex(d, N3) for d in dcols]
    rng = np.random.default_rng(3)
    pick = rng.integers(0, N3, 1_000_000)
    samples = [np.sort(c[pick]) for c in cols]
    log("table ready")
    yield types.SimpleNamespace(cols=cols, dcols=dcols, idx=idx, samples=samples)
    for ix in idx:
        ix.free()
    for d in dcols:
        d.free()


@pytest.mark.timeout(600)
def test_config3_conjunctions_match_oracle(eng, table):
    W = (N3 + 31) // 32
    rng = np.random.default_rng(33)
    out = eng.DeviceBuffer(4 * W)
    cnt = eng.DeviceBuffer(8)
    done = 0
    for ncols in (2, 3, 4):
        for j, sel in enumerate((0.01, 0.1, 0.5)):
            cs = sorted(int(c) for c in rng.choice(4, ncols, replace=False))
            rs = ranges_for(table.samples, cs, [sel] * ncols, rng)
            log(f"conjunction cols={cs} sel={sel}")
            truth = oracle_bitmap([table.cols[c] for c in cs], rs)
            want = O.popcount_rows(truth, N3)
            for path in ("direct", "bindex"):
                out.fill(0xA5)
                if path == "direct":
                    eng.conjunction_scan([table.dcols[c] for c in cs], [r[0] for r in rs], [r[1] for r in rs],
                                         N3, out, cnt)
                else:
                    eng.bindex_query([table.idx[c] for c in cs], [r[0] for r in rs], [r[1] for r in rs], out, cnt)
                got = out.to_numpy(np.uint32, W)
                assert np.array_equal(got, truth), (path, cs, rs)
                assert int(cnt.to_numpy(np.uint64, 1)[0]) == want, (path, cs, rs)
            done += 1
    assert done >= 6
    out.free()


@pytest.mark.timeout(900)
def test_config4_batch_of_4096_matches_one_query_path_and_oracle(eng, table):
    import bench  # the benchmark's own query generator for this configuration

    nq = 4096
    shim = types.SimpleNamespace(samples=table.samples,
                                 range_for=lambda c, sel, rng: bench.Table.range_for(shim, c, sel, rng))
    queries = bench.batch_queries(shim, nq, 44, 0.0)
    W = (N3 + 31) // 32
    stride = (W + 3) // 4 * 4
    # the planner sends the widest BinDex query-columns to the direct form at this weight
    cfg = eng.batch_config(2048, 32, 4, 1, 256, 0, 60)
    plan = eng.BatchPlan(table.dcols, table.idx, N3, nq, cfg)
    plan.set_queries(queries)
    st = plan.stats()
    log(f"planned {st}")
    assert st["auto_direct_cols"] > 0 and st["bindex_cols"] > 0  # a mixed batch
    out = eng.DeviceBuffer(4 * stride * nq)
    counts = eng.DeviceBuffer(8 * nq)
    stream = eng.Stream()
    plan.graph_launch(out, stride, counts, stream)
    stream.synchronize()
    log("batch launched")
    cnt = counts.to_numpy(np.uint64, nq)
    one = eng.DeviceBuffer(4 * W)
    c1 = eng.DeviceBuffer(8)
    row = types.SimpleNamespace(ptr=0)
    bad = []
    digests = []
    for i, q in enumerate(queries):
        row.ptr = out.ptr + 4 * stride * i
        d_batch = eng.bitmap_digest(row, N3)
        eng.conjunction_scan([table.dcols[c] for c, _, _, _ in q], [lo for _, lo, _, _ in q],
                             [hi for _, _, hi, _ in q], N3, one, c1)
        d_one = eng.bitmap_digest(one, N3)
        digests.append(d_batch)
        if d_batch != d_one or int(cnt[i]) != int(c1.to_numpy(np.uint64, 1)[0]):
            bad.append(i)
    assert not bad, bad[:8]
    log("all 4096 equal the one-query path")
    rng = np.random.default_rng(5)
    for i in sorted(rng.choice(nq, 64, replace=False)):
        q = queries[i]
        d, c = O.conjunction_digest([table.cols[cc] for cc, _, _, _ in q], [(lo, hi) for _, lo, hi, _ in q])
        assert (digests[i], int(cnt[i])) == (d, c), (i, q)
    log("64 sampled queries equal the oracle")
    plan.free()
    out.free()
    one.free()


@pytest.mark.timeout(900)
def test_config5_served_tables_match_oracle(eng):
    n1, n2 = 1 << 26, 1_000_000_000
    rng = np.random.default_rng(55)
    log("generating the served tables")
    t1 = [eng.generate_column(0, 55, c, n1) for c in range(2)]
    t1_d = [eng.DeviceBuffer.from_numpy(c) for c in t1]
    t1_ix = [eng.BinDexIndex(d, n1) for d in t1_d]
    t2, t2_d, t2_ix = [], [], []
    for c, (k, p0, p1) in enumerate(KINDS):
        t2.append(eng.generate_column(k, 56, c, n2, p0, p1))
        t2_d.append(eng.DeviceBuffer.from_numpy(t2[-1]))
        t2_ix.append(eng.BinDexIndex(t2_d[-1], n2))
        log(f"T2 column {c} indexed")
    s1 = [np.sort(c[rng.integers(0, n1, 1_000_000)]) for c in t1]
    s2 = [np.sort(c[rng.integers(0, n2, 1_000_000)]) for c in t2]

    def query(samples, ncols_max):
        k = int(rng.integers(1, ncols_max + 1))
        cs = [int(c) for c in rng.choice(ncols_max, k, replace=False)]
        sels = [float(np.exp(rng.uniform(np.log(1e-5), np.log(0.9)))) for _ in cs]
        return [(c, lo, hi, 0) for c, (lo, hi) in zip(cs, ranges_for(samples, cs, sels, rng))]

    srv = eng.Server(eng.server_config(max_batch=64, streams=3, pool_bytes=8 << 30))
    tid1 = srv.add_table(t1_d, t1_ix, n1)
    tid2 = srv.add_table(t2_d, t2_ix, n2)
    srv.start()
    nq = 400  # Poisson arrivals at 2000 q/s, half the requests on each table
    tids = [tid1 if rng.random() < 0.5 else tid2 for _ in range(nq)]
    qs = [query(s1, 2) if t == tid1 else query(s2, 4) for t in tids]
    arrivals = np.cumsum(rng.exponential(1e9 / 2000, nq)).astype(np.uint64)
    lat, cnt, elapsed = srv.replay(tids, qs, arrivals)
    log(f"replayed {nq} requests in {elapsed:.2f} s")
    checked = 0
    for tid, cols, n in ((tid1, t1, n1), (tid2, t2, n2)):
        sample = [i for i in range(nq) if tids[i] == tid][:16]
        assert len(sample) == 16
        # the same requests again as a burst of concurrent submissions: their bitmaps
        tickets = [srv.submit(tid, qs[i]) for i in sample]
        for i, t in zip(sample, tickets):
            st, ptr, c, _lat = srv.wait(t)
            assert st == eng.REQ_DONE
            d = eng.bitmap_digest(types.SimpleNamespace(ptr=ptr), n)
            srv.release(t)
            want = O.conjunction_digest([cols[cc] for cc, _, _, _ in qs[i]], [(lo, hi) for _, lo, hi, _ in qs[i]])
            assert (d, c) == want, (tid, i, qs[i])
            assert int(cnt[i]) == want[1], (tid, i, qs[i])
            checked += 1
        log(f"table {tid}: 16 results equal the oracle")
    assert checked == 32
    srv.destroy()
    for ix in t1_ix + t2_ix:
        ix.free()
    for d in t1_d + t2_d:
        d.free()
