#!/usr/bin/env python3
"""Benchmark of the MI355X range-predicate scan engine (one GPU).

Default workload (the headline metric's configuration): the plain GPU scan at scale, one
uint32 column of 1e8 uniform rows, 1000 single-range predicates with selectivity uniform
in [0, 1], one query per launch (kernel a, rts_scan_range: bitmap + popcount).  A step is
one predicate's scan of the whole column; --steps K times K of them back to back on one
stream, bracketed by HIP events on that stream.

Prints ONE JSON line: rows scanned per second, the dominant kernel's HBM roofline
fraction (algorithmic bytes per launch / measured time per launch), and the CPU baseline
(the oracle's multithreaded port of the reference raw_scan, timed on this host on a
bounded sample of the same predicates).  `extra` carries the other configurations: the
BinDex query of the same predicates, conjunctions over the 1e8 x 4 table, 4096 batched
queries, the paths either side of the scan (remap encoders, result checker and digest,
pinned result egress, index snapshots), and the request loop over resident 2^26 x 2 and
1e9 x 4 tables (p50 / p99).
"""
import argparse
import gc
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "rows scanned/s & % HBM peak (1e8 uint32); batched queries/s; p50/p99 latency"
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s
N_ROWS = 100_000_000
N_PRED = 1000
DATA_SEED = 2
PRED_SEED = 22
CPU_BASELINES = True  # --no-cpu: skip the CPU baselines of the extra legs (profiling runs)


def config2_predicates(count=N_PRED, seed=PRED_SEED):
    """Closed ranges with selectivity uniform in [0, 1]: width = round(sel * 2^32), lo
    uniform on [0, 2^32 - width] (so hi never clips)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        width = max(1, int(round(float(rng.random()) * 2.0 ** 32)))
        lo = int(rng.integers(0, 2 ** 32 - width + 1))
        out.append((lo, lo + width - 1))
    return out


def load_traffic(name):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE, scaled by
    the copy-kernel calibration), from the committed profile summary, or None."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.isfile(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            try:
                return float(line.split(":", 1)[1].strip())
            except ValueError:
                return None
    return None


def reference_predicate(lo, hi):
    """The reference query line for the closed range [lo, hi]: bt (open) inside the domain,
    le / ge at its edges."""
    import oracle

    if lo == 0 and hi == 0xFFFFFFFF:
        return oracle.GE, 0, 0
    if lo == 0:
        return oracle.LE, hi, 0
    if hi == 0xFFFFFFFF:
        return oracle.GE, lo, 0
    return oracle.BT, lo - 1, hi + 1


def progress(msg):
    """One line per finished step of the run, on stderr (stdout carries only the JSON line)."""
    print("[bench] %s" % msg, file=sys.stderr, flush=True)


def host_info():
    """The host the CPU baselines run on: CPU model, the SIMD level of the timed port, the
    host's CPUs and the CPUs this process may run on (its affinity mask)."""
    import oracle

    model, flags = "unknown", set()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name") and model == "unknown":
                    model = line.split(":", 1)[1].strip()
                elif line.startswith("flags") and not flags:
                    flags = set(line.split(":", 1)[1].split())
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"cpu_model": model, "isa": "avx512f" if "avx512f" in flags else "avx2" if "avx2" in flags else "sse",
            "port_simd": "avx2" if oracle.simd_level() == 2 else "scalar",
            "host_cpus": os.cpu_count(), "usable_cpus": usable}


def cpu_threads():
    return host_info()["usable_cpus"]


class quiet_stdout:
    """Sends file descriptor 1 to /dev/null while the reference's compiled code runs (it
    prints timings); the benchmark's one JSON line stays the only stdout output."""

    def __enter__(self):
        sys.stdout.flush()
        self.saved = os.dup(1)
        self.null = os.open(os.devnull, os.O_WRONLY)
        os.dup2(self.null, 1)
        return self

    def __exit__(self, *exc):
        os.dup2(self.saved, 1)
        os.close(self.saved)
        os.close(self.null)
        return False


def port_time(columns, queries, threads, budget_s, max_queries):
    """The oracle's vectorised port of the checker's conjunction (orc_conjunction_scan: raw_scan
    + AND merge, bindex.cpp:1851-1892 and 2265-2289) on `threads` host threads over the
    queries (each a list of (column, lo, hi)) until ~budget_s of wall time.  Returns
    (queries done, seconds, popcounts)."""
    import oracle

    n = columns[0].size
    bm = np.empty((n + 31) // 32, np.uint32)
    done, t_total, counts = 0, 0.0, []
    for q in queries[:max_queries]:
        cols = [columns[c] for c, _, _ in q]
        rs = [(lo, hi) for _, lo, hi in q]
        t0 = time.perf_counter()
        oracle.conjunction_scan(cols, rs, threads=threads, bitmap=bm)
        t_total += time.perf_counter() - t0
        counts.append(oracle.popcount_rows(bm, n))
        done += 1
        if t_total >= budget_s:
            break
    return done, t_total, counts


def cpu_baseline(x, order, threads, budget_s=8.0):
    """The timed CPU port on the headline workload: the predicates in `order` (the GPU
    loop's first, then the rest of the 1000) until ~budget_s of wall time."""
    done, t_total, counts = port_time([x], [[(0, lo, hi)] for lo, hi in order], threads, budget_s, len(order))
    return {"rows_per_s": done * x.size / t_total, "preds": done, "wall_s": t_total,
            "counts": counts, "order": order[:done], "reference": reference_raw_scan(x, order, counts)}


class RefBindexSet:
    """The reference's own CPU BinDex, compiled from bindex.cpp at N = 1e8 into oracle/_ref,
    over host columns of 1e8 rows, built one column after another as its driver does (its
    init_bindex keeps the sorted copy in a global, bindex.cpp:598-676, so two builds must
    not run at once; each build uses the reference's own threads).  Queries are timed over
    the reference driver's timed region (timer slot 11, bindex.cpp:2190-2262: the BinDex scan
    of every column, then the AND merge on THREAD_NUM threads)."""

    def __init__(self, columns):
        import oracle

        self.lib = oracle.RefBindexLib(N_ROWS)  # FileNotFoundError when not built
        self.columns = columns
        t0 = time.perf_counter()
        self.b = []
        for i, c in enumerate(columns):
            with quiet_stdout():
                self.b.append(self.lib.bindex(c))
            progress("reference CPU BinDex built for column %d of %d" % (i + 1, len(columns)))
        self.build_s = time.perf_counter() - t0
        self.mins = [int(c.min()) for c in columns]
        self.maxs = [int(c.max()) for c in columns]
        W = (N_ROWS + 31) // 32
        self.bitmaps = []
        for _ in range(4):
            raw = np.empty(W + 8, np.uint32)
            off = (-raw.ctypes.data % 32) // 4
            self.bitmaps.append(raw[off:off + W])

    def usable(self, q):
        """The reference asserts (and, built without asserts, reads out of bounds) for bounds
        outside the data (bindex.cpp:1521-1523, 1583): time only ranges inside it."""
        return all(self.mins[c] < lo and hi < self.maxs[c] for c, lo, hi in q)

    def time_queries(self, queries, budget_s, max_queries):
        import oracle

        done, ms, agree = 0, 0.0, 0
        with quiet_stdout():
            for q in queries:
                if done >= max_queries or ms / 1e3 >= budget_s:
                    break
                if not self.usable(q):
                    continue
                preds = [reference_predicate(lo, hi) for _, lo, hi in q]
                ms += self.lib.query_slot11([self.b[c] for c, _, _ in q], preds, self.bitmaps[:len(q)])
                want = oracle.conjunction_scan([self.columns[c] for c, _, _ in q], [(lo, hi) for _, lo, hi in q])
                agree += int(np.array_equal(oracle.mask_tail(self.bitmaps[0], N_ROWS), want))
                done += 1
        return {"queries": done, "ms_per_query": ms / max(done, 1), "queries_per_s": done / (ms / 1e3) if ms else 0.0,
                "threads": self.lib.threads, "kind": "reference",
                "agree_with_raw_scan": agree}


def reference_raw_scan(x, preds, port_counts, budget_s=1.0, max_preds=4):
    """The reference's own raw_scan (bindex.cpp:1851-1892, compiled from its sources into
    oracle/_ref; one host thread, as the reference runs it) on the first predicates over
    the full column, until ~budget_s of wall time; None when that library is not built."""
    import oracle

    try:
        ref = oracle.RefBindexLib(524288)  # raw_scan reads the row count at run time
    except (FileNotFoundError, OSError):
        return None
    n = x.size
    done, t_total, same = 0, 0.0, True
    for i, (lo, hi) in enumerate(preds[:max_preds]):
        op, t1, t2 = reference_predicate(lo, hi)
        t0 = time.perf_counter()
        bm = ref.raw_scan(x, op, t1, t2)
        t_total += time.perf_counter() - t0
        if i < len(port_counts):
            same = same and oracle.popcount_rows(bm, n) == port_counts[i]
        done += 1
        if t_total >= budget_s:
            break
    return {"rows_per_s": done * n / t_total, "preds": done, "wall_s": t_total, "threads": 1,
            "counts_match_port": bool(same)}


class _Slot:
    """A device pointer into a DeviceBuffer (one count slot per step)."""

    def __init__(self, buf, offset):
        self.ptr = buf.ptr + offset


def timed_loop(R, stream, steps, warmup, launch):
    """Run launch(i) for warmup + steps iterations on `stream`; HIP events on that stream
    bracket the timed steps.  Returns milliseconds for the timed steps."""
    e0, e1 = R.Event(), R.Event()
    for i in range(warmup):
        launch(i)
    stream.synchronize()
    R.synchronize()
    e0.record(stream)
    for i in range(steps):
        launch(i)
    e1.record(stream)
    e1.synchronize()
    R.synchronize()
    return e0.elapsed_ms(e1)


def bench_scan(args, R):
    n = N_ROWS
    t0 = time.time()
    x = R.generate_column(R.DIST_UNIFORM, DATA_SEED, 0, n)
    dx = R.DeviceBuffer.from_numpy(x)
    preds = config2_predicates()
    W = (n + 31) // 32
    out = R.DeviceBuffer(4 * W)
    steps, warmup = args.steps, args.warmup
    counts = R.DeviceBuffer(8 * max(steps, 1))
    stream = R.Stream()

    # the predicate of step i: the steps stride through all 1000 predicates (20 steps take
    # every 50th), so any step count samples the whole selectivity range
    def pred(i):
        return preds[(i * len(preds)) // steps % len(preds)] if steps < len(preds) else preds[i % len(preds)]

    # headline: kernel (a) alone, one predicate per launch, bitmap output
    def scan(i):
        lo, hi = pred(i)
        R.scan_range(dx, n, lo, hi, out, None, stream)

    ms = timed_loop(R, stream, steps, warmup, scan)
    progress("headline scan: %.2f us per launch" % (ms * 1e3 / steps))
    per_launch_s = ms / 1e3 / steps
    alg_bytes = n * 4 + W * 4  # column read + bitmap write
    achieved = alg_bytes / per_launch_s / 1e9

    # same predicates with the popcount (a one-workgroup finalize launch per query)
    def scan_count(i):
        lo, hi = pred(i)
        R.scan_range(dx, n, lo, hi, out, _Slot(counts, 8 * (i % max(steps, 1))), stream)

    ms_count = timed_loop(R, stream, steps, warmup, scan_count)
    gpu_counts = counts.to_numpy(np.uint64, steps)

    # BinDex path (kernel b): 128 equal-frequency bins, precomputed filter vectors
    e0, e1 = R.Event(), R.Event()
    e0.record(stream)
    ix = R.BinDexIndex(dx, n, stream)
    e1.record(stream)
    e1.synchronize()
    build_ms = e0.elapsed_ms(e1)
    # the same build again: the first one also pays the runtime's one-time costs (loading
    # the build kernels, first allocations of these sizes)
    e0.record(stream)
    ix_again = R.BinDexIndex(dx, n, stream)
    e1.record(stream)
    e1.synchronize()
    build_again_ms = e0.elapsed_ms(e1)
    del ix_again
    bcounts = R.DeviceBuffer(8 * max(steps, 1))

    def bquery_with(index):
        def bquery(i):
            lo, hi = pred(i)
            R.bindex_query([index], [lo], [hi], out, _Slot(bcounts, 8 * (i % max(steps, 1))), stream)
        return bquery

    ms_b = timed_loop(R, stream, steps, warmup, bquery_with(ix))
    b_counts = bcounts.to_numpy(np.uint64, steps)
    bq_ok = bool(np.array_equal(b_counts, gpu_counts))
    # algorithmic bytes of one BinDex query: two filter vectors + the two boundary bins'
    # refine lists (8 B per entry) + the bitmap write
    b_alg = 3 * W * 4 + 2 * (ix.refine_entries() / ix.bins) * 8
    # the same queries over a 512-bin index of the column (lists a quarter as long)
    ix512 = R.BinDexIndex(dx, n, stream, bins=512)
    ms_b512 = timed_loop(R, stream, steps, warmup, bquery_with(ix512))
    b512_ok = bool(np.array_equal(bcounts.to_numpy(np.uint64, steps), gpu_counts))
    b512_alg = 3 * W * 4 + 2 * (ix512.refine_entries() / ix512.bins) * 8
    ix512.free()

    progress("BinDex queries timed (128 and 512 bins)")
    # all 1000 predicates through rts_scan_ranges: the column read once per 64 predicates
    # (16 calls), the 64 bitmaps of a call written to one 800 MB buffer
    los, his = [p[0] for p in preds], [p[1] for p in preds]
    mout = R.DeviceBuffer(4 * W * 64)
    mcnt = R.DeviceBuffer(8 * len(preds))

    def multi(i, counts=None):
        for p0 in range(0, len(preds), 64):
            R.scan_ranges(dx, n, los[p0:p0 + 64], his[p0:p0 + 64], mout, W,
                          _Slot(counts, 8 * p0) if counts else None, stream)

    ms_multi = timed_loop(R, stream, 2, 1, multi)
    multi(0, mcnt)
    stream.synchronize()
    multi_counts = mcnt.to_numpy(np.uint64, len(preds))
    multi_alg = (len(preds) + 63) // 64 * 4 * n + len(preds) * 4 * W
    mout.free()
    cores = cpu_threads()
    first = [pred(i) for i in range(min(steps, len(preds)))]
    seen = set(first)
    cpu = cpu_baseline(x, first + [p for p in preds if p not in seen], cores)
    # the port's counts against the GPU's, for the predicates both ran
    gpu_by_pred = {pred(i): int(gpu_counts[i]) for i in range(steps)}
    ok = all(gpu_by_pred.get(p, c) == c for p, c in zip(cpu["order"], cpu["counts"]))
    multi_by_pred = {p: int(c) for p, c in zip(preds, multi_counts)}
    multi_ok = all(multi_by_pred[p] == c for p, c in zip(cpu["order"], cpu["counts"])) and \
        all(multi_by_pred[pred(i)] == int(gpu_counts[i]) for i in range(steps))
    # the reference's own CPU BinDex on the same column and predicates (slot-11 region)
    progress("CPU port timed: %.3g rows/s" % cpu["rows_per_s"])
    ref_b = {"skipped": "--no-cpu"}
    if CPU_BASELINES:
        try:
            ref = RefBindexSet([x])
            ref_b = ref.time_queries([[(0, lo, hi)] for lo, hi in cpu["order"]], 10.0, 200)
            ref_b["build_s"] = ref.build_s
            del ref
        except (FileNotFoundError, OSError) as e:
            ref_b = {"unavailable": str(e)}

    return {
        "metric": METRIC,
        "value": n * steps / (ms / 1e3),
        "unit": "rows/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": ms / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (seeded uniform uint32 column, seed %d)" % DATA_SEED,
        "config": {
            "workload": "plain GPU scan: 1 uint32 column x 1e8 rows, 1000 single-range "
                        "predicates with selectivity U[0,1], one query per launch "
                        "(rts_scan_range, kernel k_scan_tiles, bitmap output)",
            "rows": n,
            "predicates": N_PRED,
            "kernel": "k_scan_tiles<1,4,0,false>",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": load_traffic("scan_range_1e8_traffic.txt"),
        },
        "cpu_baseline": {
            "value": cpu["rows_per_s"],
            "unit": "rows/s",
            "cores": cores,
            "kind": "port",
            "sample": "oracle's vectorised raw_scan port (orc_conjunction_scan, %s; "
                      "bindex.cpp:1851-1892), %d threads, %d of the 1000 predicates (striding "
                      "through all of them) over the full 1e8-row column (%.2f s wall)"
                      % (host_info()["port_simd"], cores, cpu["preds"], cpu["wall_s"]),
            "host": host_info(),
        },
        "extra": {
            "scan_with_popcount_ms_per_step": ms_count / steps,
            "bindex_build_ms": build_ms,
            "bindex_build_again_ms": build_again_ms,
            "bindex_query_ms_per_step": ms_b / steps,
            "bindex_query_rows_per_s": n * steps / (ms_b / 1e3),
            "bindex_query_roofline": {"bound": "hbm", "alg_bytes": b_alg,
                                      "traffic": load_traffic("bindex_query_1e8_kernel_stats.txt"),
                                      "achieved_GBps": b_alg / (ms_b / 1e3 / steps) / 1e9,
                                      "frac": b_alg / (ms_b / 1e3 / steps) / 1e9 / HBM_PEAK_GBS},
            "bindex_query_512bins_ms_per_step": ms_b512 / steps,
            "bindex_query_512bins_roofline": {"alg_bytes": b512_alg,
                                              "frac": b512_alg / (ms_b512 / 1e3 / steps) / 1e9 / HBM_PEAK_GBS},
            "bindex_device_bytes": ix.device_bytes(),
            "ranges_in_passes": {
                "call": "rts_scan_ranges: the 1000 predicates in 16 calls of up to 64, the "
                        "column read once per call (k_scan_multi), no counts in the timed region",
                "ms_all_predicates": ms_multi / 2, "us_per_predicate": ms_multi / 2 / len(preds) * 1e3,
                "row_predicates_per_s": n * len(preds) / (ms_multi / 2 / 1e3),
                "roofline": {"bound": "hbm", "alg_bytes": multi_alg,
                             "achieved_GBps": multi_alg / (ms_multi / 2 / 1e3) / 1e9,
                             "frac": multi_alg / (ms_multi / 2 / 1e3) / 1e9 / HBM_PEAK_GBS},
                "counts_match": multi_ok},
            # the reference's own compiled code on this host, beside the port above: its
            # raw_scan (one thread, as it runs) and its CPU BinDex query (slot-11 region)
            "cpu_reference_raw_scan": cpu["reference"],
            "cpu_reference_bindex": ref_b,
        },
        "parity": {"checked_predicates": len(cpu["counts"]), "popcounts_match_oracle": bool(ok),
                   "bindex_counts_match_scan": bq_ok, "bindex_512bins_counts_match_scan": b512_ok},
        "setup_s": round(time.time() - t0, 2),
    }


TABLE_SEED = 3
TABLE_KINDS = [  # (distribution, p0, p1): uniform, normal(2^31, 2^28), Zipf 1.1 and 1.5
    (0, 0.0, 0.0), (1, 2.0 ** 31, 2.0 ** 28), (2, 1.1, 2.0 ** 20), (2, 1.5, 2.0 ** 20)]


class Table:
    """The 1e8-row x 4-column table of the conjunction and batched workloads: uniform,
    normal (mean 2^31, sd 2^28, clipped), Zipf s=1.1 and s=1.5 over 2^20 ranks mapped to
    random uint32 keys; resident in HBM with a BinDex per column."""

    def __init__(self, R, n=N_ROWS, seed=TABLE_SEED, stream=None):
        self.n = n
        self.cols = [R.generate_column(k, seed, c, n, p0, p1)
                     for c, (k, p0, p1) in enumerate(TABLE_KINDS)]
        self.dcols = [R.DeviceBuffer.from_numpy(c) for c in self.cols]
        e0, e1 = R.Event(), R.Event()
        st = stream or R.Stream()
        e0.record(st)
        self.idx = [R.BinDexIndex(d, n, st) for d in self.dcols]
        e1.record(st)
        e1.synchronize()
        self.build_ms = e0.elapsed_ms(e1)
        rng = np.random.default_rng(seed)
        pick = rng.integers(0, n, 1_000_000)
        self.samples = [np.sort(c[pick]) for c in self.cols]  # for quantile endpoints
        self.R = R
        self.sets = {}
        self._ref = None

    def index_set(self, bins):
        """The table's columns indexed with `bins` engine bins (built on first use)."""
        if bins == 128:
            return self.idx
        if bins not in self.sets:
            self.sets[bins] = [self.R.BinDexIndex(d, self.n, bins=bins) for d in self.dcols]
        return self.sets[bins]

    def ref(self):
        """The reference's own CPU BinDex over the table's columns (None if not built)."""
        if not CPU_BASELINES:
            return None
        if self._ref is None:
            try:
                self._ref = RefBindexSet(self.cols)
            except (FileNotFoundError, OSError) as e:
                self._ref = str(e)
        return self._ref if not isinstance(self._ref, str) else None

    def range_for(self, c, sel, rng):
        """Closed range on column c whose selectivity is about `sel`: endpoints at sample
        quantiles q0 and q0 + sel."""
        s = self.samples[c]
        q0 = float(rng.random()) * (1.0 - sel)
        lo = int(s[int(q0 * (s.size - 1))])
        hi = int(s[min(s.size - 1, int((q0 + sel) * (s.size - 1)))])
        return lo, hi


def bench_conj(R, table, nq=30, seed=33):
    """Config-3 conjunctions over 2, 3 and 4 columns, per-column selectivity in {0.01, 0.1,
    0.5}: the fused direct kernel (k_scan_tiles<C>), the fused BinDex kernel
    (k_bindex_query<C>), and the reference's unfused pattern (one scan per column, then a
    GPUbitAndWithCuda chain as merge_with_GPU does, bindex_cuda.cpp:1693-1702)."""
    n = table.n
    W = (n + 31) // 32
    rng = np.random.default_rng(seed)
    stream = R.Stream()
    out = R.DeviceBuffer(4 * W)
    tmp = [R.DeviceBuffer(4 * W) for _ in range(4)]
    cnts = R.DeviceBuffer(8 * nq)  # one count slot per query
    slot = [_Slot(cnts, 8 * i) for i in range(nq)]
    res = {}
    for k in (2, 3, 4):
        qs = []
        for _ in range(nq):
            cs = list(rng.choice(4, k, replace=False))
            rngs = [table.range_for(c, [0.01, 0.1, 0.5][int(rng.integers(3))], rng) for c in cs]
            qs.append((cs, rngs))
        counts = {}

        def run(kind, q, cnt):
            cs, rr = q
            lo = [a for a, _ in rr]
            hi = [b for _, b in rr]
            if kind == "direct":
                R.conjunction_scan([table.dcols[c] for c in cs], lo, hi, n, out, cnt, stream)
            elif kind == "bindex":
                R.bindex_query([table.idx[c] for c in cs], lo, hi, out, cnt, stream)
            else:
                for j, c in enumerate(cs):
                    R.scan_range(table.dcols[c], n, lo[j], hi[j], tmp[j], None, stream)
                stream.synchronize()
                for j in range(1, len(cs)):
                    R.check(R.lib().GPUbitAndWithCuda(tmp[0].ptr, tmp[j].ptr, n), "and")
                R.bitmap_popcount(tmp[0], n, cnt, stream)

        entry = {}
        for kind in ("direct", "bindex", "unfused"):
            for q in qs[:2]:
                run(kind, q, slot[0])
            stream.synchronize()
            R.synchronize()
            if kind == "unfused":
                # the reference's pattern synchronises on every call: host wall clock
                t0 = time.perf_counter()
                for i, q in enumerate(qs):
                    run(kind, q, slot[i])
                    stream.synchronize()
                dt_ms = (time.perf_counter() - t0) * 1e3
            else:
                # queries back to back on one stream, timed with HIP events on that stream
                e0, e1 = R.Event(), R.Event()
                e0.record(stream)
                for i, q in enumerate(qs):
                    run(kind, q, slot[i])
                e1.record(stream)
                e1.synchronize()
                dt_ms = e0.elapsed_ms(e1)
            stream.synchronize()
            counts[kind] = [int(c) for c in cnts.to_numpy(np.uint64, nq)]
            entry[kind + "_us_per_query"] = dt_ms * 1e3 / nq
        entry["direct_alg_GBps"] = (k * 4 * n + 4 * W) / (entry["direct_us_per_query"] * 1e-6) / 1e9
        entry["counts_agree"] = counts["direct"] == counts["bindex"] == counts["unfused"]
        entry["realised_selectivity_median"] = float(np.median(counts["direct"])) / n
        # rooflines: the direct kernel reads every column once and writes the bitmap; the
        # BinDex kernel reads two filter vectors and two bins' list slices per column
        dir_bytes = k * 4 * n + 4 * W
        bx_bytes = float(np.mean([sum(2 * 4 * W + 2 * table.idx[c].refine_entries() / table.idx[c].bins * 8
                                      for c in cs) + 4 * W for cs, _ in qs]))
        entry["roofline"] = {
            "bound": "hbm", "peak_GBps": HBM_PEAK_GBS,
            "direct": {"alg_bytes": dir_bytes,
                       "frac": dir_bytes / (entry["direct_us_per_query"] * 1e-6) / 1e9 / HBM_PEAK_GBS},
            "bindex": {"alg_bytes": bx_bytes,
                       "frac": bx_bytes / (entry["bindex_us_per_query"] * 1e-6) / 1e9 / HBM_PEAK_GBS}}
        # CPU baselines on the same queries: the vectorised port (all usable host threads)
        # and the reference's own CPU BinDex (slot-11 region, its THREAD_NUM threads)
        cq = [[(int(c), lo, hi) for c, (lo, hi) in zip(cs, rr)] for cs, rr in qs]
        threads = cpu_threads()
        done, secs, pc = port_time(table.cols, cq, threads, 3.0, 12 if CPU_BASELINES else 1)
        entry["cpu_baseline"] = {
            "port": {"us_per_query": secs / done * 1e6, "queries_per_s": done / secs, "cores": threads,
                     "kind": "port", "queries": done, "counts_match_gpu": pc == counts["direct"][:done]}}
        ref = table.ref()
        if ref is not None:
            entry["cpu_baseline"]["reference_bindex"] = ref.time_queries(cq, 3.0, 12)
        res["cols%d" % k] = entry
        progress("conjunctions of %d columns timed" % k)
    # bitmaps, not only counts: one query per column count through the three paths, compared
    # on the device by digest (each path's parity with the oracle is pinned by the tests)
    ok = True
    for k in (2, 4):
        q = ([c for c in range(k)], [table.range_for(c, 0.5, rng) for c in range(k)])
        digests = []
        for kind in ("direct", "bindex", "unfused"):
            run(kind, q, slot[0])
            stream.synchronize()
            digests.append(R.bitmap_digest(tmp[0] if kind == "unfused" else out, n, stream))
        ok = ok and digests[0] == digests[1] == digests[2]
    res["bitmaps_agree"] = ok
    return res


def config1_predicates(n_pred=64, seed=20240101):
    """Config 1's 64 closed ranges: width = selectivity x 2^32 with the selectivity cycling
    through {0.001, 0.01, 0.1, 0.5, 0.9}, lo uniform on [0, 2^32 - width] (hi never clips)."""
    rng = np.random.default_rng(seed)
    sels = [0.001, 0.01, 0.1, 0.5, 0.9]
    out = []
    for i in range(n_pred):
        width = int(round(sels[i % 5] * 2.0 ** 32))
        lo = int(rng.integers(0, 2 ** 32 - width + 1))
        out.append((lo, lo + width - 1))
    return out


def bench_config1(R, n=1 << 20, reps=20):
    """Config 1: one 2^20-row uniform column, 64 closed ranges, one query per launch with the
    popcount (bitmap + count); every bitmap checked against the oracle's raw_scan."""
    import oracle

    x = R.generate_column(R.DIST_UNIFORM, 1, 0, n)
    dx = R.DeviceBuffer.from_numpy(x)
    W = (n + 31) // 32
    preds = config1_predicates()
    stream = R.Stream()
    out = R.DeviceBuffer(4 * W * len(preds))
    cnt = R.DeviceBuffer(8 * len(preds))

    def launch(i):
        j = i % len(preds)
        lo, hi = preds[j]
        R.scan_range(dx, n, lo, hi, _Slot(out, 4 * W * j), _Slot(cnt, 8 * j), stream)

    steps = reps * len(preds)
    ms = timed_loop(R, stream, steps, len(preds), launch)
    got = out.to_numpy(np.uint32, W * len(preds)).reshape(len(preds), W)
    counts = cnt.to_numpy(np.uint64, len(preds))
    ok = True
    for j, (lo, hi) in enumerate(preds):
        op, t1, t2 = reference_predicate(lo, hi)
        truth = oracle.raw_scan(x, op, t1, t2, threads=cpu_threads())
        ok = ok and bool(np.array_equal(got[j], truth)) and int(counts[j]) == oracle.popcount_rows(truth, n)
    us = ms * 1e3 / steps
    alg = 4 * n + 4 * W
    # the same 64 ranges in one pass over the column (rts_scan_ranges: bitmaps and counts)
    los, his = [p[0] for p in preds], [p[1] for p in preds]
    out.fill(0)
    ms_one = timed_loop(R, stream, reps, 2, lambda i: R.scan_ranges(dx, n, los, his, out, W, cnt, stream))
    got1 = out.to_numpy(np.uint32, W * len(preds)).reshape(len(preds), W)
    one_ok = bool(np.array_equal(got1, got)) and bool(np.array_equal(cnt.to_numpy(np.uint64, len(preds)), counts))
    us_one = ms_one * 1e3 / reps
    alg_one = 4 * n + 4 * W * len(preds)  # the column once + 64 bitmaps
    threads = cpu_threads()
    done, secs, _ = port_time([x], [[(0, lo, hi)] for lo, hi in preds] * 20, threads, 2.0, 64 * 20)
    ref = reference_raw_scan(x, preds, [], budget_s=1.0, max_preds=64)
    out.free()
    dx.free()
    return {"rows": n, "predicates": len(preds), "us_per_query": us, "rows_per_s": n / (us * 1e-6),
            "roofline": {"bound": "hbm (launch-bound at this size)", "alg_bytes": alg,
                         "frac": alg / (us * 1e-6) / 1e9 / HBM_PEAK_GBS},
            "cpu_baseline": {"port": {"rows_per_s": done * n / secs, "cores": threads, "kind": "port"},
                             "reference_raw_scan": ref},
            "one_pass": {"call": "rts_scan_ranges, all 64 ranges in one call (k_scan_multi + counts)",
                         "us_per_call": us_one, "us_per_query": us_one / len(preds),
                         "rows_per_s": n * len(preds) / (us_one * 1e-6),
                         "roofline": {"bound": "hbm", "alg_bytes": alg_one,
                                      "frac": alg_one / (us_one * 1e-6) / 1e9 / HBM_PEAK_GBS},
                         "equal_to_one_per_launch": one_ok},
            "bitmaps_and_counts_match_oracle": ok}


def bench_aux(R, n=N_ROWS, snap_rows=1 << 24, reps=3):
    """The paths either side of the scan on the 1e8-row uniform column: the remap encoders
    (rank and dispersed forms of normalEncode, remap.cpp:110-241; wall clock of the C entry
    point, which returns after the map size is known), the device result checker and digest
    over a 1e8-row bitmap (HIP events), result egress into pinned host memory (HIP events),
    and BinDex snapshot save / load for a 2^24-row column (wall clock, a temporary file)."""
    import ctypes

    L = R.lib()
    x = R.generate_column(R.DIST_UNIFORM, DATA_SEED, 0, n)
    dx = R.DeviceBuffer.from_numpy(x)
    del x
    stream = R.Stream()
    res = {}
    codes, mv, ms = (R.DeviceBuffer(4 * n) for _ in range(3))
    m = ctypes.c_int64()

    def wall(fn):
        fn()  # warm-up
        t = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            t.append(time.perf_counter() - t0)
        return float(np.median(t)) * 1e3

    enc_ms = wall(lambda: R.check(L.rts_encode_column(dx.ptr, n, 0, 0xFFFFFFFF, codes.ptr, None, None,
                                                      mv.ptr, ms.ptr, ctypes.byref(m), stream.handle),
                                  "rts_encode_column"))
    res["encode_rank"] = {"ms": enc_ms, "rows_per_s": n / enc_ms * 1e3, "distinct": m.value}
    disp_ms = wall(lambda: R.check(L.rts_encode_column_dispersed(dx.ptr, n, 0, 0xFFFFFFFF, 7, codes.ptr,
                                                                 mv.ptr, ms.ptr, ctypes.byref(m),
                                                                 stream.handle),
                                   "rts_encode_column_dispersed"))
    res["encode_dispersed"] = {"ms": disp_ms, "rows_per_s": n / disp_ms * 1e3}
    for b in (codes, mv, ms):
        b.free()

    W = (n + 31) // 32
    a, b = R.DeviceBuffer(4 * W), R.DeviceBuffer(4 * W)
    R.scan_range(dx, n, 1 << 30, 3 << 30, a, None, stream)
    R.scan_range(dx, n, 1 << 30, 3 << 30, b, None, stream)
    stream.synchronize()
    steps = 20
    dg = ctypes.c_ulonglong()
    cr = R._CompareResult()
    dig = timed_loop(R, stream, steps, 2,
                     lambda i: R.check(L.rts_bitmap_digest(a.ptr, n, ctypes.byref(dg), stream.handle),
                                       "rts_bitmap_digest")) / steps
    cmp_ = timed_loop(R, stream, steps, 2,
                      lambda i: R.check(L.rts_bitmap_compare(a.ptr, b.ptr, n, ctypes.byref(cr),
                                                             stream.handle), "rts_bitmap_compare")) / steps
    res["digest"] = {"us": dig * 1e3, "GBps": 4 * W / (dig * 1e-3) / 1e9}
    res["compare"] = {"us": cmp_ * 1e3, "GBps": 8 * W / (cmp_ * 1e-3) / 1e9, "equal": cr.first_mismatch == -1}
    host = ctypes.c_void_p()
    R.check(L.rts_host_alloc(ctypes.byref(host), 4 * W), "rts_host_alloc")
    egress = timed_loop(R, stream, 10, 2,
                        lambda i: R.check(L.rts_memcpy_dtoh_async(host.value, a.ptr, 4 * W, stream.handle),
                                          "rts_memcpy_dtoh_async")) / 10
    L.rts_host_free(host.value)
    res["egress_pinned"] = {"ms": egress, "GBps": 4 * W / (egress * 1e-3) / 1e9, "bytes": 4 * W}
    a.free()
    b.free()

    ix = R.BinDexIndex(_Slot(dx, 0), snap_rows)
    nbytes = ix.device_bytes()
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "index.rtsbdx")
        t0 = time.perf_counter()
        ix.save(path)
        save_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        ix2 = R.BinDexIndex(_Slot(dx, 0), snap_rows, snapshot=path)
        load_s = time.perf_counter() - t0
        fsize = os.path.getsize(path)
        ix2.free()
    ix.free()
    res["snapshot"] = {"rows": snap_rows, "file_bytes": fsize, "device_bytes": nbytes,
                       "save_GBps": fsize / save_s / 1e9, "load_GBps": fsize / load_s / 1e9}
    dx.free()
    return res


def batch_queries(table, nq, seed, direct_share):
    """Config-4 queries: 1..4 columns chosen uniformly, per-column selectivity log-uniform in
    [1e-5, 0.9]; a share of the query-columns is evaluated directly (mode 2)."""
    rng = np.random.default_rng(seed)
    qs = []
    for _ in range(nq):
        k = int(rng.integers(1, 5))
        cs = rng.choice(4, k, replace=False)
        q = []
        for c in cs:
            sel = float(np.exp(rng.uniform(np.log(1e-5), np.log(0.9))))
            lo, hi = table.range_for(int(c), sel, rng)
            q.append((int(c), lo, hi, 2 if rng.random() < direct_share else 0))
        qs.append(q)
    return qs


def bench_batch(R, table, nq=4096, seed=44, steps=5, tune_cache=None, config=None):
    """Config 4: nq mixed queries in one persistent hipGraph-captured launch.  The launch
    shape comes from the autotuner in two stages: tile / group / workgroup shape on the
    table's 128-bin indexes, then the bin count (index sets of 128, 256 and 512 bins) and the
    planner's direct weight, which decides per query-column between the direct and the BinDex
    form.  Compared with the plain one-query-per-launch paths (direct conjunction, and BinDex)
    on the same GPU and with the CPU baselines on a sample of the same queries."""
    n = table.n
    W = (n + 31) // 32
    stride = (W + 3) // 4 * 4
    qs = batch_queries(table, nq, seed, 0.0)
    stream = R.Stream()
    out = R.DeviceBuffer(4 * stride * nq)
    counts = R.DeviceBuffer(8 * nq)
    plan = R.BatchPlan(table.dcols, table.idx, n, nq, None, stream)
    for bins in (256, 512):
        plan.add_indexes(table.index_set(bins))
    plan.set_queries(qs, stream)
    t0 = time.perf_counter()
    if config:  # fixed launch shape (profiling runs)
        cands = [R.batch_config(*config)]
        chosen, tune_ms = plan.autotune(out, stride, cands, reps=2, cache_path=tune_cache, stream=stream)
        stages = [len(cands)]
    else:
        # launch shape x bin count jointly, then the planner's direct weight on the winner
        shapes = [R.batch_config(t, g, 4, 1, nt, bins=b) for b in (128, 256, 512) for t in (1024, 2048)
                  for g in (32, 64) for nt in (256, 512)]
        shape, _ = plan.autotune(out, stride, shapes, reps=2, stream=stream)
        s = shape.astuple()[:6]
        plans = [R.batch_config(*s, direct_weight=w) for w in (0, 40, 80)]
        chosen, tune_ms = plan.autotune(out, stride, plans, reps=2, cache_path=tune_cache, stream=stream)
        stages = [len(shapes), len(plans)]
    tune_s = time.perf_counter() - t0
    progress("batch autotuned: %s" % (chosen.astuple(),))
    plan.graph_launch(out, stride, counts, stream)  # capture + first run
    stream.synchronize()
    ms = timed_loop(R, stream, steps, 1,
                    lambda i: plan.graph_launch(out, stride, counts, stream))
    batch_counts = counts.to_numpy(np.uint64, nq)
    st = plan.stats()
    # host planning time (bound lookups, cost model, grouping, upload), which a served batch
    # pays before its launch: the whole batch and a 256-query serving batch
    plan_ms = {}
    for m in (nq, 256):
        t = []
        for _ in range(3):
            t0 = time.perf_counter()
            plan.set_queries(qs[:m], stream)
            t.append((time.perf_counter() - t0) * 1e3)
        plan_ms[str(m)] = float(np.median(t))
    plan.set_queries(qs, stream)

    # plain one-query-per-launch paths over the same queries
    one = R.DeviceBuffer(4 * W)
    c1 = R.DeviceBuffer(8 * nq)

    class Slot:
        def __init__(self, i):
            self.ptr = c1.ptr + 8 * i

    def plain_direct(i):
        q = qs[i]
        R.conjunction_scan([table.dcols[c] for c, _, _, _ in q], [lo for _, lo, _, _ in q],
                           [hi for _, _, hi, _ in q], n, one, Slot(i), stream)

    progress("batch timed: %.2f ms per launch" % (ms / steps))
    ms_plain = timed_loop(R, stream, nq, 8, plain_direct)
    plain_counts = c1.to_numpy(np.uint64, nq)

    def plain_bindex(i):
        q = qs[i]
        R.bindex_query([table.idx[c] for c, _, _, _ in q], [lo for _, lo, _, _ in q],
                       [hi for _, _, hi, _ in q], one, Slot(i), stream)

    ms_plain_b = timed_loop(R, stream, nq, 8, plain_bindex)
    plain_b_counts = c1.to_numpy(np.uint64, nq)
    # full bitmaps of a sample of queries: batched == one-query-per-launch
    same = True
    for i in range(0, nq, max(1, nq // 8)):
        plain_direct(i)
        stream.synchronize()
        a = one.to_numpy(np.uint32, W)
        b = np.empty(W, np.uint32)
        R.check(R.lib().rts_memcpy_dtoh(b.ctypes.data, out.ptr + 4 * stride * i, 4 * W), "dtoh")
        same = same and bool(np.array_equal(a, b))
    qps = nq * steps / (ms / 1e3)
    qps_plain = nq / (ms_plain / 1e3)
    qps_plain_b = nq / (ms_plain_b / 1e3)
    out_bytes = nq * W * 4
    floor_bytes = out_bytes + len(table.dcols) * 4 * n  # every bitmap written + the table read once
    plan.free()
    # CPU baselines on a sample of the same queries
    cq = [[(c, lo, hi) for c, lo, hi, _ in q] for q in qs]
    threads = cpu_threads()
    done, secs, pc = port_time(table.cols, cq, threads, 4.0, 32 if CPU_BASELINES else 1)
    cpu = {"port": {"queries_per_s": done / secs, "cores": threads, "kind": "port", "queries": done,
                    "counts_match_gpu": [int(v) for v in batch_counts[:done]] == pc}}
    ref = table.ref()
    if ref is not None:
        cpu["reference_bindex"] = ref.time_queries(cq, 4.0, 32)
    total_cols = st["bindex_cols"] + st["direct_cols"]
    return {
        "queries": nq,
        "batch_ms": ms / steps,
        "batch_queries_per_s": qps,
        "plain_direct_queries_per_s": qps_plain,
        "plain_bindex_queries_per_s": qps_plain_b,
        "speedup_vs_plain_direct": qps / qps_plain,
        "speedup_vs_plain_bindex": qps / qps_plain_b,
        "bitmap_write_GBps": out_bytes / (ms / steps / 1e3) / 1e9,
        "roofline": {"bound": "hbm", "alg_bytes": floor_bytes,
                     "achieved_GBps": floor_bytes / (ms / steps / 1e3) / 1e9,
                     "frac": floor_bytes / (ms / steps / 1e3) / 1e9 / HBM_PEAK_GBS,
                     "traffic": load_traffic("batch_4096q_1e8x4_kernel_stats.txt")},
        "cpu_baseline": cpu,
        "autotune": {"chosen": list(chosen.astuple()), "best_ms": tune_ms, "search_s": tune_s,
                     "candidates": stages},
        "chosen_bins": chosen.bins or 128,
        "planner": st,
        "host_planning_ms": plan_ms,
        "planner_direct_share": st["direct_cols"] / max(1, total_cols),
        "counts_match_plain": bool(np.array_equal(batch_counts, plain_counts)) and
                              bool(np.array_equal(batch_counts, plain_b_counts)),
        "sampled_bitmaps_match_plain": same,
    }


def serve_queries(rng, table_cols, nq, ncols_max):
    """Config-4 query shapes for the serving workload: 1..ncols_max columns chosen uniformly,
    per-column selectivity log-uniform in [1e-5, 0.9], endpoints from sample quantiles."""
    qs = []
    for _ in range(nq):
        k = int(rng.integers(1, ncols_max + 1))
        q = []
        for c in rng.choice(ncols_max, k, replace=False):
            sel = float(np.exp(rng.uniform(np.log(1e-5), np.log(0.9))))
            s = table_cols[c]
            q0 = float(rng.random()) * (1.0 - sel)
            lo = int(s[int(q0 * (s.size - 1))])
            hi = int(s[min(s.size - 1, int((q0 + sel) * (s.size - 1)))])
            q.append((int(c), lo, hi, 0))
        qs.append(q)
    return qs


def bench_serve(R, rates=(1000, 10000, 50000), seconds=(2.0, 1.0, 0.5), n2=1_000_000_000,
                seed=55, pool_gb=None, batch_config=None, bins=256):
    """Config 5: two resident tables, T1 = 2^26 rows x 2 uniform columns, T2 = n2 rows x 4
    columns (uniform, normal, Zipf 1.1, Zipf 1.5), a BinDex of `bins` bins per column (the
    count the autotuner picks for the config-4 batch); Poisson arrivals at
    each rate, every query on T1 or T2 with probability 1/2, shapes as config 4; admission by
    the result pool's free HBM (a pool of up to 64 GB of what the tables leave free, 8 GB kept
    back: 512 of the 125 MB bitmaps of the 1e9-row table); continuous batching of up to 512
    requests over 3 streams."""
    t_setup = time.time()
    n1 = 1 << 26
    import oracle

    t1_cols = [R.generate_column(0, seed, c, n1) for c in range(2)]
    t1_d = [R.DeviceBuffer.from_numpy(c) for c in t1_cols]
    t1_ix = [R.BinDexIndex(d, n1, bins=bins) for d in t1_d]
    t2_d, t2_ix, t2_samples, t2_cols = [], [], [], []
    rng = np.random.default_rng(seed)
    for c, (k, p0, p1) in enumerate(TABLE_KINDS):
        col = R.generate_column(k, seed + 1, c, n2, p0, p1)
        t2_samples.append(np.sort(col[rng.integers(0, n2, 1_000_000)]))
        t2_d.append(R.DeviceBuffer.from_numpy(col))
        t2_cols.append(col)  # kept on the host for the CPU baseline and the oracle checks
        t2_ix.append(R.BinDexIndex(t2_d[-1], n2, bins=bins))
    t1_samples = [np.sort(c[rng.integers(0, n1, 1_000_000)]) for c in t1_cols]
    setup_s = time.time() - t_setup
    if pool_gb is None:
        pool_gb = int(max(8, min(64, (R.device_free_bytes() >> 30) - 8)))
    cfg = R.server_config(max_batch=512, streams=3, pool_bytes=pool_gb << 30)
    if batch_config is not None:
        cfg.batch = R.batch_config(*batch_config)
    srv = R.Server(cfg)
    tid1 = srv.add_table(t1_d, t1_ix, n1)
    tid2 = srv.add_table(t2_d, t2_ix, n2)
    srv.start()
    progress("served tables ready (%.1f s)" % setup_s)
    out = {"setup_s": round(setup_s, 1), "rows": [n1, n2], "pool_gb": pool_gb, "bins": bins,
           "batch_config": list(cfg.batch.astuple()), "rates": []}
    for rate, secs in zip(rates, seconds):
        nq = int(rate * secs)
        tids = [tid1 if rng.random() < 0.5 else tid2 for _ in range(nq)]
        qs1 = serve_queries(rng, t1_samples, nq, 2)
        qs2 = serve_queries(rng, t2_samples, nq, 4)
        qs = [qs1[i] if tids[i] == tid1 else qs2[i] for i in range(nq)]
        arrivals = np.cumsum(rng.exponential(1e9 / rate, nq)).astype(np.uint64)
        st0 = srv.stats()
        lat, cnt, elapsed = srv.replay(tids, qs, arrivals)
        st1 = srv.stats()
        lat_ms = lat.astype(np.float64) / 1e6
        # bit-exactness: counts of a sample of queries against the oracle's conjunction
        check = list(range(0, nq, max(1, nq // 16)))
        ok = True
        for i in check:
            cols = t1_cols if tids[i] == tid1 else t2_cols
            _d, c = oracle.conjunction_digest([cols[cc] for cc, _, _, _ in qs[i]],
                                              [(lo, hi) for _, lo, hi, _ in qs[i]], threads=cpu_threads())
            ok = ok and c == int(cnt[i])
        out_bytes = sum((n1 if t == tid1 else n2) + 31 for t in tids) // 32 * 4
        out["rates"].append({
            "offered_qps": rate, "queries": nq,
            "throughput_qps": nq / elapsed,
            "p50_ms": float(np.percentile(lat_ms, 50)), "p99_ms": float(np.percentile(lat_ms, 99)),
            "mean_ms": float(lat_ms.mean()),
            "batches": st1["batches"] - st0["batches"], "max_batch": st1["max_batch_seen"],
            "pool_peak_gb": st1["pool_peak"] / 2 ** 30,
            "bitmap_write_GBps": out_bytes / elapsed / 1e9,
            "roofline_frac": out_bytes / elapsed / 1e9 / HBM_PEAK_GBS,
            "sampled_counts_match_oracle": bool(ok), "checked": len(check),
        })
        last_qs, last_tids = qs, tids
        progress("served %d q/s offered: %.0f q/s" % (rate, nq / elapsed))
    srv.destroy()
    # CPU baseline: the vectorised port, one query after another over both tables (the
    # reference has no request loop; its driver runs one query at a time)
    threads = cpu_threads()
    t_cpu, done = 0.0, 0
    for i in range(min(32 if CPU_BASELINES else 1, len(last_qs))):
        cols = t1_cols if last_tids[i] == tid1 else t2_cols
        q = [(cc, lo, hi) for cc, lo, hi, _ in last_qs[i]]
        d1, s1, _ = port_time(cols, [q], threads, 1e9, 1)
        t_cpu += s1
        done += d1
        if t_cpu > 10.0:
            break
    out["cpu_baseline"] = {"queries_per_s": done / t_cpu, "cores": threads, "kind": "port", "queries": done,
                           "sample": "vectorised port, the last rate's first queries in arrival order"}
    sustained = [r for r in out["rates"] if r["throughput_qps"] >= 0.95 * r["offered_qps"]]
    out["max_sustained_offered_qps"] = max((r["offered_qps"] for r in sustained), default=0)
    return out


def bench_readprobe(R, nbytes=4_000_000_000, reps=5):
    """The device's streaming-read ceiling with the scan kernels' load pattern (rts_probe_read:
    non-temporal uint4 per lane), swept over grid sizes and loads in flight; the best sweep
    point is the practical ceiling the scan's bandwidth is set against."""
    buf = R.DeviceBuffer(nbytes)
    R.check(R.lib().rts_memset(buf.ptr, 1, nbytes), "memset")
    sink = R.DeviceBuffer(16)
    stream = R.Stream()
    best, points = None, []
    for bpc in (4, 8, 16, 32):
        for unroll in (1, 2, 4, 8):
            def launch(i):
                R.check(R.lib().rts_probe_read(buf.ptr, nbytes, bpc, unroll, sink.ptr, stream.handle),
                        "rts_probe_read")
            ms = timed_loop(R, stream, reps, 1, launch) / reps
            gbps = nbytes / (ms * 1e-3) / 1e9
            points.append([bpc, unroll, round(gbps, 1)])
            if best is None or gbps > best[2]:
                best = [bpc, unroll, gbps]
    buf.free()
    return {"bytes": nbytes, "best_GBps": best[2], "best_blocks_per_cu": best[0], "best_unroll": best[1],
            "frac_of_8TBps": best[2] / HBM_PEAK_GBS, "sweep": points}


def bench_copycal(args, R):
    """Calibration for rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE: five device-to-device
    copies of exactly 400,000,000 bytes (the HIP runtime's plain copy kernel)."""
    nbytes = 400_000_000
    a = R.DeviceBuffer(nbytes).fill(1)
    b = R.DeviceBuffer(nbytes)
    for _ in range(5):
        R.check(R.lib().rts_memcpy_dtod(b.ptr, a.ptr, nbytes), "dtod")
    R.synchronize()
    return {"workload": "copycal", "copies": 5, "bytes_per_copy": nbytes}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="scan",
                    choices=["scan", "copycal", "conj", "batch", "serve", "aux", "config1", "readprobe"])
    ap.add_argument("--serve-rows", type=int, default=1_000_000_000,
                    help="rows of the large serving table")
    ap.add_argument("--batch-config", default=None,
                    help="tile_words,queries_per_group,blocks_per_cu[,order[,threads[,bins[,direct_weight]]]] for the "
                         "batch workload (skips the autotuner) or the serve workload's batches")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the conjunction, batched and serving legs of the default run")
    ap.add_argument("--no-cpu", action="store_true",
                    help="skip the CPU baselines of the conj / batch / serve legs (profiling runs)")
    args = ap.parse_args()
    global CPU_BASELINES
    CPU_BASELINES = not args.no_cpu
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this engine runs on one GPU only; --gpus %d is not supported "
              "(row sharding across GPUs is designed, not built: every GPU would take a "
              "contiguous range of rows and its own index)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    import rtscan_amd as R

    if R.lib().rts_device_count() < 1:
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.workload == "copycal":
        res = bench_copycal(args, R)
    elif args.workload == "aux":
        res = {"aux": bench_aux(R)}
    elif args.workload == "config1":
        res = {"config1": bench_config1(R)}
    elif args.workload == "readprobe":
        res = {"readprobe": bench_readprobe(R)}
    elif args.workload == "serve":
        cfg = tuple(int(v) for v in args.batch_config.split(",")) if args.batch_config else None
        res = {"serve": bench_serve(R, n2=args.serve_rows, batch_config=cfg)}
    elif args.workload in ("conj", "batch"):
        table = Table(R)
        res = {"table_build_ms": table.build_ms}
        if args.workload == "conj":
            res["conjunction"] = bench_conj(R, table)
        else:
            cfg = tuple(int(v) for v in args.batch_config.split(",")) if args.batch_config else None
            res["batch"] = bench_batch(R, table, steps=max(1, min(args.steps, 20)), config=cfg)
    else:
        res = bench_scan(args, R)
        if not args.no_extra:
            try:
                res["extra"]["config1"] = bench_config1(R)
            except Exception as e:  # noqa: BLE001
                res["extra"]["config1_error"] = "%s: %s" % (type(e).__name__, e)
            try:
                res["extra"]["read_ceiling"] = bench_readprobe(R)
            except Exception as e:  # noqa: BLE001
                res["extra"]["read_ceiling_error"] = "%s: %s" % (type(e).__name__, e)
            # config 5 first: the reference-named wrappers (the unfused conjunction pattern
            # below) and the snapshot load and save run on the calling thread's per-thread
            # stream, and a request loop started after that in the same process sustained
            # 10-20% fewer requests at saturation (its three streams then share the
            # process's hardware queues with that stream's); a serving process runs the
            # loop alone.  Launch shape: the request loop's default (2048-word tiles,
            # 256-thread workgroups, the shape the autotuner picks for the config-4 batch).
            try:
                res["extra"]["serve"] = bench_serve(R, n2=args.serve_rows, bins=256)
            except Exception as e:  # noqa: BLE001
                res["extra"]["serve_error"] = "%s: %s" % (type(e).__name__, e)
            gc.collect()  # the served tables and the result pool leave HBM
            # config 3 and config 4 legs; a failure here is reported, not fatal to the line
            try:
                table = Table(R)
                res["extra"]["table_1e8x4_bindex_build_ms"] = table.build_ms
                res["extra"]["conjunction"] = bench_conj(R, table)
                res["extra"]["batch"] = bench_batch(R, table)
                del table
            except Exception as e:  # noqa: BLE001
                res["extra"]["error"] = "%s: %s" % (type(e).__name__, e)
            gc.collect()
            try:
                res["extra"]["aux"] = bench_aux(R)
            except Exception as e:  # noqa: BLE001
                res["extra"]["aux_error"] = "%s: %s" % (type(e).__name__, e)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
