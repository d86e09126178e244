"""GPU parity of the BinDex path (kernel b, fused with the multi-column AND of kernel c) and
of the reference's BinDex call surface, through the C ABI, against the CPU oracle."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from _util import MAXU, closed_form, interval_mask, pack_rows

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def upload(eng, arr):
    return eng.DeviceBuffer.from_numpy(np.ascontiguousarray(arr))


def datasets(eng, n):
    rng = np.random.default_rng(n)
    out = {
        "uniform": eng.generate_column(eng.DIST_UNIFORM, 3, 0, n),
        "normal": eng.generate_column(eng.DIST_NORMAL, 3, 1, n, 2.0 ** 31, 2.0 ** 28),
        "zipf11": eng.generate_column(eng.DIST_ZIPF, 3, 2, n, 1.1, 2.0 ** 20),
        "zipf15": eng.generate_column(eng.DIST_ZIPF, 3, 3, n, 1.5, 2.0 ** 20),
        "small_range": rng.integers(0, 50, n, dtype=np.uint32),
    }
    heavy = rng.integers(0, 2 ** 32, n, dtype=np.uint32)
    heavy[rng.choice(n, 3 * (n // 128) + 11, replace=False)] = 0x9000_0000
    out["split_heavy"] = heavy
    return out


def query_bounds(x, sv, rng, count=40):
    """Closed ranges that hit the edge cases: bin starts and their neighbours, values
    present in the data, the data min/max, 0 and 2^32-1, random ranges."""
    u = np.unique(x)
    pts = [0, 1, MAXU - 1, MAXU, int(u[0]), int(u[-1])]
    pts += [int(v) for v in sv[[0, 1, 2, 63, 64, 126, 127]]]
    pts += [int(v) for v in rng.choice(u, min(8, len(u)))]
    pts = sorted({(p + d) & MAXU for p in pts for d in (-1, 0, 1)})
    out = [(0, MAXU), (0, 0), (MAXU, MAXU)]
    for _ in range(count):
        a, b = sorted(int(v) for v in rng.choice(pts, 2))
        out.append((a, b))
    for _ in range(8):
        a, b = sorted(int(v) for v in rng.integers(0, 2 ** 32, 2))
        out.append((a, b))
    # both bounds inside one bin (a single pass over that bin's refine list)
    for k in (0, 5, 64, 120, 126):
        s0, s1 = int(sv[k]), int(sv[k + 1])
        if s1 - s0 > 3:
            out += [(s0 + 1, s1 - 1), (s0 + 1, s0 + 1), (s0 + 2, s1 - 2)]
    return out


@pytest.mark.parametrize("n", [128, 1000, 70001, 1 << 20])
def test_build_matches_oracle(eng, n):
    """Area starts, area counts, sorted arrays, data min/max and all 127 filter vectors
    equal the oracle's init_bindex restatement (itself pinned to the compiled reference)."""
    for name, x in datasets(eng, n).items():
        dx = upload(eng, x)
        ix = eng.BinDexIndex(dx, n)
        ob = O.Bindex(x)
        assert np.array_equal(ix.start_values(), ob.start_values), name
        assert np.array_equal(ix.area_counts(), ob.area_counts), name
        sv, sp = ix.sorted()
        assert np.array_equal(sv, ob.sorted_val), name
        assert np.array_equal(sp, ob.sorted_pos), name
        assert ix.s.data_min == int(x.min()) and ix.s.data_max == int(x.max())
        fv = ob.fvs()
        for k in range(127):
            assert np.array_equal(ix.filter_vector(k), fv[k]), (name, k)
        # areas / blocks drive in_which_block: compare both lookups with the oracle
        L = eng.lib()
        for c in [int(v) for v in x[:6]] + [0, MAXU] + [int(v) for v in ob.start_values[[5, 77]]]:
            a = L.in_which_area(ix.ptr, c)
            assert a == ob.in_which_area(c), (name, c)
            if a >= 0:
                area_ptr = ctypes.cast(ix.s.areas[a], ctypes.c_void_p).value
                assert L.in_which_block(area_ptr, c) == ob.in_which_block(a, c), (name, c)
        ix.free()


@pytest.mark.parametrize("n", [1000, 70001, 1 << 20, 1000003])
def test_single_column_query_matches_raw_scan(eng, n):
    rng = np.random.default_rng(n + 1)
    out = eng.DeviceBuffer(4 * ((n + 31) // 32))
    cnt = eng.DeviceBuffer(8)
    for name, x in datasets(eng, n).items():
        dx = upload(eng, x)
        ix = eng.BinDexIndex(dx, n)
        for lo, hi in query_bounds(x, ix.start_values(), rng):
            out.fill(0xCD)
            eng.bindex_query([ix], [lo], [hi], out, cnt)
            want = pack_rows(interval_mask(x, lo, hi))
            got = out.to_numpy(np.uint32, (n + 31) // 32)
            assert np.array_equal(got, want), (name, lo, hi)
            assert int(cnt.to_numpy(np.uint64, 1)[0]) == int(interval_mask(x, lo, hi).sum())
        ix.free()


def engine_bin_starts(ix, sv, bins):
    """The index's engine bin starts: area a (sorted slots [s0, s0 + len)) cut into
    bins / 128 equal-count bins, bin j starting at the value in slot s0 + j * len // sub."""
    ends = ix.area_counts()
    sub = bins // 128
    out = []
    for a in range(128):
        s0 = 0 if a == 0 else int(ends[a - 1])
        ln = int(ends[a]) - s0
        out += [int(sv[s0 + j * ln // sub]) for j in range(sub)]
    return out


@pytest.mark.parametrize("n", [128, 257, 1000, 70001, (1 << 20) + 3])
def test_one_column_counts_at_every_bin_count(eng, n):
    """A counted one-column query takes its count from the sorted column (rank(hi + 1) -
    rank(lo), a search inside each bound's bin) while the other workgroups write the
    bitmap: counts and bitmaps equal numpy at 128, 256 and 1024 bins, for bounds on, next
    to and between the engine's bin starts, at the data's ends and beyond them."""
    rng = np.random.default_rng(n + 7)
    W = (n + 31) // 32
    out = eng.DeviceBuffer(4 * W)
    cnt = eng.DeviceBuffer(8)
    sets = datasets(eng, n)
    for name in ("uniform", "zipf15", "small_range", "split_heavy"):
        x = sets[name]
        dx = upload(eng, x)
        sv = np.sort(x)
        for bins in (128, 256, 1024):
            if n < bins:
                continue
            ix = eng.BinDexIndex(dx, n, bins=bins)
            assert ix.bins == bins
            es = engine_bin_starts(ix, sv, bins)
            pts = sorted({(e + d) & MAXU for e in rng.choice(es, 12) for d in (-1, 0, 1)})
            qs = query_bounds(x, ix.start_values(), rng, count=12)
            qs += [tuple(sorted(int(v) for v in rng.choice(pts, 2))) for _ in range(16)]
            qs += [(int(sv[0]), int(sv[-1])), (int(sv[-1]), MAXU), (0, int(sv[0]))]
            for lo, hi in qs:
                out.fill(0xA5)
                cnt.fill(0xEE)
                eng.bindex_query([ix], [lo], [hi], out, cnt)
                m = interval_mask(x, lo, hi)
                assert int(cnt.to_numpy(np.uint64, 1)[0]) == int(m.sum()), (name, bins, lo, hi)
                assert np.array_equal(out.to_numpy(np.uint32, W), pack_rows(m)), (name, bins, lo, hi)
            ix.free()


def test_split_heavy_value_is_exact(eng):
    """The inputs where the reference BinDex misses rows (a value split mid-run across
    areas, queried just above it) are exact here."""
    n = 1 << 20
    rng = np.random.default_rng(5)
    x = rng.integers(0, 2 ** 32, n, dtype=np.uint32)
    c = 0x9000_0000
    x[rng.choice(n, 3 * (n // 128) + 17, replace=False)] = c
    x[x == c + 1] = c + 2
    dx = upload(eng, x)
    ix = eng.BinDexIndex(dx, n)
    assert (ix.start_values() == c).sum() >= 2
    out = eng.DeviceBuffer(4 * (n // 32))
    for lo, hi in [(0, c), (c + 1, MAXU), (c, c), (c - 5, c + 5), (0, c - 1)]:
        eng.bindex_query([ix], [lo], [hi], out)
        assert np.array_equal(out.to_numpy(np.uint32, n // 32), pack_rows(interval_mask(x, lo, hi)))
    ob = O.Bindex(x)
    ref_style = O.mask_tail(ob.scan(O.LT, c + 1), n)
    assert not np.array_equal(ref_style, pack_rows(x <= c))  # the reference's D1 defect


@pytest.mark.parametrize("ncols", [2, 3, 4, 5])
def test_multi_column_query(eng, ncols):
    """Config-3 style conjunctions (uniform, normal, Zipf 1.1, Zipf 1.5 columns) with
    per-column selectivity in {0.01, 0.1, 0.5}, endpoints from column quantiles."""
    n = 1 << 20
    kinds = [(eng.DIST_UNIFORM, 0, 0), (eng.DIST_NORMAL, 2.0 ** 31, 2.0 ** 28),
             (eng.DIST_ZIPF, 1.1, 2.0 ** 20), (eng.DIST_ZIPF, 1.5, 2.0 ** 20),
             (eng.DIST_UNIFORM, 0, 0)]
    cols = [eng.generate_column(kinds[c][0], 9, c, n, kinds[c][1], kinds[c][2])
            for c in range(ncols)]
    dcols = [upload(eng, c) for c in cols]
    idx = [eng.BinDexIndex(d, n) for d in dcols]
    sorted_cols = [np.sort(c) for c in cols]
    rng = np.random.default_rng(ncols)
    out = eng.DeviceBuffer(4 * (n // 32))
    cnt = eng.DeviceBuffer(8)
    outd = eng.DeviceBuffer(4 * (n // 32))
    for trial in range(12):
        los, his, mask = [], [], np.ones(n, bool)
        for c in range(ncols):
            sel = [0.01, 0.1, 0.5][int(rng.integers(3))]
            q0 = float(rng.random()) * (1 - sel)
            lo = int(sorted_cols[c][int(q0 * (n - 1))])
            hi = int(sorted_cols[c][int((q0 + sel) * (n - 1))])
            los.append(lo)
            his.append(hi)
            mask &= interval_mask(cols[c], lo, hi)
        eng.bindex_query(idx, los, his, out, cnt)
        want = pack_rows(mask)
        assert np.array_equal(out.to_numpy(np.uint32, n // 32), want), trial
        assert int(cnt.to_numpy(np.uint64, 1)[0]) == int(mask.sum())
        eng.conjunction_scan(dcols, los, his, n, outd)  # direct path agrees
        assert np.array_equal(outd.to_numpy(np.uint32, n // 32), want)


def test_copy_filter_vector_in_gpu(eng):
    n = 70001
    x = eng.generate_column(eng.DIST_UNIFORM, 4, 0, n)
    dx = upload(eng, x)
    ix = eng.BinDexIndex(dx, n)
    ob = O.Bindex(x)
    fv = ob.fvs()
    W = (n - 1) // 32 + 1
    out = eng.DeviceBuffer(4 * W)
    L = eng.lib()
    for k, neg, want in [(0, False, fv[0]), (63, False, fv[63]), (126, True, ~fv[126]),
                         (-1, False, np.zeros(W, np.uint32)), (-5, True, np.zeros(W, np.uint32)),
                         (127, False, np.full(W, MAXU, np.uint32)),
                         (200, True, np.full(W, MAXU, np.uint32))]:
        L.copy_filter_vector_in_GPU(ix.ptr, out.ptr, k, neg)
        assert np.array_equal(out.to_numpy(np.uint32, W), want.astype(np.uint32)), (k, neg)


@pytest.mark.parametrize("eq", [False, True])
def test_refine_area_surface(eng, eq):
    """GPURefineAreaWithCuda / GPURefineEqAreaWithCuda equal the oracle's restatement
    (block choice by in_which_area / in_which_block, exclusive bounds, set or clear)."""
    n = 1 << 20
    cols = [eng.generate_column(eng.DIST_UNIFORM, 6, c, n) for c in range(3)]
    cols[2] = eng.generate_column(eng.DIST_ZIPF, 6, 2, n, 1.1, 2.0 ** 20)
    dcols = [upload(eng, c) for c in cols]
    idx = [eng.BinDexIndex(d, n) for d in dcols]
    obs = [O.Bindex(c) for c in cols]
    arr = (ctypes.c_void_p * 3)(*[ix.ptr for ix in idx])
    rng = np.random.default_rng(7 + eq)
    W = n // 32
    dbm = eng.DeviceBuffer(4 * W)
    L = eng.lib()
    fn = L.GPURefineEqAreaWithCuda if eq else L.GPURefineAreaWithCuda
    tested = 0
    for trial in range(16):
        sel = int(rng.integers(3))
        pred = []
        for c in range(3):
            a, b = sorted(int(v) for v in rng.choice(cols[c], 2))
            if trial % 4 == 0:
                a, b = 0, MAXU
            pred += [a, b]
        if trial % 5 == 1:  # upper bound exactly an area start
            pred[2 * sel + 1] = int(obs[sel].start_values[40])
        pred = np.array(pred, np.uint32)
        for inverse in (False, True):
            base = rng.integers(0, 2 ** 32, W, dtype=np.uint32)
            want = base.copy()
            ok = O.refine_area(obs, cols, pred, sel, inverse, want, eq=eq)
            dbm_host = base.copy()
            dbm = upload(eng, dbm_host)
            rc = fn(arr, dbm.ptr, pred.ctypes.data, sel, 3, inverse)
            if not ok:
                assert rc != 0
                continue
            assert rc == 0
            assert np.array_equal(dbm.to_numpy(np.uint32, W), want), (trial, sel, inverse)
            tested += 1
    assert tested >= 16


def test_scale_1e8_bindex(eng):
    """1e8 uniform rows: the GPU build and a few BinDex queries equal the oracle raw_scan."""
    n = 100_000_000
    x = eng.generate_column(eng.DIST_UNIFORM, 2, 0, n)
    dx = upload(eng, x)
    ix = eng.BinDexIndex(dx, n)
    out = eng.DeviceBuffer(4 * (n // 32))
    cnt = eng.DeviceBuffer(8)
    sv = ix.start_values()
    for lo, hi in [(123456789, 123456789 + 429496729), (0, 2 ** 31), (int(sv[10]), int(sv[100]) - 1),
                   (int(sv[3]) + 1, int(sv[3]) + 5000000)]:
        eng.bindex_query([ix], [lo], [hi], out, cnt)
        truth = O.raw_scan(x, O.BT, lo - 1, hi + 1, threads=16) if lo > 0 else \
            O.raw_scan(x, O.LE, hi, threads=16)
        assert np.array_equal(out.to_numpy(np.uint32, n // 32), truth)
        assert int(cnt.to_numpy(np.uint64, 1)[0]) == O.popcount_rows(truth, n)


@pytest.mark.parametrize("n", [1000, 1 << 20])
def test_pos_block_views_match_oracle(eng, n):
    """areas[k]->blocks[b] (host views, after host_blocks()) and areasInGPU[k]->blocks[b]
    (device views) hold the oracle's sorted positions and values block by block, and
    on_which_pos over the engine's blocks equals the oracle's (bindex.cpp:1201-1215)."""
    L = eng.lib()
    for name, x in datasets(eng, n).items():
        dx = upload(eng, x)
        ix = eng.BinDexIndex(dx, n)
        ob = O.Bindex(x)
        assert not ix.s.areas[0].contents.blocks  # no host copies until asked for
        ix.host_blocks()
        ix.host_blocks()  # idempotent
        nums = ob.block_nums()
        starts = ob.area_start
        for k in list(range(0, 128, 9)) + [127]:
            a = ix.s.areas[k].contents
            assert a.blockNum == nums[k], (name, k)
            slot = int(starts[k])
            for b in range(a.blockNum):
                pos, val = ix.block(k, b)
                want_len = min(3276, int(a.length) - 3276 * b)
                assert len(pos) == want_len, (name, k, b)
                assert np.array_equal(val, ob.sorted_val[slot:slot + want_len]), (name, k, b)
                assert np.array_equal(pos, ob.sorted_pos[slot:slot + want_len]), (name, k, b)
                if b in (0, a.blockNum - 1):
                    gpos, gval = ix.block(k, b, gpu=True)
                    assert np.array_equal(gpos, pos) and np.array_equal(gval, val)
                for c in {int(val[0]), int(val[-1]), int(val[len(val) // 2]) + 1}:
                    pb = ctypes.cast(a.blocks[b], ctypes.c_void_p).value
                    assert L.on_which_pos(pb, c) == ob.on_which_pos(k, b, c), (name, k, b, c)
                slot += want_len
        ix.free()


def test_reference_shaped_caller_runs_on_the_engine(eng):
    """A caller written like the reference's BinDex drivers (area->blocks[b]->pos,
    in_which_block, on_which_pos, copy_filter_vector_in_GPU, host refine), compiled
    against include/ and linked to the engine, answers `x < c` exactly."""
    exe = os.path.join(ROOT, "tests", "ref_caller", "ref_shaped_lt")
    assert os.path.isfile(exe), "build the test caller first (make -C tests/ref_caller)"
    print("ref_shaped_lt gpu 1000003", flush=True)
    r = subprocess.run([exe, "gpu", "1000003"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    assert r.stdout.startswith("ok ")


@pytest.mark.parametrize("n", [(1 << 24) + 12345, (1 << 24) + 32 * 1024 * 3 + 7, (1 << 25) + 32 * 1024 * 5 + 9])
def test_wide_workgroup_queries_on_large_columns(eng, n):
    """Columns of at least 2^24 rows run the one- and two-column query kernels with
    2048-word workgroups (two directory tiles each); ragged lengths leave the last
    workgroup a partial tile or one directory tile. Single- and two-column queries equal
    numpy, at 128 and 512 bins."""
    rng = np.random.default_rng(n % 1000)
    xs = [eng.generate_column(eng.DIST_UNIFORM, 9, 0, n),
          eng.generate_column(eng.DIST_ZIPF, 9, 2, n, 1.1, 2.0 ** 20)]
    ds = [upload(eng, x) for x in xs]
    W = (n + 31) // 32
    out = eng.DeviceBuffer(4 * W)
    cnt = eng.DeviceBuffer(8)
    for bins in (128, 512):
        ixs = [eng.BinDexIndex(d, n, bins=bins) for d in ds]
        for _ in range(6):
            cs = [0] if rng.random() < 0.5 else [0, 1]
            rs = []
            for c in cs:
                s = np.sort(xs[c][rng.integers(0, n, 10000)])
                a, b = sorted(int(v) for v in rng.integers(0, s.size, 2))
                rs.append((int(s[a]), int(s[b])))
            out.fill(0x5C)
            eng.bindex_query([ixs[c] for c in cs], [r[0] for r in rs], [r[1] for r in rs], out, cnt)
            m = np.ones(n, bool)
            for c, (lo, hi) in zip(cs, rs):
                m &= interval_mask(xs[c], lo, hi)
            assert np.array_equal(out.to_numpy(np.uint32, W), pack_rows(m)), (bins, cs, rs)
            assert int(cnt.to_numpy(np.uint64, 1)[0]) == int(m.sum())
        for ix in ixs:
            ix.free()
