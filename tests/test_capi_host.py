"""Host-side checks of the C-ABI library (no GPU needed): it loads, exports every symbol the
headers in include/ declare, and its host logic (predicate normalisation, seeded
generators) is right."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

import rtscan_amd as R
from _util import MAXU, closed_form

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for path in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(path).read()
        # C++-only inline wrappers over the C ABI are not library symbols
        src = re.sub(r"#if defined\(__cplusplus\) && !defined\(\w+\).*?#endif /\* __cplusplus[^\n]*",
                     " ", src, flags=re.S)
        src = re.sub(r"/\*.*?\*/", " ", src, flags=re.S)
        src = re.sub(r"//[^\n]*", " ", src)
        src = "\n".join(l for l in src.splitlines()
                        if not l.strip().startswith("#") and l.strip() not in ('extern "C" {', "}"))
        # drop struct / enum bodies
        while True:
            new = re.sub(r"\{[^{}]*\}", " ", src)
            if new == src:
                break
            src = new
        for stmt in src.split(";"):
            stmt = " ".join(stmt.split())
            if "(" not in stmt or stmt.split(" ")[0] in ("typedef", "extern", "inline"):
                continue
            m = re.search(r"(\w+)\s*\(", stmt)
            if m:
                names.add(m.group(1))
    return names


def test_headers_declare_the_reference_surface():
    names = declared_functions()
    for ref_name in ["GPUbitAndWithCuda", "GPUbitOrWithCuda", "GPUbitCopyWithCuda",
                     "GPUbitCopyNegationWithCuda", "GPUbitCopySIMDWithCuda",
                     "GPUbitMemsetWithCuda", "CPUbitAndWithCuda", "refineWithCuda", "raw_scan"]:
        assert ref_name in names


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(R.LIB_PATH)
    missing = [n for n in sorted(declared_functions()) if not hasattr(L, n)]
    assert not missing, missing


def test_python_front_end_covers_every_declared_symbol():
    missing = sorted(declared_functions() - set(R.SIGNATURES))
    assert not missing, missing
    R.lib()  # binds every signature


EDGES = [0, 1, 2, 1000, 0x7FFFFFFF, 0x80000000, MAXU - 2, MAXU - 1, MAXU]


@pytest.mark.parametrize("op", [R.EQ, R.LT, R.GT, R.LE, R.GE, R.BT])
def test_predicate_interval_equals_reference_predicate(op):
    """The closed interval selects exactly the rows raw_scan's predicate selects
    (bindex.cpp:1857-1874), including at 0 and 2^32-1 and for swapped bt bounds."""
    xs = np.array(sorted(set(EDGES + [e + d for e in EDGES for d in (-1, 1) if 0 <= e + d <= MAXU])),
                  np.uint32)
    for t1 in EDGES:
        for t2 in (EDGES if op == R.BT else [0]):
            a, b = (min(t1, t2), max(t1, t2)) if op == R.BT else (t1, t2)
            want = closed_form(xs, op, a, b)
            iv = R.predicate_interval(op, t1, t2)
            if iv is None:
                assert not want.any()
                # empty means no uint32 at all qualifies
                if op == R.LT:
                    assert t1 == 0
                elif op == R.GT:
                    assert t1 == MAXU
                elif op == R.BT:
                    assert max(t1, t2) <= min(t1, t2) + 1
                else:
                    raise AssertionError("le/ge/eq are never empty")
            else:
                lo, hi = iv
                assert lo <= hi
                assert np.array_equal((xs >= lo) & (xs <= hi), want), (op, t1, t2)


def test_generator_is_deterministic_across_thread_counts():
    for dist, p0, p1 in ((R.DIST_UNIFORM, 0, 0), (R.DIST_NORMAL, 2.0 ** 31, 2.0 ** 28),
                         (R.DIST_ZIPF, 1.1, 2.0 ** 20)):
        a = R.generate_column(dist, 42, 3, 100003, p0, p1, threads=1)
        b = R.generate_column(dist, 42, 3, 100003, p0, p1, threads=7)
        assert np.array_equal(a, b)
        c = R.generate_column(dist, 43, 3, 1000, p0, p1)
        assert not np.array_equal(a[:1000], c)


def test_generator_distributions():
    n = 400000
    u = R.generate_column(R.DIST_UNIFORM, 1, 0, n)
    assert abs(u.astype(np.float64).mean() / 2 ** 32 - 0.5) < 0.005
    assert abs((u < 2 ** 31).mean() - 0.5) < 0.005
    g = R.generate_column(R.DIST_NORMAL, 1, 1, n, 2.0 ** 31, 2.0 ** 28).astype(np.float64)
    assert abs(g.mean() - 2 ** 31) < 0.01 * 2 ** 28
    assert abs(g.std() / 2 ** 28 - 1) < 0.01
    for s in (1.1, 1.5):
        z = R.generate_column(R.DIST_ZIPF, 1, 2, n, s, 2.0 ** 20)
        vals, cnt = np.unique(z, return_counts=True)
        assert len(vals) <= 2 ** 20
        h = np.sum(np.arange(1, 2 ** 20 + 1, dtype=np.float64) ** -s)
        p_top = 1.0 / h
        assert abs(cnt.max() / n - p_top) < 0.01, (s, cnt.max() / n, p_top)


def test_generator_rejects_bad_parameters():
    with pytest.raises(ValueError):
        R.generate_column(R.DIST_ZIPF, 1, 0, 10, -1.0, 100)
    with pytest.raises(ValueError):
        R.generate_column(7, 1, 0, 10)


def test_query_driver_is_built_and_prints_usage():
    """The reference-style query driver links against the library and parses its command
    line without a GPU (-h prints the usage and exits 0; a bad column count exits 2)."""
    import subprocess
    drv = os.path.join(ROOT, "rtscan_amd", "rts_bindex_driver")
    assert os.path.isfile(drv)
    p = subprocess.run([drv, "-h"], capture_output=True, text=True, timeout=30)
    assert p.returncode == 0 and "-p <scan-cmd-file>" in p.stdout
    p = subprocess.run([drv, "-b", "9", "-n", "1000"], capture_output=True, text=True, timeout=30)
    assert p.returncode == 2


HIP_ERROR_INVALID_VALUE = 1


def test_entry_points_reject_bad_arguments_before_any_device_work():
    """Argument errors come back as hipErrorInvalidValue with a message, before the entry
    point touches the device (so this runs without a GPU)."""
    L = R.lib()
    # snapshot that does not exist
    from rtscan_amd.bindex_struct import BinDex
    bx = BinDex()
    assert L.rts_bindex_load(ctypes.byref(bx), b"/nonexistent/dir/x.rtsbdx", None) == HIP_ERROR_INVALID_VALUE
    assert b"rts_bindex_load" in L.rts_last_error()
    # batch plan: bad shape, no columns
    cfg = R.BatchConfig()
    L.rts_batch_default_config(ctypes.byref(cfg))
    plan = ctypes.c_void_p()
    cols = (ctypes.c_void_p * 1)(None)
    for bad in ((300, 16, 2, 1), (1024, 0, 2, 1), (1024, 16, 0, 1), (1024, 16, 2, 7)):
        c = R.BatchConfig(*bad)
        assert L.rts_batch_plan_create(ctypes.byref(plan), cols, None, 1, 1000, 4, ctypes.byref(c),
                                       None) == HIP_ERROR_INVALID_VALUE, bad
    assert L.rts_batch_plan_create(ctypes.byref(plan), cols, None, 0, 1000, 4, ctypes.byref(cfg),
                                   None) == HIP_ERROR_INVALID_VALUE
    # server: zero batch size / streams
    sc = R.ServerConfig()
    L.rts_server_default_config(ctypes.byref(sc))
    sc.max_batch = 0
    srv = ctypes.c_void_p()
    assert L.rts_server_create(ctypes.byref(srv), ctypes.byref(sc)) == HIP_ERROR_INVALID_VALUE
    # checker and encoders: missing pointers / sizes
    assert L.rts_bitmap_compare(None, None, 10, None, None) == HIP_ERROR_INVALID_VALUE
    d = ctypes.c_ulonglong()
    assert L.rts_bitmap_digest(None, 10, ctypes.byref(d), None) == HIP_ERROR_INVALID_VALUE
    m = ctypes.c_int64()
    assert L.rts_encode_column(None, 10, 0, 10, None, None, None, None, None, ctypes.byref(m),
                               None) == HIP_ERROR_INVALID_VALUE
    assert L.rts_scan_range(None, 10, 0, 5, None, None, None) == HIP_ERROR_INVALID_VALUE
    # many ranges per pass: no column, no predicates, bitmaps overlapping (stride < words)
    lo = (ctypes.c_uint32 * 2)(1, 2)
    hi = (ctypes.c_uint32 * 2)(5, 6)
    fake = ctypes.c_void_p(256)  # never dereferenced: the checks come first
    assert L.rts_scan_ranges(None, 1000, lo, hi, 2, fake, 32, None, None) == HIP_ERROR_INVALID_VALUE
    assert L.rts_scan_ranges(fake, 1000, lo, hi, 0, fake, 32, None, None) == HIP_ERROR_INVALID_VALUE
    assert L.rts_scan_ranges(fake, 1000, None, hi, 2, fake, 32, None, None) == HIP_ERROR_INVALID_VALUE
    assert L.rts_scan_ranges(fake, 1000, lo, hi, 2, fake, 31, None, None) == HIP_ERROR_INVALID_VALUE
    assert b"rts_scan_ranges" in L.rts_last_error()


def test_batch_default_config_is_the_documented_shape():
    """The default launch shape (include/rts_batch.h, used when a plan gets no config):
    2048-word tiles, up to 32 queries per work item, 4 workgroups per CU, locality order,
    256-thread workgroups, the plan's own index set, no direct form for indexed columns;
    shapes outside the kernels' instantiations, negative weights and bin counts no index set
    has are refused."""
    cfg = R.BatchConfig()
    R.lib().rts_batch_default_config(ctypes.byref(cfg))
    assert cfg.astuple() == (2048, 32, 4, 1, 256, 0, 0)
    L = R.lib()
    plan = ctypes.c_void_p()
    cols = (ctypes.c_void_p * 1)(None)
    # invalid shapes are refused before any device work (valid ones would allocate)
    for bad in ((2048, 32, 4, 1, 128, 0, 0), (4096, 32, 4, 1, 256, 0, 0), (2048, 65, 4, 1, 256, 0, 0),
                (2048, 32, 4, 1, 256, 0, -1), (2048, 32, 4, 1, 256, 256, 0)):
        c = R.BatchConfig(*bad)
        assert L.rts_batch_plan_create(ctypes.byref(plan), cols, None, 1, 1000, 4, ctypes.byref(c),
                                       None) == HIP_ERROR_INVALID_VALUE, bad


def test_reference_shaped_caller_compiles_and_runs_on_host(tmp_path):
    """A translation unit shaped like the reference's drivers (bindex_cuda.cpp:723-770,
    bindex.cpp:1365-1411: area->blocks[b]->pos / val / length, in_which_area,
    in_which_block, on_which_pos) compiles against include/rts_bindex.h unchanged, links to
    the engine, and its host-only path (a descriptor built by hand, the library's lookups)
    answers `x < c` exactly."""
    import subprocess
    exe = str(tmp_path / "ref_shaped_lt")
    src = os.path.join(ROOT, "tests", "ref_caller", "ref_shaped_lt.cpp")
    libdir = os.path.dirname(R.LIB_PATH)
    subprocess.run(["g++", "-O1", "-std=c++17", "-Wall", "-Werror", "-D__HIP_PLATFORM_AMD__",
                    "-I/opt/rocm/include", "-I" + os.path.join(ROOT, "include"), "-o", exe, src,
                    "-L" + libdir, "-lrtscan_hip", "-L/opt/rocm/lib", "-lamdhip64",
                    "-Wl,-rpath," + libdir, "-Wl,-rpath,/opt/rocm/lib"], check=True)
    r = subprocess.run([exe, "host"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert r.stdout.startswith("ok ")


def test_snapshot_host_section_damage_is_refused_before_device_work(tmp_path):
    """rts_bindex_load checks the header and the host arrays' structure (with their digest)
    before it touches the device: each of these single-field damages is refused with
    hipErrorInvalidValue, even without a GPU."""
    import snapshot_format as S
    from rtscan_amd.bindex_struct import BinDex
    L = R.lib()
    f0, a0 = S.synthetic()
    cases = {}
    f = dict(f0); f["magic"] = b"RTSBDX\x00\x01"; cases["old format"] = (f, a0)
    f = dict(f0); f["bins"] = 200; cases["bins not 128 * sub"] = (f, a0)
    f = dict(f0); f["fv_stride"] += 32; cases["stride not 64-word aligned"] = (f, a0)
    f = dict(f0); f["nb"] = f0["n"] + 1; cases["more list entries than rows"] = (f, a0)
    f = dict(f0); f["ntiles"] += 1; cases["tile count"] = (f, a0)
    a = {k: v.copy() for k, v in a0.items()}; a["area_start"][5] += 1; cases["areas do not tile"] = (f0, a)
    a = {k: v.copy() for k, v in a0.items()}; a["area_len"][0] = 0; a["area_len"][1] += 4000
    cases["empty area"] = (f0, a)
    a = {k: v.copy() for k, v in a0.items()}; a["starts"][7] = 0; a["estarts"][7] = 0
    cases["starts decrease"] = (f0, a)
    a = {k: v.copy() for k, v in a0.items()}; a["boff"][3] = 9; cases["list offsets past nb"] = (f0, a)
    f = dict(f0); f["nblocks"] += 1
    a = {k: v.copy() for k, v in a0.items()}
    a["block_first"] = np.zeros(f["nblocks"], np.uint32); a["block_last"] = a["block_first"].copy()
    cases["block count"] = (f, a)
    for name, (f, a) in cases.items():
        p = tmp_path / "x.rtsbdx"
        p.write_bytes(S.build(f, a))
        bx = BinDex()
        assert L.rts_bindex_load(ctypes.byref(bx), str(p).encode(), None) == HIP_ERROR_INVALID_VALUE, name
    # the undamaged file passes these checks (without a GPU it then fails at the first
    # device call, with another error)
    if L.rts_device_count() < 1:
        for sub in (1, 4):
            f, a = S.synthetic(sub=sub)
            p = tmp_path / "ok.rtsbdx"
            p.write_bytes(S.build(f, a))
            bx = BinDex()
            assert L.rts_bindex_load(ctypes.byref(bx), str(p).encode(), None) not in (0, HIP_ERROR_INVALID_VALUE)
    # a damaged host digest alone is refused too
    good = bytearray(S.build(f0, a0))
    good[S.HEADER.size + 3] ^= 1
    p = tmp_path / "d.rtsbdx"
    p.write_bytes(bytes(good))
    bx = BinDex()
    assert L.rts_bindex_load(ctypes.byref(bx), str(p).encode(), None) == HIP_ERROR_INVALID_VALUE
