"""ctypes front end of the CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg import this
package: as the checker, or as the timed CPU port of the reference path.  The shipped
engine (``rtscan_amd``) never imports it.

``liboracle.so`` is the C restatement in oracle.c (each function cites the reference
file:line it follows).  ``_ref/libref_bindex_<N>.so`` are the reference's own
bindex.cpp compiled by oracle/Makefile (one per compile-time row count), present only
where the reference tree was available to build them.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

EQ, LT, GT, LE, GE, BT = 0, 1, 2, 3, 4, 5  # bindex.h:88-95
OPS = {"eq": EQ, "lt": LT, "gt": GT, "le": LE, "ge": GE, "bt": BT}
BLOCK_INIT_SIZE = 3276

_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")


class _OrcBindex(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int),
        ("K", ctypes.c_int),
        ("words", ctypes.c_int),
        ("sorted_val", ctypes.POINTER(ctypes.c_uint32)),
        ("sorted_pos", ctypes.POINTER(ctypes.c_int32)),
        ("start_values", ctypes.POINTER(ctypes.c_uint32)),
        ("area_start", ctypes.POINTER(ctypes.c_int32)),
        ("area_len", ctypes.POINTER(ctypes.c_int32)),
        ("area_counts", ctypes.POINTER(ctypes.c_int32)),
        ("fvs", ctypes.POINTER(ctypes.c_uint32)),
        ("data_min", ctypes.c_uint32),
        ("data_max", ctypes.c_uint32),
    ]


_lib = None


def build():
    """Compile liboracle.so (and the reference checker libraries when the reference tree
    is present)."""
    import subprocess

    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if os.path.isfile("/root/reference/bindex.cpp"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.isfile(path):
            build()
        L = ctypes.CDLL(path)
        L.orc_bits_num_needed.argtypes = [ctypes.c_int]
        L.orc_bits_num_needed.restype = ctypes.c_int
        L.orc_raw_scan.argtypes = [_u32p, ctypes.c_int, _u32p, ctypes.c_uint32, ctypes.c_uint32,
                                   ctypes.c_int]
        L.orc_raw_scan_mt.argtypes = L.orc_raw_scan.argtypes + [ctypes.c_int]
        L.orc_raw_scan_entry.argtypes = [_u32p, ctypes.c_int, ctypes.c_int, _u32p,
                                         ctypes.c_void_p, ctypes.c_int, _u32p, _u32p,
                                         ctypes.c_int]
        L.orc_compare_bitmap.argtypes = [_u32p, _u32p, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_int),
                                         ctypes.POINTER(ctypes.c_int),
                                         ctypes.POINTER(ctypes.c_int)]
        L.orc_popcount_rows.argtypes = [_u32p, ctypes.c_int]
        L.orc_popcount_rows.restype = ctypes.c_uint64
        L.orc_init_bindex.argtypes = [ctypes.POINTER(_OrcBindex), _u32p, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_int]
        L.orc_free_bindex.argtypes = [ctypes.POINTER(_OrcBindex)]
        for f in ("orc_in_which_area",):
            getattr(L, f).argtypes = [ctypes.POINTER(_OrcBindex), ctypes.c_uint32]
        L.orc_area_block_num.argtypes = [ctypes.POINTER(_OrcBindex), ctypes.c_int]
        L.orc_in_which_block.argtypes = [ctypes.POINTER(_OrcBindex), ctypes.c_int,
                                         ctypes.c_uint32]
        L.orc_on_which_pos.argtypes = [ctypes.POINTER(_OrcBindex), ctypes.c_int, ctypes.c_int,
                                       ctypes.c_uint32]
        L.orc_bindex_scan.argtypes = [ctypes.POINTER(_OrcBindex), _u32p, ctypes.c_int,
                                      ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int]
        L.orc_bitmap_digest.argtypes = [_u32p, ctypes.c_int64]
        L.orc_conjunction_digest.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, _u32p, _u32p,
                                             ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_conjunction_digest.restype = None
        L.orc_conjunction_scan.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, _u32p, _u32p, _u32p,
                                           ctypes.c_int]
        L.orc_conjunction_scan.restype = None
        L.orc_simd_level.restype = ctypes.c_int
        L.orc_bitmap_digest.restype = ctypes.c_uint64
        L.orc_normal_encode.argtypes = [_u32p, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32,
                                        _u32p, _i32p, _u32p, _u32p, _u32p]
        L.orc_normal_encode.restype = ctypes.c_int64
        L.orc_dispersed_map.argtypes = [_u32p, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32,
                                        _u32p, _u32p, _u32p]
        L.orc_dispersed_map.restype = ctypes.c_int64
        L.orc_encode_query.argtypes = [_u32p, _u32p, ctypes.c_int64, ctypes.c_uint32, ctypes.c_int,
                                       ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)]
        L.orc_find_key_by_value.argtypes = [_u32p, _u32p, ctypes.c_int64, ctypes.c_uint32]
        L.orc_find_key_by_value.restype = ctypes.c_uint32
        for f in ("orc_refine_area", "orc_refine_eq_area"):
            getattr(L, f).argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, _u32p,
                                      ctypes.c_int, ctypes.c_int, _u32p]
        _lib = L
    return _lib


def words_for(n):
    return 0 if n <= 0 else lib().orc_bits_num_needed(n)


def raw_scan(raw, op, t1, t2=0, bitmap=None, threads=8):
    """raw_scan (bindex.cpp:1851-1892): XOR hits into `bitmap` (zeroed when not given)."""
    raw = np.ascontiguousarray(raw, dtype=np.uint32)
    n = raw.size
    if bitmap is None:
        bitmap = np.zeros(max(words_for(n), 1), np.uint32)
    lib().orc_raw_scan_mt(raw, n, bitmap, t1 & 0xFFFFFFFF, t2 & 0xFFFFFFFF, op, threads)
    return bitmap


def raw_scan_st(raw, op, t1, t2=0):
    raw = np.ascontiguousarray(raw, dtype=np.uint32)
    bitmap = np.zeros(max(words_for(raw.size), 1), np.uint32)
    lib().orc_raw_scan(raw, raw.size, bitmap, t1 & 0xFFFFFFFF, t2 & 0xFFFFFFFF, op)
    return bitmap


def raw_scan_entry(raw, op, tl, tr=None, bitmap=None, merge=None, threads=8):
    """raw_scan_entry (bindex.cpp:1894-1957) for one column of one query."""
    raw = np.ascontiguousarray(raw, dtype=np.uint32)
    n = raw.size
    W = max(words_for(n), 1)
    tl = np.ascontiguousarray(tl, dtype=np.uint32)
    trp = None
    if tr is not None and len(tr):
        tr = np.ascontiguousarray(tr, dtype=np.uint32)
        trp = tr.ctypes.data_as(ctypes.c_void_p)
    if bitmap is None:
        bitmap = np.zeros(W, np.uint32)
    if merge is None:
        merge = bitmap
    lib().orc_raw_scan_entry(raw, n, op, tl, trp, tl.size, bitmap, merge, threads)
    return merge


def conjunction(columns, preds, threads=8):
    """The reference driver's checker for one multi-column query (bindex.cpp:2265-2289):
    raw_scan_entry per column into zeroed bitmaps, AND-merged into column 0's.
    `preds` is a list of (op, [t1...], [t2...] or None) per column."""
    n = len(columns[0])
    W = max(words_for(n), 1)
    bms = [np.zeros(W, np.uint32) for _ in columns]
    for c, (col, (op, tl, tr)) in enumerate(zip(columns, preds)):
        raw_scan_entry(col, op, tl, tr, bitmap=bms[c], merge=bms[0], threads=threads)
    return bms[0]


def conjunction_digest(columns, ranges, threads=None):
    """(digest, popcount) of the AND over `columns` of closed ranges [(lo, hi), ...] (lo > hi:
    empty), computed without the bitmap (orc_conjunction_digest)."""
    cols = [np.ascontiguousarray(c, dtype=np.uint32) for c in columns]
    n = cols[0].size
    assert all(c.size == n for c in cols) and len(ranges) == len(cols)
    ptrs = (ctypes.c_void_p * len(cols))(*[c.ctypes.data for c in cols])
    lo = np.array([r[0] & 0xFFFFFFFF for r in ranges], np.uint32)
    hi = np.array([r[1] & 0xFFFFFFFF for r in ranges], np.uint32)
    d, c = ctypes.c_uint64(), ctypes.c_uint64()
    lib().orc_conjunction_digest(ptrs, len(cols), n, lo, hi, threads or min(16, os.cpu_count() or 1),
                                 ctypes.byref(d), ctypes.byref(c))
    return d.value, c.value


def conjunction_scan(columns, ranges, threads=None, bitmap=None):
    """Bitmap of the AND over `columns` of closed ranges [(lo, hi), ...]: the vectorised,
    word-parallel port of the checker's conjunction (orc_conjunction_scan)."""
    cols = [np.ascontiguousarray(c, dtype=np.uint32) for c in columns]
    n = cols[0].size
    ptrs = (ctypes.c_void_p * len(cols))(*[c.ctypes.data for c in cols])
    lo = np.array([r[0] & 0xFFFFFFFF for r in ranges], np.uint32)
    hi = np.array([r[1] & 0xFFFFFFFF for r in ranges], np.uint32)
    if bitmap is None:
        bitmap = np.empty(max(words_for(n), 1), np.uint32)
    lib().orc_conjunction_scan(ptrs, len(cols), n, lo, hi, bitmap, threads or min(16, os.cpu_count() or 1))
    return bitmap


def simd_level():
    """2 when the host runs the AVX2 port, 0 for the scalar one."""
    return int(lib().orc_simd_level())


def compare_bitmap(truth, got, n):
    """compare_bitmap (bindex.cpp:1959-1978) -> (total_hit, true_hit, first_mismatch)."""
    a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    lib().orc_compare_bitmap(np.ascontiguousarray(truth, np.uint32),
                             np.ascontiguousarray(got, np.uint32), n,
                             ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
    return a.value, b.value, c.value


def popcount_rows(bitmap, n):
    return int(lib().orc_popcount_rows(np.ascontiguousarray(bitmap, np.uint32), n))


def mask_tail(bitmap, n):
    """Copy of `bitmap` with the bits past row n cleared."""
    b = np.array(bitmap, dtype=np.uint32, copy=True)
    W = words_for(n)
    tail = n - (W - 1) * 32
    if W and tail < 32:
        b[W - 1] &= np.uint32((0xFFFFFFFF << (32 - tail)) & 0xFFFFFFFF)
    return b[:W]


class Bindex:
    """The oracle's restatement of init_bindex (bindex.cpp:599-671)."""

    def __init__(self, data, K=128, threads=8):
        self._data = np.ascontiguousarray(data, dtype=np.uint32)
        self.s = _OrcBindex()
        rc = lib().orc_init_bindex(ctypes.byref(self.s), self._data, self._data.size, K, threads)
        if rc:
            raise ValueError("orc_init_bindex failed (n < K?)")
        self.n, self.K, self.words = self.s.n, self.s.K, self.s.words

    def __del__(self):
        if getattr(self, "s", None) is not None and _lib is not None:
            _lib.orc_free_bindex(ctypes.byref(self.s))
            self.s = None

    def _arr(self, ptr, count, dt):
        return np.ctypeslib.as_array(ptr, shape=(count,)).astype(dt, copy=True)

    @property
    def start_values(self):
        return self._arr(self.s.start_values, self.K, np.uint32)

    @property
    def area_start(self):
        return self._arr(self.s.area_start, self.K, np.int32)

    @property
    def area_counts(self):
        return self._arr(self.s.area_counts, self.K, np.int32)

    @property
    def sorted_val(self):
        return self._arr(self.s.sorted_val, self.n, np.uint32)

    @property
    def sorted_pos(self):
        return self._arr(self.s.sorted_pos, self.n, np.int32)

    def fvs(self):
        return self._arr(self.s.fvs, (self.K - 1) * self.words, np.uint32).reshape(
            self.K - 1, self.words)

    def block_nums(self):
        return np.array([lib().orc_area_block_num(ctypes.byref(self.s), k)
                         for k in range(self.K)], np.int32)

    def in_which_area(self, c):
        return lib().orc_in_which_area(ctypes.byref(self.s), c & 0xFFFFFFFF)

    def in_which_block(self, area, c):
        return lib().orc_in_which_block(ctypes.byref(self.s), area, c & 0xFFFFFFFF)

    def on_which_pos(self, area, block, c):
        return lib().orc_on_which_pos(ctypes.byref(self.s), area, block, c & 0xFFFFFFFF)

    def scan(self, op, t1, t2=0, threads=8):
        """bindex_scan_* (bindex.cpp:1365-1691); None where the reference asserts."""
        out = np.zeros(self.words, np.uint32)
        rc = lib().orc_bindex_scan(ctypes.byref(self.s), out, op, t1 & 0xFFFFFFFF,
                                   t2 & 0xFFFFFFFF, threads)
        return None if rc else out


class RefBindexLib:
    """The reference's compiled bindex.cpp (oracle/_ref), for one compile-time N."""

    def __init__(self, n):
        path = os.path.join(HERE, "_ref", "libref_bindex_%d.so" % n)
        if not os.path.isfile(path):
            raise FileNotFoundError(path)
        L = ctypes.CDLL(path)
        L.ref_init_bindex.argtypes = [_u32p, ctypes.c_int]
        L.ref_init_bindex.restype = ctypes.c_void_p
        L.ref_free_bindex.argtypes = [ctypes.c_void_p]
        L.ref_read_bindex.argtypes = [ctypes.c_void_p, _u32p, _i32p, _i32p, _u32p]
        L.ref_in_which_area.argtypes = [ctypes.c_void_p, ctypes.c_uint32]
        L.ref_in_which_block.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32]
        L.ref_on_which_pos.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_uint32]
        L.ref_bindex_scan.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32,
                                      ctypes.c_uint32, _u32p]
        L.ref_raw_scan.argtypes = [_u32p, ctypes.c_int, _u32p, ctypes.c_uint32,
                                   ctypes.c_uint32, ctypes.c_int]
        L.ref_query_slot11.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_void_p]
        L.ref_query_slot11.restype = ctypes.c_double
        L.ref_thread_num.restype = ctypes.c_int
        self.L = L
        self.n = L.ref_compiled_n()
        self.K = L.ref_compiled_k()

    def raw_scan(self, raw, op, t1, t2=0):
        raw = np.ascontiguousarray(raw, np.uint32)
        bm = np.zeros(max(words_for(raw.size), 1), np.uint32)
        self.L.ref_raw_scan(raw, raw.size, bm, t1 & 0xFFFFFFFF, t2 & 0xFFFFFFFF, op)
        return bm

    def bindex(self, data):
        return _RefBindex(self, data)

    @property
    def threads(self):
        return int(self.L.ref_thread_num())

    def query_slot11(self, bindexes, preds, bitmaps):
        """The reference driver's timed region (bindex.cpp:2190-2262) for one query: the
        reference BinDex scan of every column (preds: (op, t1, t2) per column) into
        bitmaps[c], then the multithreaded AND merge into bitmaps[0].  Returns its wall
        time in ms.  bitmaps: 32-byte aligned uint32 arrays of words_for(n) words."""
        k = len(bindexes)
        hs = (ctypes.c_void_p * k)(*[b.h for b in bindexes])
        ops = np.array([p[0] for p in preds], np.int32)
        t1 = np.array([p[1] & 0xFFFFFFFF for p in preds], np.uint32)
        t2 = np.array([p[2] & 0xFFFFFFFF for p in preds], np.uint32)
        bms = (ctypes.c_void_p * k)(*[b.ctypes.data for b in bitmaps])
        return float(self.L.ref_query_slot11(hs, k, ops.ctypes.data, t1.ctypes.data, t2.ctypes.data, bms))


class _RefBindex:
    def __init__(self, owner, data):
        self.o = owner
        self._data = np.ascontiguousarray(data, np.uint32)
        assert self._data.size == owner.n
        self.h = owner.L.ref_init_bindex(self._data, self._data.size)
        self.words = words_for(owner.n)

    def __del__(self):
        if getattr(self, "h", None):
            self.o.L.ref_free_bindex(self.h)
            self.h = None

    def read(self):
        K = self.o.K
        sv = np.zeros(K, np.uint32)
        cnt = np.zeros(K, np.int32)
        bn = np.zeros(K, np.int32)
        fv = np.zeros((K - 1) * self.words, np.uint32)
        self.o.L.ref_read_bindex(self.h, sv, cnt, bn, fv)
        return sv, cnt, bn, fv.reshape(K - 1, self.words)

    def in_which_area(self, c):
        return self.o.L.ref_in_which_area(self.h, c & 0xFFFFFFFF)

    def in_which_block(self, area, c):
        return self.o.L.ref_in_which_block(self.h, area, c & 0xFFFFFFFF)

    def on_which_pos(self, area, block, c):
        return self.o.L.ref_on_which_pos(self.h, area, block, c & 0xFFFFFFFF)

    def scan(self, op, t1, t2=0):
        out = np.zeros(self.words, np.uint32)
        self.o.L.ref_bindex_scan(self.h, op, t1 & 0xFFFFFFFF, t2 & 0xFFFFFFFF, out)
        return out


def refine_area(bindexes, columns, predicate, sel, inverse, bitmap, eq=False):
    """GPURefineAreaWithCuda / GPURefineEqAreaWithCuda restated (cuda_and.cu:510-812) over
    oracle Bindex objects; updates `bitmap` in place.  Returns False where the reference
    would read out of bounds."""
    k = len(bindexes)
    bx = (ctypes.POINTER(_OrcBindex) * k)(*[ctypes.pointer(b.s) for b in bindexes])
    cols = [np.ascontiguousarray(c, np.uint32) for c in columns]
    rp = (ctypes.c_void_p * k)(*[c.ctypes.data for c in cols])
    pred = np.ascontiguousarray(predicate, np.uint32)
    f = lib().orc_refine_eq_area if eq else lib().orc_refine_area
    return f(bx, rp, k, pred, sel, int(bool(inverse)), bitmap) == 0


def normal_encode(data, encode_min, encode_max):
    """normalEncode, 7-argument form (remap.cpp:110-203), for one column.  Returns (codes,
    sorted_pos, sorted_codes, map_values, map_starts)."""
    data = np.ascontiguousarray(data, np.uint32)
    n = data.size
    codes = np.zeros(n, np.uint32)
    pos = np.zeros(n, np.int32)
    sc = np.zeros(n, np.uint32)
    mv = np.zeros(n, np.uint32)
    ms = np.zeros(n, np.uint32)
    m = lib().orc_normal_encode(data, n, encode_min, encode_max, codes, pos, sc, mv, ms)
    return codes, pos, sc, mv[:m].copy(), ms[:m].copy()


def dispersed_map(data, encode_min, encode_max):
    """Encode map of normalEncode's 5-argument form (remap.cpp:205-241): (values, starts,
    steps)."""
    data = np.ascontiguousarray(data, np.uint32)
    n = data.size
    mv = np.zeros(n, np.uint32)
    ms = np.zeros(n, np.uint32)
    st = np.zeros(n, np.uint32)
    m = lib().orc_dispersed_map(data, n, encode_min, encode_max, mv, ms, st)
    return mv[:m].copy(), ms[:m].copy(), st[:m].copy()


def encode_query(map_values, map_starts, q, op, sentinel):
    """encodeQuery (remap.cpp:243-272); None for eq / bt."""
    out = ctypes.c_uint32()
    rc = lib().orc_encode_query(np.ascontiguousarray(map_values, np.uint32),
                                np.ascontiguousarray(map_starts, np.uint32), len(map_values),
                                q & 0xFFFFFFFF, op, sentinel & 0xFFFFFFFF, ctypes.byref(out))
    return None if rc else out.value


def find_key_by_value(map_values, map_starts, val):
    """findKeyByValue (remap.cpp:274-283)."""
    return int(lib().orc_find_key_by_value(np.ascontiguousarray(map_values, np.uint32),
                                           np.ascontiguousarray(map_starts, np.uint32),
                                           len(map_values), val & 0xFFFFFFFF))


def bitmap_digest(bitmap, n):
    """The engine's 64-bit bitmap digest, restated (not a reference function)."""
    return int(lib().orc_bitmap_digest(np.ascontiguousarray(bitmap, np.uint32), n))
