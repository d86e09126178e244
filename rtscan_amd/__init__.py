"""Python front end of the MI355X range-predicate scan engine.

A thin ctypes layer over ``librtscan_hip.so`` (the C ABI declared in include/*.h): the
work runs in the library's HIP kernels.  There is no fallback: when the library is not
built, or no GPU is present, the entry points raise.

Bitmaps are 1 bit per row, MSB-first in uint32 words (row i is bit 1 << (31 - i % 32)
of word i // 32), as in the reference (bindex.cpp:269).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librtscan_hip.so")

EQ, LT, GT, LE, GE, BT = 0, 1, 2, 3, 4, 5  # enum OPERATOR, bindex.h:88-95
OPS = {"eq": EQ, "lt": LT, "gt": GT, "le": LE, "ge": GE, "bt": BT}
DIST_UNIFORM, DIST_NORMAL, DIST_ZIPF = 0, 1, 2

_vp = ctypes.c_void_p
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64
_int = ctypes.c_int
_err = ctypes.c_int  # hipError_t

# name -> (restype, argtypes) for every function the headers in include/ declare
SIGNATURES = {
    # cuda_and.cu surface
    "GPUbitAndWithCuda": (_err, [_vp, _vp, ctypes.c_uint]),
    "GPUbitOrWithCuda": (_err, [_vp, _vp, ctypes.c_uint]),
    "GPUbitCopyWithCuda": (_err, [_vp, _vp, ctypes.c_uint]),
    "GPUbitCopyNegationWithCuda": (_err, [_vp, _vp, ctypes.c_uint]),
    "GPUbitCopySIMDWithCuda": (_err, [_vp, _vp, _vp, ctypes.c_uint]),
    "GPUbitMemsetWithCuda": (_err, [_vp, _u32, ctypes.c_uint]),
    "CPUbitAndWithCuda": (_err, [_vp, _vp, ctypes.c_uint]),
    "refineWithCuda": (_err, [_vp, ctypes.c_uint]),
    # bindex.h surface
    "raw_scan": (None, [_vp, _vp, _u32, _u32, _int, _vp, _vp]),
    # helpers
    "rts_malloc": (_err, [ctypes.POINTER(_vp), ctypes.c_size_t]),
    "rts_free": (_err, [_vp]),
    "rts_memcpy_htod": (_err, [_vp, _vp, ctypes.c_size_t]),
    "rts_memcpy_dtoh": (_err, [_vp, _vp, ctypes.c_size_t]),
    "rts_memcpy_dtod": (_err, [_vp, _vp, ctypes.c_size_t]),
    "rts_memset": (_err, [_vp, _int, ctypes.c_size_t]),
    "rts_host_alloc": (_err, [ctypes.POINTER(_vp), ctypes.c_size_t]),
    "rts_host_free": (_err, [_vp]),
    "rts_memcpy_dtoh_async": (_err, [_vp, _vp, ctypes.c_size_t, _vp]),
    "rts_memcpy_htod_async": (_err, [_vp, _vp, ctypes.c_size_t, _vp]),
    "rts_device_synchronize": (_err, []),
    "rts_stream_create": (_err, [ctypes.POINTER(_vp)]),
    "rts_stream_destroy": (_err, [_vp]),
    "rts_stream_synchronize": (_err, [_vp]),
    "rts_event_create": (_err, [ctypes.POINTER(_vp)]),
    "rts_event_destroy": (_err, [_vp]),
    "rts_event_record": (_err, [_vp, _vp]),
    "rts_event_synchronize": (_err, [_vp]),
    "rts_event_elapsed_ms": (_err, [ctypes.POINTER(ctypes.c_float), _vp, _vp]),
    "rts_device_count": (_int, []),
    "rts_device_info": (_err, [ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(_int),
                               ctypes.POINTER(ctypes.c_size_t)]),
    "rts_last_error": (ctypes.c_char_p, []),
    "rts_thread_scratch_buffers": (_int, []),
    "rts_probe_read": (_err, [_vp, ctypes.c_size_t, _int, _int, _vp, _vp]),
    "rts_device_mem_info": (_err, [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]),
    # scans
    "rts_predicate_interval": (_int, [_int, _u32, _u32, ctypes.POINTER(_u32),
                                      ctypes.POINTER(_u32)]),
    "rts_scan_range": (_err, [_vp, _u64, _u32, _u32, _vp, _vp, _vp]),
    "rts_scan_range_xor": (_err, [_vp, _u64, _u32, _u32, _vp, _vp]),
    "rts_scan_ranges": (_err, [_vp, _u64, _vp, _vp, _int, _vp, _u64, _vp, _vp]),
    "rts_conjunction_scan": (_err, [_vp, _vp, _vp, _int, _u64, _vp, _vp, _vp]),
    "rts_conjunction_scan_noalloc": (_err, [_vp, _vp, _vp, _int, _u64, _vp, _vp, _vp]),
    "rts_bitmap_and_popcount": (_err, [_vp, _int, _u64, _vp, _vp, _vp]),
    "rts_bitmap_popcount": (_err, [_vp, _u64, _vp, _vp]),
    # BinDex surface
    "copy_filter_vector_in_GPU": (None, [_vp, _vp, _int, ctypes.c_bool]),
    "GPURefineAreaWithCuda": (_err, [_vp, _vp, _vp, _int, _int, ctypes.c_bool]),
    "GPURefineEqAreaWithCuda": (_err, [_vp, _vp, _vp, _int, _int, ctypes.c_bool]),
    "in_which_area": (_int, [_vp, _u32]),
    "in_which_block": (_int, [_vp, _u32]),
    "on_which_pos": (_int, [_vp, _u32]),
    "rts_bindex_host_blocks": (_err, [_vp]),
    "rts_bindex_build": (_err, [_vp, _vp, _int, _vp]),
    "rts_bindex_build_host": (_err, [_vp, _vp, _int]),
    "rts_bindex_build_bins": (_err, [_vp, _vp, _int, _int, _vp]),
    "rts_bindex_bins": (_int, [_vp]),
    "rts_bindex_free": (None, [_vp]),
    "rts_bindex_query": (_err, [_vp, _vp, _vp, _int, _vp, _vp, _vp]),
    "rts_bindex_export_sorted": (_err, [_vp, _vp, _vp]),
    "rts_bindex_refine_entries": (ctypes.c_longlong, [_vp]),
    "rts_bindex_device_bytes": (ctypes.c_longlong, [_vp]),
    "rts_bindex_save": (_err, [_vp, ctypes.c_char_p]),
    "rts_bindex_load": (_err, [_vp, ctypes.c_char_p, _vp]),
    # query / data files
    "rts_query_file_parse": (_int, [ctypes.c_char_p, ctypes.c_size_t, _int, ctypes.POINTER(_vp),
                                    ctypes.POINTER(_int)]),
    "rts_query_file_load": (_int, [ctypes.c_char_p, _int, ctypes.POINTER(_vp),
                                   ctypes.POINTER(_int)]),
    "rts_query_file_count": (_int, [_vp]),
    "rts_query_file_columns": (_int, [_vp]),
    "rts_query_file_get": (_vp, [_vp, _int, _int]),
    "rts_query_file_free": (None, [_vp]),
    "rts_data_file_read": (_int, [ctypes.c_char_p, _int, ctypes.c_int64, _vp]),
    "rts_data_file_write": (_int, [ctypes.c_char_p, _int, ctypes.c_int64, _vp]),
    # batched queries
    "rts_batch_default_config": (None, [_vp]),
    "rts_batch_plan_create": (_err, [ctypes.POINTER(_vp), _vp, _vp, _int, ctypes.c_int64, _int,
                                     _vp, _vp]),
    "rts_batch_plan_set_config": (_err, [_vp, _vp, _vp]),
    "rts_batch_plan_set_queries": (_err, [_vp, _vp, _int, _vp]),
    "rts_batch_plan_launch": (_err, [_vp, _vp, ctypes.c_int64, _vp, _vp]),
    "rts_batch_plan_graph_launch": (_err, [_vp, _vp, ctypes.c_int64, _vp, _vp]),
    "rts_batch_plan_stats": (None, [_vp, ctypes.POINTER(_int), ctypes.POINTER(_int),
                                    ctypes.POINTER(_int), ctypes.POINTER(ctypes.c_longlong)]),
    "rts_batch_plan_info": (None, [_vp, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(_int),
                                   ctypes.POINTER(_int)]),
    "rts_batch_plan_free": (None, [_vp]),
    "rts_batch_plan_groups": (_int, [_vp, _vp]),
    "rts_batch_plan_auto_direct": (_int, [_vp]),
    "rts_batch_plan_list_cols": (_int, [_vp]),
    "rts_batch_plan_mix_digest": (ctypes.c_ulonglong, [_vp]),
    "rts_batch_plan_add_indexes": (_err, [_vp, _vp, ctypes.POINTER(_int)]),
    "rts_batch_autotune": (_err, [_vp, _vp, ctypes.c_int64, _vp, _int, _int, ctypes.c_char_p, _vp,
                                  ctypes.POINTER(ctypes.c_float), _vp]),
    "rts_batch_plan_set_queries_out": (_err, [_vp, _vp, _vp, _int, _vp]),
    # serving
    "rts_server_default_config": (None, [_vp]),
    "rts_server_create": (_err, [ctypes.POINTER(_vp), _vp]),
    "rts_server_add_table": (_err, [_vp, _vp, _vp, _int, ctypes.c_int64, ctypes.POINTER(_int)]),
    "rts_server_start": (_err, [_vp]),
    "rts_server_submit": (_err, [_vp, _int, _vp, ctypes.POINTER(ctypes.c_uint64)]),
    "rts_server_wait": (_int, [_vp, ctypes.c_uint64, _int, ctypes.POINTER(_vp),
                               ctypes.POINTER(ctypes.c_ulonglong), ctypes.POINTER(ctypes.c_uint64)]),
    "rts_server_release": (None, [_vp, ctypes.c_uint64]),
    "rts_server_stats": (None, [_vp, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64),
                                ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(_int),
                                ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]),
    "rts_server_stop": (_err, [_vp]),
    "rts_server_destroy": (None, [_vp]),
    "rts_server_replay": (_err, [_vp, _vp, _vp, _vp, _int, _vp, _vp, ctypes.POINTER(ctypes.c_double)]),
    # value encoding (remap surface)
    "rts_encode_column": (_err, [_vp, ctypes.c_int64, _u32, _u32, _vp, _vp, _vp, _vp, _vp,
                                 ctypes.POINTER(ctypes.c_int64), _vp]),
    "rts_encode_column_dispersed": (_err, [_vp, ctypes.c_int64, _u32, _u32, _u64, _vp, _vp, _vp,
                                           ctypes.POINTER(ctypes.c_int64), _vp]),
    "rts_encode_lookup": (_err, [_vp, _vp, ctypes.c_int64, _u32, _int, _u32, ctypes.POINTER(_u32)]),
    "rts_encode_key_of": (_u32, [_vp, _vp, ctypes.c_int64, _u32]),
    # device-side checking
    "rts_bitmap_compare": (_err, [_vp, _vp, ctypes.c_int64, _vp, _vp]),
    "rts_bitmap_digest": (_err, [_vp, ctypes.c_int64, ctypes.POINTER(ctypes.c_ulonglong), _vp]),
    "compare_bitmap": (None, [_vp, _vp, _int, _vp, _int]),
    # generators
    "rts_generate_column": (_int, [_int, _u64, _int, _u64, ctypes.c_double, ctypes.c_double,
                                   _vp, _int]),
}

_lib = None


class HipError(RuntimeError):
    pass


def lib():
    """Load librtscan_hip.so; raises when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.isfile(LIB_PATH):
            raise ImportError("librtscan_hip.so is not built: run __graft_entry__.build() "
                              "or make -C rtscan_amd/csrc")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(err, what=""):
    if err:
        msg = lib().rts_last_error().decode(errors="replace")
        raise HipError("%s failed: hipError %d %s" % (what, err, msg))


def gpu_available():
    try:
        return lib().rts_device_count() > 0
    except (ImportError, OSError):
        return False


def words_for(n):
    return (n + 31) // 32


def predicate_interval(op, t1, t2=0):
    """(lo, hi) closed interval of a reference predicate, or None when empty."""
    lo, hi = _u32(), _u32()
    ok = lib().rts_predicate_interval(op, t1 & 0xFFFFFFFF, t2 & 0xFFFFFFFF, ctypes.byref(lo),
                                      ctypes.byref(hi))
    return (lo.value, hi.value) if ok else None


class DeviceBuffer:
    """Device allocation owned by Python (rts_malloc / rts_free)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = _vp()
        check(lib().rts_malloc(ctypes.byref(p), max(self.nbytes, 4)), "rts_malloc")
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, arr):
        arr = np.ascontiguousarray(arr)
        b = cls(arr.nbytes)
        if arr.nbytes:
            check(lib().rts_memcpy_htod(b.ptr, arr.ctypes.data, arr.nbytes), "htod")
        return b

    def to_numpy(self, dtype=np.uint32, count=None):
        dtype = np.dtype(dtype)
        if count is None:
            count = self.nbytes // dtype.itemsize
        out = np.empty(count, dtype)
        if out.nbytes:
            check(lib().rts_memcpy_dtoh(out.ctypes.data, self.ptr, out.nbytes), "dtoh")
        return out

    def fill(self, byte_value=0):
        check(lib().rts_memset(self.ptr, byte_value, self.nbytes), "memset")
        return self

    def free(self):
        if self.ptr:
            lib().rts_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        s = _vp()
        check(lib().rts_stream_create(ctypes.byref(s)), "stream_create")
        self.handle = s.value

    def synchronize(self):
        check(lib().rts_stream_synchronize(self.handle), "stream_synchronize")

    def destroy(self):
        if self.handle:
            check(lib().rts_stream_destroy(self.handle), "stream_destroy")
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Event:
    def __init__(self):
        e = _vp()
        check(lib().rts_event_create(ctypes.byref(e)), "event_create")
        self.handle = e.value

    def record(self, stream=None):
        check(lib().rts_event_record(self.handle, stream.handle if stream else None), "record")

    def synchronize(self):
        check(lib().rts_event_synchronize(self.handle), "event_synchronize")

    def elapsed_ms(self, later):
        ms = ctypes.c_float()
        check(lib().rts_event_elapsed_ms(ctypes.byref(ms), self.handle, later.handle), "elapsed")
        return ms.value

    def __del__(self):
        try:
            if self.handle:
                lib().rts_event_destroy(self.handle)
        except Exception:
            pass


def device_info():
    name = ctypes.create_string_buffer(128)
    cus, mem = _int(), ctypes.c_size_t()
    check(lib().rts_device_info(name, 128, ctypes.byref(cus), ctypes.byref(mem)), "device_info")
    return {"arch": name.value.decode(), "cus": cus.value, "total_mem": mem.value}


def device_free_bytes():
    f, t = ctypes.c_size_t(), ctypes.c_size_t()
    check(lib().rts_device_mem_info(ctypes.byref(f), ctypes.byref(t)), "mem_info")
    return f.value


def _sh(stream):
    return stream.handle if stream is not None else None


def scan_range(col, n, lo, hi, out, count=None, stream=None):
    """Kernel (a): bitmap of rows of device column `col` with lo <= x <= hi."""
    check(lib().rts_scan_range(col.ptr, n, lo, hi, out.ptr, count.ptr if count else None,
                               _sh(stream)), "rts_scan_range")


def scan_ranges(col, n, los, his, out, stride_words, counts=None, stream=None):
    """Many closed ranges over one column, read once per 64 predicates: bitmap p at
    out + p * stride_words (words), counts[p] (device, optional) its matching rows."""
    lo = np.ascontiguousarray(los, np.uint32)
    hi = np.ascontiguousarray(his, np.uint32)
    check(lib().rts_scan_ranges(col.ptr, n, lo.ctypes.data, hi.ctypes.data, len(lo), out.ptr,
                                stride_words, counts.ptr if counts else None, _sh(stream)),
          "rts_scan_ranges")


def scan_range_xor(col, n, lo, hi, out, stream=None):
    check(lib().rts_scan_range_xor(col.ptr, n, lo, hi, out.ptr, _sh(stream)),
          "rts_scan_range_xor")


def _ptr_array(ptrs):
    arr = (_vp * len(ptrs))(*ptrs)
    return arr


def conjunction_scan(cols, los, his, n, out, count=None, stream=None):
    """Kernel (c), direct form: AND over columns of lo_c <= x_c <= hi_c, plus popcount."""
    k = len(cols)
    pa = _ptr_array([c.ptr for c in cols])
    lo = np.ascontiguousarray(los, np.uint32)
    hi = np.ascontiguousarray(his, np.uint32)
    check(lib().rts_conjunction_scan(pa, lo.ctypes.data, hi.ctypes.data, k, n, out.ptr,
                                     count.ptr if count else None, _sh(stream)),
          "rts_conjunction_scan")


def bitmap_and_popcount(bitmaps, n, out, count=None, stream=None):
    pa = _ptr_array([b.ptr for b in bitmaps])
    check(lib().rts_bitmap_and_popcount(pa, len(bitmaps), n, out.ptr,
                                        count.ptr if count else None, _sh(stream)),
          "rts_bitmap_and_popcount")


def bitmap_popcount(bitmap, n, count, stream=None):
    check(lib().rts_bitmap_popcount(bitmap.ptr, n, count.ptr, _sh(stream)), "rts_bitmap_popcount")


def synchronize():
    check(lib().rts_device_synchronize(), "device_synchronize")


def generate_column(dist, seed, column_id, n, p0=0.0, p1=0.0, threads=None):
    """Seeded synthetic column (host numpy array) from the library's generator."""
    out = np.empty(n, np.uint32)
    if threads is None:
        threads = min(16, os.cpu_count() or 1)
    rc = lib().rts_generate_column(dist, seed, column_id, n, p0, p1, out.ctypes.data, threads)
    if rc:
        raise ValueError("rts_generate_column: bad parameters")
    return out


class _CompareResult(ctypes.Structure):
    _fields_ = [("total_hit", ctypes.c_ulonglong), ("true_hit", ctypes.c_ulonglong),
                ("first_mismatch", ctypes.c_longlong), ("truth_count", ctypes.c_ulonglong),
                ("got_count", ctypes.c_ulonglong)]


def bitmap_compare(truth, got, n, stream=None):
    """compare_bitmap semantics (bindex.cpp:1959-1978) on device bitmaps: dict of
    total_hit, true_hit, first_mismatch (-1: equal), truth_count, got_count."""
    r = _CompareResult()
    check(lib().rts_bitmap_compare(truth.ptr, got.ptr, n, ctypes.byref(r), _sh(stream)),
          "rts_bitmap_compare")
    return {f: getattr(r, f) for f, _ in _CompareResult._fields_}


def bitmap_digest(bitmap, n, stream=None):
    """64-bit position-sensitive digest of the first n rows of a device bitmap."""
    d = ctypes.c_ulonglong()
    check(lib().rts_bitmap_digest(bitmap.ptr, n, ctypes.byref(d), _sh(stream)), "rts_bitmap_digest")
    return d.value


def encode_column(col, n, encode_min, encode_max, codes=None, want_sorted=False, stream=None):
    """Rank encoding of a device column (normalEncode, remap.cpp:110-203).  Returns (codes,
    map_values, map_starts[, sorted_pos, sorted_codes]) as device buffers / numpy maps; codes
    may be `col` to encode in place."""
    codes = codes if codes is not None else DeviceBuffer(4 * n)
    mv, ms = DeviceBuffer(4 * n), DeviceBuffer(4 * n)
    sp = DeviceBuffer(4 * n) if want_sorted else None
    sc = DeviceBuffer(4 * n) if want_sorted else None
    m = ctypes.c_int64()
    check(lib().rts_encode_column(col.ptr, n, encode_min, encode_max, codes.ptr,
                                  sp.ptr if sp else None, sc.ptr if sc else None, mv.ptr, ms.ptr,
                                  ctypes.byref(m), _sh(stream)), "rts_encode_column")
    check(lib().rts_stream_synchronize(_sh(stream)), "rts_stream_synchronize")
    out = (codes, mv.to_numpy(np.uint32, m.value), ms.to_numpy(np.uint32, m.value))
    return out + ((sp, sc) if want_sorted else ())


def encode_column_dispersed(col, n, encode_min, encode_max, seed, codes=None, stream=None):
    """Dispersed encoding (normalEncode 5-argument form, remap.cpp:205-241): (codes,
    map_values, map_starts)."""
    codes = codes if codes is not None else DeviceBuffer(4 * n)
    mv, ms = DeviceBuffer(4 * n), DeviceBuffer(4 * n)
    m = ctypes.c_int64()
    check(lib().rts_encode_column_dispersed(col.ptr, n, encode_min, encode_max, seed, codes.ptr,
                                            mv.ptr, ms.ptr, ctypes.byref(m), _sh(stream)),
          "rts_encode_column_dispersed")
    return codes, mv.to_numpy(np.uint32, m.value), ms.to_numpy(np.uint32, m.value)


def encode_lookup(map_values, map_starts, old_query, op, sentinel):
    """encodeQuery (remap.cpp:243-272) over a host map; raises HipError for eq / bt."""
    mv = np.ascontiguousarray(map_values, np.uint32)
    ms = np.ascontiguousarray(map_starts, np.uint32)
    out = _u32()
    check(lib().rts_encode_lookup(mv.ctypes.data, ms.ctypes.data, mv.size, old_query & 0xFFFFFFFF,
                                  op, sentinel & 0xFFFFFFFF, ctypes.byref(out)), "rts_encode_lookup")
    return out.value


def encode_key_of(map_values, map_starts, code):
    """findKeyByValue (remap.cpp:274-283) over a host map."""
    mv = np.ascontiguousarray(map_values, np.uint32)
    ms = np.ascontiguousarray(map_starts, np.uint32)
    return int(lib().rts_encode_key_of(mv.ctypes.data, ms.ctypes.data, mv.size, code & 0xFFFFFFFF))


class BinDexIndex:
    """A BinDex (include/rts_bindex.h) built on the GPU over a device column."""

    def __init__(self, column, n, stream=None, snapshot=None, bins=128):
        """Build over `column` (device, n rows) with `bins` engine bins (128 * 1, 2, 4 or 8),
        or load a snapshot written by save()."""
        from .bindex_struct import BinDex

        self.column = column  # keep the device column alive
        self.n = n
        self.s = BinDex()
        if snapshot is not None:
            check(lib().rts_bindex_load(ctypes.byref(self.s), os.fsencode(snapshot),
                                        column.ptr if column is not None else None),
                  "rts_bindex_load")
            if self.s.length != n:
                lib().rts_bindex_free(ctypes.byref(self.s))
                raise ValueError("snapshot %r indexes %d rows, the column has %d"
                                 % (snapshot, self.s.length, n))
        else:
            check(lib().rts_bindex_build_bins(ctypes.byref(self.s), column.ptr, n, bins, _sh(stream)),
                  "rts_bindex_build_bins")
        self.built = True

    def save(self, path):
        check(lib().rts_bindex_save(ctypes.byref(self.s), os.fsencode(path)), "rts_bindex_save")

    @property
    def ptr(self):
        return ctypes.addressof(self.s)

    def free(self):
        if getattr(self, "built", False):
            lib().rts_bindex_free(ctypes.byref(self.s))
            self.built = False

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @property
    def bins(self):
        return int(lib().rts_bindex_bins(ctypes.byref(self.s)))

    def start_values(self):
        return np.array(self.s.areaStartValues[:], np.uint32)

    def area_counts(self):
        return np.array(self.s.area_counts[:], np.int64)

    def filter_vector(self, k):
        W = (self.n + 31) // 32
        out = np.empty(W, np.uint32)
        check(lib().rts_memcpy_dtoh(out.ctypes.data, self.s.filterVectorsInGPU[k], 4 * W), "dtoh")
        return out

    def sorted(self):
        v = np.empty(self.n, np.uint32)
        p = np.empty(self.n, np.int32)
        check(lib().rts_bindex_export_sorted(ctypes.byref(self.s), v.ctypes.data, p.ctypes.data),
              "export_sorted")
        return v, p

    def host_blocks(self):
        """Fill the areas' host pos_block views (areas[k]->blocks[b]->pos / val)."""
        check(lib().rts_bindex_host_blocks(ctypes.byref(self.s)), "rts_bindex_host_blocks")

    def block(self, area, b, gpu=False):
        """(pos, val) of block b of an area as numpy arrays: host views (after host_blocks())
        or, with gpu=True, copied from the device views of areasInGPU."""
        a = (self.s.areasInGPU if gpu else self.s.areas)[area].contents
        if not a.blocks:
            raise ValueError("host blocks not filled: call host_blocks()")
        pb = a.blocks[b].contents
        if gpu:
            pos = np.empty(pb.length, np.int32)
            val = np.empty(pb.length, np.uint32)
            if pb.length:
                check(lib().rts_memcpy_dtoh(pos.ctypes.data, ctypes.cast(pb.pos, _vp), 4 * pb.length), "dtoh")
                check(lib().rts_memcpy_dtoh(val.ctypes.data, ctypes.cast(pb.val, _vp), 4 * pb.length), "dtoh")
            return pos, val
        return (np.ctypeslib.as_array(pb.pos, (pb.length,)).copy(),
                np.ctypeslib.as_array(pb.val, (pb.length,)).copy())

    def refine_entries(self):
        return int(lib().rts_bindex_refine_entries(ctypes.byref(self.s)))

    def device_bytes(self):
        return int(lib().rts_bindex_device_bytes(ctypes.byref(self.s)))


def bindex_query(indexes, los, his, out, count=None, stream=None):
    """Kernel (b)+(c): AND over indexed columns of lo_c <= x_c <= hi_c, plus popcount."""
    pa = (_vp * len(indexes))(*[ix.ptr for ix in indexes])
    lo = np.ascontiguousarray(los, np.uint32)
    hi = np.ascontiguousarray(his, np.uint32)
    check(lib().rts_bindex_query(pa, lo.ctypes.data, hi.ctypes.data, len(indexes), out.ptr,
                                 count.ptr if count else None, _sh(stream)), "rts_bindex_query")


class _ColumnPredicate(ctypes.Structure):
    _fields_ = [("op", ctypes.c_int), ("ntargets", ctypes.c_int),
                ("t1", ctypes.POINTER(ctypes.c_uint32)), ("t2", ctypes.POINTER(ctypes.c_uint32))]


class QueryFileError(ValueError):
    pass


def parse_query_file(text=None, path=None, ncols=1):
    """Parse the reference query-file format (rts_query.h) with the library's parser.
    Returns a list of queries, each a list of (op, [t1...], [t2...] or None) per column."""
    h = _vp()
    line = _int()
    if path is not None:
        rc = lib().rts_query_file_load(path.encode(), ncols, ctypes.byref(h), ctypes.byref(line))
    else:
        data = text.encode() if isinstance(text, str) else bytes(text)
        rc = lib().rts_query_file_parse(data, len(data), ncols, ctypes.byref(h), ctypes.byref(line))
    if rc:
        raise QueryFileError("query file error %d at line %d" % (rc, line.value))
    try:
        out = []
        for q in range(lib().rts_query_file_count(h)):
            cols = []
            for c in range(ncols):
                p = ctypes.cast(lib().rts_query_file_get(h, q, c),
                                ctypes.POINTER(_ColumnPredicate)).contents
                t1 = [p.t1[i] for i in range(p.ntargets)]
                t2 = [p.t2[i] for i in range(p.ntargets)] if p.t2 else None
                cols.append((p.op, t1, t2))
            out.append(cols)
        return out
    finally:
        lib().rts_query_file_free(h)


def write_data_file(path, columns):
    cols = [np.ascontiguousarray(c, np.uint32) for c in columns]
    arr = (_vp * len(cols))(*[c.ctypes.data for c in cols])
    rc = lib().rts_data_file_write(path.encode(), len(cols), cols[0].size, arr)
    if rc:
        raise OSError("rts_data_file_write failed: %d" % rc)


def read_data_file(path, ncols, n):
    cols = [np.empty(n, np.uint32) for _ in range(ncols)]
    arr = (_vp * ncols)(*[c.ctypes.data for c in cols])
    rc = lib().rts_data_file_read(path.encode(), ncols, n, arr)
    if rc:
        raise OSError("rts_data_file_read failed: %d" % rc)
    return cols


BATCH_MAX_COLS = 4
EVAL_AUTO, EVAL_BINDEX, EVAL_DIRECT = 0, 1, 2


class QuerySpec(ctypes.Structure):
    _fields_ = [("ncols", ctypes.c_int), ("column", ctypes.c_int * BATCH_MAX_COLS),
                ("lo", ctypes.c_uint32 * BATCH_MAX_COLS), ("hi", ctypes.c_uint32 * BATCH_MAX_COLS),
                ("mode", ctypes.c_int * BATCH_MAX_COLS)]


class BatchConfig(ctypes.Structure):
    _fields_ = [("tile_words", ctypes.c_int), ("queries_per_group", ctypes.c_int),
                ("blocks_per_cu", ctypes.c_int), ("order", ctypes.c_int), ("threads", ctypes.c_int),
                ("bins", ctypes.c_int), ("direct_weight", ctypes.c_int)]

    def astuple(self):
        return (self.tile_words, self.queries_per_group, self.blocks_per_cu, self.order,
                self.threads, self.bins, self.direct_weight)


def batch_config(tile_words=None, queries_per_group=None, blocks_per_cu=None, order=None,
                 threads=None, bins=None, direct_weight=None):
    c = BatchConfig()
    lib().rts_batch_default_config(ctypes.byref(c))
    if tile_words is not None:
        c.tile_words = tile_words
    if queries_per_group is not None:
        c.queries_per_group = queries_per_group
    if blocks_per_cu is not None:
        c.blocks_per_cu = blocks_per_cu
    if order is not None:
        c.order = order
    if threads is not None:
        c.threads = threads
    if bins is not None:
        c.bins = bins
    if direct_weight is not None:
        c.direct_weight = direct_weight
    return c


def query_specs(queries):
    """queries: list of lists of (column, lo, hi[, mode]) -> ctypes array of QuerySpec."""
    arr = (QuerySpec * max(1, len(queries)))()
    for i, q in enumerate(queries):
        if not 1 <= len(q) <= BATCH_MAX_COLS:
            raise ValueError("a query has 1..%d columns" % BATCH_MAX_COLS)
        arr[i].ncols = len(q)
        for c, item in enumerate(q):
            col, lo, hi = item[:3]
            arr[i].column[c] = col
            arr[i].lo[c] = lo
            arr[i].hi[c] = hi
            arr[i].mode[c] = item[3] if len(item) > 3 else EVAL_AUTO
    return arr


class BatchPlan:
    """Batched evaluation plan over a table (include/rts_batch.h)."""

    def __init__(self, columns, indexes, n, max_queries, config=None, stream=None):
        self.columns, self.indexes = columns, indexes  # keep device objects alive
        cols = (_vp * len(columns))(*[c.ptr for c in columns])
        idx = (_vp * len(columns))(*[(ix.ptr if ix is not None else None) for ix in indexes])
        h = _vp()
        cfg = ctypes.byref(config) if config is not None else None
        check(lib().rts_batch_plan_create(ctypes.byref(h), cols, idx, len(columns), n, max_queries,
                                          cfg, _sh(stream)), "rts_batch_plan_create")
        self.h = h.value
        self.n = n

    def set_config(self, config, stream=None):
        check(lib().rts_batch_plan_set_config(self.h, ctypes.byref(config), _sh(stream)),
              "rts_batch_plan_set_config")

    def set_queries(self, queries, stream=None):
        arr = query_specs(queries)
        check(lib().rts_batch_plan_set_queries(self.h, arr, len(queries), _sh(stream)),
              "rts_batch_plan_set_queries")

    def launch(self, out, stride_words, counts=None, stream=None):
        check(lib().rts_batch_plan_launch(self.h, out.ptr, stride_words,
                                          counts.ptr if counts else None, _sh(stream)),
              "rts_batch_plan_launch")

    def graph_launch(self, out, stride_words, counts=None, stream=None):
        check(lib().rts_batch_plan_graph_launch(self.h, out.ptr, stride_words,
                                                counts.ptr if counts else None, _sh(stream)),
              "rts_batch_plan_graph_launch")

    def autotune(self, out, stride_words, candidates, reps=3, cache_path=None, stream=None):
        arr = (BatchConfig * len(candidates))(*candidates)
        chosen = BatchConfig()
        ms = ctypes.c_float()
        check(lib().rts_batch_autotune(self.h, out.ptr, stride_words, arr, len(candidates), reps,
                                       cache_path.encode() if cache_path else None,
                                       ctypes.byref(chosen), ctypes.byref(ms), _sh(stream)),
              "rts_batch_autotune")
        return chosen, ms.value

    def stats(self):
        b, d, nr = _int(), _int(), _int()
        nt = ctypes.c_longlong()
        lib().rts_batch_plan_stats(self.h, ctypes.byref(b), ctypes.byref(d), ctypes.byref(nr),
                                   ctypes.byref(nt))
        return {"bindex_cols": b.value, "direct_cols": d.value, "narrow_cols": nr.value,
                "tiles": nt.value, "auto_direct_cols": lib().rts_batch_plan_auto_direct(self.h),
                "list_cols": lib().rts_batch_plan_list_cols(self.h),
                "groups": lib().rts_batch_plan_groups(self.h, None)}

    def config(self):
        c = BatchConfig()
        lib().rts_batch_plan_groups(self.h, ctypes.byref(c))
        return c

    def add_indexes(self, indexes):
        """Another index set of the same table (another bin count); returns its bins."""
        self.extra_sets = getattr(self, "extra_sets", []) + [indexes]  # keep them alive
        idx = (_vp * len(indexes))(*[(ix.ptr if ix is not None else None) for ix in indexes])
        b = _int()
        check(lib().rts_batch_plan_add_indexes(self.h, idx, ctypes.byref(b)), "rts_batch_plan_add_indexes")
        return b.value

    def mix_digest(self):
        return int(lib().rts_batch_plan_mix_digest(self.h))

    def free(self):
        if getattr(self, "h", None):
            lib().rts_batch_plan_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_FAILED = 0, 1, 2, 3


class ServerConfig(ctypes.Structure):
    _fields_ = [("max_batch", ctypes.c_int), ("streams", ctypes.c_int),
                ("pool_bytes", ctypes.c_size_t), ("queue_capacity", ctypes.c_int),
                ("batch", BatchConfig)]


def server_config(**kw):
    c = ServerConfig()
    lib().rts_server_default_config(ctypes.byref(c))
    for k, v in kw.items():
        if k == "batch":
            c.batch = v
        else:
            setattr(c, k, v)
    return c


class Server:
    """Request loop over resident tables (include/rts_serve.h)."""

    def __init__(self, config=None):
        h = _vp()
        check(lib().rts_server_create(ctypes.byref(h), ctypes.byref(config) if config else None),
              "rts_server_create")
        self.h = h.value
        self.tables = []

    def add_table(self, columns, indexes, n):
        cols = (_vp * len(columns))(*[c.ptr for c in columns])
        idx = (_vp * len(columns))(*[(ix.ptr if ix is not None else None) for ix in indexes])
        tid = _int()
        check(lib().rts_server_add_table(self.h, cols, idx, len(columns), n, ctypes.byref(tid)),
              "rts_server_add_table")
        self.tables.append((columns, indexes, n))
        return tid.value

    def start(self):
        check(lib().rts_server_start(self.h), "rts_server_start")

    def submit(self, table, query):
        arr = query_specs([query])
        t = ctypes.c_uint64()
        check(lib().rts_server_submit(self.h, table, arr, ctypes.byref(t)), "rts_server_submit")
        return t.value

    def wait(self, ticket, timeout_ms=-1):
        bm = _vp()
        cnt = ctypes.c_ulonglong()
        lat = ctypes.c_uint64()
        st = lib().rts_server_wait(self.h, ticket, timeout_ms, ctypes.byref(bm), ctypes.byref(cnt),
                                   ctypes.byref(lat))
        return st, bm.value, cnt.value, lat.value

    def read_bitmap(self, ptr, n):
        out = np.empty((n + 31) // 32, np.uint32)
        check(lib().rts_memcpy_dtoh(out.ctypes.data, ptr, out.nbytes), "dtoh")
        return out

    def release(self, ticket):
        lib().rts_server_release(self.h, ticket)

    def stats(self):
        a, b, c = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        m = _int()
        u, pk = ctypes.c_size_t(), ctypes.c_size_t()
        lib().rts_server_stats(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                               ctypes.byref(m), ctypes.byref(u), ctypes.byref(pk))
        return {"submitted": a.value, "completed": b.value, "batches": c.value,
                "max_batch_seen": m.value, "pool_in_use": u.value, "pool_peak": pk.value}

    def replay(self, table_ids, queries, arrival_ns):
        nq = len(queries)
        tids = np.ascontiguousarray(table_ids, np.int32)
        arr = query_specs(queries)
        arrivals = np.ascontiguousarray(arrival_ns, np.uint64)
        lat = np.zeros(nq, np.uint64)
        cnt = np.zeros(nq, np.uint64)
        el = ctypes.c_double()
        check(lib().rts_server_replay(self.h, tids.ctypes.data, arr, arrivals.ctypes.data, nq,
                                      lat.ctypes.data, cnt.ctypes.data, ctypes.byref(el)),
              "rts_server_replay")
        return lat, cnt, el.value

    def stop(self):
        check(lib().rts_server_stop(self.h), "rts_server_stop")

    def destroy(self):
        if getattr(self, "h", None):
            lib().rts_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
