"""The snapshot file layout of rts_bindex_save (format 2), written out field by field so that
tests can damage one field at a time (the loader's checks are in bindex_build.hip)."""
import struct

import numpy as np

AREAS = 128
HEADER = struct.Struct("<8sIIII6qII6Q")


def fnv1a(chunks):
    h = 0xCBF29CE484222325
    for c in chunks:
        for b in bytes(c):
            h = ((h ^ b) * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return h


def parse(raw):
    """Header fields and the host arrays of a snapshot, plus the payload offset."""
    f = dict(zip(("magic", "version", "bins", "sub", "pad0", "n", "words", "fv_stride", "ntiles", "nb",
                  "nblocks", "data_max", "pad1", "dg_host", "dg_val", "dg_pos", "dg_fvs", "dg_bent", "dg_col"),
                 HEADER.unpack_from(raw, 0)))
    off = HEADER.size
    arrays = {}
    for name, dt, count in (("starts", np.uint32, AREAS), ("area_start", np.int64, AREAS),
                            ("area_len", np.int64, AREAS), ("estarts", np.uint32, f["bins"]),
                            ("boff", np.uint32, f["bins"] + 1), ("block_first", np.uint32, f["nblocks"]),
                            ("block_last", np.uint32, f["nblocks"])):
        a = np.frombuffer(raw, dt, count, off).copy()
        arrays[name] = a
        off += a.nbytes
    return f, arrays, off


def build(f, arrays, payload=b""):
    """A snapshot from header fields and host arrays, with a correct host digest."""
    order = ("starts", "area_start", "area_len", "estarts", "boff", "block_first", "block_last")
    f = dict(f)
    f["dg_host"] = fnv1a([arrays[k].tobytes() for k in order])
    head = HEADER.pack(*[f[k] for k in ("magic", "version", "bins", "sub", "pad0", "n", "words", "fv_stride",
                                       "ntiles", "nb", "nblocks", "data_max", "pad1", "dg_host", "dg_val",
                                       "dg_pos", "dg_fvs", "dg_bent", "dg_col")])
    return head + b"".join(arrays[k].tobytes() for k in order) + payload


def synthetic(n=128 * 4000, sub=1):
    """Header and host arrays of a plausible index over n = 128 * m rows of distinct values
    (equal areas, no refine lists), as the builder would write them."""
    bins = AREAS * sub
    per = n // AREAS
    words = (n + 31) // 32
    starts = (np.arange(AREAS, dtype=np.uint64) * per * 16).astype(np.uint32)
    area_start = np.arange(AREAS, dtype=np.int64) * per
    area_len = np.full(AREAS, per, np.int64)
    estarts = np.array([starts[b // sub] + (b % sub) * (per // sub) * 16 for b in range(bins)], np.uint32)
    nblk = -(-per // 3276) if per > 3276 else 1
    block_first = np.zeros(AREAS * nblk, np.uint32)
    f = {"magic": b"RTSBDX\x00\x02", "version": 2, "bins": bins, "sub": sub, "pad0": 0, "n": n,
         "words": words, "fv_stride": (words + 63) // 64 * 64, "ntiles": (n + 32767) // 32768, "nb": 0,
         "nblocks": AREAS * nblk, "data_max": int(starts[-1]) + per * 16, "pad1": 0, "dg_host": 0,
         "dg_val": 0, "dg_pos": 0, "dg_fvs": 0, "dg_bent": 0, "dg_col": 0}
    arrays = {"starts": starts, "area_start": area_start, "area_len": area_len, "estarts": estarts,
              "boff": np.zeros(bins + 1, np.uint32), "block_first": block_first,
              "block_last": block_first.copy()}
    return f, arrays
