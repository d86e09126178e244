"""Shared helpers for the tests: closed-form predicates and bitmap packing in numpy."""
import numpy as np

MAXU = 0xFFFFFFFF


def pack_rows(mask):
    """Pack a boolean row mask into MSB-first uint32 words (row i -> bit 31 - i % 32 of word
    i // 32), tail bits 0.  Independent of the oracle: numpy packbits in big-endian order."""
    mask = np.asarray(mask, dtype=bool)
    n = mask.size
    W = (n + 31) // 32
    padded = np.zeros(W * 32, dtype=bool)
    padded[:n] = mask
    return np.packbits(padded, bitorder="big").view(">u4").astype(np.uint32)


def unpack_rows(words, n):
    b = np.asarray(words, np.uint32).astype(">u4").view(np.uint8)
    return np.unpackbits(b, bitorder="big")[:n].astype(bool)


def closed_form(x, op, t1, t2=0):
    """The reference predicate of raw_scan (bindex.cpp:1857-1874) as numpy uint32 compares."""
    x = np.asarray(x, np.uint32)
    t1 = np.uint32(t1 & MAXU)
    t2 = np.uint32(t2 & MAXU)
    if op == 0:
        return x == t1
    if op == 1:
        return x < t1
    if op == 2:
        return x > t1
    if op == 3:
        return x <= t1
    if op == 4:
        return x >= t1
    if op == 5:
        return (x > t1) & (x < t2)
    raise ValueError(op)


def interval_mask(x, lo, hi):
    x = np.asarray(x, np.uint32)
    return (x >= np.uint32(lo)) & (x <= np.uint32(hi))


def config1_predicates(n_pred=64, seed=20240101):
    """Closed ranges [lo, hi] whose width gives selectivity cycling through
    {0.001, 0.01, 0.1, 0.5, 0.9}; lo uniform on [0, 2^32 - width] so hi never clips."""
    rng = np.random.default_rng(seed)
    sels = [0.001, 0.01, 0.1, 0.5, 0.9]
    out = []
    for i in range(n_pred):
        width = int(round(sels[i % 5] * 2.0 ** 32))
        lo = int(rng.integers(0, 2 ** 32 - width + 1))
        out.append((lo, lo + width - 1))
    return out
