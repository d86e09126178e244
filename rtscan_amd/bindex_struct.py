"""ctypes mirror of the BinDex / Area descriptors declared in include/rts_bindex.h."""
import ctypes

VAREA_N = 128  # RTS_VAREA_N, K of the reference (bindex.h:35-37)
BLOCK_INIT_SIZE = 3276  # RTS_BLOCK_INIT_SIZE (bindex.h:104)


class PosBlock(ctypes.Structure):
    _fields_ = [
        ("pos", ctypes.POINTER(ctypes.c_int32)),
        ("val", ctypes.POINTER(ctypes.c_uint32)),
        ("length", ctypes.c_int),
    ]


class Area(ctypes.Structure):
    _fields_ = [
        ("blocks", ctypes.POINTER(ctypes.POINTER(PosBlock))),
        ("blockNum", ctypes.c_int),
        ("length", ctypes.c_int),
        ("start", ctypes.c_int),
        ("blockFirst", ctypes.POINTER(ctypes.c_uint32)),
        ("blockLast", ctypes.POINTER(ctypes.c_uint32)),
    ]


class BinDex(ctypes.Structure):
    _fields_ = [
        ("areas", ctypes.POINTER(Area) * VAREA_N),
        ("areaStartValues", ctypes.c_uint32 * VAREA_N),
        ("filterVectors", ctypes.c_void_p * (VAREA_N - 1)),
        ("filterVectorsInGPU", ctypes.c_void_p * (VAREA_N - 1)),
        ("area_counts", ctypes.c_int * VAREA_N),
        ("length", ctypes.c_int),
        ("data_min", ctypes.c_uint32),
        ("data_max", ctypes.c_uint32),
        ("areasInGPU", ctypes.POINTER(Area) * VAREA_N),
        ("rawDataInGPU", ctypes.c_void_p),
        ("dev", ctypes.c_void_p),
    ]


def new_bindex_struct(length=0):
    b = BinDex()
    b.length = length
    return b
