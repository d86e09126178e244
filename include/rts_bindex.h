/*
 * rts_bindex.h -- the BinDex descriptor of the MI355X engine (drop-in for bindex.h:104-138).
 *
 * The reference keeps one host pos_block (and one cudaMalloc'd device copy) per 3276-row
 * block, ~30k per column at 1e8 rows (bindex_cuda.cpp:125-152).  This engine keeps each
 * column's index as a few flat device arrays owned by the library (value-sorted values
 * and row positions, K-1 filter vectors in one allocation, per-bin refine lists in row
 * order) behind `dev`.  The fields the reference's callers read keep their names and
 * meaning: length, areaStartValues, area_counts, data_min, data_max, filterVectorsInGPU,
 * rawDataInGPU, and areas[k] (block counts and block boundary values, host side, enough
 * for in_which_block).
 */
#ifndef RTS_BINDEX_H_
#define RTS_BINDEX_H_

#include "rtscan_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define RTS_VAREA_N 128          /* K, bindex.h:35-37 */
#define RTS_BLOCK_INIT_SIZE 3276 /* blockInitSize, bindex.h:104 */

/* Area, bindex.h:116-120: blockNum blocks of RTS_BLOCK_INIT_SIZE rows (the last holds the
 * rest) over the area's slice [start, start + length) of the value-sorted column. */
typedef struct Area {
  int blockNum;
  int length;
  POSTYPE start;
  CODE *blockFirst; /* host, blockNum: smallest value of each block */
  CODE *blockLast;  /* host, blockNum: largest value of each block */
} Area;

struct rts_bindex_dev;

/* BinDex, bindex.h:126-138 */
typedef struct BinDex {
  Area *areas[RTS_VAREA_N];
  CODE areaStartValues[RTS_VAREA_N];
  BITS *filterVectors[RTS_VAREA_N - 1];      /* host copies; NULL unless requested */
  BITS *filterVectorsInGPU[RTS_VAREA_N - 1]; /* device: fv[k] bit i = (x_i < areaStartValues[k+1]) */
  POSTYPE area_counts[RTS_VAREA_N];          /* rows in areas 0..k */
  POSTYPE length;
  CODE data_min;
  CODE data_max;
  Area *areasInGPU[RTS_VAREA_N]; /* always NULL: replaced by the flat device index */
  CODE *rawDataInGPU;            /* device column (owned by the caller or by the library) */
  struct rts_bindex_dev *dev;    /* engine-owned device index */
} BinDex;

#ifdef __cplusplus
}
#endif
#endif /* RTS_BINDEX_H_ */
