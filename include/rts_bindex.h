/*
 * rts_bindex.h -- the BinDex descriptor of the MI355X engine (drop-in for bindex.h:104-138).
 *
 * The reference keeps one host pos_block (and one cudaMalloc'd device copy) per 3276-row
 * block, ~30k per column at 1e8 rows (bindex_cuda.cpp:125-152).  This engine keeps each
 * column's index as a few flat device arrays owned by the library (value-sorted values
 * and row positions, K-1 filter vectors in one allocation, per-bin refine lists in row
 * order) behind `dev`.  The fields the reference's callers read keep their names and
 * meaning: length, areaStartValues, area_counts, data_min, data_max, filterVectorsInGPU,
 * rawDataInGPU, areas[k] and areasInGPU[k] with their blocks[b]->pos / val / length
 * (views into the flat sorted arrays: host copies for areas, device memory for
 * areasInGPU), plus per-block boundary values for in_which_block.
 */
#ifndef RTS_BINDEX_H_
#define RTS_BINDEX_H_

#include "rtscan_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define RTS_VAREA_N 128          /* K, bindex.h:35-37 */
#define RTS_BLOCK_INIT_SIZE 3276 /* blockInitSize, bindex.h:104 */

/* pos_block, bindex.h:112-117: one block of an area's value-sorted rows.  Here a view
 * into the column's flat value-sorted arrays (ties in row order). */
typedef struct pos_block {
  POSTYPE *pos; /* row positions of the block's rows, in value order */
  CODE *val;    /* their values, ascending */
  int length;
} pos_block;

/* Area, bindex.h:118-122: blockNum blocks of RTS_BLOCK_INIT_SIZE rows (the last holds the
 * rest) over the area's slice [start, start + length) of the value-sorted column.  The
 * reference's fixed array blocks[blockNumMax] (blockNumMax fixed by the compile-time row
 * count) is a pointer to blockNum block pointers here: area->blocks[b]->pos reads the same.
 * In BinDex::areas the blocks hold host pointers and exist once rts_bindex_host_blocks has
 * run (init_bindex_in_GPU runs it; blocks is NULL before); in BinDex::areasInGPU they
 * hold device pointers and always exist. */
typedef struct Area {
  pos_block **blocks;
  int blockNum;
  int length;
  POSTYPE start;    /* first sorted slot of the area */
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
  Area *areasInGPU[RTS_VAREA_N]; /* blocks[b]->pos / val point into device memory */
  CODE *rawDataInGPU;            /* device column (owned by the caller or by the library) */
  struct rts_bindex_dev *dev;    /* engine-owned device index */
} BinDex;

/* ------------------------------------------------ reference surface (bindex.h:145-183) */

/* copy_filter_vector_in_GPU (bindex.h:145, bindex_cuda.cpp:486-504): dev_bitmap =
 * fv[k] (negation: ~fv[k]) over bits_num_needed(length) words; k < 0 gives all zeros and
 * k >= K-1 all ones.  Synchronous. */
void copy_filter_vector_in_GPU(BinDex *bindex, BITS *dev_bitmap, int k,
                               bool negation RTS_DEFAULT(false));

/* GPURefineAreaWithCuda (bindex.h:180, cuda_and.cu:510-640).  predicate: host array of
 * column_num (lo, hi) pairs, exclusive bounds (lo < x < hi).  Refines the blocks of the
 * selected column's boundary area: blocks [0, blockIdx] set the bit of every row inside all
 * bounds, or with inverse blocks [blockIdx, blockNum) clear it.  The reference launches one
 * kernel per 3276-row block and leaks a device buffer per call; this is one launch over the
 * contiguous slot range, no allocation.  Synchronous. */
hipError_t GPURefineAreaWithCuda(BinDex **bindexs, BITS *dev_result_bitmap, CODE *predicate,
                                 int selected_id, int column_num RTS_DEFAULT(3),
                                 bool inverse RTS_DEFAULT(false));
/* GPURefineEqAreaWithCuda (bindex.h:181, cuda_and.cu:642-812): the same test over the slots
 * from the block of the selected column's smaller bound to the block of its larger bound. */
hipError_t GPURefineEqAreaWithCuda(BinDex **bindexs, BITS *dev_result_bitmap, CODE *predicate,
                                   int selected_id, int column_num RTS_DEFAULT(3),
                                   bool inverse RTS_DEFAULT(false));
/* in_which_area (bindex.h:182, bindex.cpp:1138-1151) */
int in_which_area(BinDex *bindex, CODE compare);
/* in_which_block (bindex.h:183, bindex.cpp:1153-1176) */
int in_which_block(Area *area, CODE compare);
/* on_which_pos (bindex.cpp:1201-1215, bindex_cuda.cpp:756-770): index of the first value
 * >= compare in a host block (pb->length when every value is below it).  The reference
 * asserts compare >= pb->val[0]; this returns 0 there. */
int on_which_pos(pos_block *pb, CODE compare);

/* ------------------------------------------------------------------ additive BinDex API */

/* Build the index of a device column (init_bindex, bindex.cpp:599-671) on the GPU.  The
 * column is referenced, not copied (BinDex::rawDataInGPU), and must outlive the index.
 * n >= RTS_VAREA_N.  Synchronous. */
hipError_t rts_bindex_build(BinDex *bindex, const CODE *dev_column, POSTYPE n, hipStream_t stream);
/* As rts_bindex_build with `bins` engine bins (128, 256, 512 or 1024): each of the
 * reference's 128 areas is cut into bins / 128 equal-count bins, each with its filter
 * vector and refine list, so a query bound refines a list bins / 128 times shorter for
 * (bins - 1) * n / 8 bytes of filter vectors.  The reference-facing fields (areas, area
 * starts, filterVectorsInGPU[k] = x < areaStartValues[k + 1]) are the same for every bin
 * count; rts_bindex_query and the batched launch use the engine bins. */
hipError_t rts_bindex_build_bins(BinDex *bindex, const CODE *dev_column, POSTYPE n, int bins,
                                 hipStream_t stream);
/* Engine bins of a built index (-1 if none). */
int rts_bindex_bins(const BinDex *bindex);
/* As rts_bindex_build for a host column, uploaded into a library-owned device copy. */
hipError_t rts_bindex_build_host(BinDex *bindex, const CODE *host_column, POSTYPE n);
/* Host copies of the value-sorted values and positions behind BinDex::areas[k]->blocks
 * (8 bytes per row of host memory), as the reference keeps them (bindex.cpp:300-309).
 * Idempotent.  Synchronous. */
hipError_t rts_bindex_host_blocks(BinDex *bindex);
/* Release the device index (and the library-owned column copy, if any). */
void rts_bindex_free(BinDex *bindex);
/* Index persistence (no reference counterpart; the reference rebuilds at every start,
 * bindex_cuda.cpp:1860-1890): write a built index to a binary snapshot file, and load one
 * into device memory without sorting or rebuilding the filter vectors.  `dev_column` is the
 * indexed column already on the device (kept as rawDataInGPU for the refine surface; may
 * be NULL; when given it must hold the snapshot's n rows).  A snapshot whose header, host
 * arrays, refine-list rows or per-array digests do not check out, or whose indexed column
 * differs from `dev_column`, is refused with hipErrorInvalidValue before the index can be
 * used.
 * Synchronous. */
hipError_t rts_bindex_save(const BinDex *bindex, const char *path);
hipError_t rts_bindex_load(BinDex *bindex, const char *path, const CODE *dev_column);

/* Kernel (b) fused with kernel (c): out = AND over ncols indexed columns of
 * (lo[c] <= x_c <= hi[c]) evaluated from filter vectors and boundary-bin refine lists,
 * popcount into `count` (device, may be NULL).  All columns have the same length; lo and
 * hi are host arrays.  Asynchronous on `stream`. */
hipError_t rts_bindex_query(const BinDex *const *bindexes, const CODE *lo, const CODE *hi,
                            int ncols, BITS *out, unsigned long long *count, hipStream_t stream);

/* Host copies of the value-sorted column and its row positions (ties in row order). */
hipError_t rts_bindex_export_sorted(const BinDex *bindex, CODE *host_sorted_val,
                                    int32_t *host_sorted_pos);
/* Number of refine-list entries (rows strictly inside a bin, plus the rows of each bin start
 * value that has at most 64 rows), or -1. */
long long rts_bindex_refine_entries(const BinDex *bindex);
/* Device bytes held by the index (excluding the column), or -1. */
long long rts_bindex_device_bytes(const BinDex *bindex);

#ifdef __cplusplus
}
#endif

#ifdef __cplusplus
/* init_bindex_in_GPU (bindex_cuda.cpp:164-282): the reference driver's build call, with its
 * signature.  base_data grows by the device bytes of the uploaded column, as there; the
 * areas' host blocks are filled, as the reference's host build leaves them. */
inline void init_bindex_in_GPU(BinDex *bindex, CODE *data, POSTYPE n, size_t &base_data) {
  if (rts_bindex_build_host(bindex, data, n) == hipSuccess && rts_bindex_host_blocks(bindex) == hipSuccess)
    base_data += (size_t)n * sizeof(CODE);
}
#endif
#endif /* RTS_BINDEX_H_ */
