/*
 * oracle.h -- CPU restatement of the reference's scan / BinDex path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped engine (rtscan_amd/, include/)
 * includes, links or calls this code.  It is used by tests/ (as the checker),
 * by __graft_entry__.smoke() (as the checker) and by bench.py's cpu_baseline leg
 * (as the timed CPU port of the reference path).
 *
 * Every function cites the reference file:line it restates.  Reference paths are
 * relative to the RTScan repository root.
 *
 * Pinning (what checks each function against something other than itself):
 *   orc_raw_scan / _mt / _entry   closed-form numpy predicates on random and edge inputs,
 *                                 hand-written bit-order fixtures, and the compiled
 *                                 reference raw_scan (oracle/_ref) on the same inputs.
 *   orc_compare_bitmap            hand-computed counts on crafted bitmaps.
 *   orc_init_bindex               compiled reference init_bindex (oracle/_ref): area start
 *                                 values, area counts and all 127 filter vectors bit-exact;
 *                                 closed form fv[k] = (x < start[k+1]) in numpy.
 *   orc_in_which_area/_block,     compiled reference on the same BinDex.
 *   orc_on_which_pos
 *   orc_conjunction_scan          orc_raw_scan_entry (bt / le / ge per column, AND merge) on
 *                                 ragged row counts, 1..4 columns, empty and full ranges.
 *   orc_conjunction_digest        orc_raw_scan_entry (bt / le / ge per column, AND merge) and
 *                                 orc_bitmap_digest on the same inputs, including empty
 *                                 ranges and ragged row counts.
 *   orc_bindex_scan               compiled reference bindex_scan_{lt,le,gt,ge,bt,eq}
 *                                 bit-exact on every tested query (including the inputs
 *                                 where the reference disagrees with its own raw_scan),
 *                                 and raw_scan on the inputs where the reference is right.
 */
#ifndef RTSCAN_ORACLE_H_
#define RTSCAN_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* bindex.h:88-95 */
enum orc_op { ORC_EQ = 0, ORC_LT = 1, ORC_GT = 2, ORC_LE = 3, ORC_GE = 4, ORC_BT = 5 };

/* bindex.h:104 (blockInitSize) */
#define ORC_BLOCK_INIT_SIZE 3276

int orc_bits_num_needed(int n);

void orc_raw_scan(const uint32_t *raw, int n, uint32_t *bitmap, uint32_t t1, uint32_t t2, int op);
void orc_raw_scan_mt(const uint32_t *raw, int n, uint32_t *bitmap, uint32_t t1, uint32_t t2, int op,
                     int threads);
void orc_raw_scan_entry(const uint32_t *raw, int n, int op, const uint32_t *tl, const uint32_t *tr,
                        int ntargets, uint32_t *bitmap, uint32_t *merge_bitmap, int threads);
void orc_compare_bitmap(const uint32_t *truth, const uint32_t *got, int n, int *total_hit,
                        int *true_hit, int *first_mismatch);
uint64_t orc_popcount_rows(const uint32_t *bitmap, int n);

/* Flat restatement of the reference BinDex (bindex.cpp:275-298): areas are contiguous
 * ranges of the value-sorted arrays, blocks are ORC_BLOCK_INIT_SIZE-row slices of an area. */
typedef struct {
  int n;
  int K;
  int words;              /* bits_num_needed(n) */
  uint32_t *sorted_val;   /* n, ascending (stable argsort) */
  int32_t *sorted_pos;    /* n */
  uint32_t *start_values; /* K: area_start_value(areas[k]) */
  int32_t *area_start;    /* K: areaStartIdx */
  int32_t *area_len;      /* K */
  int32_t *area_counts;   /* K: prefix sums (bindex.cpp:649-652) */
  uint32_t *fvs;          /* (K-1) * words: fv[k] = x < start_values[k+1] */
  uint32_t data_min;
  uint32_t data_max;
} orc_bindex;

int orc_init_bindex(orc_bindex *bx, const uint32_t *data, int n, int K, int threads);
void orc_free_bindex(orc_bindex *bx);

int orc_area_block_num(const orc_bindex *bx, int area);
int orc_in_which_area(const orc_bindex *bx, uint32_t c);
int orc_in_which_block(const orc_bindex *bx, int area, uint32_t c);
int orc_on_which_pos(const orc_bindex *bx, int area, int block, uint32_t c);

/* One reference BinDex scan (bindex_scan_lt/le/gt/ge/bt/eq), written into `result`.
 * Returns 0, or -1 where the reference asserts (bt with c2 <= c1, bt whose upper
 * bound is below the data, eq outside the data, eq whose xor path asserts). */
int orc_bindex_scan(const orc_bindex *bx, uint32_t *result, int op, uint32_t t1, uint32_t t2,
                    int threads);

/* GPURefineAreaWithCuda / GPURefineEqAreaWithCuda (cuda_and.cu:510-812) with refineBlock
 * (cuda_and.cu:74-99): set (inverse: clear) the bit of every row in the chosen blocks of the
 * selected column's area(s) whose value in every column lies strictly inside its
 * (lo, hi) pair of `predicate`.  `bxs[c]` indexes `raws[c]`.  Returns -1 where the
 * reference would read out of bounds (bound below the data). */
int orc_refine_area(const orc_bindex *const *bxs, const uint32_t *const *raws, int ncols,
                    const uint32_t *predicate, int sel, int inverse, uint32_t *bitmap);
int orc_refine_eq_area(const orc_bindex *const *bxs, const uint32_t *const *raws, int ncols,
                       const uint32_t *predicate, int sel, int inverse, uint32_t *bitmap);

/* Bitmap digest (the engine's rts_bitmap_digest): sum mod 2^64 over words i of
 * mix64((i << 32) | word_i), bits past row n cleared, mix64(x) = splitmix64 finaliser of
 * x ^ 0x9E3779B97F4A7C15.  Not a reference function: a checksum shared by both sides. */
uint64_t orc_bitmap_digest(const uint32_t *bitmap, int64_t n);

/* The checker's conjunction (raw_scan_entry per column + AND merge, bindex.cpp:1851-1892,
 * 2265-2289) for closed ranges lo[c] <= x <= hi[c] over 1..4 columns, vectorised (AVX2 when
 * the host has it: orc_simd_level() == 2) and word-parallel: the CPU baseline's timed port.
 * Writes all ceil(n/32) words, tail bits 0.  Pinned against orc_raw_scan_entry on ragged n. */
void orc_conjunction_scan(const uint32_t *const *cols, int ncols, int64_t n, const uint32_t *lo,
                          const uint32_t *hi, uint32_t *bitmap, int threads);
int orc_simd_level(void);

/* Digest (orc_bitmap_digest) and popcount of the AND over ncols columns of n rows of the
 * closed ranges lo[c] <= x <= hi[c] (lo > hi: no row), computed word by word without the
 * bitmap (bindex.cpp:2265-2289's checker, reduced).  Pinned against orc_raw_scan_entry +
 * orc_bitmap_digest on small n. */
void orc_conjunction_digest(const uint32_t *const *cols, int ncols, int64_t n, const uint32_t *lo,
                            const uint32_t *hi, int threads, uint64_t *digest, uint64_t *count);

/* normalEncode, 7-argument form (remap.cpp:110-203), for one column: rank-order codes
 * code(row at sorted index i) = encode_min + i * interval (uint32 wrap), interval =
 * (int)((encode_max - encode_min) / n) in double; the encode map gives every distinct value
 * the code of its first sorted occurrence.  Ties follow the stable sort (the reference's
 * parallel argsort leaves their order unspecified).  codes[n]; sorted_pos[n];
 * sorted_codes[n]; map_values / map_starts receive up to n entries; returns the map size.
 * parity of the tie order is unpinned; everything else is pinned by the invariants in
 * tests/test_oracle.py. */
int64_t orc_normal_encode(const uint32_t *data, int n, uint32_t encode_min, uint32_t encode_max,
                          uint32_t *codes, int32_t *sorted_pos, uint32_t *sorted_codes,
                          uint32_t *map_values, uint32_t *map_starts);
/* normalEncode, 5-argument form (remap.cpp:205-241), its encode map: distinct values, start
 * codes and each value's code range width step = (int)(D * count), D = (max - min) / n in
 * double; start = encode_min + sum of earlier steps (uint32 wrap).  Row codes are drawn by
 * rand() in the reference (remap.cpp:234) and are not restated: a row's code lies in
 * [start, start + step - 1) of its value.  Returns the map size. */
int64_t orc_dispersed_map(const uint32_t *data, int n, uint32_t encode_min, uint32_t encode_max,
                          uint32_t *map_values, uint32_t *map_starts, uint32_t *map_steps);
/* encodeQuery (remap.cpp:243-272) over a map: op lt/ge -> first value >= q, le/gt -> first
 * value > q; its start code (le: minus 1), or `sentinel` (the reference returns N) when
 * there is none.  Returns -1 for eq / bt (the reference exits). */
int orc_encode_query(const uint32_t *map_values, const uint32_t *map_starts, int64_t m, uint32_t q,
                     int op, uint32_t sentinel, uint32_t *out);
/* findKeyByValue (remap.cpp:274-283): the last key whose code is <= val, 0 when val is
 * below every code or at/above the last one. */
uint32_t orc_find_key_by_value(const uint32_t *map_values, const uint32_t *map_starts, int64_t m,
                               uint32_t val);

#ifdef __cplusplus
}
#endif
#endif
