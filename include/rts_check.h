/*
 * rts_check.h -- result checking on the device (replaces the host-side compare_bitmap of
 * bindex.cpp:1959-1978 / bindex_cuda.cpp:1288-1311) and a bitmap digest for at-scale parity.
 *
 * Bitmaps are device arrays of ceil(n / 32) MSB-first words; only the first n rows count
 * (bits past row n are ignored, as the reference walks exactly `len` rows).
 */
#ifndef RTS_CHECK_H_
#define RTS_CHECK_H_

#include "rtscan_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rts_compare_result {
  unsigned long long total_hit;   /* compare_bitmap's total_hit: truth rows set, rows [0, stop] */
  unsigned long long true_hit;    /* compare_bitmap's true_hit: rows set in both, rows [0, stop] */
  long long first_mismatch;       /* first row whose bits differ (stop), -1 when the bitmaps agree */
  unsigned long long truth_count; /* rows set in `truth` over all n rows */
  unsigned long long got_count;   /* rows set in `got` over all n rows */
} rts_compare_result;

/* compare_bitmap semantics on device bitmaps: the reference walks rows 0.. in order,
 * counting truth hits (total_hit) and hits present in both (true_hit), and stops after the
 * first row whose bits differ (that row included).  Returns after the work has finished. */
hipError_t rts_bitmap_compare(const BITS *truth, const BITS *got, int64_t n, rts_compare_result *res,
                              hipStream_t stream);

/* 64-bit digest of the first n rows of a device bitmap: the sum (mod 2^64) over words i of
 * mix64((i << 32) | word_i), word_i with the bits past row n cleared, mix64 the splitmix64
 * finaliser of (x ^ 0x9E3779B97F4A7C15).  Position-sensitive; independent of how the sum is
 * split across workgroups, so the CPU restatement gives the same value.  Returns after the
 * work has finished. */
hipError_t rts_bitmap_digest(const BITS *bitmap, int64_t n, unsigned long long *digest,
                             hipStream_t stream);

/* compare_bitmap (bindex.cpp:1959-1978), same signature: HOST bitmaps of len rows, checked
 * on the GPU.  Prints the reference's lines: on the first mismatch "[ERROR] check error in
 * raw_data[i]=" with the row's values of the first min(bindex_num, 3) columns of raw_data
 * (host arrays; skipped when NULL) and "the correct is .., but we have ..", then
 * "[CHECK]hit true/total". */
void compare_bitmap(BITS *bitmap_a, BITS *bitmap_b, int len, CODE **raw_data, int bindex_num);

#ifdef __cplusplus
}
#endif
#endif /* RTS_CHECK_H_ */
