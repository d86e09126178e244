/*
 * rtscan_hip.h -- C ABI of the MI355X range-predicate scan engine (librtscan_hip.so).
 *
 * Drop-in surface: the names, argument order and units of the reference's GPU call
 * surface (bindex.h:145-183, cuda_and.cu, rt.h:142-145), with hipError_t in place of
 * cudaError_t.  Additive surface: rts_* entry points (direct scan, fused conjunction,
 * AND+popcount, device-memory and timing helpers, seeded column generators).
 *
 * Conventions shared by every entry point:
 *   - Columns are uint32 (CODE) arrays in device memory, one value per row.
 *   - Bitmaps are 1 bit per row, MSB-first inside 32-bit words: row i is bit
 *     (1u << (31 - i % 32)) of word i / 32 (bindex.cpp:269).  rts_* entry points write
 *     ceil(n / 32) words and leave the bits past row n at 0.
 *   - Reference-named wrappers run on the calling thread's default stream
 *     (hipStreamPerThread: ordered after work on the legacy null stream) and return after
 *     the work has finished (the reference ends every wrapper with cudaDeviceSynchronize,
 *     cuda_and.cu:276-287); calls from different host threads may overlap on the device.
 *     They use the calling thread's current device instead of forcing device 0
 *     (cuda_and.cu:109).  rts_* entry points are asynchronous on the stream they are given.
 *   - Errors are returned, never fatal: no entry point exits the process.
 *   - Entry points may be called from several host threads at once, also on one stream
 *     (the reference calls its wrappers from per-column threads, bindex_cuda.cpp:1901-1914):
 *     internal scratch is kept per (device, stream, host thread).
 */
#ifndef RTSCAN_HIP_H_
#define RTSCAN_HIP_H_

#include <hip/hip_runtime_api.h>
#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

/* Default arguments of the reference's C++ declarations (bindex.h:145-183), kept for C++
 * callers; C callers pass every argument. */
#ifdef __cplusplus
#define RTS_DEFAULT(value) = value
#else
#define RTS_DEFAULT(value)
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- types (bindex.h:59-100) */
typedef uint32_t CODE;     /* bindex.h:59 */
typedef int POSTYPE;       /* bindex.h:86 */
typedef unsigned int BITS; /* bindex.h:87 */

/* bindex.h:88-95 */
enum OPERATOR {
  EQ = 0, /* x == a */
  LT,     /* x <  a */
  GT,     /* x >  a */
  LE,     /* x <= a */
  GE,     /* x >= a */
  BT      /* a < x < b (open interval; a > b is swapped by the query driver) */
};

/* ------------------------------------------------------ cuda_and.cu surface (bitmap ops) */

/* cuda_and.cu:255-296: a[i] &= b[i] over bits_num_needed(n) words; n counts ROWS. */
hipError_t GPUbitAndWithCuda(BITS *dev_bitmap_a, BITS *dev_bitmap_b, unsigned int n);
/* cuda_and.cu:298-338: a[i] |= b[i]; n counts ROWS. */
hipError_t GPUbitOrWithCuda(BITS *dev_bitmap_a, BITS *dev_bitmap_b, unsigned int n);
/* cuda_and.cu:340-380: a[i] = b[i]; n counts WORDS. */
hipError_t GPUbitCopyWithCuda(BITS *dev_bitmap_a, BITS *dev_bitmap_b, unsigned int n);
/* cuda_and.cu:382-422: a[i] = ~b[i]; n counts WORDS. */
hipError_t GPUbitCopyNegationWithCuda(BITS *dev_bitmap_a, BITS *dev_bitmap_b, unsigned int n);
/* cuda_and.cu:424-463: result[i] = b[i] & ~a[i]; bitnum counts WORDS. */
hipError_t GPUbitCopySIMDWithCuda(BITS *result, BITS *dev_bitmap_a, BITS *dev_bitmap_b,
                                  unsigned int bitnum);
/* cuda_and.cu:468-508 (defined, not declared, in the reference): a[i] = value over
 * bits_num_needed(n) words; n counts ROWS. */
hipError_t GPUbitMemsetWithCuda(BITS *dev_bitmap_a, BITS value, unsigned int n);
/* cuda_and.cu:168-253: host bitmaps; a &= b over bits_num_needed(size) words. */
hipError_t CPUbitAndWithCuda(BITS *bitmap_a, BITS *bitmap_b, unsigned int size);
/* cuda_and.cu:101-166: flips rows 0..size-1 of a host bitmap of `size` words.  The
 * reference kernel (cuda_and.cu:3-10) races 32 threads on one word without atomics; this
 * one flips every bit exactly once. */
hipError_t refineWithCuda(CODE *bitmap, unsigned int size);

/* ------------------------------------------------------------- bindex.h surface (scan) */

/* raw_scan (bindex.h:146, bindex.cpp:1851-1892): the reference's ground-truth scan.  Every
 * row of the HOST column `raw_data` (length bindex->length) that satisfies
 * `OP` (t1, t2) XOR-toggles its bit in the HOST bitmap `bitmap`.  With compare_bitmap
 * non-NULL the reference only inspects and leaves `bitmap` unchanged; so does this one.
 * Evaluated on the GPU (upload, scan kernel in XOR mode, download).  The BinDex argument
 * is only read for its `length`; see rts_bindex.h for the struct. */
struct BinDex;
void raw_scan(struct BinDex *bindex, BITS *bitmap, CODE target1, CODE target2,
              enum OPERATOR OP, CODE *raw_data, BITS *compare_bitmap RTS_DEFAULT(NULL));

/* ---------------------------------------------------------- additive: device helpers */

/* Synchronous helpers: each returns after its work has finished, so work then enqueued on
 * any stream (non-blocking ones included) sees its result. */
hipError_t rts_malloc(void **ptr, size_t bytes);
hipError_t rts_free(void *ptr);
hipError_t rts_memcpy_htod(void *dst, const void *src, size_t bytes);
hipError_t rts_memcpy_dtoh(void *dst, const void *src, size_t bytes);
hipError_t rts_memcpy_dtod(void *dst, const void *src, size_t bytes);
hipError_t rts_memset(void *dst, int byte_value, size_t bytes);
/* Result egress: page-locked host buffers (hipHostMalloc) and copies queued on a stream, so
 * a bitmap's download overlaps the next query's kernels (the reference allocates pinned
 * memory and copies synchronously, bindex_cuda.cpp:1933-1946). */
hipError_t rts_host_alloc(void **ptr, size_t bytes);
hipError_t rts_host_free(void *ptr);
hipError_t rts_memcpy_dtoh_async(void *dst, const void *src, size_t bytes, hipStream_t stream);
hipError_t rts_memcpy_htod_async(void *dst, const void *src, size_t bytes, hipStream_t stream);
hipError_t rts_device_synchronize(void);
hipError_t rts_stream_create(hipStream_t *stream);
hipError_t rts_stream_destroy(hipStream_t stream);
hipError_t rts_stream_synchronize(hipStream_t stream);
hipError_t rts_event_create(hipEvent_t *event);
hipError_t rts_event_destroy(hipEvent_t event);
hipError_t rts_event_record(hipEvent_t event, hipStream_t stream);
hipError_t rts_event_synchronize(hipEvent_t event);
hipError_t rts_event_elapsed_ms(float *ms, hipEvent_t start, hipEvent_t stop);
/* Device count, name and compute-unit count of the current device. */
int rts_device_count(void);
hipError_t rts_device_info(char *name, size_t name_len, int *cu_count, size_t *total_mem);
/* Free and total device memory of the current device (hipMemGetInfo). */
hipError_t rts_device_mem_info(size_t *free_bytes, size_t *total_bytes);
/* Message of the last error recorded by this library on the calling thread. */
const char *rts_last_error(void);
/* Scratch buffers (count partials, checker parts) the calling thread holds: one per
 * (kind, device, stream) it has used, at most 16 (the oldest is freed first); a thread's
 * buffers are freed when it exits.  Counted and checking calls allocate on their first use
 * of a stream, so that first call is not for graph capture (rts_conjunction_scan_noalloc
 * is). */
int rts_thread_scratch_buffers(void);

/* ------------------------------------------------------------ additive: scan kernels */

/* Normalise a reference predicate (OPERATOR, t1, t2) to a closed interval [*lo, *hi];
 * returns 0 when no uint32 satisfies it (empty), 1 otherwise.  BT is the open interval of
 * the swapped bounds (bindex.cpp:1903-1905, 1857-1874). */
int rts_predicate_interval(enum OPERATOR op, CODE t1, CODE t2, CODE *lo, CODE *hi);

/* Kernel (a): direct streaming compare of one device column against the closed interval
 * [lo, hi] (lo > hi means empty), writing the n-row bitmap `out` (ceil(n/32) words, tail
 * bits 0).  `count` (device, may be NULL) receives the number of matching rows. */
hipError_t rts_scan_range(const CODE *column, uint64_t n, CODE lo, CODE hi, BITS *out,
                          unsigned long long *count, hipStream_t stream);

/* Many closed-range predicates over one device column, the column read once per 64
 * predicates (the reference scans one predicate per call, raw_scan bindex.cpp:1851-1892,
 * and its driver loops over the query file's predicates, bindex_cuda.cpp:1829-1935):
 * bitmap p, ceil(n/32) words at out + p * out_stride_words, gets the rows with
 * lo[p] <= x <= hi[p] (lo[p] > hi[p]: no row), tail bits 0; `counts` (device, npred
 * entries, may be NULL) the number of matching rows of each.  `lo`, `hi` are host arrays
 * of npred >= 1 entries; out_stride_words >= ceil(n/32) when npred > 1. */
hipError_t rts_scan_ranges(const CODE *column, uint64_t n, const CODE *lo, const CODE *hi, int npred,
                           BITS *out, uint64_t out_stride_words, unsigned long long *counts,
                           hipStream_t stream);

/* As rts_scan_range, but XOR-toggles the matching rows into `out` (raw_scan semantics). */
hipError_t rts_scan_range_xor(const CODE *column, uint64_t n, CODE lo, CODE hi, BITS *out,
                              hipStream_t stream);

/* Kernel (c), direct form: AND over `ncols` columns of closed-interval predicates and a
 * popcount in one pass (no per-column bitmaps).  `columns`, `lo`, `hi` are host arrays of
 * ncols entries (device column pointers, bounds).  ncols >= 1. */
hipError_t rts_conjunction_scan(const CODE *const *columns, const CODE *lo, const CODE *hi,
                                int ncols, uint64_t n, BITS *out, unsigned long long *count,
                                hipStream_t stream);

/* No counterpart in the reference.  For callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): rts_conjunction_scan with
 * the count's per-workgroup partials in a fixed 256 KB device array of this library (one per
 * GPU, held until the process ends) instead of per-thread scratch, so no call allocates.
 * n <= RTS_NOALLOC_MAX_ROWS (2^40 rows; more is refused with hipErrorInvalidValue).  Every
 * call on one GPU shares that GPU's array and each counted call rewrites it before reading
 * it: calls that overlap on one GPU (from several streams or host threads) race, so order
 * them.  Asynchronous on `stream`, like rts_conjunction_scan. */
#define RTS_NOALLOC_MAX_ROWS (1ull << 40)
hipError_t rts_conjunction_scan_noalloc(const CODE *const *columns, const CODE *lo, const CODE *hi,
                                        int ncols, uint64_t n, BITS *out, unsigned long long *count,
                                        hipStream_t stream);

/* Kernel (c), bitmap form: out = AND of `nbitmaps` n-row device bitmaps, popcount of the
 * first n rows into `count` (may be NULL), one pass.  Replaces the chain of
 * GPUbitAndWithCuda calls in merge_with_GPU (bindex_cuda.cpp:1693-1702) plus the host
 * hit count.  `bitmaps` is a host array; `out` may alias bitmaps[0]. */
hipError_t rts_bitmap_and_popcount(const BITS *const *bitmaps, int nbitmaps, uint64_t n,
                                   BITS *out, unsigned long long *count, hipStream_t stream);

/* Popcount of the first n rows of a device bitmap into *count (device). */
hipError_t rts_bitmap_popcount(const BITS *bitmap, uint64_t n, unsigned long long *count,
                               hipStream_t stream);

/* -------------------------------------------------- additive: seeded column generators */

/* Distributions of the synthetic workloads. */
enum rts_dist {
  RTS_DIST_UNIFORM = 0, /* uniform on [0, 2^32) */
  RTS_DIST_NORMAL = 1,  /* round(N(p0, p1)) clipped to [0, 2^32-1] */
  RTS_DIST_ZIPF = 2     /* Zipf(s = p0) over ranks 1..p1, rank -> distinct random key */
};

/* Fill host array `out` (n values) with column `column_id` of a seeded table.  The value
 * of row i depends only on (dist, seed, column_id, i, p0, p1), so any thread split gives
 * the same column.  Returns 0, or -1 on bad parameters. */
int rts_generate_column(int dist, uint64_t seed, int column_id, uint64_t n, double p0, double p1,
                        CODE *out, int threads);

#ifdef __cplusplus
}
#endif
#endif /* RTSCAN_HIP_H_ */
