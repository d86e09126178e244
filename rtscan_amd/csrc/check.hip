// check.hip -- bitmap comparison (compare_bitmap semantics) and digest on the device.
//
// One pass over the words: each workgroup owns a contiguous chunk and reports its popcounts,
// its first mismatching word and the popcounts of its words up to that mismatch; the host
// combines the (few hundred) per-workgroup records in order.
#include <hip/hip_runtime.h>

#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <thread>
#include <tuple>
#include <vector>

#include "common.h"
#include "rts_check.h"

namespace rts {
namespace {

constexpr int kCheckBlock = 256;
constexpr unsigned kCheckMaxBlocks = 4096;
constexpr uint64_t kNone = ~0ull;

struct ComparePart {
  unsigned long long truth, both, got;  // whole chunk
  unsigned long long truth_upto, both_upto;  // chunk words up to the first mismatch, its row included
  unsigned long long first_word;        // first mismatching word of the chunk, kNone: none
};

__device__ __forceinline__ uint32_t word_at(const BITS *b, uint64_t w, uint64_t words, uint32_t last) {
  const uint32_t v = b[w];
  return w == words - 1 ? v & last : v;
}

__device__ unsigned long long block_sum(unsigned long long v, unsigned long long *sh) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  __syncthreads();
  if ((threadIdx.x & (kWave - 1)) == 0) sh[threadIdx.x / kWave] = v;
  __syncthreads();
  unsigned long long t = 0;
  for (int i = 0; i < kCheckBlock / kWave; i++) t += sh[i];
  return t;
}

__device__ unsigned long long block_min(unsigned long long v, unsigned long long *sh) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long u = __shfl_xor(v, o, kWave);
    v = u < v ? u : v;
  }
  __syncthreads();
  if ((threadIdx.x & (kWave - 1)) == 0) sh[threadIdx.x / kWave] = v;
  __syncthreads();
  unsigned long long t = kNone;
  for (int i = 0; i < kCheckBlock / kWave; i++) t = sh[i] < t ? sh[i] : t;
  return t;
}

__global__ __launch_bounds__(kCheckBlock) void k_compare(const BITS *truth, const BITS *got, uint64_t words,
                                                         uint32_t last, uint64_t chunk, ComparePart *parts) {
  __shared__ unsigned long long sh[kCheckBlock / kWave];
  const uint64_t w0 = (uint64_t)blockIdx.x * chunk;
  const uint64_t w1 = w0 + chunk < words ? w0 + chunk : words;
  unsigned long long pt = 0, pb = 0, pg = 0, fm = kNone;
#pragma unroll 4
  for (uint64_t w = w0 + threadIdx.x; w < w1; w += kCheckBlock) {
    const uint32_t a = word_at(truth, w, words, last), b = word_at(got, w, words, last);
    pt += __popc(a);
    pb += __popc(a & b);
    pg += __popc(b);
    if (a != b && w < fm) fm = w;
  }
  const unsigned long long t = block_sum(pt, sh), bo = block_sum(pb, sh), g = block_sum(pg, sh);
  const unsigned long long first = block_min(fm, sh);
  unsigned long long tu = 0, bu = 0;
  if (first != kNone) {  // recount the chunk's words before the mismatch (now in cache)
    unsigned long long ut = 0, ub = 0;
    for (uint64_t w = w0 + threadIdx.x; w < first; w += kCheckBlock) {
      const uint32_t a = word_at(truth, w, words, last), b = word_at(got, w, words, last);
      ut += __popc(a);
      ub += __popc(a & b);
    }
    tu = block_sum(ut, sh);
    bu = block_sum(ub, sh);
    // the mismatch word itself, rows up to and including the first differing one
    const uint32_t a = word_at(truth, first, words, last), b = word_at(got, first, words, last);
    const uint32_t upto = ~0u << (31 - __clz(a ^ b));  // MSB-first: top (row + 1) bits
    tu += __popc(a & upto);
    bu += __popc(a & b & upto);
  }
  if (threadIdx.x == 0) parts[blockIdx.x] = ComparePart{t, bo, g, tu, bu, first};
}

// The per-block parts of k_compare folded into compare_bitmap's result on the device, so
// that only the result crosses to the host: counts over all blocks; hits up to the first
// block with a mismatch (that block up to its first differing row) or over all blocks; the
// first differing row from its word pair.  One workgroup of kFinalThreads.
constexpr int kFinalThreads = 1024;

__device__ unsigned long long final_sum(unsigned long long v, unsigned long long *sh) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  __syncthreads();
  if ((threadIdx.x & (kWave - 1)) == 0) sh[threadIdx.x / kWave] = v;
  __syncthreads();
  unsigned long long t = 0;
  for (int i = 0; i < kFinalThreads / kWave; i++) t += sh[i];
  return t;
}

__device__ unsigned long long final_min(unsigned long long v, unsigned long long *sh) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long u = __shfl_xor(v, o, kWave);
    v = u < v ? u : v;
  }
  __syncthreads();
  if ((threadIdx.x & (kWave - 1)) == 0) sh[threadIdx.x / kWave] = v;
  __syncthreads();
  unsigned long long t = kNone;
  for (int i = 0; i < kFinalThreads / kWave; i++) t = sh[i] < t ? sh[i] : t;
  return t;
}

__global__ __launch_bounds__(kFinalThreads) void k_compare_finalize(const ComparePart *parts, unsigned blocks,
                                                                    const BITS *truth, const BITS *got,
                                                                    uint64_t words, uint32_t last,
                                                                    rts_compare_result *out) {
  __shared__ unsigned long long sh[kFinalThreads / kWave];
  unsigned long long t = 0, g = 0, fb = kNone;
  for (unsigned i = threadIdx.x; i < blocks; i += kFinalThreads) {
    t += parts[i].truth;
    g += parts[i].got;
    if (parts[i].first_word != kNone && i < fb) fb = i;
  }
  const unsigned long long truth_count = final_sum(t, sh), got_count = final_sum(g, sh);
  const unsigned long long first_block = final_min(fb, sh);
  const unsigned long long upto = first_block == kNone ? blocks : first_block;  // whole blocks
  unsigned long long ht = 0, hb = 0;
  for (unsigned i = threadIdx.x; i < upto; i += kFinalThreads) {
    ht += parts[i].truth;
    hb += parts[i].both;
  }
  const unsigned long long total_hit = final_sum(ht, sh), true_hit = final_sum(hb, sh);
  if (threadIdx.x == 0) {
    rts_compare_result r{total_hit, true_hit, -1, truth_count, got_count};
    if (first_block != kNone) {
      const ComparePart &p = parts[first_block];
      r.total_hit += p.truth_upto;
      r.true_hit += p.both_upto;
      const uint32_t a = word_at(truth, p.first_word, words, last), b = word_at(got, p.first_word, words, last);
      r.first_mismatch = (long long)(p.first_word * 32 + (uint64_t)__clz(a ^ b));
    }
    *out = r;
  }
}

__device__ __forceinline__ unsigned long long mix64(unsigned long long x) {
  x ^= 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

__global__ __launch_bounds__(kCheckBlock) void k_digest(const BITS *bm, uint64_t words, uint32_t last,
                                                        uint64_t chunk, unsigned long long *parts) {
  __shared__ unsigned long long sh[kCheckBlock / kWave];
  const uint64_t w0 = (uint64_t)blockIdx.x * chunk;
  const uint64_t w1 = w0 + chunk < words ? w0 + chunk : words;
  unsigned long long d = 0;
#pragma unroll 4
  for (uint64_t w = w0 + threadIdx.x; w < w1; w += kCheckBlock)
    d += mix64((w << 32) | word_at(bm, w, words, last));
  d = block_sum(d, sh);
  if (threadIdx.x == 0) parts[blockIdx.x] = d;
}

void grid_for(uint64_t words, unsigned *blocks, uint64_t *chunk) {
  const uint64_t want = (words + 2047) / 2048;  // 8 words per thread: enough workgroups in flight
  *blocks = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(kCheckMaxBlocks, want));
  *chunk = (words + *blocks - 1) / *blocks;
  *blocks = (unsigned)((words + *chunk - 1) / *chunk);
}

// Per-block partials of one check, on the device and in pinned host memory, kept per
// (device, stream, host thread) by thread_scratch, so that a check allocates nothing after
// its first call (hipFree would also synchronise the device) and two threads never share
// them.
struct CheckScratch {
  void *dev = nullptr;
  void *host = nullptr;
};

// the parts of one check, then (compare) its folded result
constexpr size_t kCheckScratchBytes = kCheckMaxBlocks * sizeof(ComparePart) + sizeof(rts_compare_result);

hipError_t check_scratch(hipStream_t stream, CheckScratch *out) {
  return thread_scratch(kScratchCheck, stream, kCheckScratchBytes, kCheckScratchBytes, &out->dev, &out->host);
}

}  // namespace

hipError_t bitmap_compare(const BITS *truth, const BITS *got, int64_t n, rts_compare_result *res, hipStream_t s) {
  if (!truth || !got || !res || n < 1) return hipErrorInvalidValue;
  const uint64_t words = words_for_rows((uint64_t)n);
  unsigned blocks;
  uint64_t chunk;
  grid_for(words, &blocks, &chunk);
  CheckScratch sc;
  hipError_t e = check_scratch(s, &sc);
  if (e != hipSuccess) return e;
  ComparePart *d_parts = static_cast<ComparePart *>(sc.dev);
  rts_compare_result *d_res = reinterpret_cast<rts_compare_result *>(d_parts + kCheckMaxBlocks);
  hipLaunchKernelGGL(k_compare, dim3(blocks), dim3(kCheckBlock), 0, s, truth, got, words, tail_mask((uint64_t)n),
                     chunk, d_parts);
  e = hipGetLastError();
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_compare_finalize, dim3(1), dim3(kFinalThreads), 0, s, d_parts, blocks, truth, got, words,
                       tail_mask((uint64_t)n), d_res);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(sc.host, d_res, sizeof(rts_compare_result), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return e;
  *res = *static_cast<const rts_compare_result *>(sc.host);
  return hipSuccess;
}

hipError_t bitmap_digest(const BITS *bm, int64_t n, unsigned long long *digest, hipStream_t s) {
  if (!bm || !digest || n < 1) return hipErrorInvalidValue;
  const uint64_t words = words_for_rows((uint64_t)n);
  unsigned blocks;
  uint64_t chunk;
  grid_for(words, &blocks, &chunk);
  CheckScratch sc;
  hipError_t e = check_scratch(s, &sc);
  if (e != hipSuccess) return e;
  unsigned long long *d_parts = static_cast<unsigned long long *>(sc.dev);
  hipLaunchKernelGGL(k_digest, dim3(blocks), dim3(kCheckBlock), 0, s, bm, words, tail_mask((uint64_t)n), chunk,
                     d_parts);
  e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(sc.host, d_parts, blocks * sizeof(unsigned long long), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return e;
  const std::vector<unsigned long long> parts(static_cast<const unsigned long long *>(sc.host),
                                              static_cast<const unsigned long long *>(sc.host) + blocks);
  unsigned long long d = 0;
  for (unsigned long long p : parts) d += p;
  *digest = d;
  return hipSuccess;
}

}  // namespace rts

extern "C" hipError_t rts_bitmap_compare(const BITS *truth, const BITS *got, int64_t n, rts_compare_result *res,
                                         hipStream_t stream) {
  return rts::set_error(rts::bitmap_compare(truth, got, n, res, stream), "rts_bitmap_compare");
}

extern "C" hipError_t rts_bitmap_digest(const BITS *bitmap, int64_t n, unsigned long long *digest,
                                        hipStream_t stream) {
  return rts::set_error(rts::bitmap_digest(bitmap, n, digest, stream), "rts_bitmap_digest");
}

extern "C" void compare_bitmap(BITS *bitmap_a, BITS *bitmap_b, int len, CODE **raw_data, int bindex_num) {
  if (!bitmap_a || !bitmap_b || len < 1) {
    printf("[CHECK]hit 0/0\n");
    return;
  }
  const size_t bytes = rts::words_for_rows((uint64_t)len) * sizeof(BITS);
  BITS *da = nullptr, *db = nullptr;
  rts_compare_result r{};
  hipError_t e = hipMalloc((void **)&da, bytes);
  if (e == hipSuccess) e = hipMalloc((void **)&db, bytes);
  if (e == hipSuccess) e = hipMemcpy(da, bitmap_a, bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(db, bitmap_b, bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = rts::bitmap_compare(da, db, len, &r, rts::ref_stream());
  (void)hipFree(da);
  (void)hipFree(db);
  if (e != hipSuccess) {
    (void)rts::set_error(e, "compare_bitmap");
    fprintf(stderr, "[ERROR]compare_bitmap: %s\n", hipGetErrorString(e));
    return;
  }
  if (r.first_mismatch >= 0) {
    const long long i = r.first_mismatch;
    const uint32_t bit = 1u << (31 - (uint32_t)(i % 32));
    printf("[ERROR] check error in raw_data[%lld]=", i);
    if (raw_data) {
      const int cols = bindex_num < 3 ? bindex_num : 3;
      for (int c = 0; c < cols; c++) printf(c ? " / %u" : " %u", raw_data[c][i]);
    }
    printf(" \n");
    printf("the correct is %#x, but we have %#x\n", bitmap_a[i / 32] & bit, bitmap_b[i / 32] & bit);
  }
  printf("[CHECK]hit %llu/%llu\n", r.true_hit, r.total_hit);
  fflush(stdout);
}
