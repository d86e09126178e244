// common.h -- shared device helpers for the range-scan kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rtscan_hip.h"

namespace rts {

constexpr int kWave = 64;
constexpr int kBlock = 256;  // threads per workgroup: 4 waves

// Row i of a bitmap is bit (1u << (31 - i % 32)) of word i / 32 (MSB-first, bindex.cpp:269).
__host__ __device__ inline uint64_t words_for_rows(uint64_t n) { return (n + 31) / 32; }

// bits_num_needed (bindex.h:140-143): one word even for n == 0.
__host__ __device__ inline uint64_t ref_words(int64_t n) { return n <= 0 ? 1 : (uint64_t)((n - 1) / 32 + 1); }

// Mask of the valid (row < n) bits of the last word of an n-row bitmap.
__host__ __device__ inline uint32_t tail_mask(uint64_t n) {
  uint32_t t = (uint32_t)(n & 31);
  return t == 0 ? 0xFFFFFFFFu : ~(0xFFFFFFFFu >> t);
}

// OR of a value over each aligned group of 8 lanes, result in all 8 lanes.  Three DPP
// steps: quad_perm [1,0,3,2] (xor 1), quad_perm [2,3,0,1] (xor 2), row_half_mirror
// (lane i <- lane 7-i within 8 lanes, which pairs each quad with the other).
__device__ __forceinline__ uint32_t or_lanes8(uint32_t x) {
  x |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0xB1, 0xF, 0xF, false);
  x |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x4E, 0xF, 0xF, false);
  x |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x141, 0xF, 0xF, false);
  return x;
}

// Sum over the 64 lanes of a wave (result valid in every lane).
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// 16-byte streaming load (non-temporal: the column is read once per scan).
__device__ __forceinline__ uint4 load_stream(const uint4 *p) {
  const u32x4 v = __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(p));
  return make_uint4(v.x, v.y, v.z, v.w);
}

// Match bits of the 4 rows held by one lane (row j of the uint4 -> bit 3-j), for the
// closed interval [lo, lo + span]: (x - lo) <= span in uint32 arithmetic.
__device__ __forceinline__ uint32_t nibble4(uint4 v, uint32_t lo, uint32_t span) {
  uint32_t b0 = (v.x - lo) <= span;
  uint32_t b1 = (v.y - lo) <= span;
  uint32_t b2 = (v.z - lo) <= span;
  uint32_t b3 = (v.w - lo) <= span;
  return (b0 << 3) | (b1 << 2) | (b2 << 1) | b3;
}

// Store modes for a produced bitmap word.
enum StoreMode : int { kStoreSet = 0, kStoreXor = 1, kStoreAnd = 2 };

// Stores v into word w under MODE and returns the word's new value.
template <int MODE>
__device__ __forceinline__ uint32_t store_word(BITS *out, uint64_t w, uint32_t v) {
  if (MODE == kStoreXor) v ^= out[w];
  if (MODE == kStoreAnd) v &= out[w];
  out[w] = v;
  return v;
}

// Per-stream device scratch for popcount results.  A kernel's blocks add their partial
// counts into `acc` and take a ticket; the last block moves the total to the caller's
// output and resets both fields, so the scratch is zero again when the kernel ends and
// the output needs no clearing launch.  Launches on one stream are ordered, so one
// scratch per stream suffices.
struct CountScratch {
  unsigned long long acc;
  unsigned int ticket;
  unsigned int pad;
};

struct CountSink {
  unsigned long long *out;  // nullptr: no count wanted
  CountScratch *scr;
};

// Adds this block's threads' counts `pc` into the sink; the last block of the grid writes
// the total.  Every block of the grid must call it exactly once.
__device__ __forceinline__ void count_finish(uint32_t pc, CountSink sink) {
  __shared__ uint32_t part[kBlock / kWave];
  __shared__ bool last;
  pc = wave_sum(pc);
  const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
  if (lane == 0) part[wv] = pc;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long s = 0;
    for (int i = 0; i < (int)(blockDim.x / kWave); i++) s += part[i];
    if (s) atomicAdd(&sink.scr->acc, s);
    __threadfence();
    const unsigned int t = atomicAdd(&sink.scr->ticket, 1u);
    last = t == gridDim.x - 1;
  }
  __syncthreads();
  if (last && threadIdx.x == 0) {
    __threadfence();
    const unsigned long long total = atomicExch(&sink.scr->acc, 0ull);
    atomicExch(&sink.scr->ticket, 0u);
    *sink.out = total;
  }
}

// Number of compute units of the current device (cached per device).
int device_cu_count();

// The count scratch of (current device, stream), allocated and zeroed on first use.
CountScratch *stream_count_scratch(hipStream_t stream);

// Record `e` as the calling thread's last error; returns e.
hipError_t set_error(hipError_t e, const char *what);

}  // namespace rts
