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

// Sum over the 64 lanes of a wave by DPP (rows of 16 by quad permutes and rotates, then
// the row broadcasts), the total read from lane 63: a scalar.
__device__ __forceinline__ uint32_t wave_total(uint32_t v) {
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, true);   // quad_perm [1,0,3,2]
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, true);   // quad_perm [2,3,0,1]
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x124, 0xF, 0xF, true);  // row_ror:4
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x128, 0xF, 0xF, true);  // row_ror:8
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false); // row_bcast:15
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false); // row_bcast:31
  return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
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

// Popcounts of a launch, two ways.  Accumulators (the default): every workgroup adds
// (1 << 44) | its count to its group's 64-bit accumulator in one atomic (group =
// blockIdx.x % kCountGroups, the accumulators 128 bytes apart so the groups' atomics go to
// different places); the workgroup whose ticket completes its group adds the group's total
// to the final accumulator the same way, and the one completing that writes the count and
// leaves every accumulator at zero for the next launch.  Every exchange is one atomic, so
// no fence is needed and no second kernel runs.  Partials (the allocation-free entry
// point, whose fixed scratch cannot be trusted to start at zero): every workgroup stores
// its count in its own slot and a one-workgroup kernel sums the slots.
constexpr int kMaxCountBlocks = 1 << 16;
constexpr int kCountGroups = 64;
constexpr int kCountAccStride = 16;  // 128 bytes between accumulators
constexpr size_t kCountAccBytes = (kCountGroups + 1) * kCountAccStride * sizeof(unsigned long long);
constexpr int kCountTicketShift = 44;  // counts below 2^44, at most 2^20 workgroups per group

struct CountSink {
  unsigned long long *out;  // nullptr: no count wanted
  uint32_t *partials;       // kMaxCountBlocks slots, or nullptr
  unsigned long long *acc;  // the accumulators (zero between launches), or nullptr
};

// Block-wide sum of the threads' counts into the launch's count.  Every block of the grid
// (of kBlock threads) must call it exactly once; the grid must not exceed kMaxCountBlocks
// blocks.
__device__ __forceinline__ void count_block_partial(uint32_t pc, CountSink sink) {
  __shared__ uint32_t part[kBlock / kWave];
  pc = wave_total(pc);
  const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
  if (lane == 0) part[wv] = pc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t s = 0;
    for (int i = 0; i < kBlock / kWave; i++) s += part[i];
    if (!sink.acc) {
      sink.partials[blockIdx.x] = s;
      return;
    }
    constexpr unsigned long long kLow = (1ull << kCountTicketShift) - 1;
    const unsigned g = blockIdx.x % kCountGroups;
    const unsigned in_group = (gridDim.x - g - 1) / kCountGroups + 1;  // blocks of group g
    unsigned long long *ga = sink.acc + (size_t)g * kCountAccStride;
    const unsigned long long old = atomicAdd(ga, (1ull << kCountTicketShift) | s);
    if ((old >> kCountTicketShift) + 1 != in_group) return;
    // this block completes group g: every block of the group has added, so the group's
    // accumulator can be cleared for the next launch
    const unsigned long long gsum = (old & kLow) + s;
    atomicExch(ga, 0ull);
    const unsigned groups = gridDim.x < (unsigned)kCountGroups ? gridDim.x : (unsigned)kCountGroups;
    unsigned long long *fa = sink.acc + (size_t)kCountGroups * kCountAccStride;
    const unsigned long long old2 = atomicAdd(fa, (1ull << kCountTicketShift) | gsum);
    if ((old2 >> kCountTicketShift) + 1 != groups) return;
    *sink.out = (old2 & kLow) + gsum;
    atomicExch(fa, 0ull);
  }
}

// Stream of the reference-named wrappers: the calling thread's default stream.  It is a
// blocking stream (ordered after work on the legacy null stream, as the reference's
// wrappers are), and calls from different host threads can run at the same time.  Every
// wrapper synchronises it before returning, as the reference ends with
// cudaDeviceSynchronize (cuda_and.cu:276-287).
inline hipStream_t ref_stream() { return hipStreamPerThread; }

// Scratch of (current device, stream, calling host thread), allocated on first use and
// kept by the calling thread (a bounded number per thread, freed when the thread exits).
// Not for use under graph capture: the first use on a stream allocates.
enum ScratchKind : int { kScratchCheck = 1, kScratchCountAcc = 2, kScratchMulti = 3 };
hipError_t thread_scratch(int kind, hipStream_t stream, size_t dev_bytes, size_t host_bytes, void **dptr,
                          void **hptr, bool *created = nullptr);
size_t thread_scratch_entries();
// The count sink of a counted launch on `stream` writing *out: the calling thread's
// accumulators for the stream (zeroed when first made, on the stream), or sink.acc ==
// nullptr when they cannot be had (the launch then fails with hipErrorOutOfMemory).
CountSink stream_count_sink(unsigned long long *out, hipStream_t stream);

// For a partials sink: sum partials[0, nblocks) into *out on `stream` (one small launch).
// For an accumulator sink: nothing (the counting kernel wrote the count).
hipError_t launch_count_finalize(const CountSink &sink, unsigned nblocks, hipStream_t stream);

// Number of compute units of the current device (cached per device).
int device_cu_count();


// Record `e` as the calling thread's last error; returns e.
hipError_t set_error(hipError_t e, const char *what);

}  // namespace rts
