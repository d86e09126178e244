// scan.hip -- kernel (a) direct streaming range scan, kernel (c) fused conjunction /
// AND+popcount, and the bitmap operations of the cuda_and.cu surface.
//
// Layout: a wave64 reads 256 consecutive rows per load instruction as one uint4 (4 rows)
// per lane, so lane l holds rows 4l..4l+3 of the 256.  Output word g (rows 32g..32g+31)
// is spread over the 8 lanes 8g..8g+7, 4 rows each: each lane places its 4 match bits at
// bit 31-4*(l%8)-j, an OR over the 8 lanes (DPP, no LDS) assembles the MSB-first word,
// and one lane per group stores it.
#include <algorithm>

#include "common.h"

namespace rts {

template <int C>
struct ColPreds {
  const uint32_t *col[C];
  uint32_t lo[C];
  uint32_t span[C];  // hi - lo
};

// One row per lane for rows [r0, r0 + 64): 2 MSB-first words (ballot + bit reverse),
// stored by lanes 0 and 1; rows >= n contribute 0 bits.  Returns the popcount the lane
// stored (0 on other lanes).
template <int C, int MODE>
__device__ __forceinline__ uint32_t scan_chunk64(const ColPreds<C> &p, uint64_t r0, uint64_t n,
                                                 BITS *out, int lane) {
  const uint64_t r = r0 + lane;
  bool m = r < n;
  if (m) {
#pragma unroll
    for (int c = 0; c < C; c++) m = m && (p.col[c][r] - p.lo[c]) <= p.span[c];
  }
  const uint64_t b = __ballot(m);
  uint32_t pc = 0;
  if (lane < 2) {
    const uint64_t w = r0 / 32 + lane;
    if (w * 32 < n) {
      const uint32_t v = __brev(lane == 0 ? (uint32_t)b : (uint32_t)(b >> 32));
      pc = __popc(store_word<MODE>(out, w, v));
    }
  }
  return pc;
}

// Whole tiles of 256*U rows with 128-bit loads (all columns 16-byte aligned); the rows
// past the last whole tile (fewer than 256*U) are done by the grid's last block one row
// per lane.  One launch per scan, count included.
template <int C, int U, int MODE, bool COUNT>
__global__ __launch_bounds__(kBlock) void k_scan_tiles(ColPreds<C> p, uint64_t n, BITS *out,
                                                       CountSink sink) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t wave = (uint64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
  const uint64_t nwaves = (uint64_t)gridDim.x * (kBlock / kWave);
  const uint64_t ntiles = n / (uint64_t)(256 * U);
  const uint32_t shift = 28 - 4 * (lane & 7);
  const bool writer = (lane & 7) == 0;
  uint32_t pc = 0;
  for (uint64_t t = wave; t < ntiles; t += nwaves) {
    const uint64_t row0 = t * (uint64_t)(256 * U);
    uint32_t nib[U];
#pragma unroll
    for (int u = 0; u < U; u++) nib[u] = 0xF;
#pragma unroll
    for (int c = 0; c < C; c++) {
      const uint4 *src = reinterpret_cast<const uint4 *>(p.col[c] + row0) + lane;
      uint4 v[U];
#pragma unroll
      for (int u = 0; u < U; u++) v[u] = load_stream(src + 64 * u);
#pragma unroll
      for (int u = 0; u < U; u++) nib[u] &= nibble4(v[u], p.lo[c], p.span[c]);
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const uint32_t w = or_lanes8(nib[u] << shift);
      if (writer) {
        const uint32_t stored = store_word<MODE>(out, row0 / 32 + 8 * u + lane / 8, w);
        if (COUNT) pc += __popc(stored);
      }
    }
  }
  if (blockIdx.x == gridDim.x - 1) {
    const uint64_t begin = ntiles * (uint64_t)(256 * U);
    for (uint64_t r0 = begin + 64 * (threadIdx.x / kWave); r0 < n; r0 += kBlock)
      pc += scan_chunk64<C, MODE>(p, r0, n, out, lane);
  }
  if (COUNT) count_block_partial(pc, sink);
}

// Any alignment: one row per lane over all n rows.
template <int C, int MODE, bool COUNT>
__global__ __launch_bounds__(kBlock) void k_scan_rows(ColPreds<C> p, uint64_t n, BITS *out,
                                                      CountSink sink) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t wave = (uint64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
  const uint64_t nwaves = (uint64_t)gridDim.x * (kBlock / kWave);
  const uint64_t nchunks = (n + 63) / 64;
  uint32_t pc = 0;
  for (uint64_t ch = wave; ch < nchunks; ch += nwaves)
    pc += scan_chunk64<C, MODE>(p, ch * 64, n, out, lane);
  if (COUNT) count_block_partial(pc, sink);
}

__global__ void k_zero_u64(unsigned long long *p) { *p = 0ull; }

// ------------------------------------------------------------------ bitmap operations

enum BitOp : int { kCopy = 0, kNot = 1, kAndNot = 2, kAnd = 3, kOr = 4, kFill = 5, kXorWith = 6 };

template <int OP>
__device__ __forceinline__ uint32_t bitop(uint32_t d, uint32_t a, uint32_t b, uint32_t v) {
  if (OP == kCopy) return a;
  if (OP == kNot) return ~a;
  if (OP == kAndNot) return b & ~a;
  if (OP == kAnd) return d & a;
  if (OP == kOr) return d | a;
  if (OP == kFill) return v;
  return d ^ a;
}

template <int OP>
__device__ __forceinline__ uint4 bitop4(uint4 d, uint4 a, uint4 b, uint32_t v) {
  return make_uint4(bitop<OP>(d.x, a.x, b.x, v), bitop<OP>(d.y, a.y, b.y, v),
                    bitop<OP>(d.z, a.z, b.z, v), bitop<OP>(d.w, a.w, b.w, v));
}

constexpr bool reads_dst(int op) { return op == kAnd || op == kOr || op == kXorWith; }
constexpr bool reads_a(int op) { return op != kFill; }
constexpr bool reads_b(int op) { return op == kAndNot; }

// dst[i] = OP(dst[i], a[i], b[i]) over `words` words; VEC: all pointers 16-byte aligned.
template <int OP, bool VEC>
__global__ __launch_bounds__(kBlock) void k_bitop(BITS *dst, const BITS *a, const BITS *b,
                                                  uint64_t words, uint32_t value) {
  const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t nth = (uint64_t)gridDim.x * blockDim.x;
  uint64_t done = 0;
  if (VEC) {
    const uint64_t n4 = words / 4;
    uint4 *d4 = reinterpret_cast<uint4 *>(dst);
    const uint4 *a4 = reinterpret_cast<const uint4 *>(a);
    const uint4 *b4 = reinterpret_cast<const uint4 *>(b);
    for (uint64_t i = tid; i < n4; i += nth) {
      uint4 dv = reads_dst(OP) ? d4[i] : make_uint4(0, 0, 0, 0);
      uint4 av = reads_a(OP) ? a4[i] : make_uint4(0, 0, 0, 0);
      uint4 bv = reads_b(OP) ? b4[i] : make_uint4(0, 0, 0, 0);
      d4[i] = bitop4<OP>(dv, av, bv, value);
    }
    done = n4 * 4;
  }
  for (uint64_t i = done + tid; i < words; i += nth) {
    uint32_t dv = reads_dst(OP) ? dst[i] : 0u;
    uint32_t av = reads_a(OP) ? a[i] : 0u;
    uint32_t bv = reads_b(OP) ? b[i] : 0u;
    dst[i] = bitop<OP>(dv, av, bv, value);
  }
}

// Toggle rows [0, rows) of a bitmap (refineWithCuda's intent, cuda_and.cu:3-10): whole
// words are negated, the partial last word only in its first rows % 32 bits.
__global__ __launch_bounds__(kBlock) void k_flip_rows(BITS *bm, uint64_t rows) {
  const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t nth = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t full = rows / 32;
  for (uint64_t i = tid; i < full; i += nth) bm[i] = ~bm[i];
  if (tid == 0 && (rows & 31)) bm[full] ^= ~(0xFFFFFFFFu >> (rows & 31));
}

// ---------------------------------------------------------------- AND + popcount

constexpr int kMaxAnd = 8;
struct BitmapSet {
  const BITS *bm[kMaxAnd];
};

// out = AND of nb bitmaps over `words` words; the last word is masked by `last_mask`.
// out == nullptr: count only.
template <bool VEC, bool COUNT>
__global__ __launch_bounds__(kBlock) void k_and_popcount(BitmapSet s, int nb, uint64_t words,
                                                         uint32_t last_mask, BITS *out,
                                                         CountSink sink) {
  const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t nth = (uint64_t)gridDim.x * blockDim.x;
  uint32_t pc = 0;
  uint64_t done = 0;
  if (VEC) {
    const uint64_t n4 = words / 4;
    for (uint64_t i = tid; i < n4; i += nth) {
      uint4 acc = reinterpret_cast<const uint4 *>(s.bm[0])[i];
      for (int k = 1; k < nb; k++) {
        uint4 v = reinterpret_cast<const uint4 *>(s.bm[k])[i];
        acc.x &= v.x;
        acc.y &= v.y;
        acc.z &= v.z;
        acc.w &= v.w;
      }
      if (4 * i + 3 == words - 1) acc.w &= last_mask;
      if (out) reinterpret_cast<uint4 *>(out)[i] = acc;
      if (COUNT) pc += __popc(acc.x) + __popc(acc.y) + __popc(acc.z) + __popc(acc.w);
    }
    done = n4 * 4;
  }
  for (uint64_t i = done + tid; i < words; i += nth) {
    uint32_t acc = s.bm[0][i];
    for (int k = 1; k < nb; k++) acc &= s.bm[k][i];
    if (i == words - 1) acc &= last_mask;
    if (out) out[i] = acc;
    if (COUNT) pc += __popc(acc);
  }
  if (COUNT) count_block_partial(pc, sink);
}

// ------------------------------------------------- many predicates over one column

constexpr int kMultiPreds = 64;  // predicates per pass
constexpr int kMultiLaneRows = 32;                      // rows per lane: one bitmap word
constexpr int kMultiTileRows = kWave * kMultiLaneRows;  // 2048 rows per wave tile
constexpr unsigned kMultiMaxGrid = 4096;

struct MultiPreds {
  uint32_t lo[kMultiPreds];
  uint32_t span[kMultiPreds];  // hi - lo
  uint64_t empty;              // bit p: predicate p matches no row
};

// Up to kMultiPreds closed ranges over one column in one pass.  A wave takes a tile of
// 2048 rows, lane l the 32 rows of bitmap word l (eight 16-byte loads of its own 128 bytes;
// the wave's loads cover the tile's 8 KB once), kept in registers; for every predicate p
// the lane shifts its rows' match bits into one MSB-first word (subtract, compare, add
// with carry per row: no cross-lane assembly) and the wave stores its 64 consecutive
// words of bitmap p (out + p * stride) with one instruction.  Rows past the whole tiles
// (or every row, for a column that is not 16-byte aligned) go one row per lane, 64 rows per
// wave step.  The predicates are split over gridDim.y groups of `pper` (workgroup y takes
// [y * pper, y * pper + pper)), so that a short column still fills the device; the groups
// read the same tiles, the later ones from L2.  COUNT: per-workgroup totals per predicate
// (DPP wave totals added in LDS) into partials[p * gridDim.x + blockIdx.x], summed by
// k_multi_counts.
template <bool COUNT>
__global__ __launch_bounds__(kBlock) void k_scan_multi(const uint32_t *col, uint64_t n, uint64_t ntiles,
                                                       MultiPreds mp, int np, int pper, BITS *out,
                                                       uint64_t stride, uint32_t *partials) {
  __shared__ uint32_t cnt[kMultiPreds];
  const int lane = threadIdx.x & (kWave - 1);
  const int pbeg = blockIdx.y * pper, pend = min(np, pbeg + pper);
  if (COUNT) {
    if (threadIdx.x < kMultiPreds) cnt[threadIdx.x] = 0;
    __syncthreads();
  }
  const uint64_t wave = (uint64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
  const uint64_t nwaves = (uint64_t)gridDim.x * (kBlock / kWave);
  for (uint64_t t = wave; t < ntiles; t += nwaves) {
    const uint64_t row0 = t * (uint64_t)kMultiTileRows;
    const uint4 *src = reinterpret_cast<const uint4 *>(col + row0 + (uint64_t)kMultiLaneRows * lane);
    uint32_t x[kMultiLaneRows];
#pragma unroll
    for (int k = 0; k < kMultiLaneRows / 4; k++) {
      const uint4 v = src[k];
      x[4 * k] = v.x;
      x[4 * k + 1] = v.y;
      x[4 * k + 2] = v.z;
      x[4 * k + 3] = v.w;
    }
    BITS *o = out + row0 / 32 + lane;
    for (int p = pbeg; p < pend; p++) {
      uint32_t w = 0;
      if (!((mp.empty >> p) & 1)) {
        const uint32_t lo = mp.lo[p], span = mp.span[p];
#pragma unroll
        for (int r = 0; r < kMultiLaneRows; r++) w = w + w + (uint32_t)(x[r] - lo <= span);
      }
      o[p * stride] = w;
      if (COUNT) {
        const uint32_t c = wave_total(__popc(w));
        if (lane == 0 && c) atomicAdd(&cnt[p], c);
      }
    }
  }
  // rows past the whole tiles: one row per lane, 64 rows per wave step
  const uint64_t begin = ntiles * (uint64_t)kMultiTileRows;
  const uint64_t nchunks = (n - begin + 63) / 64;
  for (uint64_t c = wave; c < nchunks; c += nwaves) {
    const uint64_t r0 = begin + c * 64, r = r0 + lane;
    const uint32_t x = r < n ? col[r] : 0u;
    const uint64_t wd = r0 / 32 + (lane & 1);
    for (int p = pbeg; p < pend; p++) {
      const bool m = r < n && !((mp.empty >> p) & 1) && (x - mp.lo[p]) <= mp.span[p];
      const uint64_t b = __ballot(m);
      const uint32_t word = __brev((lane & 1) ? (uint32_t)(b >> 32) : (uint32_t)b);
      if (lane < 2 && wd * 32 < n) {
        out[p * stride + wd] = word;
        if (COUNT && word) atomicAdd(&cnt[p], (uint32_t)__popc(word));
      }
    }
  }
  if (COUNT) {
    __syncthreads();
    if ((int)threadIdx.x >= pbeg && (int)threadIdx.x < pend)
      partials[(uint64_t)threadIdx.x * gridDim.x + blockIdx.x] = cnt[threadIdx.x];
  }
}

// counts[p] = sum of partials[p * nblocks .. + nblocks); one workgroup per predicate.
__global__ __launch_bounds__(kBlock) void k_multi_counts(const uint32_t *partials, unsigned nblocks,
                                                         unsigned long long *counts) {
  __shared__ unsigned long long part[kBlock / kWave];
  const uint32_t *src = partials + (uint64_t)blockIdx.x * nblocks;
  unsigned long long s = 0;
  for (unsigned i = threadIdx.x; i < nblocks; i += kBlock) s += src[i];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, kWave);
  if ((threadIdx.x & (kWave - 1)) == 0) part[threadIdx.x / kWave] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int i = 0; i < kBlock / kWave; i++) t += part[i];
    counts[blockIdx.x] = t;
  }
}

// ------------------------------------------------------------------ host launchers

static inline bool aligned16(const void *p) { return ((uintptr_t)p & 15) == 0; }

static unsigned grid_for(uint64_t work_items, unsigned per_block, unsigned cap) {
  uint64_t g = (work_items + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

static unsigned stream_grid_cap() { return (unsigned)std::max(1, device_cu_count()) * 8u; }

// The scan kernel runs one 1024-row tile per wave without a grid-stride loop up to this
// many workgroups (16 waves per CU of 256 CUs in flight with room to spare), which measured
// faster than a persistent grid of 8 workgroups per CU.
static unsigned scan_grid_cap() { return (unsigned)std::min<int>(kMaxCountBlocks, std::max(1, device_cu_count()) * 128); }

// A counted scan runs 32 workgroups per compute unit over the tiles instead: each workgroup
// then ends with one count reduction and one accumulator add for ~3 of the uncounted grid's
// workgroups.  On uniform columns scanned back to back: 1e8-row counted scan 68.9 -> 65.2
// us, 4-column conjunction 262 -> 254 us (8 per CU: 67.2 and 257); on the mixed-distribution
// 1e8 x 4 table, queries interleaved with other paths, the conjunctions were unchanged.
static unsigned counted_scan_grid_cap() { return (unsigned)std::min<int>(kMaxCountBlocks, std::max(1, device_cu_count()) * 32); }

template <int OP>
static hipError_t launch_bitop(BITS *dst, const BITS *a, const BITS *b, uint64_t words,
                               uint32_t value, hipStream_t s) {
  if (words == 0) return hipSuccess;
  const bool vec = aligned16(dst) && (!reads_a(OP) || aligned16(a)) && (!reads_b(OP) || aligned16(b));
  const unsigned grid = grid_for(vec ? (words + 3) / 4 : words, kBlock, stream_grid_cap());
  if (vec)
    hipLaunchKernelGGL((k_bitop<OP, true>), dim3(grid), dim3(kBlock), 0, s, dst, a, b, words, value);
  else
    hipLaunchKernelGGL((k_bitop<OP, false>), dim3(grid), dim3(kBlock), 0, s, dst, a, b, words, value);
  return hipGetLastError();
}

constexpr int kTileU = 4;  // uint4 loads per lane per column per tile: 1024 rows per wave tile

template <int C, int MODE, bool COUNT>
static hipError_t launch_scan_c(const ColPreds<C> &p, uint64_t n, BITS *out, CountSink sink,
                                hipStream_t s) {
  bool vec = true;
  for (int c = 0; c < C; c++) vec = vec && aligned16(p.col[c]);
  unsigned grid;
  if (vec) {
    const uint64_t ntiles = n / (256ull * kTileU);
    grid = grid_for(std::max<uint64_t>(ntiles, 1), kBlock / kWave, sink.out ? counted_scan_grid_cap() : scan_grid_cap());
    hipLaunchKernelGGL((k_scan_tiles<C, kTileU, MODE, COUNT>), dim3(grid), dim3(kBlock), 0, s, p,
                       n, out, sink);
  } else {
    grid = grid_for((n + 63) / 64, kBlock / kWave, stream_grid_cap());
    hipLaunchKernelGGL((k_scan_rows<C, MODE, COUNT>), dim3(grid), dim3(kBlock), 0, s, p, n, out,
                       sink);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || !COUNT) return e;
  return launch_count_finalize(sink, grid, s);
}

template <int C, int MODE>
static hipError_t launch_scan(const ColPreds<C> &p, uint64_t n, BITS *out,
                              unsigned long long *count, hipStream_t s, uint32_t *partials) {
  if (count) {
    const CountSink sink = partials ? CountSink{count, partials, nullptr} : stream_count_sink(count, s);
    if (!sink.partials && !sink.acc) return hipErrorOutOfMemory;
    return launch_scan_c<C, MODE, true>(p, n, out, sink, s);
  }
  return launch_scan_c<C, MODE, false>(p, n, out, CountSink{nullptr, nullptr, nullptr}, s);
}

// Empty predicate: all-zero bitmap (set/and modes), zero count.
static hipError_t launch_empty(uint64_t n, BITS *out, unsigned long long *count, int mode,
                               hipStream_t s) {
  if (count) {
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, s, count);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (mode == kStoreXor) return hipSuccess;
  return launch_bitop<kFill>(out, nullptr, nullptr, words_for_rows(n), 0u, s);
}

template <int C>
static hipError_t scan_group(const CODE *const *cols, const CODE *lo, const CODE *hi, uint64_t n,
                             BITS *out, unsigned long long *count, int mode, hipStream_t s,
                             uint32_t *partials) {
  ColPreds<C> p;
  for (int c = 0; c < C; c++) {
    p.col[c] = cols[c];
    p.lo[c] = lo[c];
    p.span[c] = hi[c] - lo[c];
  }
  if (mode == kStoreSet) return launch_scan<C, kStoreSet>(p, n, out, count, s, partials);
  if (mode == kStoreXor) return launch_scan<C, kStoreXor>(p, n, out, count, s, partials);
  return launch_scan<C, kStoreAnd>(p, n, out, count, s, partials);
}

// Conjunction over any number of columns: groups of up to 4 columns per pass, the first
// pass in `mode`, later passes AND into the result; the popcount on the last pass, its
// per-workgroup partials in `partials` (kMaxCountBlocks slots) when given, else in the
// calling thread's scratch for the stream.
hipError_t scan_conjunction(const CODE *const *cols, const CODE *lo, const CODE *hi, int ncols,
                            uint64_t n, BITS *out, unsigned long long *count, int mode,
                            hipStream_t s, uint32_t *partials) {
  if (ncols < 1 || !cols || !lo || !hi) return hipErrorInvalidValue;
  if (mode == kStoreXor && ncols > 4) return hipErrorInvalidValue;
  if (n == 0) return count ? launch_empty(0, out, count, kStoreXor, s) : hipSuccess;
  if (!out) return hipErrorInvalidValue;
  for (int c = 0; c < ncols; c++) {
    if (!cols[c]) return hipErrorInvalidValue;
    if (lo[c] > hi[c]) return launch_empty(n, out, count, mode, s);
  }
  int done = 0;
  while (done < ncols) {
    const int g = std::min(4, ncols - done);
    const bool last = done + g == ncols;
    const int m = done == 0 ? mode : (int)kStoreAnd;
    unsigned long long *cnt = last ? count : nullptr;
    hipError_t e;
    switch (g) {
      case 1: e = scan_group<1>(cols + done, lo + done, hi + done, n, out, cnt, m, s, partials); break;
      case 2: e = scan_group<2>(cols + done, lo + done, hi + done, n, out, cnt, m, s, partials); break;
      case 3: e = scan_group<3>(cols + done, lo + done, hi + done, n, out, cnt, m, s, partials); break;
      default: e = scan_group<4>(cols + done, lo + done, hi + done, n, out, cnt, m, s, partials); break;
    }
    if (e != hipSuccess) return e;
    done += g;
  }
  return hipSuccess;
}

hipError_t scan_ranges(const CODE *col, uint64_t n, const CODE *lo, const CODE *hi, int np, BITS *out,
                       uint64_t stride, unsigned long long *counts, hipStream_t s) {
  if (!col || !lo || !hi || np < 1 || (n && !out)) return hipErrorInvalidValue;
  const uint64_t words = words_for_rows(n);
  if (np > 1 && stride < words) return hipErrorInvalidValue;
  if (n == 0) {
    if (counts) return hipMemsetAsync(counts, 0, (size_t)np * sizeof(unsigned long long), s);
    return hipSuccess;
  }
  const uint64_t ntiles = aligned16(col) ? n / kMultiTileRows : 0;
  const uint64_t waves = std::max<uint64_t>(ntiles, (n - ntiles * kMultiTileRows + 63) / 64);
  const unsigned grid = grid_for(waves, kBlock / kWave, std::min<unsigned>(kMultiMaxGrid, stream_grid_cap() * 2));
  // predicate groups: enough workgroups for 4 per compute unit, at least 8 predicates each
  const unsigned want = (unsigned)std::max(1, device_cu_count()) * 4;
  uint32_t *partials = nullptr;
  if (counts) {
    void *pp = nullptr;
    const hipError_t e = thread_scratch(kScratchMulti, s, (size_t)kMultiPreds * kMultiMaxGrid * sizeof(uint32_t), 0,
                                        &pp, nullptr);
    if (e != hipSuccess) return e;
    partials = static_cast<uint32_t *>(pp);
  }
  for (int p0 = 0; p0 < np; p0 += kMultiPreds) {  // one pass over the column per 64 predicates
    const int k = std::min(kMultiPreds, np - p0);
    MultiPreds mp{};
    for (int p = 0; p < k; p++) {
      if (lo[p0 + p] > hi[p0 + p]) {
        mp.empty |= 1ull << p;
      } else {
        mp.lo[p] = lo[p0 + p];
        mp.span[p] = hi[p0 + p] - lo[p0 + p];
      }
    }
    BITS *o = out + (uint64_t)p0 * stride;
    const int groups = (int)std::max<unsigned>(1, std::min<unsigned>((k + 7) / 8, (want + grid - 1) / grid));
    const int pper = (k + groups - 1) / groups;
    const dim3 g2(grid, (unsigned)((k + pper - 1) / pper));
    if (counts) {
      hipLaunchKernelGGL(k_scan_multi<true>, g2, dim3(kBlock), 0, s, col, n, ntiles, mp, k, pper, o, stride,
                         partials);
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) return e;
      hipLaunchKernelGGL(k_multi_counts, dim3(k), dim3(kBlock), 0, s, partials, grid, counts + p0);
    } else {
      hipLaunchKernelGGL(k_scan_multi<false>, g2, dim3(kBlock), 0, s, col, n, ntiles, mp, k, pper, o, stride,
                         nullptr);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t bitmap_and_popcount(const BITS *const *bms, int nb, uint64_t n, BITS *out,
                               unsigned long long *count, hipStream_t s) {
  if (nb < 1 || !bms || (!out && nb > kMaxAnd)) return hipErrorInvalidValue;
  const uint64_t words = words_for_rows(n);
  if (words == 0) {
    if (!count) return hipSuccess;
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, s, count);
    return hipGetLastError();
  }
  const CountSink counted = count ? stream_count_sink(count, s) : CountSink{nullptr, nullptr, nullptr};
  if (count && !counted.acc) return hipErrorOutOfMemory;
  // Chains of more than kMaxAnd inputs: AND the first kMaxAnd into `out`, continue from it.
  int done = 0;
  while (true) {
    BitmapSet set;
    int k = 0;
    if (done > 0) set.bm[k++] = out;
    while (k < kMaxAnd && done < nb) {
      if (!bms[done]) return hipErrorInvalidValue;
      set.bm[k++] = bms[done++];
    }
    const bool last = done == nb;
    bool vec = !out || aligned16(out);
    for (int i = 0; i < k; i++) vec = vec && aligned16(set.bm[i]);
    const unsigned grid = grid_for(vec ? (words + 3) / 4 : words, kBlock, stream_grid_cap());
    const uint32_t lm = tail_mask(n);
    const CountSink sink = last ? counted : CountSink{nullptr, nullptr, nullptr};
#define RTS_LAUNCH_AND(V, C)                                                                    \
  hipLaunchKernelGGL((k_and_popcount<V, C>), dim3(grid), dim3(kBlock), 0, s, set, k, words, lm, \
                     out, sink)
    if (vec) {
      if (sink.out) RTS_LAUNCH_AND(true, true); else RTS_LAUNCH_AND(true, false);
    } else {
      if (sink.out) RTS_LAUNCH_AND(false, true); else RTS_LAUNCH_AND(false, false);
    }
#undef RTS_LAUNCH_AND
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && sink.out) e = launch_count_finalize(sink, grid, s);
    if (e != hipSuccess || last) return e;
  }
}

hipError_t bitop(int op, BITS *dst, const BITS *a, const BITS *b, uint64_t words, uint32_t value,
                 hipStream_t s) {
  switch (op) {
    case kCopy: return launch_bitop<kCopy>(dst, a, b, words, value, s);
    case kNot: return launch_bitop<kNot>(dst, a, b, words, value, s);
    case kAndNot: return launch_bitop<kAndNot>(dst, a, b, words, value, s);
    case kAnd: return launch_bitop<kAnd>(dst, a, b, words, value, s);
    case kOr: return launch_bitop<kOr>(dst, a, b, words, value, s);
    case kFill: return launch_bitop<kFill>(dst, a, b, words, value, s);
    case kXorWith: return launch_bitop<kXorWith>(dst, a, b, words, value, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t flip_rows(BITS *bm, uint64_t rows, hipStream_t s) {
  if (rows == 0) return hipSuccess;
  const unsigned grid = grid_for(rows / 32 + 1, kBlock, stream_grid_cap());
  hipLaunchKernelGGL(k_flip_rows, dim3(grid), dim3(kBlock), 0, s, bm, rows);
  return hipGetLastError();
}

}  // namespace rts
