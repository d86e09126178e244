// bindex_query.hip -- kernel (b): BinDex sieve + boundary refine, fused with the
// multi-column AND and popcount (kernel c), plus the reference's BinDex call surface
// (copy_filter_vector_in_GPU, GPURefineAreaWithCuda, GPURefineEqAreaWithCuda,
// in_which_area, in_which_block).
//
// Query of one column for the closed range [a, b]: rows(a <= x <= b) = L(b+1) & ~L(a) with
// L(v) = {x < v}.  Each L(v) is a filter vector with the rows of one bin's B list that
// have x >= v removed (see bindex_dev.h).  A workgroup owns a tile of kTileWords output
// words: it marks the tile's B-list rows to remove in LDS (LDS atomics, no global
// atomics), then streams the filter-vector words once with 128-bit loads and writes each
// output word once.  Traffic per column: two filter vectors and two bins' list slices.
#include <string.h>

#include <algorithm>

#include "bindex_dev.h"

namespace rts {

hipError_t bindex_build(BinDex *bx, const CODE *col, int64_t n, CODE *owned, int sub, hipStream_t s);
void bindex_free(BinDex *bx);
hipError_t bindex_save(const BinDex *bx, const char *path);
hipError_t bindex_load(BinDex *bx, const char *path, const CODE *col);
hipError_t bindex_host_blocks(BinDex *bx);
hipError_t bitop(int op, BITS *dst, const BITS *a, const BITS *b, uint64_t words, uint32_t value,
                 hipStream_t s);

enum : uint32_t { kBaseZero = 0, kBaseOnes = 1, kBaseFv = 2 };

struct BoundDev {
  const BITS *fv;       // kBaseFv
  const uint32_t *dir;  // tile directory row of the bin (nullptr: no refine)
  const uint2 *bent;    // (row, value) entries
  CODE thr, span;       // list rows with thr <= x <= thr + span are marked
  uint32_t base;
};

template <int C>
struct QueryArgs {
  BoundDev lo[C];  // L(a)
  BoundDev hi[C];  // L(b + 1)
};

__device__ __forceinline__ uint4 base_words(const BoundDev &b, int64_t w, int64_t words) {
  if (b.base == kBaseZero) return make_uint4(0, 0, 0, 0);
  if (b.base == kBaseOnes) return make_uint4(~0u, ~0u, ~0u, ~0u);
  if (w + 3 < words) return *reinterpret_cast<const uint4 *>(b.fv + w);
  uint4 v = make_uint4(0, 0, 0, 0);
  v.x = b.fv[w];
  if (w + 1 < words) v.y = b.fv[w + 1];
  if (w + 2 < words) v.z = b.fv[w + 2];
  return v;
}

// Count of a one-column query [a, b] without its bitmap: rank(b + 1) - rank(a) with
// rank(v) = #{x < v}, each the first slot of the sorted column holding a value >= v.  The
// host narrows each rank to a slot interval [L, R] of fewer than 256 slots (rank_interval,
// from its every-256th-slot samples); one wave then counts the slots of both intervals
// below their bounds: one round of loads.  (Wider intervals take more rounds: each samples
// 256 evenly spaced slots, and the number below v, a prefix since the values are sorted,
// narrows the interval 256-fold.)
struct RankArgs {
  const CODE *sorted;       // nullptr: no rank count
  unsigned long long *out;  // the count
  int64_t L[2], R[2];       // bound 0: rank(a), bound 1: rank(b + 1)
  CODE v[2];
};

__device__ void rank_count(const RankArgs &r) {
  const int lane = threadIdx.x & (kWave - 1);
  int64_t L[2] = {r.L[0], r.L[1]}, R[2] = {r.R[0], r.R[1]};
  // every round leaves an interval of fewer than len / 256 + 1 slots, or none: n < 2^31
  // needs at most four
  for (int round = 0; round < 6 && (L[0] < R[0] || L[1] < R[1]); round++) {
    int64_t step[2];
    CODE x[2][4];
#pragma unroll
    for (int b = 0; b < 2; b++) {
      const int64_t len = R[b] - L[b];
      step[b] = (len + 255) / 256;
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int64_t q = L[b] + (int64_t)(u * kWave + lane) * step[b];
        x[b][u] = len > 0 && q < R[b] ? r.sorted[q] : 0xFFFFFFFFu;
      }
    }
#pragma unroll
    for (int b = 0; b < 2; b++) {
      const int64_t len = R[b] - L[b];
      if (len <= 0) continue;
      int64_t c = 0;  // sampled slots holding a value < v: samples 0 .. c-1
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int64_t q = L[b] + (int64_t)(u * kWave + lane) * step[b];
        c += __popcll(__ballot(q < R[b] && x[b][u] < r.v[b]));
      }
      if (step[b] == 1) {  // every slot sampled: the rank itself
        L[b] += c;
        R[b] = L[b];
      } else if (c == 0) {  // the interval's first value is >= v
        R[b] = L[b];
      } else {  // sample c-1 is < v and sample c (or R) is >= v
        const int64_t nl = L[b] + (c - 1) * step[b] + 1;
        R[b] = std::min<int64_t>(L[b] + c * step[b], R[b]);
        L[b] = nl;
      }
    }
  }
  if (lane == 0) *r.out = (unsigned long long)(L[1] - L[0]);
}

constexpr int kListU = 2;  // list entries per thread per list per round

// A workgroup owns M directory tiles: TQ = M * kTileWords output words, kPer = TQ / kBlock
// words per thread (4 or 8, one or two uint4 per filter vector).
// With rank.sorted set, workgroup 0 writes the count by rank_count (one wave) and the
// others take the tiles; COUNT (popcount of the stored words) is then false.
template <int C, int MODE, bool COUNT, int M>
__global__ __launch_bounds__(kBlock) void k_bindex_query(QueryArgs<C> q, int64_t n, int64_t words,
                                                         int64_t ndir, BITS *out, bool out_vec,
                                                         CountSink sink, RankArgs rank) {
  // One toggle array per column for both bounds: with L(v) = fv ^ t_v the rows below v,
  // L(lo) is inside L(hi + 1), so L(hi + 1) & ~L(lo) = (fv_hi ^ fv_lo) ^ (t_lo ^ t_hi); the two
  // bounds' lists cover different bins (or, both bounds in one bin, a single list) and
  // toggle disjoint rows of the same array.
  constexpr int TQ = M * kTileWords;
  constexpr int kPer = TQ / kBlock;  // 4 or 8 words per thread
  constexpr int kVec = kPer / 4;     // uint4 per filter vector per thread
  static_assert(kPer % 4 == 0, "whole uint4 of output words per thread");
  __shared__ alignas(16) uint32_t tog[C][TQ];
  const uint32_t last_mask = tail_mask((uint64_t)n);
  const int64_t ntiles = (words + TQ - 1) / TQ;
  uint32_t pc = 0;
  int64_t t0 = blockIdx.x, tstride = gridDim.x;
  if (rank.sorted) {
    if (blockIdx.x == 0) {
      if (threadIdx.x < kWave) rank_count(rank);
      return;
    }
    t0 -= 1;
    tstride -= 1;
  }
  for (int64_t t = t0; t < ntiles; t += tstride) {
    // The filter-vector words and the lists' directory words do not depend on the refine:
    // issue them first, so that one round trip covers both and the second round trip is
    // the refine entries.
    const int lw = threadIdx.x * 4;  // the thread's uint4 within each 1024-word part of the tile
    const int64_t d0 = t * M, d1 = d0 + M < ndir ? d0 + M : ndir;  // directory tiles covered
    uint4 lo_w[C][kVec], hi_w[C][kVec];
    uint32_t lo_e[C][2], hi_e[C][2];
#pragma unroll
    for (int c = 0; c < C; c++) {
#pragma unroll
      for (int k = 0; k < kVec; k++) {
        const int64_t w = t * TQ + k * kTileWords + lw;
        if (w < words) {
          lo_w[c][k] = base_words(q.lo[c], w, words);
          hi_w[c][k] = base_words(q.hi[c], w, words);
        }
      }
      lo_e[c][0] = lo_e[c][1] = hi_e[c][0] = hi_e[c][1] = 0;
      if (q.lo[c].dir) {
        lo_e[c][0] = q.lo[c].dir[d0];
        lo_e[c][1] = q.lo[c].dir[d1];
      }
      if (q.hi[c].dir) {
        hi_e[c][0] = q.hi[c].dir[d0];
        hi_e[c][1] = q.hi[c].dir[d1];
      }
    }
    for (int i = threadIdx.x; i < C * TQ; i += kBlock) (&tog[0][0])[i] = 0;
    __syncthreads();
    // The 2C lists' tile slices: every list's next kListU entries per thread are loaded
    // before any of them is applied, so one round trip covers all lists (a tile's slice of
    // one list is about M * kTileRows / bins entries: one or two rounds in practice).
    const uint32_t row0 = (uint32_t)(t * TQ * 32);
    for (uint32_t off = 0;; off += kListU * kBlock) {
      uint2 ent[2 * C][kListU];
      bool more = false;
#pragma unroll
      for (int l = 0; l < 2 * C; l++) {
        const BoundDev &b = (l & 1) ? q.hi[l / 2] : q.lo[l / 2];
        const uint32_t e0 = (l & 1) ? hi_e[l / 2][0] : lo_e[l / 2][0];
        const uint32_t e1 = (l & 1) ? hi_e[l / 2][1] : lo_e[l / 2][1];
#pragma unroll
        for (int u = 0; u < kListU; u++) {
          const uint32_t e = e0 + off + u * kBlock + threadIdx.x;
          ent[l][u] = e < e1 ? b.bent[e] : make_uint2(0xFFFFFFFFu, 0);
        }
        more = more || e0 + off + kListU * kBlock < e1;  // the same in every thread
      }
#pragma unroll
      for (int l = 0; l < 2 * C; l++) {
        const BoundDev &b = (l & 1) ? q.hi[l / 2] : q.lo[l / 2];
#pragma unroll
        for (int u = 0; u < kListU; u++) {
          const uint2 pv = ent[l][u];
          // rows of the slice lie in [row0, row0 + TQ * 32): the word index is < TQ
          if (pv.x != 0xFFFFFFFFu && pv.y - b.thr <= b.span)
            atomicXor(&tog[l / 2][(pv.x - row0) >> 5], 0x80000000u >> (pv.x & 31));
        }
      }
      if (!more) break;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kVec; k++) {
      const int lwk = k * kTileWords + lw;  // this uint4's first word in the tile
      const int64_t w = t * TQ + lwk;
      if (w >= words) continue;
      uint4 acc = make_uint4(~0u, ~0u, ~0u, ~0u);
#pragma unroll
      for (int c = 0; c < C; c++) {
        const uint4 lo = lo_w[c][k], hi = hi_w[c][k];
        const uint4 tg = *reinterpret_cast<const uint4 *>(&tog[c][lwk]);
        acc.x &= hi.x ^ lo.x ^ tg.x;
        acc.y &= hi.y ^ lo.y ^ tg.y;
        acc.z &= hi.z ^ lo.z ^ tg.z;
        acc.w &= hi.w ^ lo.w ^ tg.w;
      }
      uint32_t v[4] = {acc.x, acc.y, acc.z, acc.w};
#pragma unroll
      for (int j = 0; j < 4; j++) {
        if (w + j == words - 1) v[j] &= last_mask;
        if (w + j >= words) v[j] = 0;
      }
      if (MODE == kStoreAnd) {
#pragma unroll
        for (int j = 0; j < 4; j++)
          if (w + j < words) v[j] &= out[w + j];
      }
      if (out_vec && w + 3 < words) {
        *reinterpret_cast<uint4 *>(out + w) = make_uint4(v[0], v[1], v[2], v[3]);
      } else {
#pragma unroll
        for (int j = 0; j < 4; j++)
          if (w + j < words) out[w + j] = v[j];
      }
      if (COUNT) pc += __popc(v[0]) + __popc(v[1]) + __popc(v[2]) + __popc(v[3]);
    }
    __syncthreads();
  }
  if (COUNT) count_block_partial(pc, sink);
}

static BoundDev make_bound(const rts_bindex_dev *d, const BoundSpec &s) {
  BoundDev b{};
  if (s.fv < 0) {
    b.base = kBaseZero;
  } else if (s.fv >= d->bins - 1) {
    b.base = kBaseOnes;
  } else {
    b.base = kBaseFv;
    b.fv = d->fvs + (size_t)s.fv * d->fv_stride;
  }
  if (s.list >= 0) {
    b.dir = d->bdir + (size_t)s.list * (d->ntiles + 1);
    b.bent = d->bent;
    b.thr = s.thr;
    b.span = 0xFFFFFFFFu - s.thr;  // value >= thr
  }
  return b;
}

// Two directory tiles (2048 words, two uint4 per filter vector per thread) per workgroup on
// columns of at least 2^24 rows, for one or two columns: half the workgroups, each with
// twice the loads in flight (1e8 rows: 14.5 -> 13.8 us at 128 bins, 13.9 -> 12.4 us at 512).
// With three or four columns the doubled registers cost occupancy and it measured slower
// (28.3 -> 30.3 us and 33.5 -> 37.3 us), so those keep one tile per workgroup. Four tiles
// for one column measured no faster than two (12.9 vs 12.6 us at 128 bins, 11.4 vs 11.2 at
// 512).
constexpr int64_t kWideQueryRows = (int64_t)1 << 24;

template <int C, int MODE, int M>
static hipError_t launch_query_m(const QueryArgs<C> &q, const rts_bindex_dev *d, BITS *out,
                                 unsigned long long *count, const RankArgs &rank, hipStream_t s) {
  const bool vec = ((uintptr_t)out & 15) == 0;
  const int64_t ntiles = (d->words + M * kTileWords - 1) / (M * kTileWords);
  const unsigned grid = (unsigned)std::min<int64_t>(ntiles, (int64_t)device_cu_count() * 32);
  const CountSink none{nullptr, nullptr, nullptr};
  if (rank.sorted) {  // one more workgroup, for the count
    hipLaunchKernelGGL((k_bindex_query<C, MODE, false, M>), dim3(grid + 1), dim3(kBlock), 0, s, q, d->n,
                       d->words, d->ntiles, out, vec, none, rank);
  } else if (count) {
    const CountSink sink = stream_count_sink(count, s);
    if (!sink.acc) return hipErrorOutOfMemory;
    hipLaunchKernelGGL((k_bindex_query<C, MODE, true, M>), dim3(grid), dim3(kBlock), 0, s, q, d->n,
                       d->words, d->ntiles, out, vec, sink, rank);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_count_finalize(sink, grid, s);
  } else {
    hipLaunchKernelGGL((k_bindex_query<C, MODE, false, M>), dim3(grid), dim3(kBlock), 0, s, q, d->n,
                       d->words, d->ntiles, out, vec, none, rank);
  }
  return hipGetLastError();
}

template <int C, int MODE>
static hipError_t launch_query(const QueryArgs<C> &q, const rts_bindex_dev *d, BITS *out,
                               unsigned long long *count, const RankArgs &rank, hipStream_t s) {
  if (C <= 2 && d->n >= kWideQueryRows) return launch_query_m<C, MODE, 2>(q, d, out, count, rank, s);
  return launch_query_m<C, MODE, 1>(q, d, out, count, rank, s);
}

template <int C>
static hipError_t query_group(const BinDex *const *bxs, const CODE *lo, const CODE *hi, int mode,
                              BITS *out, unsigned long long *count, hipStream_t s) {
  QueryArgs<C> q;
  for (int c = 0; c < C; c++) {
    const rts_bindex_dev *d = bxs[c]->dev;
    const BoundSpec sl = lo[c] == 0 ? BoundSpec{-1, -1, 0} : bound_spec(d, lo[c], false);
    const BoundSpec sh = bound_spec(d, hi[c] + 1u, hi[c] == 0xFFFFFFFFu);
    q.lo[c] = make_bound(d, sl);
    q.hi[c] = make_bound(d, sh);
    int b0, b1;
    if (sl.list >= 0 && sl.list == sh.list) {
      // both bounds in one bin: (1 ^ 1) ^ t = t with t the bin's rows in [lo, hi], one pass
      // over the list and no filter vector
      q.lo[c].base = q.hi[c].base = kBaseOnes;
      q.lo[c].span = hi[c] - lo[c];
      q.hi[c].dir = nullptr;
    } else if (list_form(d, sl, sh, &b0, &b1)) {
      // one or two bins whose lists hold the column's rows: (0 & ~0) ^ t = t with t their
      // rows in [lo, hi] (disjoint bins, so the XOR of the two lists is their union)
      for (int j = 0; j < 2; j++) {
        BoundDev &b = j ? q.hi[c] : q.lo[c];
        b = BoundDev{};
        b.base = kBaseZero;
        if (j == 0 || b1 != b0) {
          b.dir = d->bdir + (size_t)(j ? b1 : b0) * (d->ntiles + 1);
          b.bent = d->bent;
          b.thr = lo[c];
          b.span = hi[c] - lo[c];
        }
      }
    }
  }
  RankArgs rank{};
  if (mode == kStoreAnd) return launch_query<C, kStoreAnd>(q, bxs[0]->dev, out, count, rank, s);
  if (C == 1 && count) {
    // one column: the count from the sorted column, beside the bitmap instead of after it
    const rts_bindex_dev *d = bxs[0]->dev;
    rank.sorted = d->sorted_val;
    rank.out = count;
    rank.v[0] = lo[0];
    rank.v[1] = hi[0] + 1u;
    rank_interval(d, rank.v[0], false, &rank.L[0], &rank.R[0]);
    rank_interval(d, rank.v[1], hi[0] == 0xFFFFFFFFu, &rank.L[1], &rank.R[1]);
  }
  return launch_query<C, kStoreSet>(q, bxs[0]->dev, out, count, rank, s);
}

__global__ void k_zero_count(unsigned long long *p) { *p = 0ull; }

hipError_t bindex_query(const BinDex *const *bxs, const CODE *lo, const CODE *hi, int ncols,
                        BITS *out, unsigned long long *count, hipStream_t s) {
  if (!bxs || !lo || !hi || ncols < 1 || !out) return hipErrorInvalidValue;
  const rts_bindex_dev *d0 = bxs[0] ? bxs[0]->dev : nullptr;
  if (!d0) return hipErrorInvalidValue;
  for (int c = 0; c < ncols; c++) {
    if (!bxs[c] || !bxs[c]->dev || bxs[c]->dev->n != d0->n) return hipErrorInvalidValue;
  }
  for (int c = 0; c < ncols; c++) {
    if (lo[c] > hi[c]) {  // empty predicate: no row qualifies
      if (count) {
        hipLaunchKernelGGL(k_zero_count, dim3(1), dim3(1), 0, s, count);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
      }
      return bitop(5 /* fill */, out, nullptr, nullptr, d0->words, 0u, s);
    }
  }
  int done = 0;
  while (done < ncols) {
    const int g = std::min(4, ncols - done);
    const int mode = done == 0 ? kStoreSet : kStoreAnd;
    unsigned long long *cnt = done + g == ncols ? count : nullptr;
    hipError_t e;
    switch (g) {
      case 1: e = query_group<1>(bxs + done, lo + done, hi + done, mode, out, cnt, s); break;
      case 2: e = query_group<2>(bxs + done, lo + done, hi + done, mode, out, cnt, s); break;
      case 3: e = query_group<3>(bxs + done, lo + done, hi + done, mode, out, cnt, s); break;
      default: e = query_group<4>(bxs + done, lo + done, hi + done, mode, out, cnt, s); break;
    }
    if (e != hipSuccess) return e;
    done += g;
  }
  return hipSuccess;
}

// ------------------------------------------------- refineBlock (cuda_and.cu:74-99)

constexpr int kMaxRefineCols = 16;
struct RefineCols {
  const CODE *raw[kMaxRefineCols];
  CODE lo[kMaxRefineCols];  // exclusive bounds: lo < x < hi
  CODE hi[kMaxRefineCols];
};

// For every sorted slot in [s0, s1) of the selected column: if every column's value at the
// slot's row lies strictly inside its bounds, set (or, inverse, clear) the row's bit.
__global__ __launch_bounds__(kBlock) void k_refine_slots(RefineCols rc, int ncols,
                                                         const int32_t *sorted_pos, int64_t s0,
                                                         int64_t s1, BITS *bitmap, bool inverse) {
  for (int64_t s = s0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < s1;
       s += (int64_t)gridDim.x * blockDim.x) {
    const int32_t p = sorted_pos[s];
    bool ok = true;
    for (int j = 0; j < ncols && ok; j++) {
      const CODE x = rc.raw[j][p];
      ok = rc.lo[j] < x && x < rc.hi[j];
    }
    if (!ok) continue;
    const uint32_t bit = 0x80000000u >> (p & 31);
    if (inverse)
      atomicAnd(&bitmap[p >> 5], ~bit);
    else
      atomicOr(&bitmap[p >> 5], bit);
  }
}

static hipError_t refine_slots(BinDex **bxs, int ncols, int sel, const CODE *predicate,
                               int64_t s0, int64_t s1, BITS *bitmap, bool inverse) {
  if (ncols < 1 || ncols > kMaxRefineCols) return hipErrorInvalidValue;
  RefineCols rc;
  for (int j = 0; j < ncols; j++) {
    if (!bxs[j] || !bxs[j]->rawDataInGPU) return hipErrorInvalidValue;
    rc.raw[j] = bxs[j]->rawDataInGPU;
    rc.lo[j] = predicate[2 * j];
    rc.hi[j] = predicate[2 * j + 1];
  }
  if (s1 <= s0) return hipSuccess;
  const int64_t items = s1 - s0;
  const unsigned grid =
      (unsigned)std::min<int64_t>((items + kBlock - 1) / kBlock, (int64_t)device_cu_count() * 16);
  hipLaunchKernelGGL(k_refine_slots, dim3(grid), dim3(kBlock), 0, ref_stream(), rc, ncols,
                     bxs[sel]->dev->sorted_pos, s0, s1, bitmap, inverse);
  return hipGetLastError();
}

static inline int64_t block_slot(const Area *a, int block) {
  return (int64_t)a->start + (int64_t)block * RTS_BLOCK_INIT_SIZE;
}

static inline int64_t block_end_slot(const Area *a, int block) {
  return std::min<int64_t>(block_slot(a, block) + RTS_BLOCK_INIT_SIZE, (int64_t)a->start + a->length);
}

}  // namespace rts

using namespace rts;

extern "C" {

// in_which_area, bindex.cpp:1138-1151 (dup bindex_cuda.cpp:723-731)
int in_which_area(BinDex *bindex, CODE compare) {
  if (compare < bindex->areaStartValues[0]) return -1;
  for (int i = 0; i < kAreas; i++) {
    const CODE sv = bindex->areaStartValues[i];
    if (sv == compare) return i;
    if (sv > compare) return i - 1;
  }
  return kAreas - 1;
}

// in_which_block, bindex.cpp:1153-1176: the first block starting at `compare`, else the
// last block starting below it; one block earlier if that block ends with `compare`.
int in_which_block(Area *area, CODE compare) {
  int res = area->blockNum - 1;
  for (int i = 0; i < area->blockNum; i++) {
    const CODE sv = area->blockFirst[i];
    if (sv == compare) {
      res = i;
      break;
    }
    if (sv > compare) {
      res = i - 1;
      break;
    }
  }
  if (res > 0 && area->blockLast[res - 1] == compare) res--;
  return res;
}

// on_which_pos, bindex.cpp:1201-1215 (dup bindex_cuda.cpp:756-770): lower bound of
// `compare` in the block's ascending values.
int on_which_pos(pos_block *pb, CODE compare) {
  int lo = 0, hi = pb->length;
  while (lo < hi) {
    const int mid = (lo + hi) / 2;
    if (pb->val[mid] >= compare)
      hi = mid;
    else
      lo = mid + 1;
  }
  return lo;
}

// copy_filter_vector_in_GPU, bindex_cuda.cpp:486-504 (declared bindex.h:145)
void copy_filter_vector_in_GPU(BinDex *bindex, BITS *dev_bitmap, int k, bool negation) {
  if (!bindex || !dev_bitmap) return;
  const uint64_t words = ref_words(bindex->length);
  hipError_t e;
  if (k < 0)
    e = bitop(5, dev_bitmap, nullptr, nullptr, words, 0u, ref_stream());
  else if (k >= kAreas - 1)
    e = bitop(5, dev_bitmap, nullptr, nullptr, words, 0xFFFFFFFFu, ref_stream());
  else
    e = bitop(negation ? 1 : 0, dev_bitmap, bindex->filterVectorsInGPU[k], nullptr, words, 0u, ref_stream());
  if (e == hipSuccess) e = hipStreamSynchronize(ref_stream());
  (void)set_error(e, "copy_filter_vector_in_GPU");
}

// GPURefineAreaWithCuda, cuda_and.cu:510-640.  predicate: host array of 2 * column_num
// exclusive bounds (lo, hi) per column.  The boundary area of the selected column is found
// from its upper bound (its lower bound when the upper bound starts an area), then the
// slots of blocks [0, blockIdx] (inverse: [blockIdx, blockNum)) are refined: rows whose
// every column lies strictly inside its bounds get their bit set (inverse: cleared).
// One launch over the contiguous slot range instead of one launch per block.
hipError_t GPURefineAreaWithCuda(BinDex **bindexs, BITS *dev_result_bitmap, CODE *predicate,
                                 int selected_id, int column_num, bool inverse) {
  if (!bindexs || !dev_result_bitmap || !predicate || selected_id < 0 ||
      selected_id >= column_num || !bindexs[selected_id] || !bindexs[selected_id]->dev)
    return set_error(hipErrorInvalidValue, "GPURefineAreaWithCuda");
  BinDex *b = bindexs[selected_id];
  CODE compare = predicate[2 * selected_id + 1];
  int area = in_which_area(b, compare);
  if (area >= 0 && b->areaStartValues[area] == compare) {
    compare = predicate[2 * selected_id];
    area = in_which_area(b, compare);
  }
  if (area < 0) return set_error(hipErrorInvalidValue, "GPURefineAreaWithCuda: bound below data");
  const Area *a = b->areas[area];
  const int block = in_which_block(b->areas[area], compare);
  int64_t s0, s1;
  if (!inverse) {
    s0 = block_slot(a, 0);
    s1 = block_end_slot(a, block);
  } else {
    s0 = block_slot(a, block);
    s1 = block_end_slot(a, a->blockNum - 1);
  }
  hipError_t e = refine_slots(bindexs, column_num, selected_id, predicate, s0, s1,
                              dev_result_bitmap, inverse);
  if (e == hipSuccess) e = hipStreamSynchronize(ref_stream());
  return set_error(e, "GPURefineAreaWithCuda");
}

// GPURefineEqAreaWithCuda, cuda_and.cu:642-812: refines from the block of the selected
// column's smaller bound to the block of its larger bound, across the areas in between.
hipError_t GPURefineEqAreaWithCuda(BinDex **bindexs, BITS *dev_result_bitmap, CODE *predicate,
                                   int selected_id, int column_num, bool inverse) {
  if (!bindexs || !dev_result_bitmap || !predicate || selected_id < 0 ||
      selected_id >= column_num || !bindexs[selected_id] || !bindexs[selected_id]->dev)
    return set_error(hipErrorInvalidValue, "GPURefineEqAreaWithCuda");
  BinDex *b = bindexs[selected_id];
  CODE left = predicate[2 * selected_id], right = predicate[2 * selected_id + 1];
  if (left > right) std::swap(left, right);
  const int al = in_which_area(b, left), ar = in_which_area(b, right);
  if (ar < 0) return set_error(hipErrorInvalidValue, "GPURefineEqAreaWithCuda: bound below data");
  const int br = in_which_block(b->areas[ar], right);
  int64_t s0;
  if (al >= 0) {
    const int bl = in_which_block(b->areas[al], left);
    s0 = block_slot(b->areas[al], bl);
  } else {
    s0 = block_slot(b->areas[0], 0);
  }
  const int64_t s1 = block_end_slot(b->areas[ar], br);
  hipError_t e = refine_slots(bindexs, column_num, selected_id, predicate, s0, s1,
                              dev_result_bitmap, inverse);
  if (e == hipSuccess) e = hipStreamSynchronize(ref_stream());
  return set_error(e, "GPURefineEqAreaWithCuda");
}

// ------------------------------------------------------------ additive BinDex API

hipError_t rts_bindex_build(BinDex *bindex, const CODE *dev_column, POSTYPE n, hipStream_t stream) {
  return rts_bindex_build_bins(bindex, dev_column, n, RTS_VAREA_N, stream);
}

hipError_t rts_bindex_build_bins(BinDex *bindex, const CODE *dev_column, POSTYPE n, int bins,
                                 hipStream_t stream) {
  if (bins % RTS_VAREA_N != 0) return set_error(hipErrorInvalidValue, "rts_bindex_build_bins: bins");
  hipError_t e = bindex_build(bindex, dev_column, n, nullptr, bins / RTS_VAREA_N, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  return set_error(e, "rts_bindex_build_bins");
}

int rts_bindex_bins(const BinDex *bindex) { return bindex && bindex->dev ? bindex->dev->bins : -1; }

hipError_t rts_bindex_save(const BinDex *bindex, const char *path) {
  return set_error(bindex_save(bindex, path), "rts_bindex_save");
}

hipError_t rts_bindex_load(BinDex *bindex, const char *path, const CODE *dev_column) {
  return set_error(bindex_load(bindex, path, dev_column), "rts_bindex_load");
}

hipError_t rts_bindex_build_host(BinDex *bindex, const CODE *host_column, POSTYPE n) {
  if (!bindex || !host_column || n <= 0) return set_error(hipErrorInvalidValue, "rts_bindex_build_host");
  CODE *d = nullptr;
  hipError_t e = hipMalloc((void **)&d, (size_t)n * sizeof(CODE));
  if (e == hipSuccess) e = hipMemcpy(d, host_column, (size_t)n * sizeof(CODE), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = bindex_build(bindex, d, n, d, 1, ref_stream());
  if (e == hipSuccess) e = hipStreamSynchronize(ref_stream());
  if (e != hipSuccess && d) (void)hipFree(d);
  return set_error(e, "rts_bindex_build_host");
}

void rts_bindex_free(BinDex *bindex) { bindex_free(bindex); }

hipError_t rts_bindex_host_blocks(BinDex *bindex) {
  return set_error(bindex_host_blocks(bindex), "rts_bindex_host_blocks");
}

hipError_t rts_bindex_query(const BinDex *const *bindexes, const CODE *lo, const CODE *hi,
                            int ncols, BITS *out, unsigned long long *count, hipStream_t stream) {
  return set_error(bindex_query(bindexes, lo, hi, ncols, out, count, stream), "rts_bindex_query");
}

hipError_t rts_bindex_export_sorted(const BinDex *bindex, CODE *host_sorted_val,
                                    int32_t *host_sorted_pos) {
  if (!bindex || !bindex->dev) return set_error(hipErrorInvalidValue, "rts_bindex_export_sorted");
  const rts_bindex_dev *d = bindex->dev;
  hipError_t e = hipSuccess;
  if (host_sorted_val)
    e = hipMemcpy(host_sorted_val, d->sorted_val, d->n * sizeof(CODE), hipMemcpyDeviceToHost);
  if (e == hipSuccess && host_sorted_pos)
    e = hipMemcpy(host_sorted_pos, d->sorted_pos, d->n * sizeof(int32_t), hipMemcpyDeviceToHost);
  return set_error(e, "rts_bindex_export_sorted");
}

long long rts_bindex_refine_entries(const BinDex *bindex) {
  return bindex && bindex->dev ? (long long)bindex->dev->nb : -1;
}

long long rts_bindex_device_bytes(const BinDex *bindex) {
  if (!bindex || !bindex->dev) return -1;
  const rts_bindex_dev *d = bindex->dev;
  return (long long)(d->n * 8 + (int64_t)(d->bins - 1) * d->fv_stride * 4 + d->nb * 8 +
                     (int64_t)(d->bins + 1) * 4 + (int64_t)d->bins * (d->ntiles + 1) * 4);
}

}  // extern "C"
