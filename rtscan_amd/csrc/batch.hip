// batch.hip -- batched evaluation of many range queries in one persistent launch.
//
// A query is a conjunction of closed ranges over 1..4 columns of a table (all columns
// have n rows).  Each query-column is evaluated either from the column's BinDex (filter
// vectors + the tile slices of two bins' refine lists, as in k_bindex_query) or directly
// from the column's values (the wave64 DPP word assembly of k_scan_tiles).
//
// Work item = (row tile t, query group g): the workgroup writes the tile's output words of
// every query of the group.  Items are assigned statically so that the items of one tile
// run on the same XCD one after another (workgroup b runs on XCD b % 8 on MI355X, and XCD x
// takes tiles t with t % 8 == x, tile-major): the tile's filter-vector and list slices are
// read from HBM once and re-read by the other query groups from that XCD's L2.  Popcounts
// go to one partial per (query, tile), summed by a second small kernel.
#include <string.h>

#include <algorithm>
#include <vector>

#include "batch.h"
#include "bindex_dev.h"

namespace rts {

enum : uint32_t { kBaseZero = 0, kBaseOnes = 1, kBaseFv = 2 };

struct BatchBound {
  const BITS *fv;       // kBaseFv
  const uint32_t *dir;  // tile directory row of the bin, nullptr: no refine list
  const uint32_t *bpos;
  const CODE *bval;
  CODE thr;
  uint32_t base;
};

struct BatchCol {
  BatchBound lo, hi;  // BinDex form: (fv_hi & ~fv_lo) ^ T, T = refine toggles of both bounds
  const CODE *col;    // direct form when non-null: T = (x - dlo) <= dspan
  CODE dlo, dspan;
};

struct BatchQuery {
  int ncols;
  int empty;  // some column's range is empty: all-zero result
  BITS *out;  // this query's bitmap; nullptr: out + q * out_stride of the launch
  BatchCol c[kBatchMaxCols];
};

// A work item's query group: queries [q0, q0 + nq), their columns numbered as slots in
// order; slot s belongs to query q0 + sq[s], column sc[s].
struct BatchGroup {
  int q0, nq, nslots, pad;
  uint8_t sq[kBatchMaxSlots];
  uint8_t sc[kBatchMaxSlots];
};

// LDS of one workgroup: one toggle array per slot (t_lo ^ t_hi of its two bounds, or the
// predicate words of a direct slot) and the flattened list table of the current item.
template <int TW>
struct BatchLds {
  static constexpr int kSlots = kBatchLdsWords / TW;
  uint32_t tog[kSlots][TW];
  const uint32_t *lpos[2 * kSlots];
  const CODE *lval[2 * kSlots];
  uint32_t lstart[2 * kSlots];
  uint32_t lpref[2 * kSlots + 1];
  CODE lthr[2 * kSlots];
  const BITS *fvlo[kSlots];
  const BITS *fvhi[kSlots];
  uint32_t lobase[kSlots], hibase[kSlots];
  uint32_t slot_q[kSlots];
  uint32_t qcount[kBatchMaxGroup];
};

// Direct form: the predicate words of this tile into `dst` (LDS).  Each wave covers TW/4
// words, 256 rows (8 words) per step, with the DPP assembly of k_scan_tiles.  Rows past n
// give 0 bits.
template <int TW>
__device__ __forceinline__ void batch_direct(const BatchCol &c, int64_t t, int64_t n,
                                             uint32_t *dst) {
  const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
  const uint32_t shift = 28 - 4 * (lane & 7);
  constexpr int kWordsPerWave = TW / (kBlock / kWave);
  const int64_t tile_row0 = t * (int64_t)TW * 32;
  for (int s = 0; s < kWordsPerWave / 8; s++) {
    const int w_local = wv * kWordsPerWave + s * 8;  // first word of this 256-row step
    const int64_t r0 = tile_row0 + (int64_t)w_local * 32;
    uint32_t nib = 0;
    if (r0 + 256 <= n) {
      const uint4 v = load_stream(reinterpret_cast<const uint4 *>(c.col + r0) + lane);
      nib = nibble4(v, c.dlo, c.dspan);
    } else {
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int64_t r = r0 + 4 * lane + j;
        if (r < n && (c.col[r] - c.dlo) <= c.dspan) nib |= 8u >> j;
      }
    }
    const uint32_t w = or_lanes8(nib << shift);
    if ((lane & 7) == 0) dst[w_local + lane / 8] = w;
  }
}

template <int KP>
struct Words {
  uint32_t v[KP];
};

template <int KP>
__device__ __forceinline__ Words<KP> base_words(const BITS *fv, uint32_t base, int64_t w,
                                                int64_t words) {
  Words<KP> r;
  if (base != kBaseFv) {
#pragma unroll
    for (int k = 0; k < KP; k++) r.v[k] = base == kBaseOnes ? ~0u : 0u;
    return r;
  }
  if (w + KP <= words) {
#pragma unroll
    for (int k = 0; k < KP; k += 4) {
      const uint4 x = *reinterpret_cast<const uint4 *>(fv + w + k);
      r.v[k] = x.x;
      r.v[k + 1] = x.y;
      r.v[k + 2] = x.z;
      r.v[k + 3] = x.w;
    }
  } else {
#pragma unroll
    for (int k = 0; k < KP; k++) r.v[k] = w + k < words ? fv[w + k] : 0u;
  }
  return r;
}

template <int TW>
__global__ __launch_bounds__(kBlock) void k_batch(const BatchQuery *__restrict__ qs,
                                                  const BatchGroup *__restrict__ groups,
                                                  int ngroups, int64_t n, int64_t words,
                                                  int64_t ntiles, BITS *out, int64_t out_stride,
                                                  uint32_t *partials) {
  constexpr int kPer = TW / kBlock;  // output words per thread
  static_assert(kPer == 4 || kPer == 8, "tile words per thread: a multiple of 4");
  using Lds = BatchLds<TW>;
  constexpr int kSlots = Lds::kSlots;
  __shared__ Lds lds;
  for (int i = threadIdx.x; i < kSlots * TW; i += kBlock) (&lds.tog[0][0])[i] = 0;
  const int xcd = blockIdx.x % kXcds, nblk_x = gridDim.x / kXcds, j = blockIdx.x / kXcds;
  const int64_t tiles_x = ntiles > xcd ? (ntiles - 1 - xcd) / kXcds + 1 : 0;
  const int64_t items_x = tiles_x * ngroups;
  const uint32_t last_mask = tail_mask((uint64_t)n);
  const int lw = threadIdx.x * kPer;
  __syncthreads();
  for (int64_t it = j; it < items_x; it += nblk_x) {
    const int64_t t = xcd + kXcds * (it / ngroups);
    const BatchGroup &g = groups[it % ngroups];
    const int nslots = g.nslots, nlists = 2 * nslots;
    const uint32_t row0 = (uint32_t)(t * (int64_t)TW * 32);

    // 1. per-slot bases and the tile's list slices
    if (threadIdx.x < nlists) {
      const int s = threadIdx.x >> 1, side = threadIdx.x & 1;
      const BatchCol &bc = qs[g.q0 + g.sq[s]].c[g.sc[s]];
      const BatchBound &b = side ? bc.hi : bc.lo;
      uint32_t st = 0, cnt = 0;
      if (!bc.col && b.dir) {
        st = b.dir[t];
        cnt = b.dir[t + 1] - st;
      }
      lds.lstart[threadIdx.x] = st;
      lds.lpref[threadIdx.x + 1] = cnt;
      lds.lpos[threadIdx.x] = b.bpos;
      lds.lval[threadIdx.x] = b.bval;
      lds.lthr[threadIdx.x] = b.thr;
      if (side) {
        lds.fvhi[s] = bc.col ? nullptr : b.fv;
        lds.hibase[s] = bc.col ? kBaseZero : b.base;
      } else {
        lds.fvlo[s] = bc.col ? nullptr : b.fv;
        lds.lobase[s] = bc.col ? kBaseZero : b.base;
        lds.slot_q[s] = g.sq[s];
      }
    }
    __syncthreads();
    if (threadIdx.x < kWave) {  // inclusive scan of the list lengths (nlists <= 64)
      const int l = threadIdx.x;
      uint32_t v = l < nlists ? lds.lpref[l + 1] : 0;
#pragma unroll
      for (int o = 1; o < kWave; o <<= 1) {
        const uint32_t u = __shfl_up(v, o, kWave);
        if (l >= o) v += u;
      }
      if (l < nlists) lds.lpref[l + 1] = v;
      if (l == 0) lds.lpref[0] = 0;
    }
    __syncthreads();

    // 2. refine toggles of every list of the group, one flattened pass; direct slots
    const uint32_t total = lds.lpref[nlists];
    for (uint32_t e = threadIdx.x; e < total; e += kBlock) {
      int lo = 0, hi = nlists - 1;  // last list with lpref[l] <= e
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (lds.lpref[mid] <= e)
          lo = mid;
        else
          hi = mid - 1;
      }
      const uint32_t idx = lds.lstart[lo] + (e - lds.lpref[lo]);
      const uint32_t p = lds.lpos[lo][idx];
      if (lds.lval[lo][idx] >= lds.lthr[lo])
        atomicXor(&lds.tog[lo >> 1][(p - row0) >> 5], 0x80000000u >> (p & 31));
    }
    for (int s = 0; s < nslots; s++) {
      const BatchCol &bc = qs[g.q0 + g.sq[s]].c[g.sc[s]];
      if (bc.col) batch_direct<TW>(bc, t, n, lds.tog[s]);
    }
    __syncthreads();

    // 3. every slot's base words (all loads issued together), then each query's AND
    const int64_t w = t * TW + lw;
    Words<kPer> base[kSlots];
#pragma unroll
    for (int s = 0; s < kSlots; s++) {
      if (s < nslots && w < words) {
        const Words<kPer> h = base_words<kPer>(lds.fvhi[s], lds.hibase[s], w, words);
        const Words<kPer> l = base_words<kPer>(lds.fvlo[s], lds.lobase[s], w, words);
#pragma unroll
        for (int k = 0; k < kPer; k++) base[s].v[k] = h.v[k] & ~l.v[k];
      }
    }
    if (threadIdx.x < kBatchMaxGroup) lds.qcount[threadIdx.x] = 0;
    __syncthreads();
    for (int qi = 0; qi < g.nq; qi++) {
      const int q = g.q0 + qi;
      uint32_t acc[kPer];
#pragma unroll
      for (int k = 0; k < kPer; k++) acc[k] = qs[q].empty ? 0u : ~0u;
#pragma unroll
      for (int s = 0; s < kSlots; s++) {
        if (s < nslots && (int)lds.slot_q[s] == qi) {
#pragma unroll
          for (int k = 0; k < kPer; k++) acc[k] &= base[s].v[k] ^ lds.tog[s][lw + k];
        }
      }
      uint32_t pc = 0;
      if (w < words) {
        BITS *o = (qs[q].out ? qs[q].out : out + (int64_t)q * out_stride) + w;
#pragma unroll
        for (int k = 0; k < kPer; k++) {
          if (w + k == words - 1) acc[k] &= last_mask;
          if (w + k >= words) acc[k] = 0;
          pc += __popc(acc[k]);
        }
        if (w + kPer <= words) {
#pragma unroll
          for (int k = 0; k < kPer; k += 4)
            *reinterpret_cast<uint4 *>(o + k) = make_uint4(acc[k], acc[k + 1], acc[k + 2], acc[k + 3]);
        } else {
#pragma unroll
          for (int k = 0; k < kPer; k++)
            if (w + k < words) o[k] = acc[k];
        }
      }
      pc = wave_sum(pc);
      if ((threadIdx.x & (kWave - 1)) == 0 && pc) atomicAdd(&lds.qcount[qi], pc);
    }
    for (int s = 0; s < nslots; s++)
#pragma unroll
      for (int k = 0; k < kPer; k++) lds.tog[s][lw + k] = 0;
    __syncthreads();
    if (threadIdx.x < g.nq) partials[(int64_t)(g.q0 + threadIdx.x) * ntiles + t] = lds.qcount[threadIdx.x];
  }
}

// counts[q] = sum over tiles of partials[q * ntiles + t]; one workgroup per query.
__global__ __launch_bounds__(kBlock) void k_batch_counts(const uint32_t *partials, int64_t ntiles,
                                                         unsigned long long *counts) {
  __shared__ unsigned long long part[kBlock / kWave];
  const uint32_t *p = partials + (int64_t)blockIdx.x * ntiles;
  unsigned long long s = 0;
  for (int64_t i = threadIdx.x; i < ntiles; i += kBlock) s += p[i];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, kWave);
  if ((threadIdx.x & (kWave - 1)) == 0) part[threadIdx.x / kWave] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int i = 0; i < kBlock / kWave; i++) t += part[i];
    counts[blockIdx.x] = t;
  }
}

static BatchBound make_batch_bound(const rts_bindex_dev *d, const BoundSpec &s) {
  BatchBound b{};
  if (s.fv < 0) {
    b.base = kBaseZero;
  } else if (s.fv >= kBins - 1) {
    b.base = kBaseOnes;
  } else {
    b.base = kBaseFv;
    b.fv = d->fvs + (size_t)s.fv * d->fv_stride;
  }
  if (s.list >= 0) {
    b.bpos = d->bpos;
    b.bval = d->bval;
    b.thr = s.thr;
  }
  return b;
}

}  // namespace rts

using namespace rts;

struct rts_batch_plan {
  const BinDex *const *index = nullptr;  // per table column (may hold nullptr: direct only)
  std::vector<const BinDex *> index_store;
  std::vector<const CODE *> columns;
  int ncols = 0;
  int64_t n = 0, words = 0;
  int max_queries = 0;
  int nq = 0;
  rts_batch_config cfg{};
  int64_t ntiles = 0;
  BatchQuery *d_queries = nullptr;
  BatchGroup *d_groups = nullptr;
  int ngroups = 0;
  uint32_t *d_partials = nullptr;
  std::vector<BatchQuery> h_queries;
  std::vector<BatchGroup> h_groups;
  std::vector<uint32_t *> dir_base;  // per column: bdir re-tiled for cfg.tile_words
  std::vector<int64_t> dir_tiles;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  BITS *graph_out = nullptr;
  unsigned long long *graph_counts = nullptr;
  int graph_nq = -1;
  int64_t graph_stride = 0;
  int stats_bindex_cols = 0, stats_direct_cols = 0, stats_narrow_cols = 0;
  std::vector<rts_query_spec> specs;  // the current queries, re-planned on a new tiling
  std::vector<BITS *> outs;           // their per-query outputs, if given
  bool has_outs = false;
};

namespace rts {

// Tile directory of one column for tiles of `tile_rows` rows (the index's own directory is
// for kTileRows): first B-list entry of bin k with row >= t * tile_rows.
__global__ void k_retile_directory(const uint32_t *bpos, const uint32_t *boff, int64_t ntiles,
                                   int64_t tile_rows, uint32_t *bdir) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t per = ntiles + 1;
  if (id >= per * kBins) return;
  const int k = (int)(id / per);
  const uint64_t row = (uint64_t)(id % per) * (uint64_t)tile_rows;
  uint32_t lo = boff[k], hi = boff[k + 1];
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (bpos[mid] < row)
      lo = mid + 1;
    else
      hi = mid;
  }
  bdir[id] = lo;
}

static bool valid_config(const rts_batch_config &c) {
  return (c.tile_words == 1024 || c.tile_words == 2048) && c.queries_per_group >= 1 &&
         c.queries_per_group <= kBatchMaxGroup && c.blocks_per_cu >= 1 && c.blocks_per_cu <= 16;
}

static void free_dirs(rts_batch_plan *p) {
  for (uint32_t *d : p->dir_base) (void)hipFree(d);
  p->dir_base.clear();
  p->dir_tiles.clear();
}

static hipError_t build_dirs(rts_batch_plan *p, hipStream_t s) {
  free_dirs(p);
  const int64_t tile_rows = (int64_t)p->cfg.tile_words * 32;
  const int64_t ntiles = (p->n + tile_rows - 1) / tile_rows;
  p->dir_base.assign(p->ncols, nullptr);
  p->dir_tiles.assign(p->ncols, ntiles);
  for (int c = 0; c < p->ncols; c++) {
    const BinDex *b = p->index[c];
    if (!b || !b->dev) continue;
    const rts_bindex_dev *d = b->dev;
    const int64_t items = (int64_t)kBins * (ntiles + 1);
    hipError_t e = hipMalloc((void **)&p->dir_base[c], items * sizeof(uint32_t));
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_retile_directory, dim3((unsigned)((items + kBlock - 1) / kBlock)), dim3(kBlock),
                       0, s, d->bpos, d->boff, ntiles, tile_rows, p->dir_base[c]);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  p->ntiles = ntiles;
  return hipSuccess;
}

}  // namespace rts

namespace rts {

// Workgroups of k_batch<TW> that fit on one compute unit at once (LDS-bound).
template <int TW>
static int batch_resident_per_cu() {
  static int cached = 0;
  if (!cached) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_batch<TW>, kBlock, 0) != hipSuccess || nb < 1)
      nb = 1;
    cached = nb;
  }
  return cached;
}

static hipError_t launch_batch(rts_batch_plan *p, BITS *out, int64_t out_stride,
                               unsigned long long *counts, hipStream_t s) {
  const int cus = std::max(1, device_cu_count());
  // a persistent grid larger than what is resident would leave its statically assigned
  // work items to a second round: clamp to the occupancy limit
  const int resident = p->cfg.tile_words == 2048 ? batch_resident_per_cu<2048>() : batch_resident_per_cu<1024>();
  unsigned grid = (unsigned)(cus * std::min(p->cfg.blocks_per_cu, resident));
  grid = std::max<unsigned>(kXcds, grid / kXcds * kXcds);
  switch (p->cfg.tile_words) {
    case 2048:
      hipLaunchKernelGGL(k_batch<2048>, dim3(grid), dim3(kBlock), 0, s, p->d_queries, p->d_groups,
                         p->ngroups, p->n, p->words, p->ntiles, out, out_stride, p->d_partials);
      break;
    default:
      hipLaunchKernelGGL(k_batch<1024>, dim3(grid), dim3(kBlock), 0, s, p->d_queries, p->d_groups,
                         p->ngroups, p->n, p->words, p->ntiles, out, out_stride, p->d_partials);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || !counts) return e;
  hipLaunchKernelGGL(k_batch_counts, dim3((unsigned)p->nq), dim3(kBlock), 0, s, p->d_partials, p->ntiles,
                     counts);
  return hipGetLastError();
}

}  // namespace rts

extern "C" {

void rts_batch_default_config(rts_batch_config *cfg) {
  cfg->tile_words = 1024;
  cfg->queries_per_group = 16;
  cfg->blocks_per_cu = 4;
}

hipError_t rts_batch_plan_create(rts_batch_plan **out, const CODE *const *columns,
                                 const BinDex *const *indexes, int ncols, int64_t n,
                                 int max_queries, const rts_batch_config *cfg, hipStream_t stream) {
  if (!out || !columns || ncols < 1 || n < 1 || max_queries < 1 || n > 0x7FFFFFFFll)
    return set_error(hipErrorInvalidValue, "rts_batch_plan_create");
  auto *p = new rts_batch_plan();
  p->ncols = ncols;
  p->n = n;
  p->words = (n + 31) / 32;
  p->max_queries = max_queries;
  p->columns.assign(columns, columns + ncols);
  p->index_store.assign(ncols, nullptr);
  for (int c = 0; c < ncols; c++) {
    const BinDex *b = indexes ? indexes[c] : nullptr;
    if (b && (!b->dev || b->dev->n != n)) {
      delete p;
      return set_error(hipErrorInvalidValue, "rts_batch_plan_create: index length");
    }
    p->index_store[c] = b;
  }
  p->index = p->index_store.data();
  if (cfg) p->cfg = *cfg; else rts_batch_default_config(&p->cfg);
  if (!valid_config(p->cfg)) {
    delete p;
    return set_error(hipErrorInvalidValue, "rts_batch_plan_create: config");
  }
  hipError_t e = hipMalloc((void **)&p->d_queries, (size_t)max_queries * sizeof(BatchQuery));
  if (e == hipSuccess) e = hipMalloc((void **)&p->d_groups, (size_t)max_queries * sizeof(BatchGroup));
  if (e == hipSuccess) e = build_dirs(p, stream);
  if (e == hipSuccess)
    e = hipMalloc((void **)&p->d_partials, (size_t)max_queries * p->ntiles * sizeof(uint32_t));
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (e != hipSuccess) {
    rts_batch_plan_free(p);
    return set_error(e, "rts_batch_plan_create");
  }
  *out = p;
  return hipSuccess;
}

hipError_t rts_batch_plan_set_config(rts_batch_plan *p, const rts_batch_config *cfg,
                                     hipStream_t stream) {
  if (!p || !cfg || !valid_config(*cfg)) return set_error(hipErrorInvalidValue, "rts_batch_plan_set_config");
  const bool retile = cfg->tile_words != p->cfg.tile_words;
  p->cfg = *cfg;
  if (p->exec) {
    (void)hipGraphExecDestroy(p->exec);
    p->exec = nullptr;
  }
  if (p->graph) {
    (void)hipGraphDestroy(p->graph);
    p->graph = nullptr;
  }
  if (!retile) return hipSuccess;
  hipError_t e = build_dirs(p, stream);
  if (e == hipSuccess) {
    (void)hipFree(p->d_partials);
    p->d_partials = nullptr;
    e = hipMalloc((void **)&p->d_partials, (size_t)p->max_queries * p->ntiles * sizeof(uint32_t));
  }
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (e != hipSuccess) return set_error(e, "rts_batch_plan_set_config");
  // the planned queries hold directory rows of the old tiling: plan them again
  const std::vector<rts_query_spec> specs = p->specs;
  const std::vector<BITS *> outs = p->outs;
  return rts_batch_plan_set_queries_out(p, specs.data(), p->has_outs ? outs.data() : nullptr,
                                        (int)specs.size(), stream);
}

hipError_t rts_batch_plan_set_queries(rts_batch_plan *p, const rts_query_spec *queries, int nq,
                                      hipStream_t stream) {
  return rts_batch_plan_set_queries_out(p, queries, nullptr, nq, stream);
}

hipError_t rts_batch_plan_set_queries_out(rts_batch_plan *p, const rts_query_spec *queries,
                                          BITS *const *outs, int nq, hipStream_t stream) {
  if (!p || (!queries && nq) || nq < 0 || nq > p->max_queries)
    return set_error(hipErrorInvalidValue, "rts_batch_plan_set_queries");
  for (int q = 0; outs && q < nq; q++)
    if (!outs[q] || ((uintptr_t)outs[q] & 15))
      return set_error(hipErrorInvalidValue, "rts_batch_plan_set_queries: outputs must be 16-byte aligned");
  p->has_outs = outs != nullptr;
  p->outs.assign(outs ? outs : (BITS *const *)nullptr, outs ? outs + nq : (BITS *const *)nullptr);
  p->h_queries.assign(nq, BatchQuery{});
  for (int q = 0; outs && q < nq; q++) p->h_queries[q].out = outs[q];
  p->stats_bindex_cols = p->stats_direct_cols = p->stats_narrow_cols = 0;
  for (int q = 0; q < nq; q++) {
    const rts_query_spec &in = queries[q];
    BatchQuery &bq = p->h_queries[q];
    if (in.ncols < 1 || in.ncols > kBatchMaxCols)
      return set_error(hipErrorInvalidValue, "rts_batch_plan_set_queries: ncols");
    bq.ncols = in.ncols;
    for (int c = 0; c < in.ncols; c++) {
      const int col = in.column[c];
      if (col < 0 || col >= p->ncols) return set_error(hipErrorInvalidValue, "rts_batch_plan_set_queries: column");
      const CODE lo = in.lo[c], hi = in.hi[c];
      if (lo > hi) bq.empty = 1;
      BatchCol &bc = bq.c[c];
      const BinDex *b = p->index[col];
      const bool direct = in.mode[c] == RTS_EVAL_DIRECT || !b;
      if (in.mode[c] == RTS_EVAL_BINDEX && !b)
        return set_error(hipErrorInvalidValue, "rts_batch_plan_set_queries: column has no index");
      if (direct) {
        if (((uintptr_t)p->columns[col] & 15) != 0)
          return set_error(hipErrorInvalidValue, "rts_batch_plan_set_queries: direct column not 16-byte aligned");
        bc.col = p->columns[col];
        bc.dlo = lo;
        bc.dspan = hi - lo;
        p->stats_direct_cols++;
        continue;
      }
      const rts_bindex_dev *d = b->dev;
      const BoundSpec sl = lo == 0 ? BoundSpec{-1, -1, 0} : bound_spec(d, lo, false);
      const BoundSpec sh = bound_spec(d, hi + 1u, hi == 0xFFFFFFFFu);
      bc.lo = make_batch_bound(d, sl);
      bc.hi = make_batch_bound(d, sh);
      const int64_t per = p->ntiles + 1;
      if (sl.list >= 0) bc.lo.dir = p->dir_base[col] + (size_t)sl.list * per;
      if (sh.list >= 0) bc.hi.dir = p->dir_base[col] + (size_t)sh.list * per;
      // Both bounds in one bin: the result is the bin's list rows in [lo, hi] alone,
      // (1 ^ t_hi) & ~(1 ^ t_lo) = t_lo & ~t_hi, so no filter vector is read.
      if (sl.fv == sh.fv && sl.list >= 0 && sl.list == sh.list) {
        bc.lo.base = bc.hi.base = kBaseOnes;
        bc.lo.fv = bc.hi.fv = nullptr;
        p->stats_narrow_cols++;
      }
      p->stats_bindex_cols++;
    }
  }
  p->nq = nq;
  p->specs.assign(queries, queries + nq);
  // Work-item groups: consecutive queries while their columns fit the LDS slots of one tile
  // and the configured group size.
  const int slots = kBatchLdsWords / p->cfg.tile_words;
  p->h_groups.clear();
  for (int q = 0; q < nq;) {
    BatchGroup g{};
    g.q0 = q;
    while (q < nq && g.nq < p->cfg.queries_per_group) {
      const int need = p->h_queries[q].empty ? 0 : p->h_queries[q].ncols;
      if (g.nslots + need > slots) break;
      for (int c = 0; c < need; c++) {
        g.sq[g.nslots] = (uint8_t)g.nq;
        g.sc[g.nslots] = (uint8_t)c;
        g.nslots++;
      }
      g.nq++;
      q++;
    }
    p->h_groups.push_back(g);
  }
  p->ngroups = (int)p->h_groups.size();
  if (nq == 0) return hipSuccess;
  hipError_t e = hipMemcpyAsync(p->d_queries, p->h_queries.data(), (size_t)nq * sizeof(BatchQuery),
                                hipMemcpyHostToDevice, stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(p->d_groups, p->h_groups.data(), (size_t)p->ngroups * sizeof(BatchGroup),
                       hipMemcpyHostToDevice, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);  // h_queries may change on the next call
  return set_error(e, "rts_batch_plan_set_queries");
}

static bool valid_output(const rts_batch_plan *p, const BITS *out, int64_t stride) {
  if (p && !out) return p->has_outs;  // per-query outputs given with the queries
  return p && stride >= p->words && stride % 4 == 0 && ((uintptr_t)out & 15) == 0;
}

hipError_t rts_batch_plan_launch(rts_batch_plan *p, BITS *out, int64_t out_stride_words,
                                 unsigned long long *counts, hipStream_t stream) {
  if (!valid_output(p, out, out_stride_words))
    return set_error(hipErrorInvalidValue, "rts_batch_plan_launch: output must be 16-byte aligned, stride a multiple of 4 words");
  if (p->nq == 0) return hipSuccess;
  return set_error(launch_batch(p, out, out_stride_words, counts, stream), "rts_batch_plan_launch");
}

hipError_t rts_batch_plan_graph_launch(rts_batch_plan *p, BITS *out, int64_t out_stride_words,
                                       unsigned long long *counts, hipStream_t stream) {
  if (!valid_output(p, out, out_stride_words) || !stream)
    return set_error(hipErrorInvalidValue, "rts_batch_plan_graph_launch");
  if (p->nq == 0) return hipSuccess;
  hipError_t e = hipSuccess;
  if (!p->exec || p->graph_out != out || p->graph_counts != counts || p->graph_nq != p->nq ||
      p->graph_stride != out_stride_words) {
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    p->exec = nullptr;
    p->graph = nullptr;
    e = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return set_error(e, "rts_batch_plan_graph_launch: capture");
    hipError_t le = launch_batch(p, out, out_stride_words, counts, stream);
    e = hipStreamEndCapture(stream, &p->graph);
    if (le != hipSuccess) e = le;
    if (e == hipSuccess) e = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) return set_error(e, "rts_batch_plan_graph_launch: instantiate");
    p->graph_out = out;
    p->graph_counts = counts;
    p->graph_nq = p->nq;
    p->graph_stride = out_stride_words;
  }
  return set_error(hipGraphLaunch(p->exec, stream), "rts_batch_plan_graph_launch");
}

void rts_batch_plan_stats(const rts_batch_plan *p, int *bindex_cols, int *direct_cols,
                          int *narrow_cols, long long *ntiles) {
  if (!p) return;
  if (bindex_cols) *bindex_cols = p->stats_bindex_cols;
  if (direct_cols) *direct_cols = p->stats_direct_cols;
  if (narrow_cols) *narrow_cols = p->stats_narrow_cols;
  if (ntiles) *ntiles = p->ntiles;
}

void rts_batch_plan_info(const rts_batch_plan *p, long long *n, int *ncols, int *nq) {
  if (!p) return;
  if (n) *n = p->n;
  if (ncols) *ncols = p->ncols;
  if (nq) *nq = p->nq;
}

void rts_batch_plan_free(rts_batch_plan *p) {
  if (!p) return;
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  free_dirs(p);
  (void)hipFree(p->d_queries);
  (void)hipFree(p->d_groups);
  (void)hipFree(p->d_partials);
  delete p;
}

}  // extern "C"
