// batch.hip -- batched evaluation of many range queries in one persistent launch.
//
// A query is a conjunction of closed ranges over 1..4 columns of a table (all columns
// have n rows).  Each query-column ("slot") is evaluated either from the column's BinDex
// (two filter vectors and the tile slices of up to two bins' refine lists, as in
// k_bindex_query; or, for a range over one or two bins whose lists hold all their rows,
// those lists alone) or directly from the column's values (the wave64 DPP word assembly
// of k_scan_tiles).
//
// Work item = (row tile t, query group g): the workgroup writes the tile's output words of
// every query of the group.  The items of one tile run on the same XCD one after another
// (workgroup b runs on XCD b % 8 on MI355X, and XCD x takes tiles t with t % 8 == x,
// tile-major), handed out by a per-XCD counter, so that the XCD's workgroups stay on the
// same tile or two: the tile's filter-vector and list slices are read from HBM once and,
// as far as the 4 MB L2 holds them, re-read by the other query groups from there.  (A
// static round-robin of the same order lets the workgroups drift over several tiles, whose
// slices together overflow the L2; it ran 4-18% slower.)  Popcounts go to one partial per
// (query, tile), summed by a second small kernel.
//
// An item is a chain of dependent memory round trips (group record, directory words, list
// entries and filter-vector words) and two barriers, so the launch is bound by latency:
// each wave reads its own lists' records and directory words (wave-uniform addresses, no
// exchange through LDS); refine entries are list-major (wave v of W owns lists v, v + W,
// ...), 12-16 entries per lane in flight; the filter-vector words go out with the first
// refine round and are XORed into the LDS toggles as they arrive; the query descriptors
// reach LDS before the barrier that precedes the AND phase, whose per-query work needs no
// LDS round trip of its own beyond one descriptor read and two slots' words at a time.
#include <string.h>

#include <math.h>

#include <algorithm>
#include <array>
#include <map>
#include <set>
#include <vector>

#include "batch.h"
#include "bindex_dev.h"

namespace rts {

enum : uint32_t { kBaseZero = 0, kBaseOnes = 1, kBaseFv = 2 };

// Workgroups of 512 threads (8 waves, 64 KB of toggles, 2 per CU) or 256 threads (4 waves,
// 32 KB, up to 4 per CU: more independent work items in flight per CU).
template <int NT>
struct BatchShape {
  static constexpr int kThreads = NT;
  static constexpr int kWaves = NT / kWave;
  static constexpr int kLdsWords = kBatchLdsWords / 512 * NT;  // toggle words per workgroup
};

// One query-column of a group.  BinDex form: (fv_hi & ~fv_lo) ^ T, with T the toggles of
// the slot's refine lists; bases kBaseZero / kBaseOnes need no filter vector.  Direct form
// (col non-null): the words of (x - dlo) <= dspan.
struct BatchSlot {
  const BITS *fvhi;  // the filter-vector fields first: one 32-byte scalar load
  const BITS *fvlo;
  const CODE *col;
  uint32_t hibase, lobase;
  CODE dlo, dspan;
};

// One refine list of a group: the rows of a bin's B list (in row order, tile directory
// `dir`) whose value lies in [thr, thr + span] flip bits of slot `slot` (span = 2^32 - 1 -
// thr for a bound, value >= thr; the whole range of a narrow column, both of whose
// bounds fall in this bin, in one pass over the list).  Unused entries point at a zero
// directory (no rows), so the kernel loads every entry without first reading the count.
struct BatchList {
  const uint32_t *dir;
  const uint2 *bent;  // (row, value) entries of the column's B lists
  CODE thr, span;
  uint32_t slot, pad;
};

// One query of a group: its own bitmap (or out + id * out_stride of the launch) and its
// index `id` in the caller's order (bitmap row and count index).
struct BatchQuery {
  BITS *out;
  int empty;  // some column's range is empty: all-zero result
  int id;
};

// A work item's query group, stored contiguously so that one round of loads brings in all
// of it: queries query[0..nq), query i owning slots [qslot[i], qslot[i + 1]), and the refine
// lists of those slots.
struct BatchGroup {
  int q0, nq, nslots, nlists;
  uint8_t qslot[80];
  BatchSlot slot[kBatchMaxSlots];
  BatchList list[2 * kBatchMaxSlots];
  BatchQuery query[kBatchMaxGroup];
};

// A query of the current item as the AND phase reads it: one 16-byte LDS read.
struct alignas(16) BatchQueryLds {
  BITS *out;
  uint32_t slots;  // first slot | one past its last slot << 8
  uint32_t empty;
};

// LDS of one workgroup: one toggle array per slot and the current item's query descriptors.
template <int TW, int NT>
struct BatchLds {
  static constexpr int kSlots = BatchShape<NT>::kLdsWords / TW;
  static constexpr int kLists = 2 * kSlots;
  alignas(16) uint32_t tog[kSlots][TW];
  BatchQueryLds q[kBatchMaxGroup];
  uint32_t qid[kBatchMaxGroup];
  uint32_t qcount[kBatchMaxGroup];
  unsigned long long next;  // the workgroup's next work item
};

constexpr int ilog2(int v) { return v <= 1 ? 0 : 1 + ilog2(v / 2); }

// Device memory reached through pointers read from the group records (columns, filter
// vectors, list entries and directories, result bitmaps) is global memory.  Saying so turns
// the generic (flat) accesses the compiler would emit for such pointers into global ones:
// flat accesses also count against the LDS wait counter, so every LDS wait in the kernel
// would wait for outstanding loads and result stores as well.
#define RTS_GLOBAL __attribute__((address_space(1)))
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
template <class T>
__device__ __forceinline__ const RTS_GLOBAL T *gmem(const T *p) {
  return (const RTS_GLOBAL T *)p;
}
template <class T>
__device__ __forceinline__ RTS_GLOBAL T *gmem(T *p) {
  return (RTS_GLOBAL T *)p;
}
// Memory that no kernel writes while a batch runs (the lists' tile directories), read
// through the constant address space: a wave-uniform address then becomes a scalar load,
// whose wait does not also wait for the wave's outstanding result stores.
#define RTS_CONST __attribute__((address_space(4)))
template <class T>
__device__ __forceinline__ const RTS_CONST T *cmem(const T *p) {
  return (const RTS_CONST T *)p;
}

// Toggle layout: thread t owns tile words t * kPer .. t * kPer + kPer - 1 (kPer = TW / NT)
// as kPer / U units of U words (U = 2 when kPer >= 2: 64-bit LDS accesses), stored
// transposed: unit k of thread t at unit position k * NT + t.  The accesses of each
// thread's k-th unit (filter-vector XORs, the AND phase, the clearing) are consecutive
// across lanes and free of bank conflicts, with half the instructions of 32-bit words.
template <int TW, int NT>
struct TogLayout {
  static constexpr int kPer = TW / NT;
  static constexpr int kUnit = kPer >= 2 ? 2 : 1;  // words per unit
  static constexpr int kUnits = kPer / kUnit;      // units per thread
};

template <int TW, int NT>
__device__ __forceinline__ uint32_t tog_index(uint32_t w) {
  using T = TogLayout<TW, NT>;
  const uint32_t u = w >> ilog2(T::kUnit), r = w & (T::kUnit - 1);
  const uint32_t pos = ((u & (T::kUnits - 1)) << ilog2(NT)) | (u >> ilog2(T::kUnits));
  return (pos << ilog2(T::kUnit)) | r;
}

// Direct form: the predicate words of this tile into `dst` (LDS).  Each wave covers TW / 8
// words, 256 rows (8 words) per step, 4 steps' loads in flight, with the DPP assembly of
// k_scan_tiles.  Rows past n give 0 bits.
template <int TW, int NT>
__device__ __forceinline__ void batch_direct(const BatchSlot &c, int64_t t, int64_t n, uint32_t *dst) {
  const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
  const uint32_t shift = 28 - 4 * (lane & 7);
  constexpr int kWordsPerWave = TW / BatchShape<NT>::kWaves;
  constexpr int kSteps = kWordsPerWave / 8;
  constexpr int kU = 4;
  static_assert(kSteps % kU == 0, "steps per wave");
  const int64_t tile_row0 = t * (int64_t)TW * 32;
  const int64_t wave_row0 = tile_row0 + (int64_t)wv * kWordsPerWave * 32;
  if (wave_row0 + (int64_t)kWordsPerWave * 32 <= n) {
    const RTS_GLOBAL u32x4 *src = reinterpret_cast<const RTS_GLOBAL u32x4 *>(gmem(c.col) + wave_row0) + lane;
#pragma unroll 1
    for (int s0 = 0; s0 < kSteps; s0 += kU) {
      uint4 v[kU];
#pragma unroll
      for (int u = 0; u < kU; u++) {
        const u32x4 x = __builtin_nontemporal_load(src + (s0 + u) * kWave);
        v[u] = make_uint4(x.x, x.y, x.z, x.w);
      }
#pragma unroll
      for (int u = 0; u < kU; u++) {
        const uint32_t w = or_lanes8(nibble4(v[u], c.dlo, c.dspan) << shift);
        if ((lane & 7) == 0) dst[tog_index<TW, NT>(wv * kWordsPerWave + (s0 + u) * 8 + lane / 8)] = w;
      }
    }
    return;
  }
#pragma unroll 1
  for (int s = 0; s < kSteps; s++) {  // the tile holding row n - 1
    const int64_t r0 = wave_row0 + (int64_t)s * 256;
    uint32_t nib = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int64_t r = r0 + 4 * lane + j;
      if (r < n && (gmem(c.col)[r] - c.dlo) <= c.dspan) nib |= 8u >> j;
    }
    const uint32_t w = or_lanes8(nib << shift);
    if ((lane & 7) == 0) dst[tog_index<TW, NT>(wv * kWordsPerWave + s * 8 + lane / 8)] = w;
  }
}

template <int KP>
struct Words {
  uint32_t v[KP];
};

// KP filter-vector words from w on, every one of them inside the bitmap: vector loads only.
template <int KP>
__device__ __forceinline__ Words<KP> base_words_whole(const BITS *fv_any, uint32_t base, int64_t w) {
  const RTS_GLOBAL BITS *fv = gmem(fv_any);
  Words<KP> r;
  if (base != kBaseFv) {
#pragma unroll
    for (int k = 0; k < KP; k++) r.v[k] = base == kBaseOnes ? ~0u : 0u;
    return r;
  }
  if constexpr (KP == 1) {
    r.v[0] = fv[w];
  } else if constexpr (KP == 2) {
    const u32x2 x = *reinterpret_cast<const RTS_GLOBAL u32x2 *>(fv + w);
    r.v[0] = x.x;
    r.v[1] = x.y;
  } else {
#pragma unroll
    for (int k = 0; k < KP; k += 4) {
      const u32x4 x = *reinterpret_cast<const RTS_GLOBAL u32x4 *>(fv + w + k);
      r.v[k] = x.x;
      r.v[k + 1] = x.y;
      r.v[k + 2] = x.z;
      r.v[k + 3] = x.w;
    }
  }
  return r;
}

// As base_words_whole, with the words past the bitmap's end read as 0 (the last tile).
template <int KP>
__device__ __forceinline__ Words<KP> base_words(const BITS *fv_any, uint32_t base, int64_t w, int64_t words) {
  const RTS_GLOBAL BITS *fv = gmem(fv_any);
  Words<KP> r;
  if (base != kBaseFv) {
#pragma unroll
    for (int k = 0; k < KP; k++) r.v[k] = base == kBaseOnes ? ~0u : 0u;
    return r;
  }
  if (w + KP <= words) {
    if constexpr (KP == 1) {
      r.v[0] = fv[w];
    } else if constexpr (KP == 2) {
      const u32x2 x = *reinterpret_cast<const RTS_GLOBAL u32x2 *>(fv + w);
      r.v[0] = x.x;
      r.v[1] = x.y;
    } else {
#pragma unroll
      for (int k = 0; k < KP; k += 4) {
        const u32x4 x = *reinterpret_cast<const RTS_GLOBAL u32x4 *>(fv + w + k);
        r.v[k] = x.x;
        r.v[k + 1] = x.y;
        r.v[k + 2] = x.z;
        r.v[k + 3] = x.w;
      }
    }
  } else {
#pragma unroll
    for (int k = 0; k < KP; k++) r.v[k] = w + k < words ? fv[w + k] : 0u;
  }
  return r;
}


// A value that is the same in every lane, moved to a scalar register (the builtin takes and
// returns 32-bit values: a pointer goes one half at a time).
__device__ __forceinline__ uint32_t uniform_u32(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}
__device__ __forceinline__ uint64_t uniform_u64(uint64_t v) {
  return ((uint64_t)uniform_u32((uint32_t)(v >> 32)) << 32) | uniform_u32((uint32_t)v);
}
template <class T>
__device__ __forceinline__ T *uniform_ptr(T *p) {
  const uint64_t a = (uint64_t)p;
  return (T *)(((uint64_t)uniform_u32((uint32_t)(a >> 32)) << 32) | uniform_u32((uint32_t)a));
}

// Result words are written once and not read again by the launch: non-temporal stores keep
// them from evicting the filter-vector and list slices that the tile's other query groups
// re-read from L2.
template <int KP>
__device__ __forceinline__ void store_words(BITS *o, const uint32_t *acc) {
  if constexpr (KP == 1) {
    __builtin_nontemporal_store(acc[0], gmem(o));
  } else if constexpr (KP == 2) {
    u32x2 v = {acc[0], acc[1]};
    __builtin_nontemporal_store(v, reinterpret_cast<RTS_GLOBAL u32x2 *>(gmem(o)));
  } else {
#pragma unroll
    for (int k = 0; k < KP; k += 4) {
      u32x4 v = {acc[k], acc[k + 1], acc[k + 2], acc[k + 3]};
      __builtin_nontemporal_store(v, reinterpret_cast<RTS_GLOBAL u32x4 *>(gmem(o) + k));
    }
  }
}

// Filter-vector words of the group's BinDex slots, XORed into their toggles: 32 words per
// thread in flight at a time (16 at 512-word tiles), XORed as they arrive (every toggle
// update is an XOR, so the order does not matter and the words need not stay in
// registers).  kWhole: every word of the tile lies inside the bitmap.
template <int TW, int NT, bool kWhole, class Lds>
__device__ __forceinline__ void xor_filter_vectors(Lds &lds, const BatchGroup *g, int nslots, int64_t w,
                                                   int64_t words, int tid) {
  constexpr int kPer = TW / NT;
  using TL = TogLayout<TW, NT>;
  constexpr int kSlots = Lds::kSlots;
  constexpr int kFvBatch = kPer == 1 ? 8 : 16 / kPer;  // slots whose words are in flight together
#pragma unroll
  for (int s0 = 0; s0 < kSlots; s0 += kFvBatch) {
    Words<kPer> h[kFvBatch], l[kFvBatch];
    bool use[kFvBatch];
#pragma unroll
    for (int u = 0; u < kFvBatch; u++) {
      const BatchSlot S = g->slot[s0 + u];
      use[u] = s0 + u < nslots && !uniform_ptr(S.col) && (kWhole || w < words);
      if (use[u]) {
        if constexpr (kWhole) {
          h[u] = base_words_whole<kPer>(uniform_ptr(S.fvhi), uniform_u32(S.hibase), w);
          l[u] = base_words_whole<kPer>(uniform_ptr(S.fvlo), uniform_u32(S.lobase), w);
        } else {
          h[u] = base_words<kPer>(uniform_ptr(S.fvhi), uniform_u32(S.hibase), w, words);
          l[u] = base_words<kPer>(uniform_ptr(S.fvlo), uniform_u32(S.lobase), w, words);
        }
      }
    }
#pragma unroll
    for (int u = 0; u < kFvBatch; u++) {
      if (use[u]) {
        if constexpr (TL::kUnit == 2) {
#pragma unroll
          for (int k = 0; k < TL::kUnits; k++) {
            const unsigned long long v = (unsigned long long)(h[u].v[2 * k] & ~l[u].v[2 * k]) |
                                         ((unsigned long long)(h[u].v[2 * k + 1] & ~l[u].v[2 * k + 1]) << 32);
            atomicXor(reinterpret_cast<unsigned long long *>(&lds.tog[s0 + u][0]) + k * NT + tid, v);
          }
        } else {
          atomicXor(&lds.tog[s0 + u][tid], h[u].v[0] & ~l[u].v[0]);
        }
      }
    }
  }
}

template <int TW, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4))) void k_batch(
    const BatchGroup *__restrict__ groups, const int *__restrict__ ngroups_ptr, int64_t n,
    int64_t words, int64_t ntiles, BITS *out, int64_t out_stride, uint32_t *partials,
    unsigned long long *sched) {
  constexpr int kBatchThreads = NT, kBatchWaves = BatchShape<NT>::kWaves;
  constexpr int kPer = TW / kBatchThreads;  // output words per thread
  static_assert(kPer == 1 || kPer == 2 || kPer == 4 || kPer == 8, "tile words per thread");
  using TL = TogLayout<TW, NT>;

  using Lds = BatchLds<TW, NT>;
  constexpr int kSlots = Lds::kSlots, kLists = Lds::kLists;
  static_assert(kLists <= kWave && kSlots <= kWave, "descriptor loads by one wave each");
  __shared__ Lds lds;
  const int tid = threadIdx.x;
  for (int i = tid; i < kSlots * TW; i += kBatchThreads) (&lds.tog[0][0])[i] = 0;
  if (tid < kBatchMaxGroup) lds.qcount[tid] = 0;
  // the group count is read from device memory, so a captured launch stays valid when the
  // queries are planned again
  const int ngroups = *ngroups_ptr;
  const int xcd = blockIdx.x % kXcds, nblk_x = gridDim.x / kXcds, j = blockIdx.x / kXcds;
  const int64_t tiles_x = ntiles > xcd ? (ntiles - 1 - xcd) / kXcds + 1 : 0;
  const int64_t items_x = tiles_x * ngroups;
  const uint32_t last_mask = tail_mask((uint64_t)n);
  const int lw = tid * kPer;
  unsigned long long *const next_item = sched + xcd * kSchedStride;
  // item -> (tile, group) without a 64-bit division per item: while the XCD's items fit 32
  // bits, q = hi32(it * floor(2^32 / ngroups)) is floor(it / ngroups) or one less, fixed
  // by one compare (2^32 mod ngroups < ngroups keeps the error under one)
  const bool items32 = items_x <= (int64_t)0xFFFFFFFFu;
  const uint32_t ng = (uint32_t)ngroups;
  const uint32_t ng_recip = uniform_u32(ng > 1 ? (uint32_t)((1ull << 32) / ng) : 0u);
  __syncthreads();
  // the first item of each workgroup is fixed (j); later ones come from the XCD's counter,
  // so the workgroups of one XCD stay within a few items of each other and share the
  // current tile's slices in L2 instead of drifting over several tiles
  for (int64_t it = j; it < items_x;) {
    int64_t tq;
    uint32_t gi;
    if (items32) {  // the item index is the same in every lane: scalar arithmetic
      const uint32_t u = uniform_u32((uint32_t)it);
      uint32_t q = ng > 1 ? __umulhi(u, ng_recip) : u;
      uint32_t r = u - q * ng;
      if (r >= ng) {
        q++;
        r -= ng;
      }
      tq = q;
      gi = r;
    } else {
      tq = it / ngroups;
      gi = (uint32_t)(it % ngroups);
    }
    // both paths leave scalar values, so the group record and directory loads are scalar
    const int64_t t = xcd + kXcds * (int64_t)uniform_u64((uint64_t)tq);
    const BatchGroup *g = groups + uniform_u32(gi);
    const uint32_t row0 = (uint32_t)(t * (int64_t)TW * 32);

    // 1. Descriptors.  Each wave reads the records of its own refine lists and the slots
    //    itself (wave-uniform addresses, no exchange through LDS, so no barrier and no wait
    //    on another wave's directory loads); the query descriptors, needed only by the AND
    //    phase, go to LDS before this item's barrier.  Only the used slot and query entries
    //    are read, so a tile's group records take less L2.
    const int nslots = g->nslots, nq = g->nq;
    unsigned long long nxt = 0;
    if (tid == 0) nxt = atomicAdd(next_item, 1ull);  // returns while this item runs
    const int qi = tid - 2 * kWave;
    const bool qload = tid >= 2 * kWave && qi < nq;
    BatchQuery Q{};
    uint32_t qs0 = 0, qs1 = 0;
    if (qload) {
      Q = g->query[qi];
      qs0 = g->qslot[qi];
      qs1 = g->qslot[qi + 1];
    }

    const int64_t w = t * TW + lw;
    const bool interior = (t + 1) * TW < words;  // the last bitmap word lies in a later tile
    // 2. Filter-vector words of the BinDex slots (xor_filter_vectors) and refine entries:
    //    wave v of W owns lists v, v + W, ...; a list's slice base is uniform across the
    //    wave, each lane takes kListUnroll entries of each of its lists per round, and the
    //    first round's loads go out together with the filter-vector loads.
    //    Two lists per wave take 6 entries per lane each (384 entries per list and round:
    //    a 2048-word tile's slice of a 256-bin list is ~256), measured faster than 4, 5, 7
    //    or 8 on the 4096-query batch.
    constexpr int kListsPerWave = kLists / kBatchWaves;
    constexpr int kListUnroll = kListsPerWave == 2 ? 6 : 16 / kListsPerWave;
    static_assert(kLists % kBatchWaves == 0, "lists per wave");
    const int lane = tid & (kWave - 1), wv = (int)uniform_u32(tid / kWave);
    const uint2 *lbase[kListsPerWave];
    uint32_t llen[kListsPerWave], lthr[kListsPerWave], lspan[kListsPerWave], lslot[kListsPerWave];
    uint32_t lmax = 0;
#pragma unroll
    for (int i = 0; i < kListsPerWave; i++) {
      // the list's tile slice: the same in every lane, kept in scalar registers
      const BatchList LR = g->list[wv + i * kBatchWaves];
      const uint32_t st = cmem(LR.dir)[t], en = cmem(LR.dir)[t + 1];
      lbase[i] = uniform_ptr(LR.bent + st);
      llen[i] = uniform_u32(en - st);
      lthr[i] = uniform_u32(LR.thr);
      lspan[i] = uniform_u32(LR.span);
      lslot[i] = uniform_u32(LR.slot);
      lmax = llen[i] > lmax ? llen[i] : lmax;
    }
    for (uint32_t off = 0, first = 1; off < lmax || first; off += kWave * kListUnroll, first = 0) {
      uint2 ent[kListsPerWave][kListUnroll];
#pragma unroll
      for (int i = 0; i < kListsPerWave; i++)
#pragma unroll
        for (int u = 0; u < kListUnroll; u++) {
          const uint32_t e = off + u * kWave + lane;
          u32x2 x = {0u, 0u};
          if (e < llen[i]) x = reinterpret_cast<const RTS_GLOBAL u32x2 *>(gmem(lbase[i]))[e];
          ent[i][u] = make_uint2(x.x, x.y);
        }
      if (first) {
        // a wave-uniform choice: interior tiles load whole vectors with no per-lane bounds
        if (interior)
          xor_filter_vectors<TW, NT, true>(lds, g, nslots, w, words, tid);
        else
          xor_filter_vectors<TW, NT, false>(lds, g, nslots, w, words, tid);
      }
#pragma unroll
      for (int i = 0; i < kListsPerWave; i++) {
        const uint32_t thr = lthr[i], span = lspan[i];
        uint32_t *tg = lds.tog[lslot[i]];
#pragma unroll
        for (int u = 0; u < kListUnroll; u++) {
          // a block of kWave entries past the slice's end is skipped by a wave-uniform
          // branch (slices are a few hundred entries, so most rounds leave blocks unused)
          if (off + u * kWave < llen[i]) {
            const uint32_t e = off + u * kWave + lane;
            // rows of the slice lie in [row0, row0 + TW * 32): the word index is < TW
            if (e < llen[i] && ent[i][u].y - thr <= span)
              atomicXor(&tg[tog_index<TW, NT>((ent[i][u].x - row0) >> 5)], 0x80000000u >> (ent[i][u].x & 31));
          }
        }
      }
    }
    for (int s = 0; s < nslots; s++)
      if (g->slot[s].col) batch_direct<TW, NT>(g->slot[s], t, n, lds.tog[s]);
    if (qload) {
      lds.q[qi] = BatchQueryLds{Q.out ? Q.out : out + (int64_t)Q.id * out_stride, qs0 | qs1 << 8, (uint32_t)Q.empty};
      lds.qid[qi] = (uint32_t)Q.id;
    }
    // the next item's index, here rather than after the AND phase: waiting for the atomic
    // there would also wait for all of this item's bitmap stores (one counter covers both)
    if (tid == 0) lds.next = nblk_x + nxt;
    __syncthreads();
    // thread 0 writes lds.next again only in the next item, past this item's end barrier
    const int64_t it_next = (int64_t)lds.next;

    // 3. each query: AND of its slots, stored with its popcount.  Every thread touches only
    //    its own words from here on (and clears them for the next item).  In the last
    //    tile, the thread's words [w, w + kPer) hold nvalid words of the bitmap, the last
    //    of them the bitmap's last word when holds_last (32-bit values, worked out once
    //    per item).
    int nvalid = kPer;
    bool holds_last = false;
    if (!interior) {
      const int64_t left = words - w;
      nvalid = left <= 0 ? 0 : left >= kPer ? kPer : (int)left;
      holds_last = nvalid > 0 && left <= kPer;
    }
    // this thread's words of slot s, read and cleared for the next item (each slot belongs
    // to exactly one query, so each is read once)
    auto take_slot = [&](uint32_t s, uint32_t(&v)[kPer]) {
      if constexpr (TL::kUnit == 2) {
#pragma unroll
        for (int k = 0; k < TL::kUnits; k++) {
          uint2 *u = reinterpret_cast<uint2 *>(&lds.tog[s][0]) + k * kBatchThreads + tid;
          const uint2 x = *u;
          *u = make_uint2(0, 0);
          v[2 * k] = x.x;
          v[2 * k + 1] = x.y;
        }
      } else {
        v[0] = lds.tog[s][tid];
        lds.tog[s][tid] = 0;
      }
    };
    for (int qi = 0; qi < nq; qi++) {
      const BatchQueryLds D = lds.q[qi];
      const uint32_t slots = uniform_u32(D.slots);
      const uint32_t s0 = slots & 0xFF, s1 = slots >> 8;
      uint32_t acc[kPer];
      const uint32_t fill = D.empty ? 0u : ~0u;
#pragma unroll
      for (int k = 0; k < kPer; k++) acc[k] = fill;
      // a query has at most kBatchMaxCols slots, taken two at a time: the words of both
      // are read before either is used
#pragma unroll
      for (int j = 0; j < kBatchMaxCols; j += 2) {
        if (s0 + j < s1) {
          uint32_t a[kPer], b[kPer];
          take_slot(s0 + j, a);
          if (s0 + j + 1 < s1) {
            take_slot(s0 + j + 1, b);
          } else {
#pragma unroll
            for (int k = 0; k < kPer; k++) b[k] = ~0u;
          }
#pragma unroll
          for (int k = 0; k < kPer; k++) acc[k] &= a[k] & b[k];
        }
      }
      uint32_t pc = 0;
      if (interior) {  // every word of the tile is a full word of the bitmap
#pragma unroll
        for (int k = 0; k < kPer; k++) pc += __popc(acc[k]);
        store_words<kPer>(D.out + w, acc);
      } else if (nvalid > 0) {
        BITS *o = D.out + w;
#pragma unroll
        for (int k = 0; k < kPer; k++) {
          if (k >= nvalid) acc[k] = 0;
          else if (k == nvalid - 1 && holds_last) acc[k] &= last_mask;
          pc += __popc(acc[k]);
        }
        if (nvalid == kPer) {
          store_words<kPer>(o, acc);
        } else {
#pragma unroll
          for (int k = 0; k < kPer; k++)
            if (k < nvalid) gmem(o)[k] = acc[k];
        }
      }
      pc = wave_total(pc);  // DPP steps: no LDS round trips inside the AND phase
      if ((tid & (kWave - 1)) == 0 && pc) atomicAdd(&lds.qcount[qi], pc);
    }
    // past the barrier, faster waves may already store the next item's query descriptors:
    // read qid before it
    const uint32_t qid = tid < nq ? lds.qid[tid] : 0;
    __syncthreads();
    if (tid < nq) {  // read and cleared by the same thread: the next item starts at 0
      partials[(int64_t)qid * ntiles + t] = lds.qcount[tid];
      lds.qcount[tid] = 0;
    }
    it = it_next;
  }
  // the last workgroup to finish resets the counters for the next launch on the stream:
  // every other workgroup has taken its last item by then
  if (tid == 0) {
    __threadfence();
    if (atomicAdd(sched + kXcds * kSchedStride, 1ull) == gridDim.x - 1) {
      for (int x = 0; x < kXcds; x++) atomicExch(sched + x * kSchedStride, 0ull);
      atomicExch(sched + kXcds * kSchedStride, 0ull);
    }
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

// Planned form of one query-column before grouping.
struct PlannedBound {
  uint32_t base = kBaseZero;
  int fvk = -1;  // filter-vector index for kBaseFv
  const BITS *fv = nullptr;
  int list = -1;  // bin of the refine list, -1: none
  CODE thr = 0;
};

struct PlannedCol {
  int col = 0;
  bool direct = false;
  bool narrow = false;  // both bounds in one bin: one list pass over [lo_v, hi_v]
  int list_b0 = -1, list_b1 = -1;  // list form: bins list_b0..list_b1, rows in [lo_v, hi_v]
  CODE lo_v = 0, hi_v = 0;
  CODE dlo = 0, dspan = 0;
  PlannedBound lo, hi;
};

static PlannedBound plan_bound(const rts_bindex_dev *d, const BoundSpec &s) {
  PlannedBound b;
  if (s.fv < 0) {
    b.base = kBaseZero;
  } else if (s.fv >= d->bins - 1) {
    b.base = kBaseOnes;
  } else {
    b.base = kBaseFv;
    b.fvk = s.fv;
    b.fv = d->fvs + (size_t)s.fv * d->fv_stride;
  }
  b.list = s.list;
  b.thr = s.thr;
  return b;
}

// Planner cost model of one query-column, in bytes per row of the column.  BinDex form:
// the filter vectors it reads (4 bytes per 32 rows each) plus its refine lists' entries
// (8 bytes each, spread over the column); a narrow column (both bounds in one bin) reads
// one list and no filter vector.  Direct form: the column's 4 bytes per row, scaled by
// direct_weight / 1000 (the batched kernel reads a direct column's tile once for all its
// query-columns from L2 but spends VALU work on every row of every one).  Weight 0: never
// direct.
static bool direct_cheaper(const rts_bindex_dev *d, const BoundSpec &sl, const BoundSpec &sh, int direct_weight) {
  if (direct_weight <= 0) return false;
  const bool narrow = sl.list >= 0 && sl.list == sh.list;
  double fvs = 0, entries = 0;
  if (narrow) {
    entries = (double)list_entries(d, sl.list);
  } else {
    fvs = (sl.fv >= 0 && sl.fv < d->bins - 1) + (sh.fv >= 0 && sh.fv < d->bins - 1);
    entries = (double)list_entries(d, sl.list) + (double)list_entries(d, sh.list);
  }
  const double bindex_cost = fvs / 8.0 + entries * 8.0 / (double)d->n;
  const double direct_cost = 4.0 * direct_weight / 1000.0;
  return direct_cost < bindex_cost;
}

}  // namespace rts

using namespace rts;

struct rts_batch_plan {
  const BinDex *const *index = nullptr;  // the active set, per table column (nullptr: direct only)
  std::vector<std::vector<const BinDex *>> index_sets;  // alternative indexes of the table
  std::vector<int> set_bins;                             // engine bins of each set
  int active_set = 0;
  std::vector<const CODE *> columns;
  int ncols = 0;
  int64_t n = 0, words = 0;
  int max_queries = 0;
  int nq = 0;
  rts_batch_config cfg{};
  int64_t ntiles = 0;
  BatchGroup *d_groups = nullptr;  // max_queries records
  int *d_ngroups = nullptr;        // device copy of ngroups
  uint32_t *d_zero_dir = nullptr;  // ntiles + 1 zeros: the directory of an unused list
  int ngroups = 0;
  uint32_t *d_partials = nullptr;
  unsigned long long *d_sched = nullptr;  // per-XCD work-item counters + finished workgroups
  std::vector<BatchGroup> h_groups;
  std::vector<uint32_t *> dir_base;  // per column: bdir re-tiled for cfg.tile_words
  std::vector<int64_t> dir_tiles;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  BITS *graph_out = nullptr;
  unsigned long long *graph_counts = nullptr;
  int graph_nq = -1;
  int64_t graph_stride = 0;
  int stats_bindex_cols = 0, stats_direct_cols = 0, stats_narrow_cols = 0, stats_auto_direct = 0;
  int stats_list_cols = 0;
  std::vector<rts_query_spec> specs;  // the current queries, re-planned on a new tiling
  std::vector<BITS *> outs;           // their per-query outputs, if given
  bool has_outs = false;
};

namespace rts {

static bool valid_config(const rts_batch_config &c) {
  return (c.tile_words == 512 || c.tile_words == 1024 || c.tile_words == 2048) && (c.order == 0 || c.order == 1) &&
         (c.threads == 256 || c.threads == 512) &&
         c.queries_per_group >= 1 &&
         c.queries_per_group <= kBatchMaxGroup && c.blocks_per_cu >= 1 && c.blocks_per_cu <= 16 &&
         c.bins >= 0 && c.direct_weight >= 0 && c.direct_weight <= 1000000;
}

// Index set a config selects: bins == 0 the first, else the set with that bin count.
static int set_for(const rts_batch_plan *p, const rts_batch_config &c) {
  if (c.bins == 0) return 0;
  for (size_t i = 0; i < p->set_bins.size(); i++)
    if (p->set_bins[i] == c.bins) return (int)i;
  return -1;
}

// Engine bins of a set of indexes (0 when no column has one, -1 when they differ or an
// index does not cover n rows).
static int bins_of_set(const BinDex *const *indexes, int ncols, int64_t n) {
  int bins = 0;
  for (int c = 0; indexes && c < ncols; c++) {
    const BinDex *b = indexes[c];
    if (!b) continue;
    if (!b->dev || b->dev->n != n) return -1;
    if (bins && b->dev->bins != bins) return -1;
    bins = b->dev->bins;
  }
  return bins;
}

static void free_dirs(rts_batch_plan *p) {
  for (uint32_t *d : p->dir_base) (void)hipFree(d);
  p->dir_base.clear();
  p->dir_tiles.clear();
  (void)hipFree(p->d_zero_dir);
  p->d_zero_dir = nullptr;
}

static hipError_t build_dirs(rts_batch_plan *p, hipStream_t s) {
  free_dirs(p);
  const int64_t tile_rows = (int64_t)p->cfg.tile_words * 32;
  const int64_t ntiles = (p->n + tile_rows - 1) / tile_rows;
  p->dir_base.assign(p->ncols, nullptr);
  p->dir_tiles.assign(p->ncols, ntiles);
  hipError_t ze = hipMalloc((void **)&p->d_zero_dir, (ntiles + 1) * sizeof(uint32_t));
  if (ze == hipSuccess) ze = hipMemsetAsync(p->d_zero_dir, 0, (ntiles + 1) * sizeof(uint32_t), s);
  if (ze != hipSuccess) return ze;
  for (int c = 0; c < p->ncols; c++) {
    const BinDex *b = p->index[c];
    if (!b || !b->dev) continue;
    const rts_bindex_dev *d = b->dev;
    const int64_t items = (int64_t)d->bins * (ntiles + 1);
    hipError_t e = hipMalloc((void **)&p->dir_base[c], items * sizeof(uint32_t));
    if (e != hipSuccess) return e;
    e = launch_tile_directory(d, ntiles, tile_rows, p->dir_base[c], s);
    if (e != hipSuccess) return e;
  }
  p->ntiles = ntiles;
  return hipSuccess;
}

}  // namespace rts

namespace rts {

// Workgroups of k_batch<TW> that fit on one compute unit at once (LDS-bound).
template <int TW, int NT>
static int batch_resident_per_cu() {
  static int cached = 0;
  if (!cached) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_batch<TW, NT>, NT, 0) != hipSuccess || nb < 1)
      nb = 1;
    cached = nb;
  }
  return cached;
}

template <int TW, int NT>
static void launch_k_batch(rts_batch_plan *p, BITS *out, int64_t out_stride, hipStream_t s) {
  const int cus = std::max(1, device_cu_count());
  // a persistent grid larger than what is resident would leave its statically assigned
  // work items to a second round: clamp to the occupancy limit
  unsigned grid = (unsigned)(cus * std::min(p->cfg.blocks_per_cu, batch_resident_per_cu<TW, NT>()));
  grid = std::max<unsigned>(kXcds, grid / kXcds * kXcds);
  hipLaunchKernelGGL((k_batch<TW, NT>), dim3(grid), dim3(NT), 0, s, p->d_groups, p->d_ngroups, p->n,
                     p->words, p->ntiles, out, out_stride, p->d_partials, p->d_sched);
}

static hipError_t launch_batch(rts_batch_plan *p, BITS *out, int64_t out_stride,
                               unsigned long long *counts, hipStream_t s) {
  const bool small = p->cfg.threads == 256;
  switch (p->cfg.tile_words) {
    case 2048:
      small ? launch_k_batch<2048, 256>(p, out, out_stride, s) : launch_k_batch<2048, 512>(p, out, out_stride, s);
      break;
    case 1024:
      small ? launch_k_batch<1024, 256>(p, out, out_stride, s) : launch_k_batch<1024, 512>(p, out, out_stride, s);
      break;
    default:
      small ? launch_k_batch<512, 256>(p, out, out_stride, s) : launch_k_batch<512, 512>(p, out, out_stride, s);
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
  cfg->tile_words = 2048;
  cfg->queries_per_group = 32;
  cfg->blocks_per_cu = 4;
  cfg->order = 1;
  cfg->threads = 256;
  cfg->bins = 0;
  cfg->direct_weight = 0;
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
  const int bins = bins_of_set(indexes, ncols, n);
  if (bins < 0) {
    delete p;
    return set_error(hipErrorInvalidValue, "rts_batch_plan_create: indexes (length or bin counts differ)");
  }
  p->index_sets.emplace_back(ncols, nullptr);
  for (int c = 0; indexes && c < ncols; c++) p->index_sets[0][c] = indexes[c];
  p->set_bins.push_back(bins);
  p->index = p->index_sets[0].data();
  if (cfg) p->cfg = *cfg; else rts_batch_default_config(&p->cfg);
  if (!valid_config(p->cfg) || set_for(p, p->cfg) < 0) {
    delete p;
    return set_error(hipErrorInvalidValue, "rts_batch_plan_create: config");
  }
  p->active_set = set_for(p, p->cfg);
  p->index = p->index_sets[p->active_set].data();
  hipError_t e = hipMalloc((void **)&p->d_groups, (size_t)max_queries * sizeof(BatchGroup));
  if (e == hipSuccess) e = hipMalloc((void **)&p->d_ngroups, sizeof(int));
  if (e == hipSuccess) e = hipMalloc((void **)&p->d_sched, kSchedWords * sizeof(unsigned long long));
  if (e == hipSuccess) e = hipMemsetAsync(p->d_sched, 0, kSchedWords * sizeof(unsigned long long), stream);
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
  if (!p || !cfg || !valid_config(*cfg) || set_for(p, *cfg) < 0)
    return set_error(hipErrorInvalidValue, "rts_batch_plan_set_config");
  const int set = set_for(p, *cfg);
  // a new tiling or index set needs new list directories
  const bool retile = cfg->tile_words != p->cfg.tile_words || set != p->active_set;
  // the grouping depends on the order, the slots a workgroup's LDS holds, the group size
  // and which query-columns the cost model sends to the direct form
  const bool reorder = cfg->order != p->cfg.order || cfg->threads != p->cfg.threads ||
                       cfg->queries_per_group != p->cfg.queries_per_group ||
                       cfg->direct_weight != p->cfg.direct_weight;
  p->cfg = *cfg;
  p->active_set = set;
  p->index = p->index_sets[set].data();
  if (p->exec) {
    (void)hipGraphExecDestroy(p->exec);
    p->exec = nullptr;
  }
  if (p->graph) {
    (void)hipGraphDestroy(p->graph);
    p->graph = nullptr;
  }
  if (!retile && !reorder) return hipSuccess;
  hipError_t e = retile ? build_dirs(p, stream) : hipSuccess;
  if (e == hipSuccess && retile) {
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

hipError_t rts_batch_plan_add_indexes(rts_batch_plan *p, const BinDex *const *indexes, int *bins) {
  if (!p || !indexes) return set_error(hipErrorInvalidValue, "rts_batch_plan_add_indexes");
  const int b = bins_of_set(indexes, p->ncols, p->n);
  if (b <= 0) return set_error(hipErrorInvalidValue, "rts_batch_plan_add_indexes: indexes");
  for (size_t i = 0; i < p->index_sets.size(); i++) {
    if (p->set_bins[i] != b) continue;
    // replace the set with that bin count (not the active one: its directories are in use)
    if ((int)i == p->active_set) return set_error(hipErrorInvalidValue, "rts_batch_plan_add_indexes: active set");
    p->index_sets[i].assign(indexes, indexes + p->ncols);
    if (bins) *bins = b;
    return hipSuccess;
  }
  p->index_sets.emplace_back(indexes, indexes + p->ncols);
  p->set_bins.push_back(b);
  p->index = p->index_sets[p->active_set].data();  // the vector of sets may have moved
  if (bins) *bins = b;
  return hipSuccess;
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
  // 1. every query-column in BinDex or direct form
  p->stats_auto_direct = 0;
  std::vector<PlannedCol> pc((size_t)nq * kBatchMaxCols);
  std::vector<int> qcols(nq, 0), qempty(nq, 0);
  int st_bindex = 0, st_direct = 0, st_narrow = 0, st_list = 0;
  for (int q = 0; q < nq; q++) {
    const rts_query_spec &in = queries[q];
    if (in.ncols < 1 || in.ncols > kBatchMaxCols)
      return set_error(hipErrorInvalidValue, "rts_batch_plan_set_queries: ncols");
    qcols[q] = in.ncols;
    for (int c = 0; c < in.ncols; c++) {
      const int col = in.column[c];
      if (col < 0 || col >= p->ncols) return set_error(hipErrorInvalidValue, "rts_batch_plan_set_queries: column");
      const CODE lo = in.lo[c], hi = in.hi[c];
      if (lo > hi) qempty[q] = 1;
      PlannedCol &P = pc[(size_t)q * kBatchMaxCols + c];
      P.col = col;
      const BinDex *b = p->index[col];
      if (in.mode[c] == RTS_EVAL_BINDEX && !b)
        return set_error(hipErrorInvalidValue, "rts_batch_plan_set_queries: column has no index");
      const bool aligned = ((uintptr_t)p->columns[col] & 15) == 0;
      BoundSpec sl{-1, -1, 0}, sh{-1, -1, 0};
      bool direct = in.mode[c] == RTS_EVAL_DIRECT || !b;
      if (!direct) {
        sl = lo == 0 ? BoundSpec{-1, -1, 0} : bound_spec(b->dev, lo, false);
        sh = bound_spec(b->dev, hi + 1u, hi == 0xFFFFFFFFu);
        int b0, b1;
        if (in.mode[c] == RTS_EVAL_AUTO && aligned && !list_form(b->dev, sl, sh, &b0, &b1) &&
            direct_cheaper(b->dev, sl, sh, p->cfg.direct_weight)) {
          direct = true;
          p->stats_auto_direct++;
        }
      }
      if (direct) {
        if (!aligned)
          return set_error(hipErrorInvalidValue, "rts_batch_plan_set_queries: direct column not 16-byte aligned");
        P.direct = true;
        P.dlo = lo;
        P.dspan = hi - lo;
        st_direct++;
        continue;
      }
      const rts_bindex_dev *d = b->dev;
      P.lo = plan_bound(d, sl);
      P.hi = plan_bound(d, sh);
      // Both bounds in one bin: the result is the bin's list rows in [lo, hi] alone,
      // (1 ^ t_hi) & ~(1 ^ t_lo) = t_lo & ~t_hi, so no filter vector is read.
      int b0, b1;
      if (sl.fv == sh.fv && sl.list >= 0 && sl.list == sh.list) {
        P.lo.base = P.hi.base = kBaseOnes;
        P.lo.fv = P.hi.fv = nullptr;
        P.narrow = true;
        P.lo_v = lo;
        P.hi_v = hi;
        st_narrow++;
      } else if (list_form(d, sl, sh, &b0, &b1)) {
        P.lo.base = P.hi.base = kBaseZero;  // (0 & ~0) ^ T = T: the lists' rows alone
        P.lo.fv = P.hi.fv = nullptr;
        P.list_b0 = b0;
        P.list_b1 = b1;
        P.lo_v = lo;
        P.hi_v = hi;
        st_list++;
      }
      st_bindex++;
    }
  }
  // 2. the order queries are grouped in: the caller's, or sorted by the (column, lower and
  //    upper filter vector) of their BinDex columns taken in column order, so that the
  //    groups a tile's workgroups run together read few distinct filter vectors
  std::vector<int> order(nq);
  for (int q = 0; q < nq; q++) order[q] = q;
  if (p->cfg.order == 1) {
    std::vector<std::vector<int>> key(nq);
    for (int q = 0; q < nq; q++) {
      std::vector<std::array<int, 3>> cols;
      for (int c = 0; c < qcols[q]; c++) {
        const PlannedCol &P = pc[(size_t)q * kBatchMaxCols + c];
        cols.push_back(P.direct ? std::array<int, 3>{P.col, kMaxBins, kMaxBins}
                                : std::array<int, 3>{P.col, P.lo.fvk, P.hi.fvk});
      }
      std::sort(cols.begin(), cols.end());
      for (const auto &a : cols) key[q].insert(key[q].end(), a.begin(), a.end());
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key[a] < key[b]; });
  }
  // 3. groups, each one self-contained record of queries whose slots fit the LDS of one
  //    tile, at most the configured group size.  Work items, and their fixed costs, scale
  //    with the group count, so groups are packed full: first fit by decreasing slot count
  //    (4 | 3 + 1 | 2 + 2 for four slots), each size class taken in that order, and the
  //    groups run in the order of their first query.  Partners are then paired in order
  //    too (the i-th 3-column query with the i-th 1-column one), so neighbouring groups
  //    still share filter vectors in the locality order.
  const int slot_cap = kBatchLdsWords / 512 * p->cfg.threads / p->cfg.tile_words;
  const int64_t per = p->ntiles + 1;
  auto add_query = [&](BatchGroup &g, int q) {
    const int need = qempty[q] ? 0 : qcols[q];
    g.qslot[g.nq] = (uint8_t)g.nslots;
    for (int c = 0; c < need; c++) {
      const PlannedCol &P = pc[(size_t)q * kBatchMaxCols + c];
      BatchSlot &S = g.slot[g.nslots];
      if (P.direct) {
        S.col = p->columns[P.col];
        S.dlo = P.dlo;
        S.dspan = P.dspan;
      } else {
        S.fvhi = P.hi.fv;
        S.fvlo = P.lo.fv;
        S.hibase = P.hi.base;
        S.lobase = P.lo.base;
        const rts_bindex_dev *d = p->index[P.col]->dev;
        if (P.narrow) {  // t_lo ^ t_hi = the bin's rows with lo <= x <= hi: one pass
          BatchList &L = g.list[g.nlists++];
          L.dir = p->dir_base[P.col] + (size_t)P.lo.list * per;
          L.bent = d->bent;
          L.thr = P.lo_v;
          L.span = P.hi_v - P.lo_v;
          L.slot = (uint32_t)g.nslots;
        } else if (P.list_b0 >= 0) {  // the list form: its bins' rows with lo <= x <= hi
          for (int k = P.list_b0; k <= P.list_b1; k++) {
            BatchList &L = g.list[g.nlists++];
            L.dir = p->dir_base[P.col] + (size_t)k * per;
            L.bent = d->bent;
            L.thr = P.lo_v;
            L.span = P.hi_v - P.lo_v;
            L.slot = (uint32_t)g.nslots;
          }
        } else {
          for (const PlannedBound *B : {&P.lo, &P.hi}) {
            if (B->list < 0) continue;
            BatchList &L = g.list[g.nlists++];
            L.dir = p->dir_base[P.col] + (size_t)B->list * per;
            L.bent = d->bent;
            L.thr = B->thr;
            L.span = 0xFFFFFFFFu - B->thr;  // value >= thr
            L.slot = (uint32_t)g.nslots;
          }
        }
      }
      g.nslots++;
    }
    BatchQuery &Q = g.query[g.nq];
    Q.out = outs ? outs[q] : nullptr;
    Q.empty = qempty[q];
    Q.id = q;
    g.nq++;
    g.qslot[g.nq] = (uint8_t)g.nslots;
  };
  std::vector<int> byneed(nq);  // positions in the order, by decreasing slot count
  for (int i = 0; i < nq; i++) byneed[i] = i;
  auto need_at = [&](int oi) { return qempty[order[oi]] ? 0 : qcols[order[oi]]; };
  std::stable_sort(byneed.begin(), byneed.end(), [&](int a, int b) { return need_at(a) > need_at(b); });
  std::vector<BatchGroup> cg;
  cg.reserve(nq);
  std::vector<int> lead;                         // order position of each group's first query
  std::vector<std::set<int>> open(slot_cap + 1);  // groups with room, by free slots
  for (int oi : byneed) {
    const int need = need_at(oi);
    int k = -1, kf = 0;
    for (int f = std::max(need, 1); f <= slot_cap; f++)  // first fit: the earliest group with room
      if (!open[f].empty() && (k < 0 || *open[f].begin() < k)) k = *open[f].begin(), kf = f;
    if (k < 0) {
      k = (int)cg.size();
      cg.emplace_back();
      BatchGroup &g = cg.back();
      memset(&g, 0, sizeof(g));
      for (BatchList &L : g.list) L.dir = p->d_zero_dir;
      lead.push_back(oi);
    } else {
      open[kf].erase(k);
    }
    add_query(cg[k], order[oi]);
    const int free_slots = slot_cap - cg[k].nslots;
    if (free_slots > 0 && cg[k].nq < p->cfg.queries_per_group) open[free_slots].insert(k);
  }
  std::vector<int> gi(cg.size());
  for (size_t k = 0; k < cg.size(); k++) gi[k] = (int)k;
  std::sort(gi.begin(), gi.end(), [&](int a, int b) { return lead[a] < lead[b]; });
  std::vector<BatchGroup> groups;
  groups.reserve(cg.size());
  for (int k : gi) {
    cg[k].q0 = lead[k];
    groups.push_back(cg[k]);
  }
  p->has_outs = outs != nullptr;
  p->outs.assign(outs ? outs : (BITS *const *)nullptr, outs ? outs + nq : (BITS *const *)nullptr);
  p->stats_bindex_cols = st_bindex;
  p->stats_direct_cols = st_direct;
  p->stats_narrow_cols = st_narrow;
  p->stats_list_cols = st_list;
  p->nq = nq;
  p->specs.assign(queries, queries + nq);
  p->h_groups.swap(groups);
  p->ngroups = (int)p->h_groups.size();
  if (nq == 0) return hipSuccess;
  hipError_t e = hipMemcpyAsync(p->d_groups, p->h_groups.data(), (size_t)p->ngroups * sizeof(BatchGroup),
                                hipMemcpyHostToDevice, stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(p->d_ngroups, &p->ngroups, sizeof(int), hipMemcpyHostToDevice, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);  // the host vectors change on the next call
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

int rts_batch_plan_auto_direct(const rts_batch_plan *p) { return p ? p->stats_auto_direct : -1; }

int rts_batch_plan_list_cols(const rts_batch_plan *p) { return p ? p->stats_list_cols : -1; }

unsigned long long rts_batch_plan_mix_digest(const rts_batch_plan *p) {
  if (!p) return 0;
  // summary statistics of the batch, each coarsely quantised: mean columns per query (in
  // quarters), the share of query-columns per evaluation mode and per width class (in
  // sixteenths); width classes: empty, 1-8 of the 128 areas, 9-32, more
  int64_t qcols = 0, mode[3] = {0, 0, 0}, width[4] = {0, 0, 0, 0};
  for (const rts_query_spec &q : p->specs) {
    for (int c = 0; c < q.ncols && c < kBatchMaxCols; c++) {
      qcols++;
      mode[q.mode[c] >= 0 && q.mode[c] < 3 ? q.mode[c] : 0]++;
      const int col = q.column[c];
      const CODE lo = q.lo[c], hi = q.hi[c];
      const BinDex *b = col >= 0 && col < p->ncols ? p->index[col] : nullptr;
      double frac;
      if (lo > hi) {
        frac = 0;
      } else if (b && b->dev) {
        // areas (the reference's 128, the same for every bin count) the range touches
        const rts_bindex_dev *d = b->dev;
        const int a0 = lo <= d->starts[0] ? 0 : bin_of(d->starts, kAreas, lo);
        const int a1 = hi >= d->starts[kAreas - 1] ? kAreas - 1 : bin_of(d->starts, kAreas, hi);
        frac = (double)(a1 - a0 + 1) / kAreas;
      } else {
        frac = ((double)hi - (double)lo + 1.0) / 4294967296.0;
      }
      width[frac <= 0 ? 0 : frac <= 8.0 / kAreas ? 1 : frac <= 32.0 / kAreas ? 2 : 3]++;
    }
  }
  const size_t nq = p->specs.size();
  std::vector<int> feat;
  feat.push_back(nq ? (int)std::lround(4.0 * qcols / nq) : 0);
  for (int64_t m : mode) feat.push_back(qcols ? (int)std::lround(16.0 * m / qcols) : 0);
  for (int64_t w : width) feat.push_back(qcols ? (int)std::lround(16.0 * w / qcols) : 0);
  uint64_t h = 0xCBF29CE484222325ull;
  for (int v : feat)
    for (int i = 0; i < 4; i++) h = (h ^ (uint8_t)(v >> (8 * i))) * 0x100000001B3ull;
  return (unsigned long long)h;
}

int rts_batch_plan_groups(const rts_batch_plan *p, rts_batch_config *cfg) {
  if (!p) return -1;
  if (cfg) *cfg = p->cfg;
  return p->ngroups;
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
  (void)hipFree(p->d_groups);
  (void)hipFree(p->d_ngroups);
  (void)hipFree(p->d_partials);
  (void)hipFree(p->d_sched);
  delete p;
}

}  // extern "C"
