// bindex_dev.h -- the engine's device-resident BinDex index (behind BinDex::dev).
//
// Per column, the reference's K = 128 equal-frequency areas (init_bindex,
// bindex.cpp:599-671), each cut into `sub` equal-count bins (sub = 1, 2, 4 or 8), so the
// engine works with bins = 128 * sub bins whose boundaries include every area start:
//   sorted_val / sorted_pos   the column sorted by value (stable: ties in row order), n each
//   estarts                   bin start values (host), estarts[a * sub] = area a's start
//   fvs                       bins-1 filter vectors, fv[k] bit i = (x_i < estarts[k+1]),
//                             fv_stride words apart (stride padded to 64 words); the
//                             reference's filter vector of area a is fv[(a + 1) * sub - 1]
//   B lists                   for every bin k (value range [estarts[k], estarts[k+1])), the
//                             rows with estarts[k] < x < estarts[k+1] in row order, and the
//                             rows with x = estarts[k] too when there are at most
//                             kLightStartRows of them; one 8-byte entry per row:
//                             bent[e] = (row, value); bin k's entries are [boff[k],
//                             boff[k+1]); complete[k] says whether they are all of bin k's
//                             rows
//   bdir                      tile directory: bin k's first entry with row >= t * kTileRows
//                             is bdir[k * (ntiles + 1) + t]
//   rank_samples              every kRankSample-th sorted value (host): rank_samples[i] =
//                             sorted_val[i * kRankSample], for counts by rank
// A bound v of a query (the set L(v) = {x < v}) is then one filter vector plus the rows of
// one bin's B list with x >= v: L(v) = fv[t(v)] minus {B_t(v) rows with x >= v}, or
// fv[t(v) - 1] exactly when v is a bin start.  Rows equal to a bin start never need a
// refine (a refine threshold lies above its bin's start), so a heavily duplicated value
// (Zipf) is left out of its bin's list and costs nothing at query time; a start value of
// few rows is kept in the list, which then holds the whole bin, so that a range covering
// only a few complete bins can be answered from their lists alone.  More bins mean
// shorter refine lists per bound (2n/bins entries at most) for (bins - 1) * n / 8 bytes of
// filter vectors.
#pragma once

#include <stdint.h>

#include <algorithm>

#include <vector>

#include "common.h"
#include "rts_bindex.h"

namespace rts {

constexpr int kAreas = RTS_VAREA_N;            // the reference's K areas
constexpr int kMaxSub = 8;                     // bins per area, at most
constexpr int kMaxBins = kAreas * kMaxSub;     // 1024
constexpr int kTileWords = 1024;              // bitmap words per query-kernel tile
constexpr int kTileRows = kTileWords * 32;    // 32768 rows
constexpr int kFvStrideAlign = 64;            // words
constexpr int64_t kRankSample = 256;          // sorted slots per rank sample
constexpr int64_t kLightStartRows = 64;       // a bin start of at most this many rows is listed

}  // namespace rts

struct rts_bindex_dev {
  int64_t n = 0;
  int64_t words = 0;
  int64_t fv_stride = 0;
  int64_t ntiles = 0;
  int64_t nb = 0;  // total B-list entries
  int sub = 1;     // bins per area
  int bins = rts::kAreas;
  CODE *sorted_val = nullptr;
  int32_t *sorted_pos = nullptr;
  BITS *fvs = nullptr;
  uint2 *bent = nullptr;  // (row, value) per B-list entry
  uint32_t *boff = nullptr;  // device, bins + 1
  uint32_t *bdir = nullptr;  // device, bins * (ntiles + 1)
  CODE *owned_column = nullptr;
  // host copies
  CODE starts[rts::kAreas];  // the reference's area starts
  int64_t area_start[rts::kAreas];
  int64_t area_len[rts::kAreas];
  std::vector<CODE> estarts;         // bins
  std::vector<uint32_t> boff_host;   // bins + 1
  std::vector<CODE> rank_samples;    // ceil(n / kRankSample)
  std::vector<uint8_t> complete;     // bins: bin k's B list holds every row of bin k
  std::vector<Area> area_store;      // BinDex::areas
  std::vector<Area> area_gpu_store;  // BinDex::areasInGPU
  std::vector<CODE> block_first, block_last;
  // pos_block views (bindex.h:112-117): device views always, host views once the host
  // copies of the sorted arrays exist (rts_bindex_host_blocks)
  std::vector<pos_block> blk_dev, blk_host;
  std::vector<pos_block *> blkp_dev, blkp_host;
  std::vector<CODE> host_val;
  std::vector<int32_t> host_pos;
};

namespace rts {

// Bin of value v: the number of k in [1, bins-1] with estarts[k] <= v.
inline int bin_of(const CODE *estarts, int bins, CODE v) {
  int lo = 1, hi = bins;  // first k in [1, bins) with estarts[k] > v
  while (lo < hi) {
    const int mid = (lo + hi) / 2;
    if (estarts[mid] <= v)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo - 1;
}

// One query bound L(v) = {x < v} as (base filter vector, refine list).
struct BoundSpec {
  int fv;        // -1: no rows, bins - 1: all rows, else filter vector index
  int list;      // -1: no refine, else bin whose B-list rows with x >= thr leave the set
  CODE thr;
};

inline BoundSpec bound_spec(const rts_bindex_dev *d, CODE v, bool infinite) {
  if (infinite) return BoundSpec{d->bins - 1, -1, 0};  // v = 2^32: every row
  if (v <= d->estarts[0]) return BoundSpec{-1, -1, 0};  // v <= min: no row is < v
  const int k = bin_of(d->estarts.data(), d->bins, v);
  if (v == d->estarts[k]) return BoundSpec{k - 1, -1, 0};
  return BoundSpec{k, d->boff_host[k + 1] > d->boff_host[k] ? k : -1, v};
}

// List form of a BinDex query-column [lo, hi] (bounds sl, sh): the bins it covers are
// sl's own bin (only when sl refines: its rows above lo, which its list always holds) and
// the bins after it up to sh's bin; when these are at most two and every one but sl's own
// is complete (its list holds all its rows), the column's rows are the rows in [lo, hi]
// of those bins' lists, with no filter vector.  Two such lists are the very lists the
// filter-vector form refines, so the list form reads strictly less.
inline bool list_form(const rts_bindex_dev *d, const BoundSpec &sl, const BoundSpec &sh, int *b0, int *b1) {
  if ((int)d->complete.size() != d->bins) return false;
  const int first = sl.list >= 0 ? sl.fv : sl.fv + 1, last = sh.fv;
  if (first < 0 || last < first || last - first > 1 || last >= d->bins) return false;
  for (int k = sl.fv + 1; k <= last; k++)
    if (k >= first && !d->complete[k]) return false;
  *b0 = first;
  *b1 = last;
  return true;
}

// Tile directory of an index for tiles of `tile_rows` rows: bdir[k * (ntiles + 1) + t] =
// first B-list entry of bin k with row >= t * tile_rows (asynchronous on s).
hipError_t launch_tile_directory(const rts_bindex_dev *d, int64_t ntiles, int64_t tile_rows, uint32_t *bdir,
                                 hipStream_t s);

// d->rank_samples (host) from the device's sorted values, on s (synchronous).
hipError_t rank_samples(rts_bindex_dev *d, hipStream_t s);

// Ranks of the bin starts in the device's sorted values (synchronous): rk[2k] = #{x <
// estarts[k]}, rk[2k + 1] = #{x <= estarts[k]}.
hipError_t start_ranks(const rts_bindex_dev *d, std::vector<int64_t> &rk, hipStream_t s);

// d->complete from the start ranks and d->boff_host: bin k's list is complete when it has
// as many entries as bin k has rows.
void mark_complete_bins(rts_bindex_dev *d, const std::vector<int64_t> &rk);

// Sorted-slot interval [*L, *R] (fewer than kRankSample slots apart) that holds rank(v) =
// #{x < v}, from the host's samples (v = 2^32 when infinite: rank n).  With c samples below
// v, slot (c - 1) * kRankSample holds a value < v and slot c * kRankSample (if any) one >= v.
inline void rank_interval(const rts_bindex_dev *d, CODE v, bool infinite, int64_t *L, int64_t *R) {
  const std::vector<CODE> &smp = d->rank_samples;
  if (infinite) {
    *L = *R = d->n;
    return;
  }
  const int64_t c = std::lower_bound(smp.begin(), smp.end(), v) - smp.begin();
  if (c == 0) {  // v <= the column's minimum
    *L = *R = 0;
    return;
  }
  *L = (c - 1) * kRankSample + 1;
  *R = std::min<int64_t>(c * kRankSample, d->n);
}

// Refine-list entries of bin k.
inline int64_t list_entries(const rts_bindex_dev *d, int k) {
  return k < 0 ? 0 : (int64_t)d->boff_host[k + 1] - d->boff_host[k];
}

}  // namespace rts
