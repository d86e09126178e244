// bindex_build.hip -- GPU build of the device-resident BinDex (init_bindex,
// bindex.cpp:599-671, and its GPU-driver twin init_bindex_in_GPU, bindex_cuda.cpp:164-282).
//
// The reference sorts on the host with an unstable parallel sort, builds the 127 filter
// vectors with 127 host passes, then issues one cudaMalloc + cudaMemcpy per 3276-row block
// (~30k per column at 1e8 rows).  Here everything stays on the device: one stable radix
// sort of (value, row), a K-thread kernel for the area starts and one for the bin starts
// (each area cut into `sub` equal-count bins), one pass over the column for all bins-1
// filter vectors and the bin keys, one stable radix sort of the bin keys for the
// row-ordered B lists, and a directory kernel.
#include <hipcub/hipcub.hpp>

#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>

#include "bindex_dev.h"

namespace rts {

__global__ void k_iota(uint32_t *p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = (uint32_t)i;
}

// Area starts (bindex.cpp:613-631): start[i] = sorted[i * avg]; if it equals the previous
// start the area begins mid-run at i * avg, otherwise at the first occurrence of the value.
__global__ void k_area_starts(const CODE *sorted, int64_t n, CODE *starts, int64_t *idx) {
  const int i = threadIdx.x;
  if (i >= kAreas) return;
  const int64_t avg = n / kAreas;
  const CODE v = sorted[i * avg];
  starts[i] = v;
  if (i == 0) {
    idx[0] = 0;
    return;
  }
  if (v == sorted[(i - 1) * avg]) {
    idx[i] = i * avg;
    return;
  }
  int64_t lo = (i - 1) * avg, hi = i * avg;  // first slot in (lo, hi] holding v
  while (lo + 1 < hi) {
    const int64_t mid = lo + (hi - lo) / 2;
    if (sorted[mid] < v)
      lo = mid;
    else
      hi = mid;
  }
  idx[i] = hi;
}

// Bin starts: area a (sorted slots [idx[a], idx[a + 1])) cut into `sub` equal-count bins;
// bin a * sub + j starts at the value in slot idx[a] + j * len / sub, so bin a * sub is
// the area's own start.
__global__ void k_bin_starts(const CODE *sorted, int64_t n, const int64_t *idx, int sub, CODE *estarts) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= kAreas * sub) return;
  const int a = b / sub, j = b % sub;
  const int64_t s0 = idx[a], len = (a + 1 < kAreas ? idx[a + 1] : n) - s0;
  estarts[b] = sorted[s0 + (int64_t)j * len / sub];
}

// out[i] = sorted[i * kRankSample] for i < m.
__global__ void k_rank_samples(const CODE *sorted, int64_t m, CODE *out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) out[i] = sorted[i * kRankSample];
}

// One pass over the column: all bins-1 filter vectors (fv[k] bit i = x_i < estart[k+1],
// set_fv_val_less, bindex.cpp:553-560) and each row's list key (its bin if estart[bin] <
// x, or x = estart[bin] with light[bin] set, else `bins` = in no list).  A wave takes 2048
// consecutive rows, 64 words of every filter vector: lane l the 32 rows of word l.  The
// lane finds its rows' bins t (binary search in LDS) and keeps them bit-sliced: NB words,
// word b holding bit b of the 32 rows' t, MSB-first like the bitmaps.  Then for each k,
// fv[k] = fv[k - 1] | {rows with t = k}, the equality taken over the NB slices (two
// instructions a slice), and the 64 lanes store their words with one instruction.  A row
// past n has t = bins, in no filter vector; the words past the last (up to the 64-word
// padded stride) are written as zeros.  NB = bits of `bins` + 1 (t <= bins).
constexpr int kFvTileRows = 2048;
template <int NB>
__global__ __launch_bounds__(kBlock) void k_build_fvs(const CODE *col, int64_t n,
                                                      const CODE *starts_g, const uint8_t *light_g,
                                                      int bins, BITS *fvs, int64_t fv_stride,
                                                      uint16_t *keys) {
  __shared__ CODE starts[kMaxBins];
  __shared__ uint8_t light[kMaxBins];
  for (int i = threadIdx.x; i < bins; i += blockDim.x) {
    starts[i] = starts_g[i];
    light[i] = light_g[i];
  }
  __syncthreads();
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t wave = (int64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
  const int64_t nwaves = (int64_t)gridDim.x * (kBlock / kWave);
  const int64_t ntiles = (n + kFvTileRows - 1) / kFvTileRows;
  const bool aligned = ((uintptr_t)col & 15) == 0;  // 16-byte loads of the column
  for (int64_t tile = wave; tile < ntiles; tile += nwaves) {
    const int64_t r0 = tile * kFvTileRows + 32 * lane;  // this lane's 32 rows
    uint32_t slice[NB];
#pragma unroll
    for (int b = 0; b < NB; b++) slice[b] = 0;
    const bool whole = r0 + 32 <= n;
    uint32_t kw[16];  // the 32 rows' keys, two per word
#pragma unroll
    for (int q = 0; q < 8; q++) {
      uint4 v = make_uint4(0, 0, 0, 0);
      if (whole && aligned) {
        v = *reinterpret_cast<const uint4 *>(col + r0 + 4 * q);
      } else {
        if (r0 + 4 * q < n) v.x = col[r0 + 4 * q];
        if (r0 + 4 * q + 1 < n) v.y = col[r0 + 4 * q + 1];
        if (r0 + 4 * q + 2 < n) v.z = col[r0 + 4 * q + 2];
        if (r0 + 4 * q + 3 < n) v.w = col[r0 + 4 * q + 3];
      }
      const CODE xs[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int j = 4 * q + i;
        const CODE x = xs[i];
        int lo = 1, hi = bins;  // t = number of k in [1, bins-1] with start[k] <= x
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (starts[mid] <= x)
            lo = mid + 1;
          else
            hi = mid;
        }
        const int t0 = lo - 1;
        const bool valid = r0 + j < n;
        const uint32_t t = valid ? (uint32_t)t0 : (uint32_t)bins;
        const uint32_t key = starts[t0] < x || light[t0] ? (uint32_t)t0 : (uint32_t)bins;
        if (j & 1)
          kw[j / 2] |= key << 16;
        else
          kw[j / 2] = key;
#pragma unroll
        for (int b = 0; b < NB; b++) slice[b] = (slice[b] << 1) | ((t >> b) & 1u);
      }
    }
    if (whole) {
      uint4 *kd = reinterpret_cast<uint4 *>(keys + r0);
#pragma unroll
      for (int q = 0; q < 4; q++) kd[q] = make_uint4(kw[4 * q], kw[4 * q + 1], kw[4 * q + 2], kw[4 * q + 3]);
    } else {
#pragma unroll
      for (int j = 0; j < 32; j++)
        if (r0 + j < n) keys[r0 + j] = (uint16_t)(kw[j / 2] >> (16 * (j & 1)));
    }
    BITS *dst = fvs + tile * (kFvTileRows / 32) + lane;
    uint32_t acc = 0;  // fv[k] restricted to this lane's word
    for (int k = 0; k < bins - 1; k++) {
      uint32_t eq = ~0u;
#pragma unroll
      for (int b = 0; b < NB; b++) eq &= slice[b] ^ (((uint32_t)k >> b) & 1u ? 0u : ~0u);
      acc |= eq;
      dst[k * fv_stride] = acc;
    }
  }
}

// B-list entries: (row, value of the row).
__global__ void k_gather_entries(const CODE *col, const uint32_t *pos, uint2 *ent, int64_t m) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
       i += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t p = pos[i];
    ent[i] = make_uint2(p, col[p]);
  }
}

// boff[k] = first sorted key >= k, for k in [0, bins]
__global__ void k_bin_offsets(const uint16_t *skeys, int64_t n, int bins, uint32_t *boff) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k > bins) return;
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = lo + (hi - lo) / 2;
    if ((int)skeys[mid] < k)
      lo = mid + 1;
    else
      hi = mid;
  }
  boff[k] = (uint32_t)lo;
}

// bdir[k * (ntiles + 1) + t] = first entry of bin k with row >= t * tile_rows
__global__ void k_tile_directory(const uint2 *bent, const uint32_t *boff, int bins, int64_t ntiles,
                                 int64_t tile_rows, uint32_t *bdir) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t per = ntiles + 1;
  if (id >= per * bins) return;
  const int k = (int)(id / per);
  const uint64_t row = (uint64_t)(id % per) * (uint64_t)tile_rows;
  uint32_t lo = boff[k], hi = boff[k + 1];
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (bent[mid].x < row)
      lo = mid + 1;
    else
      hi = mid;
  }
  bdir[id] = lo;
}

hipError_t launch_tile_directory(const rts_bindex_dev *d, int64_t ntiles, int64_t tile_rows, uint32_t *bdir,
                                 hipStream_t s) {
  const int64_t items = (int64_t)d->bins * (ntiles + 1);
  hipLaunchKernelGGL(k_tile_directory, dim3((unsigned)((items + kBlock - 1) / kBlock)), dim3(kBlock), 0, s,
                     d->bent, d->boff, d->bins, ntiles, tile_rows, bdir);
  return hipGetLastError();
}

// First and last value of every 3276-row block of every area (init_area,
// bindex.cpp:410-427), for the host-side in_which_block.
__global__ void k_block_bounds(const CODE *sorted, const int64_t *blk_first_slot,
                               const int64_t *blk_last_slot, int64_t nblocks, CODE *first,
                               CODE *last) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nblocks) return;
  first[i] = sorted[blk_first_slot[i]];
  last[i] = sorted[blk_last_slot[i]];
}

static unsigned blocks_for(int64_t items, int per = kBlock) {
  int64_t g = (items + per - 1) / per;
  const int64_t cap = (int64_t)device_cu_count() * 16;
  return (unsigned)std::max<int64_t>(1, std::min(g, cap));
}

#define RTS_TRY(x)                      \
  do {                                  \
    hipError_t _e = (x);                \
    if (_e != hipSuccess) {             \
      err = _e;                         \
      goto fail;                        \
    }                                   \
  } while (0)

// rk[2k] = first sorted slot with value >= estarts[k], rk[2k + 1] = first with value >
// estarts[k] (binary searches over the sorted column, one thread per bin start).
__global__ void k_start_ranks(const CODE *sorted, int64_t n, const CODE *estarts, int bins, int64_t *rk) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= bins) return;
  const CODE v = estarts[k];
  int64_t lo = 0, hi = n;  // first slot with sorted >= v
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (sorted[mid] < v)
      lo = mid + 1;
    else
      hi = mid;
  }
  rk[2 * k] = lo;
  hi = n;  // first slot with sorted > v
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (sorted[mid] <= v)
      lo = mid + 1;
    else
      hi = mid;
  }
  rk[2 * k + 1] = lo;
}

hipError_t start_ranks(const rts_bindex_dev *d, std::vector<int64_t> &rk, hipStream_t s) {
  const int bins = d->bins;
  CODE *est = nullptr;
  int64_t *rk_d = nullptr;
  rk.assign(2 * (size_t)bins, 0);
  hipError_t e = hipMalloc((void **)&est, bins * sizeof(CODE));
  if (e == hipSuccess) e = hipMalloc((void **)&rk_d, 2 * (size_t)bins * sizeof(int64_t));
  if (e == hipSuccess) e = hipMemcpyAsync(est, d->estarts.data(), bins * sizeof(CODE), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_start_ranks, dim3((bins + kBlock - 1) / kBlock), dim3(kBlock), 0, s, d->sorted_val, d->n,
                       est, bins, rk_d);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(rk.data(), rk_d, 2 * (size_t)bins * sizeof(int64_t), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(est);
  (void)hipFree(rk_d);
  return e;
}

void mark_complete_bins(rts_bindex_dev *d, const std::vector<int64_t> &rk) {
  d->complete.assign(d->bins, 0);
  for (int k = 0; k < d->bins; k++) {
    const int64_t rows = (k + 1 < d->bins ? rk[2 * (k + 1)] : d->n) - rk[2 * k];
    d->complete[k] = (int64_t)(d->boff_host[k + 1] - d->boff_host[k]) == rows;
  }
}

hipError_t rank_samples(rts_bindex_dev *d, hipStream_t s) {
  const int64_t m = (d->n + kRankSample - 1) / kRankSample;
  CODE *smp = nullptr;
  d->rank_samples.assign(m, 0);
  hipError_t e = hipMalloc((void **)&smp, m * sizeof(CODE));
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_rank_samples, dim3((unsigned)((m + kBlock - 1) / kBlock)), dim3(kBlock), 0, s,
                       d->sorted_val, m, smp);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(d->rank_samples.data(), smp, m * sizeof(CODE), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(smp);
  return e;
}

static void free_dev(rts_bindex_dev *d) {
  if (!d) return;
  (void)hipFree(d->sorted_val);
  (void)hipFree(d->sorted_pos);
  (void)hipFree(d->fvs);
  (void)hipFree(d->bent);
  (void)hipFree(d->boff);
  (void)hipFree(d->bdir);
  (void)hipFree(d->owned_column);
  delete d;
}

// pos_block views of every block over the sorted arrays `val` / `pos` (device or host):
// block b of area k covers slots [start_k + b * 3276, ...) (init_pos_block,
// bindex.cpp:300-309).
static void make_block_views(const rts_bindex_dev *d, CODE *val, int32_t *pos, std::vector<pos_block> &blk,
                             std::vector<pos_block *> &ptrs) {
  blk.clear();
  for (int k = 0; k < kAreas; k++) {
    const int64_t len = d->area_len[k];
    const int nb = (int)(len <= RTS_BLOCK_INIT_SIZE ? 1 : (len + RTS_BLOCK_INIT_SIZE - 1) / RTS_BLOCK_INIT_SIZE);
    for (int b = 0; b < nb; b++) {
      const int64_t s0 = d->area_start[k] + (int64_t)b * RTS_BLOCK_INIT_SIZE;
      const int64_t s1 = std::min<int64_t>(s0 + RTS_BLOCK_INIT_SIZE, d->area_start[k] + len);
      blk.push_back(pos_block{(POSTYPE *)(pos + s0), val + s0, (int)(s1 - s0)});
    }
  }
  ptrs.resize(blk.size());
  for (size_t i = 0; i < blk.size(); i++) ptrs[i] = &blk[i];
}

// The reference-shaped descriptor (bindex.h:126-138) over a complete device index: areas
// and their blocks (bindex.cpp:633-652), start values, counts, filter-vector pointers.
void describe(BinDex *bx, rts_bindex_dev *d, const CODE *col) {
  d->area_store.resize(kAreas);
  d->area_gpu_store.resize(kAreas);
  make_block_views(d, d->sorted_val, d->sorted_pos, d->blk_dev, d->blkp_dev);
  int64_t b0 = 0;
  for (int k = 0; k < kAreas; k++) {
    Area &a = d->area_store[k];
    const int64_t len = d->area_len[k];
    a.blockNum = (int)(len <= RTS_BLOCK_INIT_SIZE ? 1 : (len + RTS_BLOCK_INIT_SIZE - 1) / RTS_BLOCK_INIT_SIZE);
    a.length = (int)len;
    a.start = (POSTYPE)d->area_start[k];
    a.blockFirst = d->block_first.data() + b0;
    a.blockLast = d->block_last.data() + b0;
    a.blocks = d->blkp_host.empty() ? nullptr : d->blkp_host.data() + b0;
    Area &g = d->area_gpu_store[k];
    g = a;
    g.blocks = d->blkp_dev.data() + b0;
    b0 += a.blockNum;
  }
  bx->length = (POSTYPE)d->n;
  bx->data_min = d->starts[0];
  for (int k = 0; k < kAreas; k++) {
    bx->areas[k] = &d->area_store[k];
    bx->areaStartValues[k] = d->starts[k];
    bx->area_counts[k] = (POSTYPE)(d->area_start[k] + d->area_len[k]);
    bx->areasInGPU[k] = &d->area_gpu_store[k];
  }
  for (int k = 0; k < kAreas - 1; k++) {
    bx->filterVectors[k] = nullptr;
    bx->filterVectorsInGPU[k] = d->fvs + (size_t)((k + 1) * d->sub - 1) * d->fv_stride;
  }
  bx->rawDataInGPU = const_cast<CODE *>(col);
  bx->dev = d;
}

// Host copies of the sorted arrays and the host pos_block views over them.
hipError_t bindex_host_blocks(BinDex *bx) {
  if (!bx || !bx->dev) return hipErrorInvalidValue;
  rts_bindex_dev *d = bx->dev;
  if (!d->blkp_host.empty()) return hipSuccess;
  d->host_val.resize(d->n);
  d->host_pos.resize(d->n);
  hipError_t e = hipMemcpy(d->host_val.data(), d->sorted_val, d->n * sizeof(CODE), hipMemcpyDeviceToHost);
  if (e == hipSuccess)
    e = hipMemcpy(d->host_pos.data(), d->sorted_pos, d->n * sizeof(int32_t), hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    std::vector<CODE>().swap(d->host_val);
    std::vector<int32_t>().swap(d->host_pos);
    return e;
  }
  make_block_views(d, d->host_val.data(), d->host_pos.data(), d->blk_host, d->blkp_host);
  int64_t b0 = 0;
  for (int k = 0; k < kAreas; k++) {
    d->area_store[k].blocks = d->blkp_host.data() + b0;
    b0 += d->area_store[k].blockNum;
  }
  return hipSuccess;
}

hipError_t bindex_build(BinDex *bx, const CODE *col, int64_t n, CODE *owned, int sub, hipStream_t s) {
  hipError_t err = hipSuccess;
  rts_bindex_dev *d = nullptr;
  uint32_t *iota = nullptr, *bpos_in = nullptr;
  uint16_t *keys = nullptr, *skeys = nullptr;
  void *tmp = nullptr;
  size_t tmp_bytes = 0, need = 0;
  CODE *starts_d = nullptr, *estarts_d = nullptr;
  int64_t *idx_d = nullptr, *bf_slot_d = nullptr, *bl_slot_d = nullptr;
  CODE *bfirst_d = nullptr, *blast_d = nullptr;
  std::vector<int64_t> bf_slot, bl_slot;
  int64_t nblocks = 0;
  int64_t idx_h[kAreas];
  uint32_t nb = 0;
  int bins = 0, key_bits = 0;
  std::vector<int64_t> rk;
  uint8_t light_h[kMaxBins];
  uint8_t *light_d = nullptr;

  if (!bx || !col || n < kAreas || n > 0x7FFFFFFFll || (sub != 1 && sub != 2 && sub != 4 && sub != 8))
    return hipErrorInvalidValue;
  memset(bx, 0, sizeof(*bx));
  d = new rts_bindex_dev();
  d->n = n;
  d->words = (n + 31) / 32;
  d->fv_stride = (d->words + kFvStrideAlign - 1) / kFvStrideAlign * kFvStrideAlign;
  d->ntiles = (n + kTileRows - 1) / kTileRows;
  d->owned_column = owned;
  d->sub = sub;
  d->bins = bins = kAreas * sub;
  while ((1 << key_bits) <= bins) key_bits++;  // keys 0..bins

  // 1. stable sort of (value, row)
  RTS_TRY(hipMalloc((void **)&d->sorted_val, n * sizeof(CODE)));
  RTS_TRY(hipMalloc((void **)&d->sorted_pos, n * sizeof(int32_t)));
  RTS_TRY(hipMalloc((void **)&iota, n * sizeof(uint32_t)));
  hipLaunchKernelGGL(k_iota, dim3(blocks_for(n)), dim3(kBlock), 0, s, iota, n);
  RTS_TRY(hipGetLastError());
  RTS_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, need, col, d->sorted_val, iota,
                                             (uint32_t *)d->sorted_pos, (int)n, 0, 32, s));
  tmp_bytes = need;
  RTS_TRY(hipMalloc(&tmp, tmp_bytes));
  RTS_TRY(hipcub::DeviceRadixSort::SortPairs(tmp, need, col, d->sorted_val, iota,
                                             (uint32_t *)d->sorted_pos, (int)n, 0, 32, s));

  // 2. area starts, then bin starts
  RTS_TRY(hipMalloc((void **)&starts_d, kAreas * sizeof(CODE)));
  RTS_TRY(hipMalloc((void **)&idx_d, kAreas * sizeof(int64_t)));
  RTS_TRY(hipMalloc((void **)&estarts_d, bins * sizeof(CODE)));
  hipLaunchKernelGGL(k_area_starts, dim3(1), dim3(kAreas), 0, s, d->sorted_val, n, starts_d, idx_d);
  RTS_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_bin_starts, dim3((bins + kBlock - 1) / kBlock), dim3(kBlock), 0, s, d->sorted_val, n,
                     idx_d, sub, estarts_d);
  RTS_TRY(hipGetLastError());
  d->estarts.resize(bins);
  RTS_TRY(hipMemcpyAsync(d->starts, starts_d, sizeof(d->starts), hipMemcpyDeviceToHost, s));
  RTS_TRY(hipMemcpyAsync(d->estarts.data(), estarts_d, bins * sizeof(CODE), hipMemcpyDeviceToHost, s));
  RTS_TRY(hipMemcpyAsync(idx_h, idx_d, sizeof(idx_h), hipMemcpyDeviceToHost, s));
  RTS_TRY(hipMemcpyAsync(&bx->data_max, d->sorted_val + n - 1, sizeof(CODE), hipMemcpyDeviceToHost, s));

  // 3. filter vectors + list keys; a bin start of at most kLightStartRows rows goes into
  //    its bin's list
  RTS_TRY(hipStreamSynchronize(s));  // d->estarts
  RTS_TRY(start_ranks(d, rk, s));
  for (int k = 0; k < bins; k++) light_h[k] = rk[2 * k + 1] - rk[2 * k] <= kLightStartRows;
  RTS_TRY(hipMalloc((void **)&light_d, bins));
  RTS_TRY(hipMemcpyAsync(light_d, light_h, bins, hipMemcpyHostToDevice, s));
  RTS_TRY(hipMalloc((void **)&d->fvs, (size_t)(bins - 1) * d->fv_stride * sizeof(BITS)));
  RTS_TRY(hipMalloc((void **)&keys, n * sizeof(uint16_t)));
  {
    const dim3 grid(blocks_for((n + kFvTileRows - 1) / kFvTileRows, kBlock / kWave));
    if (bins <= 128)
      hipLaunchKernelGGL(k_build_fvs<8>, grid, dim3(kBlock), 0, s, col, n, estarts_d, light_d, bins, d->fvs,
                         d->fv_stride, keys);
    else if (bins <= 256)
      hipLaunchKernelGGL(k_build_fvs<9>, grid, dim3(kBlock), 0, s, col, n, estarts_d, light_d, bins, d->fvs,
                         d->fv_stride, keys);
    else if (bins <= 512)
      hipLaunchKernelGGL(k_build_fvs<10>, grid, dim3(kBlock), 0, s, col, n, estarts_d, light_d, bins, d->fvs,
                         d->fv_stride, keys);
    else
      hipLaunchKernelGGL(k_build_fvs<11>, grid, dim3(kBlock), 0, s, col, n, estarts_d, light_d, bins, d->fvs,
                         d->fv_stride, keys);
  }
  RTS_TRY(hipGetLastError());

  // 4. B lists: stable sort of (key, row) -> rows grouped by bin in row order
  RTS_TRY(hipMalloc((void **)&skeys, n * sizeof(uint16_t)));
  RTS_TRY(hipMalloc((void **)&bpos_in, n * sizeof(uint32_t)));
  RTS_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, need, keys, skeys, iota, bpos_in, (int)n, 0,
                                             key_bits, s));
  if (need > tmp_bytes) {
    (void)hipFree(tmp);
    tmp = nullptr;
    tmp_bytes = need;
    RTS_TRY(hipMalloc(&tmp, tmp_bytes));
  }
  RTS_TRY(hipcub::DeviceRadixSort::SortPairs(tmp, need, keys, skeys, iota, bpos_in, (int)n, 0, key_bits, s));
  RTS_TRY(hipMalloc((void **)&d->boff, (bins + 1) * sizeof(uint32_t)));
  hipLaunchKernelGGL(k_bin_offsets, dim3((bins + 1 + kBlock - 1) / kBlock), dim3(kBlock), 0, s, skeys, n, bins,
                     d->boff);
  RTS_TRY(hipGetLastError());
  d->boff_host.resize(bins + 1);
  RTS_TRY(hipMemcpyAsync(d->boff_host.data(), d->boff, (bins + 1) * sizeof(uint32_t),
                         hipMemcpyDeviceToHost, s));
  RTS_TRY(hipStreamSynchronize(s));
  nb = d->boff_host[bins];
  d->nb = nb;
  mark_complete_bins(d, rk);
  RTS_TRY(hipMalloc((void **)&d->bent, std::max<uint32_t>(nb, 1) * sizeof(uint2)));
  if (nb) {
    hipLaunchKernelGGL(k_gather_entries, dim3(blocks_for(nb)), dim3(kBlock), 0, s, col, bpos_in,
                       d->bent, (int64_t)nb);
    RTS_TRY(hipGetLastError());
  }
  RTS_TRY(hipMalloc((void **)&d->bdir, (size_t)bins * (d->ntiles + 1) * sizeof(uint32_t)));
  RTS_TRY(launch_tile_directory(d, d->ntiles, kTileRows, d->bdir, s));

  // 5. host descriptor: areas (bindex.cpp:633-652) and their blocks
  for (int k = 0; k < kAreas; k++) {
    d->area_start[k] = idx_h[k];
    d->area_len[k] = (k < kAreas - 1 ? idx_h[k + 1] : n) - idx_h[k];
  }
  for (int k = 0; k < kAreas; k++) {
    const int64_t len = d->area_len[k];
    int64_t i = 0;
    while (i + RTS_BLOCK_INIT_SIZE < len) {
      bf_slot.push_back(d->area_start[k] + i);
      bl_slot.push_back(d->area_start[k] + i + RTS_BLOCK_INIT_SIZE - 1);
      i += RTS_BLOCK_INIT_SIZE;
    }
    bf_slot.push_back(d->area_start[k] + i);
    bl_slot.push_back(d->area_start[k] + len - 1);
  }
  nblocks = (int64_t)bf_slot.size();
  d->block_first.resize(nblocks);
  d->block_last.resize(nblocks);
  RTS_TRY(hipMalloc((void **)&bf_slot_d, nblocks * sizeof(int64_t)));
  RTS_TRY(hipMalloc((void **)&bl_slot_d, nblocks * sizeof(int64_t)));
  RTS_TRY(hipMalloc((void **)&bfirst_d, nblocks * sizeof(CODE)));
  RTS_TRY(hipMalloc((void **)&blast_d, nblocks * sizeof(CODE)));
  RTS_TRY(hipMemcpyAsync(bf_slot_d, bf_slot.data(), nblocks * sizeof(int64_t), hipMemcpyHostToDevice, s));
  RTS_TRY(hipMemcpyAsync(bl_slot_d, bl_slot.data(), nblocks * sizeof(int64_t), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_block_bounds, dim3((unsigned)((nblocks + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                     s, d->sorted_val, bf_slot_d, bl_slot_d, nblocks, bfirst_d, blast_d);
  RTS_TRY(hipGetLastError());
  RTS_TRY(hipMemcpyAsync(d->block_first.data(), bfirst_d, nblocks * sizeof(CODE), hipMemcpyDeviceToHost, s));
  RTS_TRY(hipMemcpyAsync(d->block_last.data(), blast_d, nblocks * sizeof(CODE), hipMemcpyDeviceToHost, s));
  RTS_TRY(hipStreamSynchronize(s));
  RTS_TRY(rank_samples(d, s));

  describe(bx, d, col);
  d = nullptr;

fail:
  (void)hipFree(iota);
  (void)hipFree(bpos_in);
  (void)hipFree(keys);
  (void)hipFree(skeys);
  (void)hipFree(tmp);
  (void)hipFree(starts_d);
  (void)hipFree(estarts_d);
  (void)hipFree(light_d);
  (void)hipFree(idx_d);
  (void)hipFree(bf_slot_d);
  (void)hipFree(bl_slot_d);
  (void)hipFree(bfirst_d);
  (void)hipFree(blast_d);
  if (d) {
    d->owned_column = nullptr;  // the caller frees what it passed in on failure
    free_dev(d);
  }
  return err;
}

void bindex_free(BinDex *bx) {
  if (!bx) return;
  free_dev(bx->dev);
  memset(bx, 0, sizeof(*bx));
}

}  // namespace rts

// ----------------------------------------------------------------- snapshot (persistence)
//
// File layout (host byte order, little-endian on MI355X hosts):
//   "RTSBDX" 0 2 | u32 version, u32 bins, u32 sub, u32 pad | i64 n, words, fv_stride, ntiles,
//   nb, nblocks | u32 data_max, u32 pad | u64 digests: host arrays, sorted_val, sorted_pos,
//   fvs, bent, the indexed column (0 when the index had none) |
//   starts[128] u32 | area_start[128] i64 | area_len[128] i64 | estarts[bins] u32 |
//   boff[bins + 1] u32 | block_first[nblocks] u32 | block_last[nblocks] u32 |
//   sorted_val[n] u32 | sorted_pos[n] i32 | fvs[(bins - 1) * fv_stride] u32 | bent[nb] (u32, u32)
// Device arrays stream through one page-locked staging buffer.  The tile directory is not
// stored: the load rebuilds it from the B lists.  A load checks the header, the host arrays'
// structure (areas tile the sorted column, block counts, starts non-decreasing, list offsets
// non-decreasing up to nb), every B-list entry as it passes through the staging buffer (row
// < n, rows increasing within a bin: the kernels index LDS by row), and the digest of every
// array; any mismatch refuses the file before the index is used.

namespace rts {

hipError_t bitmap_digest(const BITS *bm, int64_t n, unsigned long long *digest, hipStream_t s);

namespace {

constexpr char kSnapMagic[8] = {'R', 'T', 'S', 'B', 'D', 'X', 0, 2};
constexpr uint32_t kSnapVersion = 2;
constexpr size_t kStage = 64u << 20;

struct SnapHeader {
  char magic[8];
  uint32_t version, bins, sub, pad0;
  int64_t n, words, fv_stride, ntiles, nb, nblocks;
  uint32_t data_max, pad1;
  uint64_t dg_host, dg_val, dg_pos, dg_fvs, dg_bent;
  uint64_t dg_col;  // digest of the indexed column at save time (0: not known)
};

bool put(FILE *f, const void *p, size_t bytes) { return fwrite(p, 1, bytes, f) == bytes; }
bool get(FILE *f, void *p, size_t bytes) { return fread(p, 1, bytes, f) == bytes; }

// FNV-1a over host arrays, 8 bytes at a time (a few KB per index).
struct HostDigest {
  uint64_t h = 0xCBF29CE484222325ull;
  void add(const void *p, size_t bytes) {
    const unsigned char *c = static_cast<const unsigned char *>(p);
    for (size_t i = 0; i < bytes; i++) h = (h ^ c[i]) * 0x100000001B3ull;
  }
};

uint64_t host_digest(const rts_bindex_dev *d) {
  HostDigest g;
  g.add(d->starts, sizeof(d->starts));
  g.add(d->area_start, sizeof(d->area_start));
  g.add(d->area_len, sizeof(d->area_len));
  g.add(d->estarts.data(), d->estarts.size() * sizeof(CODE));
  g.add(d->boff_host.data(), d->boff_host.size() * sizeof(uint32_t));
  g.add(d->block_first.data(), d->block_first.size() * sizeof(CODE));
  g.add(d->block_last.data(), d->block_last.size() * sizeof(CODE));
  return g.h;
}

// Digest of a device array of `words` 32-bit words (the checker's bitmap digest over all
// bits of every word).
hipError_t device_digest(const void *p, int64_t words, uint64_t *out) {
  if (words == 0) {
    *out = 0;
    return hipSuccess;
  }
  unsigned long long dg = 0;
  hipError_t e = bitmap_digest(static_cast<const BITS *>(p), words * 32, &dg, ref_stream());
  *out = dg;
  return e;
}

hipError_t put_device(FILE *f, const void *dev, size_t bytes, void *stage) {
  for (size_t off = 0; off < bytes; off += kStage) {
    const size_t len = std::min(kStage, bytes - off);
    hipError_t e = hipMemcpy(stage, (const char *)dev + off, len, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return e;
    if (!put(f, stage, len)) return hipErrorUnknown;
  }
  return hipSuccess;
}

// check(stage, offset, len) sees every chunk on the host before it is uploaded.
template <class Check>
hipError_t get_device(FILE *f, void *dev, size_t bytes, void *stage, Check check) {
  for (size_t off = 0; off < bytes; off += kStage) {
    const size_t len = std::min(kStage, bytes - off);
    if (!get(f, stage, len)) return hipErrorInvalidValue;
    if (!check(stage, off, len)) return hipErrorInvalidValue;
    hipError_t e = hipMemcpy((char *)dev + off, stage, len, hipMemcpyHostToDevice);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

const auto kNoCheck = [](const void *, size_t, size_t) { return true; };

int blocks_of(int64_t len) {
  return (int)(len <= RTS_BLOCK_INIT_SIZE ? 1 : (len + RTS_BLOCK_INIT_SIZE - 1) / RTS_BLOCK_INIT_SIZE);
}

// Structure of the host arrays read from a snapshot.
bool host_arrays_valid(const rts_bindex_dev *d, int64_t nblocks) {
  int64_t slot = 0, blocks = 0;
  for (int k = 0; k < kAreas; k++) {
    if (d->area_start[k] != slot || d->area_len[k] < 1) return false;
    slot += d->area_len[k];
    blocks += blocks_of(d->area_len[k]);
    if (k && d->starts[k] < d->starts[k - 1]) return false;
  }
  if (slot != d->n || blocks != nblocks) return false;
  for (int b = 0; b < d->bins; b++) {
    if (b && d->estarts[b] < d->estarts[b - 1]) return false;
    if (b % d->sub == 0 && d->estarts[b] != d->starts[b / d->sub]) return false;
  }
  if (d->boff_host[0] != 0 || d->boff_host[d->bins] != (uint32_t)d->nb) return false;
  for (int b = 0; b < d->bins; b++)
    if (d->boff_host[b + 1] < d->boff_host[b]) return false;
  return true;
}

}  // namespace

hipError_t bindex_save(const BinDex *bx, const char *path) {
  if (!bx || !bx->dev || !path) return hipErrorInvalidValue;
  const rts_bindex_dev *d = bx->dev;
  SnapHeader h{};
  memcpy(h.magic, kSnapMagic, 8);
  h.version = kSnapVersion;
  h.bins = (uint32_t)d->bins;
  h.sub = (uint32_t)d->sub;
  h.n = d->n;
  h.words = d->words;
  h.fv_stride = d->fv_stride;
  h.ntiles = d->ntiles;
  h.nb = d->nb;
  h.nblocks = (int64_t)d->block_first.size();
  h.data_max = bx->data_max;
  h.dg_host = host_digest(d);
  const size_t fv_words = (size_t)(d->bins - 1) * d->fv_stride;
  hipError_t e = device_digest(d->sorted_val, d->n, &h.dg_val);
  if (e == hipSuccess) e = device_digest(d->sorted_pos, d->n, &h.dg_pos);
  if (e == hipSuccess) e = device_digest(d->fvs, (int64_t)fv_words, &h.dg_fvs);
  if (e == hipSuccess) e = device_digest(d->bent, 2 * d->nb, &h.dg_bent);
  if (e == hipSuccess && bx->rawDataInGPU) e = device_digest(bx->rawDataInGPU, d->n, &h.dg_col);
  if (e != hipSuccess) return e;
  FILE *f = fopen(path, "wb");
  if (!f) return hipErrorInvalidValue;
  void *stage = nullptr;
  e = hipHostMalloc(&stage, kStage, hipHostMallocDefault);
  bool ok = e == hipSuccess && put(f, &h, sizeof(h)) && put(f, d->starts, sizeof(d->starts)) &&
            put(f, d->area_start, sizeof(d->area_start)) && put(f, d->area_len, sizeof(d->area_len)) &&
            put(f, d->estarts.data(), d->bins * sizeof(CODE)) &&
            put(f, d->boff_host.data(), (d->bins + 1) * sizeof(uint32_t)) &&
            put(f, d->block_first.data(), h.nblocks * sizeof(CODE)) &&
            put(f, d->block_last.data(), h.nblocks * sizeof(CODE));
  if (ok) e = put_device(f, d->sorted_val, d->n * sizeof(CODE), stage);
  if (ok && e == hipSuccess) e = put_device(f, d->sorted_pos, d->n * sizeof(int32_t), stage);
  if (ok && e == hipSuccess) e = put_device(f, d->fvs, fv_words * sizeof(BITS), stage);
  if (ok && e == hipSuccess) e = put_device(f, d->bent, d->nb * sizeof(uint2), stage);
  if (stage) (void)hipHostFree(stage);
  if (fclose(f) != 0) ok = false;
  if (e != hipSuccess) return e;
  return ok ? hipSuccess : hipErrorUnknown;
}

hipError_t bindex_load(BinDex *bx, const char *path, const CODE *col) {
  if (!bx || !path) return hipErrorInvalidValue;
  FILE *f = fopen(path, "rb");
  if (!f) return hipErrorInvalidValue;
  SnapHeader h{};
  rts_bindex_dev *d = nullptr;
  void *stage = nullptr;
  hipError_t e = hipErrorInvalidValue;
  size_t fv_words = 0;
  uint64_t dg = 0;
  int64_t bin = 0;        // bin of the next B-list entry to check
  uint32_t prev_row = 0;  // row of the previous entry of that bin
  auto check_entries = [&](const void *p, size_t off, size_t len) {
    if (off % sizeof(uint2) || len % sizeof(uint2)) return false;
    const uint2 *ent = static_cast<const uint2 *>(p);
    for (size_t i = 0, e0 = off / sizeof(uint2); i < len / sizeof(uint2); i++) {
      const uint64_t e = e0 + i;
      while (bin < d->bins && e >= d->boff_host[bin + 1]) bin++;
      if (bin >= d->bins || ent[i].x >= (uint64_t)d->n) return false;
      if (e > d->boff_host[bin] && ent[i].x <= prev_row) return false;  // row order in a bin
      prev_row = ent[i].x;
    }
    return true;
  };
  if (!get(f, &h, sizeof(h)) || memcmp(h.magic, kSnapMagic, 8) != 0 || h.version != kSnapVersion ||
      (h.sub != 1 && h.sub != 2 && h.sub != 4 && h.sub != 8) || h.bins != (uint32_t)kAreas * h.sub ||
      h.n < kAreas || h.n > 0x7FFFFFFFll || h.words != (h.n + 31) / 32 || h.fv_stride < h.words ||
      h.fv_stride % kFvStrideAlign != 0 || h.fv_stride > h.words + kFvStrideAlign ||
      h.ntiles != (h.n + kTileRows - 1) / kTileRows || h.nb < 0 || h.nb > h.n || h.nblocks < kAreas ||
      h.nblocks > h.n)
    goto done;
  d = new rts_bindex_dev();
  d->n = h.n;
  d->words = h.words;
  d->fv_stride = h.fv_stride;
  d->ntiles = h.ntiles;
  d->nb = h.nb;
  d->sub = (int)h.sub;
  d->bins = (int)h.bins;
  d->estarts.resize(d->bins);
  d->boff_host.resize(d->bins + 1);
  d->block_first.resize(h.nblocks);
  d->block_last.resize(h.nblocks);
  if (!get(f, d->starts, sizeof(d->starts)) || !get(f, d->area_start, sizeof(d->area_start)) ||
      !get(f, d->area_len, sizeof(d->area_len)) || !get(f, d->estarts.data(), d->bins * sizeof(CODE)) ||
      !get(f, d->boff_host.data(), (d->bins + 1) * sizeof(uint32_t)) ||
      !get(f, d->block_first.data(), h.nblocks * sizeof(CODE)) ||
      !get(f, d->block_last.data(), h.nblocks * sizeof(CODE)))
    goto done;
  if (host_digest(d) != h.dg_host || !host_arrays_valid(d, h.nblocks)) goto done;
  fv_words = (size_t)(d->bins - 1) * h.fv_stride;
  e = hipHostMalloc(&stage, kStage, hipHostMallocDefault);
  if (e == hipSuccess) e = hipMalloc((void **)&d->sorted_val, h.n * sizeof(CODE));
  if (e == hipSuccess) e = hipMalloc((void **)&d->sorted_pos, h.n * sizeof(int32_t));
  if (e == hipSuccess) e = hipMalloc((void **)&d->fvs, fv_words * sizeof(BITS));
  if (e == hipSuccess) e = hipMalloc((void **)&d->bent, std::max<int64_t>(h.nb, 1) * sizeof(uint2));
  if (e == hipSuccess) e = hipMalloc((void **)&d->boff, (d->bins + 1) * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc((void **)&d->bdir, (size_t)d->bins * (h.ntiles + 1) * sizeof(uint32_t));
  if (e == hipSuccess) e = get_device(f, d->sorted_val, h.n * sizeof(CODE), stage, kNoCheck);
  if (e == hipSuccess) e = get_device(f, d->sorted_pos, h.n * sizeof(int32_t), stage, kNoCheck);
  if (e == hipSuccess) e = get_device(f, d->fvs, fv_words * sizeof(BITS), stage, kNoCheck);
  if (e == hipSuccess) e = get_device(f, d->bent, h.nb * sizeof(uint2), stage, check_entries);
  // the arrays as uploaded must carry the digests the saved index had
  if (e == hipSuccess) e = device_digest(d->sorted_val, h.n, &dg);
  if (e == hipSuccess && dg != h.dg_val) e = hipErrorInvalidValue;
  if (e == hipSuccess) e = device_digest(d->sorted_pos, h.n, &dg);
  if (e == hipSuccess && dg != h.dg_pos) e = hipErrorInvalidValue;
  if (e == hipSuccess) e = device_digest(d->fvs, (int64_t)fv_words, &dg);
  if (e == hipSuccess && dg != h.dg_fvs) e = hipErrorInvalidValue;
  if (e == hipSuccess) e = device_digest(d->bent, 2 * h.nb, &dg);
  if (e == hipSuccess && dg != h.dg_bent) e = hipErrorInvalidValue;
  // the column given with the snapshot must be the one that was indexed
  if (e == hipSuccess && col && h.dg_col) {
    e = device_digest(col, h.n, &dg);
    if (e == hipSuccess && dg != h.dg_col) e = hipErrorInvalidValue;
  }
  if (e == hipSuccess)
    e = hipMemcpy(d->boff, d->boff_host.data(), (d->bins + 1) * sizeof(uint32_t), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = launch_tile_directory(d, d->ntiles, kTileRows, d->bdir, ref_stream());
  if (e == hipSuccess) e = hipStreamSynchronize(ref_stream());
  if (e == hipSuccess) e = rank_samples(d, ref_stream());
  if (e == hipSuccess) {
    // which bins' lists hold every row of the bin (a snapshot of an index built without
    // the start rows in its lists has none)
    std::vector<int64_t> rk;
    e = start_ranks(d, rk, ref_stream());
    if (e == hipSuccess) mark_complete_bins(d, rk);
  }
  if (e == hipSuccess) {
    memset(bx, 0, sizeof(*bx));
    bx->data_max = h.data_max;
    describe(bx, d, col);
    d = nullptr;
  }
done:
  if (stage) (void)hipHostFree(stage);
  if (d) free_dev(d);
  fclose(f);
  return e;
}

}  // namespace rts
