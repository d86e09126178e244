// bindex_build.hip -- GPU build of the device-resident BinDex (init_bindex,
// bindex.cpp:599-671, and its GPU-driver twin init_bindex_in_GPU, bindex_cuda.cpp:164-282).
//
// The reference sorts on the host with an unstable parallel sort, builds the 127 filter
// vectors with 127 host passes, then issues one cudaMalloc + cudaMemcpy per 3276-row block
// (~30k per column at 1e8 rows).  Here everything stays on the device: one stable radix
// sort of (value, row), a K-thread kernel for the area starts, one pass over the column
// for all 127 filter vectors and the bin keys, one 8-bit stable radix sort for the
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
  if (i >= kBins) return;
  const int64_t avg = n / kBins;
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

// One pass over the column: all K-1 filter vectors (fv[k] bit i = x_i < start[k+1],
// set_fv_val_less, bindex.cpp:553-560) and each row's refine key (its bin if
// start[bin] < x, else kBins = no refine).  One row per lane; a 64-row chunk gives two
// MSB-first words per filter vector from one ballot.
__global__ __launch_bounds__(kBlock) void k_build_fvs(const CODE *col, int64_t n,
                                                      const CODE *starts_g, BITS *fvs,
                                                      int64_t fv_stride, uint8_t *keys) {
  __shared__ CODE starts[kBins];
  for (int i = threadIdx.x; i < kBins; i += blockDim.x) starts[i] = starts_g[i];
  __syncthreads();
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t wave = (int64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
  const int64_t nwaves = (int64_t)gridDim.x * (kBlock / kWave);
  const int64_t nchunks = (n + 63) / 64;
  for (int64_t ch = wave; ch < nchunks; ch += nwaves) {
    const int64_t r = ch * 64 + lane;
    const bool valid = r < n;
    const CODE x = valid ? col[r] : 0;
    // t = number of k in [1, K-1] with start[k] <= x (binary search in LDS)
    int lo = 1, hi = kBins;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (starts[mid] <= x)
        lo = mid + 1;
      else
        hi = mid;
    }
    const int t = lo - 1;
    if (valid) keys[r] = starts[t] < x ? (uint8_t)t : (uint8_t)kBins;
    const int64_t w0 = ch * 2;
    const bool w1_valid = (w0 + 1) * 32 < n;
    for (int k = 0; k < kBins - 1; k++) {
      const uint64_t b = __ballot(valid && k >= t);
      if (lane == 0) fvs[k * fv_stride + w0] = __brev((uint32_t)b);
      if (lane == 1 && w1_valid) fvs[k * fv_stride + w0 + 1] = __brev((uint32_t)(b >> 32));
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

// boff[k] = first sorted key >= k, for k in [0, kBins]
__global__ void k_bin_offsets(const uint8_t *skeys, int64_t n, uint32_t *boff) {
  const int k = threadIdx.x;
  if (k > kBins) return;
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

// bdir[k * (ntiles + 1) + t] = first entry of bin k with row >= t * kTileRows
__global__ void k_tile_directory(const uint2 *bent, const uint32_t *boff, int64_t ntiles,
                                 uint32_t *bdir) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t per = ntiles + 1;
  if (id >= per * kBins) return;
  const int k = (int)(id / per);
  const int64_t t = id % per;
  const uint64_t row = (uint64_t)t * kTileRows;
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
  for (int k = 0; k < kBins; k++) {
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
  d->area_store.resize(kBins);
  d->area_gpu_store.resize(kBins);
  make_block_views(d, d->sorted_val, d->sorted_pos, d->blk_dev, d->blkp_dev);
  int64_t b0 = 0;
  for (int k = 0; k < kBins; k++) {
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
  for (int k = 0; k < kBins; k++) {
    bx->areas[k] = &d->area_store[k];
    bx->areaStartValues[k] = d->starts[k];
    bx->area_counts[k] = (POSTYPE)(d->area_start[k] + d->area_len[k]);
    bx->areasInGPU[k] = &d->area_gpu_store[k];
  }
  for (int k = 0; k < kBins - 1; k++) {
    bx->filterVectors[k] = nullptr;
    bx->filterVectorsInGPU[k] = d->fvs + (size_t)k * d->fv_stride;
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
  for (int k = 0; k < kBins; k++) {
    d->area_store[k].blocks = d->blkp_host.data() + b0;
    b0 += d->area_store[k].blockNum;
  }
  return hipSuccess;
}

hipError_t bindex_build(BinDex *bx, const CODE *col, int64_t n, CODE *owned, hipStream_t s) {
  hipError_t err = hipSuccess;
  rts_bindex_dev *d = nullptr;
  uint32_t *iota = nullptr, *bpos_in = nullptr;
  uint8_t *keys = nullptr, *skeys = nullptr;
  void *tmp = nullptr;
  size_t tmp_bytes = 0, need = 0;
  CODE *starts_d = nullptr;
  int64_t *idx_d = nullptr, *bf_slot_d = nullptr, *bl_slot_d = nullptr;
  CODE *bfirst_d = nullptr, *blast_d = nullptr;
  std::vector<int64_t> bf_slot, bl_slot;
  int64_t nblocks = 0;
  int64_t idx_h[kBins];
  uint32_t nb = 0;

  if (!bx || !col || n < kBins || n > 0x7FFFFFFFll) return hipErrorInvalidValue;
  memset(bx, 0, sizeof(*bx));
  d = new rts_bindex_dev();
  d->n = n;
  d->words = (n + 31) / 32;
  d->fv_stride = (d->words + kFvStrideAlign - 1) / kFvStrideAlign * kFvStrideAlign;
  d->ntiles = (n + kTileRows - 1) / kTileRows;
  d->owned_column = owned;

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

  // 2. area starts
  RTS_TRY(hipMalloc((void **)&starts_d, kBins * sizeof(CODE)));
  RTS_TRY(hipMalloc((void **)&idx_d, kBins * sizeof(int64_t)));
  hipLaunchKernelGGL(k_area_starts, dim3(1), dim3(kBins), 0, s, d->sorted_val, n, starts_d, idx_d);
  RTS_TRY(hipGetLastError());
  RTS_TRY(hipMemcpyAsync(d->starts, starts_d, sizeof(d->starts), hipMemcpyDeviceToHost, s));
  RTS_TRY(hipMemcpyAsync(idx_h, idx_d, sizeof(idx_h), hipMemcpyDeviceToHost, s));
  RTS_TRY(hipMemcpyAsync(&bx->data_max, d->sorted_val + n - 1, sizeof(CODE), hipMemcpyDeviceToHost, s));

  // 3. filter vectors + refine keys
  RTS_TRY(hipMalloc((void **)&d->fvs, (size_t)(kBins - 1) * d->fv_stride * sizeof(BITS)));
  RTS_TRY(hipMalloc((void **)&keys, n));
  hipLaunchKernelGGL(k_build_fvs, dim3(blocks_for((n + 63) / 64, kBlock / kWave)), dim3(kBlock), 0,
                     s, col, n, starts_d, d->fvs, d->fv_stride, keys);
  RTS_TRY(hipGetLastError());

  // 4. B lists: stable 8-bit sort of (key, row) -> rows grouped by bin in row order
  RTS_TRY(hipMalloc((void **)&skeys, n));
  RTS_TRY(hipMalloc((void **)&bpos_in, n * sizeof(uint32_t)));
  RTS_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, need, keys, skeys, iota, bpos_in, (int)n, 0,
                                             8, s));
  if (need > tmp_bytes) {
    (void)hipFree(tmp);
    tmp = nullptr;
    tmp_bytes = need;
    RTS_TRY(hipMalloc(&tmp, tmp_bytes));
  }
  RTS_TRY(hipcub::DeviceRadixSort::SortPairs(tmp, need, keys, skeys, iota, bpos_in, (int)n, 0, 8, s));
  RTS_TRY(hipMalloc((void **)&d->boff, (kBins + 1) * sizeof(uint32_t)));
  hipLaunchKernelGGL(k_bin_offsets, dim3(1), dim3(kBins + 1), 0, s, skeys, n, d->boff);
  RTS_TRY(hipGetLastError());
  d->boff_host.resize(kBins + 1);
  RTS_TRY(hipMemcpyAsync(d->boff_host.data(), d->boff, (kBins + 1) * sizeof(uint32_t),
                         hipMemcpyDeviceToHost, s));
  RTS_TRY(hipStreamSynchronize(s));
  nb = d->boff_host[kBins];
  d->nb = nb;
  RTS_TRY(hipMalloc((void **)&d->bent, std::max<uint32_t>(nb, 1) * sizeof(uint2)));
  if (nb) {
    hipLaunchKernelGGL(k_gather_entries, dim3(blocks_for(nb)), dim3(kBlock), 0, s, col, bpos_in,
                       d->bent, (int64_t)nb);
    RTS_TRY(hipGetLastError());
  }
  RTS_TRY(hipMalloc((void **)&d->bdir, (size_t)kBins * (d->ntiles + 1) * sizeof(uint32_t)));
  {
    const int64_t items = (int64_t)kBins * (d->ntiles + 1);
    hipLaunchKernelGGL(k_tile_directory, dim3((unsigned)((items + kBlock - 1) / kBlock)), dim3(kBlock),
                       0, s, d->bent, d->boff, d->ntiles, d->bdir);
    RTS_TRY(hipGetLastError());
  }

  // 5. host descriptor: areas (bindex.cpp:633-652) and their blocks
  for (int k = 0; k < kBins; k++) {
    d->area_start[k] = idx_h[k];
    d->area_len[k] = (k < kBins - 1 ? idx_h[k + 1] : n) - idx_h[k];
  }
  for (int k = 0; k < kBins; k++) {
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

  describe(bx, d, col);
  d = nullptr;

fail:
  (void)hipFree(iota);
  (void)hipFree(bpos_in);
  (void)hipFree(keys);
  (void)hipFree(skeys);
  (void)hipFree(tmp);
  (void)hipFree(starts_d);
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
//   "RTSBDX" 0 1 | u32 version, u32 bins | i64 n, words, fv_stride, ntiles, nb, nblocks |
//   u32 data_max | starts[bins] u32 | area_start[bins] i64 | area_len[bins] i64 |
//   boff[bins + 1] u32 | block_first[nblocks] u32 | block_last[nblocks] u32 |
//   sorted_val[n] u32 | sorted_pos[n] i32 | fvs[(bins - 1) * fv_stride] u32 | bent[nb] (u32, u32) |
//   bdir[bins * (ntiles + 1)] u32
// Device arrays stream through one page-locked staging buffer.

namespace rts {
namespace {

constexpr char kSnapMagic[8] = {'R', 'T', 'S', 'B', 'D', 'X', 0, 1};
constexpr uint32_t kSnapVersion = 1;
constexpr size_t kStage = 64u << 20;

struct SnapHeader {
  char magic[8];
  uint32_t version, bins;
  int64_t n, words, fv_stride, ntiles, nb, nblocks;
  uint32_t data_max, pad;
};

bool put(FILE *f, const void *p, size_t bytes) { return fwrite(p, 1, bytes, f) == bytes; }
bool get(FILE *f, void *p, size_t bytes) { return fread(p, 1, bytes, f) == bytes; }

hipError_t put_device(FILE *f, const void *dev, size_t bytes, void *stage) {
  for (size_t off = 0; off < bytes; off += kStage) {
    const size_t len = std::min(kStage, bytes - off);
    hipError_t e = hipMemcpy(stage, (const char *)dev + off, len, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return e;
    if (!put(f, stage, len)) return hipErrorUnknown;
  }
  return hipSuccess;
}

hipError_t get_device(FILE *f, void *dev, size_t bytes, void *stage) {
  for (size_t off = 0; off < bytes; off += kStage) {
    const size_t len = std::min(kStage, bytes - off);
    if (!get(f, stage, len)) return hipErrorInvalidValue;
    hipError_t e = hipMemcpy((char *)dev + off, stage, len, hipMemcpyHostToDevice);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace

hipError_t bindex_save(const BinDex *bx, const char *path) {
  if (!bx || !bx->dev || !path) return hipErrorInvalidValue;
  const rts_bindex_dev *d = bx->dev;
  FILE *f = fopen(path, "wb");
  if (!f) return hipErrorInvalidValue;
  void *stage = nullptr;
  hipError_t e = hipHostMalloc(&stage, kStage, hipHostMallocDefault);
  SnapHeader h{};
  memcpy(h.magic, kSnapMagic, 8);
  h.version = kSnapVersion;
  h.bins = kBins;
  h.n = d->n;
  h.words = d->words;
  h.fv_stride = d->fv_stride;
  h.ntiles = d->ntiles;
  h.nb = d->nb;
  h.nblocks = (int64_t)d->block_first.size();
  h.data_max = bx->data_max;
  const size_t dir = (size_t)kBins * (d->ntiles + 1);
  bool ok = e == hipSuccess && put(f, &h, sizeof(h)) && put(f, d->starts, sizeof(d->starts)) &&
            put(f, d->area_start, sizeof(d->area_start)) && put(f, d->area_len, sizeof(d->area_len)) &&
            put(f, d->boff_host.data(), (kBins + 1) * sizeof(uint32_t)) &&
            put(f, d->block_first.data(), h.nblocks * sizeof(CODE)) &&
            put(f, d->block_last.data(), h.nblocks * sizeof(CODE));
  if (ok) e = put_device(f, d->sorted_val, d->n * sizeof(CODE), stage);
  if (ok && e == hipSuccess) e = put_device(f, d->sorted_pos, d->n * sizeof(int32_t), stage);
  if (ok && e == hipSuccess)
    e = put_device(f, d->fvs, (size_t)(kBins - 1) * d->fv_stride * sizeof(BITS), stage);
  if (ok && e == hipSuccess) e = put_device(f, d->bent, d->nb * sizeof(uint2), stage);
  if (ok && e == hipSuccess) e = put_device(f, d->bdir, dir * sizeof(uint32_t), stage);
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
  size_t dir = 0;
  if (!get(f, &h, sizeof(h)) || memcmp(h.magic, kSnapMagic, 8) != 0 || h.version != kSnapVersion ||
      h.bins != (uint32_t)kBins || h.n < kBins || h.n > 0x7FFFFFFFll || h.words != (h.n + 31) / 32 ||
      h.fv_stride < h.words || h.ntiles != (h.n + kTileRows - 1) / kTileRows || h.nb < 0 || h.nb > h.n ||
      h.nblocks < kBins || h.nblocks > h.n)
    goto done;
  d = new rts_bindex_dev();
  d->n = h.n;
  d->words = h.words;
  d->fv_stride = h.fv_stride;
  d->ntiles = h.ntiles;
  d->nb = h.nb;
  d->boff_host.resize(kBins + 1);
  d->block_first.resize(h.nblocks);
  d->block_last.resize(h.nblocks);
  if (!get(f, d->starts, sizeof(d->starts)) || !get(f, d->area_start, sizeof(d->area_start)) ||
      !get(f, d->area_len, sizeof(d->area_len)) || !get(f, d->boff_host.data(), (kBins + 1) * sizeof(uint32_t)) ||
      !get(f, d->block_first.data(), h.nblocks * sizeof(CODE)) ||
      !get(f, d->block_last.data(), h.nblocks * sizeof(CODE)))
    goto done;
  if (d->boff_host[kBins] != (uint32_t)h.nb) goto done;
  dir = (size_t)kBins * (h.ntiles + 1);
  e = hipHostMalloc(&stage, kStage, hipHostMallocDefault);
  if (e == hipSuccess) e = hipMalloc((void **)&d->sorted_val, h.n * sizeof(CODE));
  if (e == hipSuccess) e = hipMalloc((void **)&d->sorted_pos, h.n * sizeof(int32_t));
  if (e == hipSuccess) e = hipMalloc((void **)&d->fvs, (size_t)(kBins - 1) * h.fv_stride * sizeof(BITS));
  if (e == hipSuccess) e = hipMalloc((void **)&d->bent, std::max<int64_t>(h.nb, 1) * sizeof(uint2));
  if (e == hipSuccess) e = hipMalloc((void **)&d->boff, (kBins + 1) * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc((void **)&d->bdir, dir * sizeof(uint32_t));
  if (e == hipSuccess) e = get_device(f, d->sorted_val, h.n * sizeof(CODE), stage);
  if (e == hipSuccess) e = get_device(f, d->sorted_pos, h.n * sizeof(int32_t), stage);
  if (e == hipSuccess) e = get_device(f, d->fvs, (size_t)(kBins - 1) * h.fv_stride * sizeof(BITS), stage);
  if (e == hipSuccess) e = get_device(f, d->bent, h.nb * sizeof(uint2), stage);
  if (e == hipSuccess) e = get_device(f, d->bdir, dir * sizeof(uint32_t), stage);
  if (e == hipSuccess)
    e = hipMemcpy(d->boff, d->boff_host.data(), (kBins + 1) * sizeof(uint32_t), hipMemcpyHostToDevice);
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
