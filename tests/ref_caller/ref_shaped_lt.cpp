// A caller written the way the reference's BinDex drivers are (bindex.cpp:1365-1411,
// bindex_cuda.cpp:723-770): it walks bindex->areas[k]->blocks[b]->pos / val / length,
// calls in_which_area / in_which_block / on_which_pos, and refines a filter vector on the
// host.  It is compiled against include/rts_bindex.h and linked to librtscan_hip.so
// unchanged, which is what the drop-in boundary promises.
//
//   ref_shaped_lt host   hand-built descriptor (host blocks only), no GPU call
//   ref_shaped_lt gpu N  index built by init_bindex_in_GPU over N seeded rows
// Prints "ok <queries>" and exits 0 when every `x < c` bitmap equals brute force.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "rts_bindex.h"

// bindex_cuda.cpp:50, :106
static CODE block_start_value(pos_block *pb) { return pb->val[0]; }
static CODE area_start_value(Area *area) { return area->blocks[0]->val[0]; }

// refine (bindex.cpp:1217): toggle the bit of each listed row
static void refine_positions(BITS *bitmap, POSTYPE *pos, POSTYPE n) {
  for (POSTYPE i = 0; i < n; i++) bitmap[pos[i] >> 5] ^= 1u << (31 - pos[i] % 32);
}

// bindex_scan_lt (bindex.cpp:1365-1411) over a host copy `fv_host(k)` of filter vector k.
template <class FV>
static void scan_lt(BinDex *bindex, std::vector<BITS> &result, CODE compare, FV fv_host) {
  const int words = (bindex->length - 1) / 32 + 1;
  const int area_idx = in_which_area(bindex, compare);
  if (area_idx < 0) {
    std::fill(result.begin(), result.end(), 0u);
    return;
  }
  Area *area = bindex->areas[area_idx];
  const int block_idx = in_which_block(area, compare);
  const int pos_idx = on_which_pos(area->blocks[block_idx], compare);
  const bool is_upper_fv = block_idx < area->blockNum / 2;
  const int start_blk = is_upper_fv ? 0 : block_idx + 1;
  const int end_blk = is_upper_fv ? block_idx : area->blockNum;
  fv_host(is_upper_fv ? area_idx - 1 : area_idx, result.data(), words);
  for (int b = start_blk; b < end_blk; b++) refine_positions(result.data(), area->blocks[b]->pos, area->blocks[b]->length);
  pos_block *blk = area->blocks[block_idx];
  if (is_upper_fv)
    refine_positions(result.data(), blk->pos, pos_idx);
  else
    refine_positions(result.data(), blk->pos + pos_idx, blk->length - pos_idx);
}

static bool check(const std::vector<CODE> &x, const std::vector<BITS> &got, CODE c) {
  for (size_t i = 0; i < x.size(); i++) {
    const int bit = (got[i / 32] >> (31 - i % 32)) & 1;
    if (bit != (x[i] < c ? 1 : 0)) {
      fprintf(stderr, "row %zu value %u compare %u: got %d\n", i, x[i], c, bit);
      return false;
    }
  }
  return true;
}

static std::vector<CODE> make_data(size_t n, unsigned seed, CODE range) {
  std::vector<CODE> x(n);
  uint64_t s = seed * 0x9E3779B97F4A7C15ull + 1;
  for (size_t i = 0; i < n; i++) {
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    const CODE v = (CODE)(s >> 32);
    x[i] = range ? v % range : v;
  }
  return x;
}

static std::vector<CODE> compares_for(const std::vector<CODE> &x) {
  std::vector<CODE> cs = {0, 1, x[0], x[x.size() / 2], x.back(), x.back() + 1, 0xFFFFFFFFu};
  std::vector<CODE> s(x);
  std::sort(s.begin(), s.end());
  for (int i = 1; i < 40; i++) cs.push_back(s[s.size() * i / 40]);
  for (int i = 1; i < 10; i++) cs.push_back(s[s.size() * i / 10] + 1);
  return cs;
}

// A descriptor built on the host the way init_bindex lays it out (bindex.cpp:599-671):
// equal-count areas of value-sorted rows, 3276-row blocks, filter vectors x < start[k+1].
static int run_host() {
  const size_t n = 200000;
  const std::vector<CODE> x = make_data(n, 7, 50000);
  std::vector<POSTYPE> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](POSTYPE a, POSTYPE b) { return x[a] < x[b]; });
  std::vector<CODE> sval(n);
  for (size_t i = 0; i < n; i++) sval[i] = x[order[i]];
  BinDex bx;
  memset(&bx, 0, sizeof(bx));
  bx.length = (POSTYPE)n;
  const size_t avg = n / RTS_VAREA_N;
  std::vector<size_t> start(RTS_VAREA_N);
  for (int k = 0; k < RTS_VAREA_N; k++) {
    size_t i = k * avg;
    if (k > 0 && sval[i] != sval[(k - 1) * avg])
      while (i > (k - 1) * avg + 1 && sval[i - 1] == sval[k * avg]) i--;
    start[k] = i;
  }
  std::vector<Area> areas(RTS_VAREA_N);
  std::vector<std::vector<pos_block>> blocks(RTS_VAREA_N);
  std::vector<std::vector<pos_block *>> bptr(RTS_VAREA_N);
  std::vector<std::vector<CODE>> bfirst(RTS_VAREA_N), blast(RTS_VAREA_N);
  for (int k = 0; k < RTS_VAREA_N; k++) {
    const size_t s0 = start[k], s1 = k + 1 < RTS_VAREA_N ? start[k + 1] : n;
    for (size_t b = s0; b < s1; b += RTS_BLOCK_INIT_SIZE) {
      const size_t e = std::min(s1, b + RTS_BLOCK_INIT_SIZE);
      blocks[k].push_back(pos_block{&order[b], &sval[b], (int)(e - b)});
      bfirst[k].push_back(sval[b]);
      blast[k].push_back(sval[e - 1]);
    }
    for (pos_block &pb : blocks[k]) bptr[k].push_back(&pb);
    areas[k].blocks = bptr[k].data();
    areas[k].blockNum = (int)blocks[k].size();
    areas[k].length = (int)(s1 - s0);
    areas[k].start = (POSTYPE)s0;
    areas[k].blockFirst = bfirst[k].data();
    areas[k].blockLast = blast[k].data();
    bx.areas[k] = &areas[k];
    bx.areaStartValues[k] = area_start_value(&areas[k]);
    bx.area_counts[k] = (POSTYPE)s1;
  }
  bx.data_min = sval[0];
  bx.data_max = sval[n - 1];
  const size_t words = (n - 1) / 32 + 1;
  auto fv_host = [&](int k, BITS *dst, int w) {
    if (k < 0) {
      memset(dst, 0, w * sizeof(BITS));
      return;
    }
    if (k >= RTS_VAREA_N - 1) {
      memset(dst, 0xFF, w * sizeof(BITS));
      return;
    }
    memset(dst, 0, w * sizeof(BITS));
    for (size_t i = 0; i < n; i++)
      if (x[i] < bx.areaStartValues[k + 1]) dst[i / 32] |= 1u << (31 - i % 32);
  };
  int nq = 0;
  for (CODE c : compares_for(x)) {
    std::vector<BITS> r(words);
    scan_lt(&bx, r, c, fv_host);
    if (!check(x, r, c)) return 1;
    nq++;
  }
  // the blocks' first values are what in_which_block compares (bindex_cuda.cpp:733-748)
  for (int k = 0; k < RTS_VAREA_N; k++)
    if (block_start_value(bx.areas[k]->blocks[0]) != bx.areaStartValues[k]) return 1;
  printf("ok %d\n", nq);
  return 0;
}

static int run_gpu(size_t n) {
  int failures = 0, nq = 0;
  for (CODE range : {0u, 5000u}) {
    std::vector<CODE> x = make_data(n, 11 + range, range);
    BinDex bx;
    size_t base_data = 0;
    init_bindex_in_GPU(&bx, x.data(), (POSTYPE)n, base_data);
    if (base_data != n * sizeof(CODE) || !bx.areas[0]->blocks) {
      fprintf(stderr, "init_bindex_in_GPU failed: %s\n", rts_last_error());
      return 1;
    }
    const size_t words = (n - 1) / 32 + 1;
    BITS *dev = nullptr;
    if (hipMalloc((void **)&dev, words * sizeof(BITS)) != hipSuccess) return 1;
    auto fv_host = [&](int k, BITS *dst, int w) {
      copy_filter_vector_in_GPU(&bx, dev, k);
      (void)hipMemcpy(dst, dev, w * sizeof(BITS), hipMemcpyDeviceToHost);
    };
    // the device views of areasInGPU cover the same slots as the host views
    for (int k = 0; k < RTS_VAREA_N; k += 17) {
      Area *a = bx.areas[k], *g = bx.areasInGPU[k];
      if (a->blockNum != g->blockNum) failures++;
      for (int b = 0; b < a->blockNum; b += 7) {
        std::vector<POSTYPE> p(g->blocks[b]->length);
        (void)hipMemcpy(p.data(), g->blocks[b]->pos, p.size() * sizeof(POSTYPE), hipMemcpyDeviceToHost);
        if ((int)p.size() != a->blocks[b]->length || memcmp(p.data(), a->blocks[b]->pos, p.size() * sizeof(POSTYPE)))
          failures++;
      }
    }
    for (CODE c : compares_for(x)) {
      std::vector<BITS> r(words);
      scan_lt(&bx, r, c, fv_host);
      if (!check(x, r, c)) failures++;
      nq++;
    }
    (void)hipFree(dev);
    rts_bindex_free(&bx);
  }
  if (failures) return 1;
  printf("ok %d\n", nq);
  return 0;
}

int main(int argc, char **argv) {
  if (argc > 1 && !strcmp(argv[1], "gpu")) return run_gpu(argc > 2 ? strtoull(argv[2], nullptr, 10) : 1000000);
  return run_host();
}
