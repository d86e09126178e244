// ref_shim.cpp -- C entry points into the reference's own CPU BinDex / raw_scan code.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile's `ref` target compiles this file together
// with the reference source bindex.cpp, read where it lies (-I <reference root>), into
// oracle/_ref/libref_bindex_<N>.so.  Nothing here is copied from the reference; the
// reference's main() is renamed away and its functions are called unchanged.  The
// reference fixes its row count N and area count K at compile time (bindex.cpp:36-42),
// so one library is built per N.  Asserting paths are compiled out (-DNDEBUG); tests do
// not compare those paths.
#define main rtscan_reference_main
#include "bindex.cpp"
#undef main

#include <chrono>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <thread>

namespace {
BITS *aligned_words(int words) {
  size_t bytes = ((size_t)words * sizeof(BITS) + 63) / 64 * 64;
  return (BITS *)aligned_alloc(64, bytes);
}
}  // namespace

extern "C" {

int ref_compiled_n() { return N; }
int ref_compiled_k() { return K; }
int ref_block_init_size() { return blockInitSize; }

// init_bindex, bindex.cpp:599-671
void *ref_init_bindex(const uint32_t *data, int n) {
  BinDex *b = (BinDex *)malloc(sizeof(BinDex));
  init_bindex(b, const_cast<CODE *>(data), n);
  fflush(stdout);
  return b;
}

void ref_free_bindex(void *bx) {
  BinDex *b = (BinDex *)bx;
  for (int i = 0; i < K - 1; i++) free(b->filterVectors[i]);
  for (int i = 0; i < K; i++) free_area(b->areas[i]);
  free(b);
}

// Area start values, prefix counts, block counts and the K-1 filter vectors.
void ref_read_bindex(void *bx, uint32_t *start_values, int32_t *area_counts, int32_t *block_nums,
                     uint32_t *fvs) {
  BinDex *b = (BinDex *)bx;
  int words = bits_num_needed(b->length);
  for (int k = 0; k < K; k++) {
    start_values[k] = area_start_value(b->areas[k]);
    area_counts[k] = b->area_counts[k];
    block_nums[k] = b->areas[k]->blockNum;
  }
  for (int k = 0; k < K - 1; k++) memcpy(fvs + (size_t)k * words, b->filterVectors[k], words * 4);
}

int ref_in_which_area(void *bx, uint32_t c) { return in_which_area((BinDex *)bx, c); }

int ref_in_which_block(void *bx, int area, uint32_t c) {
  return in_which_block(((BinDex *)bx)->areas[area], c);
}

int ref_on_which_pos(void *bx, int area, int block, uint32_t c) {
  return on_which_pos(((BinDex *)bx)->areas[area]->blocks[block], c);
}

// bindex_scan_{lt,le,gt,ge,bt,eq}, bindex.cpp:1365-1691.  op follows OPERATOR (bindex.h:88-95).
void ref_bindex_scan(void *bx, int op, uint32_t t1, uint32_t t2, uint32_t *out) {
  BinDex *b = (BinDex *)bx;
  int words = bits_num_needed(b->length);
  BITS *r = aligned_words(words);
  switch (op) {
    case LT: bindex_scan_lt(b, r, t1); break;
    case LE: bindex_scan_le(b, r, t1); break;
    case GT: bindex_scan_gt(b, r, t1); break;
    case GE: bindex_scan_ge(b, r, t1); break;
    case BT: bindex_scan_bt(b, r, t1, t2); break;
    case EQ: bindex_scan_eq(b, r, t1); break;
    default: memset(r, 0, words * 4);
  }
  memcpy(out, r, (size_t)words * 4);
  free(r);
  fflush(stdout);  // the reference prints timings; flush while the caller captures them
}

// raw_scan, bindex.cpp:1851-1892, into a caller-zeroed bitmap.
void ref_raw_scan(const uint32_t *raw, int n, uint32_t *bitmap, uint32_t t1, uint32_t t2, int op) {
  BinDex b;
  b.length = n;
  raw_scan(&b, bitmap, t1, t2, (OPERATOR)op, const_cast<CODE *>(raw));
}

}  // extern "C"

extern "C" {

// The reference driver's timed region for one query (timer slot 11, bindex.cpp:2190-2262):
// bindex_scan_{lt,le,gt,ge,bt,eq} of every column into that column's bitmap, then the AND
// merge of columns 1.. into column 0's bitmap by THREAD_NUM refine_result_bitmap threads
// over word ranges, as the driver splits them.  Returns the region's wall time in ms.
// bitmaps[c] must hold bits_num_needed(N) words, 32-byte aligned.
double ref_query_slot11(void **bxs, int ncols, const int *ops, const uint32_t *t1, const uint32_t *t2,
                        uint32_t **bitmaps) {
  const auto t0 = std::chrono::steady_clock::now();
  for (int c = 0; c < ncols; c++) {
    BinDex *b = (BinDex *)bxs[c];
    switch (ops[c]) {
      case LT: bindex_scan_lt(b, bitmaps[c], t1[c]); break;
      case LE: bindex_scan_le(b, bitmaps[c], t1[c]); break;
      case GT: bindex_scan_gt(b, bitmaps[c], t1[c]); break;
      case GE: bindex_scan_ge(b, bitmaps[c], t1[c]); break;
      case BT: bindex_scan_bt(b, bitmaps[c], t1[c], t2[c]); break;
      default: bindex_scan_eq(b, bitmaps[c], t1[c]); break;
    }
  }
  const int max_idx = (N + CODEWIDTH - 1) / CODEWIDTH;
  const int stride = (max_idx + THREAD_NUM - 1) / THREAD_NUM;
  for (int c = 1; c < ncols; c++) {
    std::thread threads[THREAD_NUM];
    int start = 0, t = 0;
    while (start < max_idx && t < THREAD_NUM) {
      const int end = start + stride > max_idx ? max_idx : start + stride;
      threads[t] = std::thread(refine_result_bitmap, bitmaps[0], bitmaps[c], start, end, t);
      start += stride;
      t++;
    }
    for (int i = 0; i < t; i++) threads[i].join();
  }
  const auto t1_ = std::chrono::steady_clock::now();
  fflush(stdout);
  return std::chrono::duration<double, std::milli>(t1_ - t0).count();
}

int ref_thread_num() { return THREAD_NUM; }

}  // extern "C"
