/*
 * oracle.c -- CPU restatement of the reference's scan / BinDex path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for tests/, __graft_entry__.smoke() and the
 * timed CPU port in bench.py's cpu_baseline leg.  The shipped engine never links it.
 * See oracle.h for what pins each function.
 *
 * Bit convention (bindex.cpp:269, bindex.cpp:1217): row i is bit (1u << (31 - i%32))
 * of 32-bit word i/32 ("MSB-first").
 */
#include "oracle.h"

#include <immintrin.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define BITSWIDTH 32

static inline uint32_t row_bit(int i) { return 1u << (BITSWIDTH - 1 - (i % BITSWIDTH)); }

/* bindex.h:140-143 */
int orc_bits_num_needed(int n) { return ((n - 1) / BITSWIDTH) + 1; }

/* The per-row predicate of raw_scan, bindex.cpp:1855-1876 (uint32 compares; BT open). */
static inline int raw_hit(uint32_t x, uint32_t t1, uint32_t t2, int op) {
  switch (op) {
    case ORC_LT: return x < t1;
    case ORC_LE: return x <= t1;
    case ORC_GT: return x > t1;
    case ORC_GE: return x >= t1;
    case ORC_EQ: return x == t1;
    case ORC_BT: return x > t1 && x < t2;
    default: return 0;
  }
}

/* raw_scan, bindex.cpp:1851-1892 (dup bindex_cuda.cpp:1313-1353): every hit row XOR-toggles
 * its bit (refine(), bindex.cpp:1217) in a caller-zeroed bitmap. */
void orc_raw_scan(const uint32_t *raw, int n, uint32_t *bitmap, uint32_t t1, uint32_t t2, int op) {
  for (int i = 0; i < n; i++)
    if (raw_hit(raw[i], t1, t2, op)) bitmap[i / BITSWIDTH] ^= row_bit(i);
}

/* Same result as orc_raw_scan; rows are split on word boundaries so each thread owns its
 * words (used to time the CPU port on many cores). */
void orc_raw_scan_mt(const uint32_t *raw, int n, uint32_t *bitmap, uint32_t t1, uint32_t t2, int op,
                     int threads) {
  if (n <= 0) return;
  int words = orc_bits_num_needed(n);
  if (threads < 1) threads = 1;
#pragma omp parallel for num_threads(threads) schedule(static)
  for (int w = 0; w < words; w++) {
    int lo = w * BITSWIDTH, hi = lo + BITSWIDTH;
    if (hi > n) hi = n;
    uint32_t acc = 0;
    for (int i = lo; i < hi; i++)
      if (raw_hit(raw[i], t1, t2, op)) acc ^= row_bit(i);
    bitmap[w] ^= acc;
  }
}

/* ---- vectorised port of the checker's conjunction (timed as the CPU baseline) ----------
 * The same rows as raw_scan_entry per column + the AND merge (bindex.cpp:1851-1892,
 * 2265-2289) for closed ranges lo[c] <= x <= hi[c], written straight into `bitmap`:
 * 8 rows per AVX2 compare ((x - lo) ^ 2^31 > (hi - lo) ^ 2^31 signed = outside), the
 * movemask bits reversed into MSB-first order.  Scalar on hosts without AVX2. */
static uint8_t rev8_table[256];

static void init_rev8(void) {
  for (int b = 0; b < 256; b++) {
    int r = 0;
    for (int i = 0; i < 8; i++)
      if (b & (1 << i)) r |= 1 << (7 - i);
    rev8_table[b] = (uint8_t)r;
  }
}

static uint32_t conj_word_scalar(const uint32_t *const *cols, int ncols, const uint32_t *lo,
                                 const uint32_t *span, int64_t r0, int rows) {
  uint32_t acc = rows == 32 ? 0xFFFFFFFFu : ~(0xFFFFFFFFu >> rows);
  for (int c = 0; c < ncols; c++) {
    uint32_t m = 0;
    for (int i = 0; i < rows; i++)
      if (cols[c][r0 + i] - lo[c] <= span[c]) m |= 0x80000000u >> i;
    acc &= m;
  }
  return acc;
}

__attribute__((target("avx2"))) static void conj_words_avx2(const uint32_t *const *cols, int ncols,
                                                            const uint32_t *lo, const uint32_t *span,
                                                            int64_t w0, int64_t w1, uint32_t *bitmap) {
  const __m256i bias = _mm256_set1_epi32((int)0x80000000u);
  __m256i vlo[4], vsp[4];
  for (int c = 0; c < ncols; c++) {
    vlo[c] = _mm256_set1_epi32((int)lo[c]);
    vsp[c] = _mm256_xor_si256(_mm256_set1_epi32((int)span[c]), bias);
  }
  for (int64_t w = w0; w < w1; w++) {
    unsigned m[4] = {0xFF, 0xFF, 0xFF, 0xFF};
    for (int c = 0; c < ncols; c++) {
      const uint32_t *x = cols[c] + w * 32;
      for (int g = 0; g < 4; g++) {
        __m256i v = _mm256_loadu_si256((const __m256i *)(x + 8 * g));
        v = _mm256_xor_si256(_mm256_sub_epi32(v, vlo[c]), bias);
        const int miss = _mm256_movemask_ps(_mm256_castsi256_ps(_mm256_cmpgt_epi32(v, vsp[c])));
        m[g] &= ~(unsigned)miss & 0xFFu;
      }
    }
    bitmap[w] = ((uint32_t)rev8_table[m[0]] << 24) | ((uint32_t)rev8_table[m[1]] << 16) |
                ((uint32_t)rev8_table[m[2]] << 8) | (uint32_t)rev8_table[m[3]];
  }
}

int orc_simd_level(void) {
  __builtin_cpu_init();
  return __builtin_cpu_supports("avx2") ? 2 : 0;
}

void orc_conjunction_scan(const uint32_t *const *cols, int ncols, int64_t n, const uint32_t *lo,
                          const uint32_t *hi, uint32_t *bitmap, int threads) {
  if (n <= 0 || ncols < 1 || ncols > 4) return;
  const int64_t words = (n + 31) / 32, full = n / 32;
  uint32_t span[4];
  for (int c = 0; c < ncols; c++) {
    if (lo[c] > hi[c]) {
      memset(bitmap, 0, words * sizeof(uint32_t));
      return;
    }
    span[c] = hi[c] - lo[c];
  }
  if (!rev8_table[1]) init_rev8();
  const int avx2 = orc_simd_level() == 2;
  if (threads < 1) threads = 1;
  const int64_t chunk = 4096;
  const int64_t nchunks = (full + chunk - 1) / chunk;
#pragma omp parallel for num_threads(threads) schedule(static)
  for (int64_t k = 0; k < nchunks; k++) {
    const int64_t w0 = k * chunk, w1 = w0 + chunk < full ? w0 + chunk : full;
    if (avx2) {
      conj_words_avx2(cols, ncols, lo, span, w0, w1, bitmap);
    } else {
      for (int64_t w = w0; w < w1; w++) bitmap[w] = conj_word_scalar(cols, ncols, lo, span, w * 32, 32);
    }
  }
  if (full < words) bitmap[full] = conj_word_scalar(cols, ncols, lo, span, full * 32, (int)(n - full * 32));
}

/* raw_scan_entry, bindex.cpp:1894-1957: for every target of a query line, swap reversed
 * bt bounds (bindex.cpp:1903-1905) and XOR the hits into `bitmap`; then AND `bitmap` into
 * `merge_bitmap` unless they are the same buffer (bindex.cpp:1936-1956). */
void orc_raw_scan_entry(const uint32_t *raw, int n, int op, const uint32_t *tl, const uint32_t *tr,
                        int ntargets, uint32_t *bitmap, uint32_t *merge_bitmap, int threads) {
  for (int pi = 0; pi < ntargets; pi++) {
    uint32_t t1 = tl[pi], t2 = tr ? tr[pi] : 0;
    if (op == ORC_BT && t1 > t2) {
      uint32_t t = t1;
      t1 = t2;
      t2 = t;
    }
    orc_raw_scan_mt(raw, n, bitmap, t1, t2, op, threads);
  }
  if (merge_bitmap != bitmap && n > 0) {
    int words = orc_bits_num_needed(n);
    for (int w = 0; w < words; w++) merge_bitmap[w] &= bitmap[w];
  }
}

/* compare_bitmap, bindex.cpp:1959-1978: walk the first n rows; count truth hits and
 * matching hits; stop at the first mismatching row. */
void orc_compare_bitmap(const uint32_t *truth, const uint32_t *got, int n, int *total_hit,
                        int *true_hit, int *first_mismatch) {
  int total = 0, good = 0, bad = -1;
  for (int i = 0; i < n; i++) {
    uint32_t a = truth[i / BITSWIDTH] & row_bit(i);
    uint32_t b = got[i / BITSWIDTH] & row_bit(i);
    if (a) {
      total++;
      if (b) good++;
    }
    if (a != b) {
      bad = i;
      break;
    }
  }
  *total_hit = total;
  *true_hit = good;
  *first_mismatch = bad;
}

uint64_t orc_popcount_rows(const uint32_t *bitmap, int n) {
  if (n <= 0) return 0;
  int words = orc_bits_num_needed(n);
  uint64_t c = 0;
  for (int w = 0; w < words - 1; w++) c += (uint64_t)__builtin_popcount(bitmap[w]);
  int tail = n - (words - 1) * BITSWIDTH;
  uint32_t mask = tail == BITSWIDTH ? 0xFFFFFFFFu : ~(0xFFFFFFFFu >> tail);
  return c + (uint64_t)__builtin_popcount(bitmap[words - 1] & mask);
}

/* ---------------------------------------------------------------- BinDex build */

/* argsort (bindex.cpp:249-257) sorts row ids by value with an unstable parallel sort.
 * Area starts and filter vectors depend only on the sorted values, so a stable LSD radix
 * sort (ties in row order) gives the same build and a deterministic position order. */
static void stable_argsort(const uint32_t *v, int n, uint32_t *sval, int32_t *spos) {
  uint32_t *tv = (uint32_t *)malloc((size_t)n * sizeof(uint32_t));
  int32_t *tp = (int32_t *)malloc((size_t)n * sizeof(int32_t));
  size_t *cnt = (size_t *)malloc(65537 * sizeof(size_t));
  for (int i = 0; i < n; i++) {
    tv[i] = v[i];
    tp[i] = i;
  }
  for (int pass = 0; pass < 2; pass++) {
    int sh = pass * 16;
    uint32_t *iv = pass ? sval : tv, *ov = pass ? tv : sval;
    int32_t *ip = pass ? spos : tp, *op = pass ? tp : spos;
    memset(cnt, 0, 65537 * sizeof(size_t));
    for (int i = 0; i < n; i++) cnt[((iv[i] >> sh) & 0xFFFF) + 1]++;
    for (int d = 0; d < 65536; d++) cnt[d + 1] += cnt[d];
    for (int i = 0; i < n; i++) {
      size_t d = cnt[(iv[i] >> sh) & 0xFFFF]++;
      ov[d] = iv[i];
      op[d] = ip[i];
    }
  }
  memcpy(sval, tv, (size_t)n * sizeof(uint32_t));
  memcpy(spos, tp, (size_t)n * sizeof(int32_t));
  free(tv);
  free(tp);
  free(cnt);
}

/* set_fv_val_less + gen_less_bits, bindex.cpp:553-560, 264-273: bit i of the filter
 * vector is (data[i] < compare), MSB-first; bits past n stay 0. */
static void set_fv_val_less(uint32_t *fv, const uint32_t *data, uint32_t compare, int n) {
  int words = orc_bits_num_needed(n);
  for (int w = 0; w < words; w++) {
    int lo = w * BITSWIDTH, hi = lo + BITSWIDTH;
    if (hi > n) hi = n;
    uint32_t acc = 0;
    for (int i = lo; i < hi; i++)
      if (data[i] < compare) acc |= row_bit(i);
    fv[w] = acc;
  }
}

/* init_bindex, bindex.cpp:599-671 (GPU-driver twin bindex_cuda.cpp:164-282). */
int orc_init_bindex(orc_bindex *bx, const uint32_t *data, int n, int K, int threads) {
  memset(bx, 0, sizeof(*bx));
  if (n < K || K < 2) return -1;
  if (threads < 1) threads = 1;
  bx->n = n;
  bx->K = K;
  bx->words = orc_bits_num_needed(n);
  bx->sorted_val = (uint32_t *)malloc((size_t)n * sizeof(uint32_t));
  bx->sorted_pos = (int32_t *)malloc((size_t)n * sizeof(int32_t));
  bx->start_values = (uint32_t *)malloc((size_t)K * sizeof(uint32_t));
  bx->area_start = (int32_t *)malloc((size_t)K * sizeof(int32_t));
  bx->area_len = (int32_t *)malloc((size_t)K * sizeof(int32_t));
  bx->area_counts = (int32_t *)malloc((size_t)K * sizeof(int32_t));
  bx->fvs = (uint32_t *)malloc((size_t)(K - 1) * bx->words * sizeof(uint32_t));
  stable_argsort(data, n, bx->sorted_val, bx->sorted_pos);
  const uint32_t *ds = bx->sorted_val;
  int avg = n / K;
  bx->start_values[0] = ds[0];
  bx->area_start[0] = 0;
  /* bindex.cpp:613-631: a start equal to the previous one keeps the mid-run split;
   * otherwise walk back to the first occurrence of the start value. */
  for (int i = 1; i < K; i++) {
    bx->start_values[i] = ds[i * avg];
    int j = i * avg;
    if (bx->start_values[i] == bx->start_values[i - 1]) {
      bx->area_start[i] = j;
    } else {
      while (ds[j] == bx->start_values[i]) j--;
      bx->area_start[i] = j + 1;
    }
  }
  /* num_insert_to_area, bindex.cpp:590-596; accumulation bindex.cpp:649-652 */
  for (int k = 0; k < K; k++)
    bx->area_len[k] = (k < K - 1 ? bx->area_start[k + 1] : n) - bx->area_start[k];
  bx->area_counts[0] = bx->area_len[0];
  for (int k = 1; k < K; k++) bx->area_counts[k] = bx->area_counts[k - 1] + bx->area_len[k];
  bx->data_min = ds[0];
  bx->data_max = ds[n - 1];
  /* bindex.cpp:656-667: fv[k] compares against area_start_value(areas[k + 1]) */
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
  for (int k = 0; k < K - 1; k++)
    set_fv_val_less(bx->fvs + (size_t)k * bx->words, data, bx->start_values[k + 1], n);
  return 0;
}

void orc_free_bindex(orc_bindex *bx) {
  free(bx->sorted_val);
  free(bx->sorted_pos);
  free(bx->start_values);
  free(bx->area_start);
  free(bx->area_len);
  free(bx->area_counts);
  free(bx->fvs);
  memset(bx, 0, sizeof(*bx));
}

/* init_area, bindex.cpp:410-427: blocks of blockInitSize rows while more than that remain,
 * then one last block with the rest. */
int orc_area_block_num(const orc_bindex *bx, int area) {
  int len = bx->area_len[area];
  int nb = 0, i = 0;
  while (i + ORC_BLOCK_INIT_SIZE < len) {
    nb++;
    i += ORC_BLOCK_INIT_SIZE;
  }
  return nb + 1;
}

static inline int block_first(const orc_bindex *bx, int area, int block) {
  return bx->area_start[area] + block * ORC_BLOCK_INIT_SIZE;
}

static inline int block_len(const orc_bindex *bx, int area, int block) {
  int rest = bx->area_len[area] - block * ORC_BLOCK_INIT_SIZE;
  return rest < ORC_BLOCK_INIT_SIZE ? rest : ORC_BLOCK_INIT_SIZE;
}

/* in_which_area, bindex.cpp:1138-1151 */
int orc_in_which_area(const orc_bindex *bx, uint32_t c) {
  if (c < bx->start_values[0]) return -1;
  for (int i = 0; i < bx->K; i++) {
    uint32_t sv = bx->start_values[i];
    if (sv == c) return i;
    if (sv > c) return i - 1;
  }
  return bx->K - 1;
}

/* in_which_block, bindex.cpp:1153-1176 (the code after its first return is dead). */
int orc_in_which_block(const orc_bindex *bx, int area, uint32_t c) {
  int nb = orc_area_block_num(bx, area);
  int res = nb - 1;
  for (int i = 0; i < nb; i++) {
    uint32_t sv = bx->sorted_val[block_first(bx, area, i)];
    if (sv == c) {
      res = i;
      break;
    }
    if (sv > c) {
      res = i - 1;
      break;
    }
  }
  if (res > 0) {
    int last = block_first(bx, area, res - 1) + block_len(bx, area, res - 1) - 1;
    if (bx->sorted_val[last] == c) res--;
  }
  return res;
}

/* on_which_pos, bindex.cpp:1201-1215: first index in the block whose value is >= c. */
int orc_on_which_pos(const orc_bindex *bx, int area, int block, uint32_t c) {
  const uint32_t *val = bx->sorted_val + block_first(bx, area, block);
  int low = 0, high = block_len(bx, area, block), mid = (low + high) / 2;
  while (low < high) {
    if (val[mid] >= c)
      high = mid;
    else
      low = mid + 1;
    mid = (low + high) / 2;
  }
  return mid;
}

/* ------------------------------------------------------------ BinDex query */

/* memset_mt + memset_numa0, bindex.cpp:892-923 (byte value repeated over every word, tail included) */
static void fill_words(uint32_t *r, int words, uint32_t v, int threads) {
#pragma omp parallel for num_threads(threads) schedule(static)
  for (int w = 0; w < words; w++) r[w] = v;
}

/* copy_filter_vector, bindex.cpp:925-964 */
static void copy_fv(const orc_bindex *bx, uint32_t *r, int k, int threads) {
  int W = bx->words;
  if (k < 0) return fill_words(r, W, 0u, threads);
  if (k >= bx->K - 1) return fill_words(r, W, 0xFFFFFFFFu, threads);
  const uint32_t *f = bx->fvs + (size_t)k * W;
#pragma omp parallel for num_threads(threads) schedule(static)
  for (int w = 0; w < W; w++) r[w] = f[w];
}

/* copy_filter_vector_not, bindex.cpp:966-1003 (negation also sets the tail bits) */
static void copy_fv_not(const orc_bindex *bx, uint32_t *r, int k, int threads) {
  int W = bx->words;
  if (k < 0) return fill_words(r, W, 0xFFFFFFFFu, threads);
  if (k >= bx->K - 1) return fill_words(r, W, 0u, threads);
  const uint32_t *f = bx->fvs + (size_t)k * W;
#pragma omp parallel for num_threads(threads) schedule(static)
  for (int w = 0; w < W; w++) r[w] = ~f[w];
}

/* copy_filter_vector_bt, bindex.cpp:1005-1054: fv[kr] & ~fv[kl] with edge cases. */
static void copy_fv_bt(const orc_bindex *bx, uint32_t *r, int kl, int kr, int threads) {
  int W = bx->words, K = bx->K;
  if (kr < 0) return fill_words(r, W, 0u, threads);
  if (kr >= K - 1) return copy_fv_not(bx, r, kl, threads);
  if (kl < 0) return copy_fv(bx, r, kr, threads);
  if (kl >= K - 1) return fill_words(r, W, 0u, threads);
  const uint32_t *fl = bx->fvs + (size_t)kl * W, *fr = bx->fvs + (size_t)kr * W;
#pragma omp parallel for num_threads(threads) schedule(static)
  for (int w = 0; w < W; w++) r[w] = fr[w] & ~fl[w];
}

/* copy_filter_vector_xor, bindex.cpp:1086-1136: its SIMD part (whole groups of 8 words)
 * computes andnot (copy_bitmap_xor_simd, bindex.cpp:1079), the remainder words xor.
 * Returns -1 where the reference asserts. */
static int copy_fv_xor(const orc_bindex *bx, uint32_t *r, int kl, int kr, int threads) {
  int W = bx->words, K = bx->K;
  if (kr < 0 || kr >= K - 1) return -1;
  if (kl < 0) {
    copy_fv(bx, r, kr, threads);
    return 0;
  }
  if (kl >= K - 1) return -1;
  const uint32_t *fl = bx->fvs + (size_t)kl * W, *fr = bx->fvs + (size_t)kr * W;
  int simd_n = (W / 8) * 8;
#pragma omp parallel for num_threads(threads) schedule(static)
  for (int w = 0; w < W; w++) r[w] = w < simd_n ? (fr[w] & ~fl[w]) : (fl[w] ^ fr[w]);
  return 0;
}

/* refine / refine_positions_mt, bindex.cpp:1217-1257: XOR-toggle the bit of every given
 * sorted-array slot's row (atomic, as the reference's __sync_fetch_and_xor). */
static void toggle_slots(const orc_bindex *bx, uint32_t *r, int first, int count, int threads) {
  const int32_t *pos = bx->sorted_pos;
#pragma omp parallel for num_threads(threads) schedule(static)
  for (int s = first; s < first + count; s++) {
    int p = pos[s];
    __atomic_fetch_xor(&r[p / BITSWIDTH], row_bit(p), __ATOMIC_RELAXED);
  }
}

/* The boundary-area refine shared by lt/gt/bt/eq (bindex.cpp:1386-1416): whole blocks on
 * the chosen side of `block`, then the part of `block` before or after `pos`. */
static void refine_side(const orc_bindex *bx, uint32_t *r, int area, int block, int pos, int upper,
                        int threads) {
  int nb = orc_area_block_num(bx, area);
  int sb = upper ? 0 : block + 1, eb = upper ? block : nb;
  if (eb > sb) {
    int first = block_first(bx, area, sb);
    int last = block_first(bx, area, eb - 1) + block_len(bx, area, eb - 1);
    toggle_slots(bx, r, first, last - first, threads);
  }
  int bf = block_first(bx, area, block);
  if (upper)
    toggle_slots(bx, r, bf, pos, threads);
  else
    toggle_slots(bx, r, bf + pos, block_len(bx, area, block) - pos, threads);
}

typedef struct {
  int area, block, pos, upper;
} side_t;

/* area/block/pos lookup and the is_upper_fv estimate, bindex.cpp:1372-1390 */
static int locate(const orc_bindex *bx, uint32_t c, side_t *s) {
  s->area = orc_in_which_area(bx, c);
  if (s->area < 0) return -1;
  s->block = orc_in_which_block(bx, s->area, c);
  if (s->block < 0) return -2;
  s->pos = orc_on_which_pos(bx, s->area, s->block, c);
  s->upper = s->block < orc_area_block_num(bx, s->area) / 2;
  return 0;
}

/* bindex_scan_lt, bindex.cpp:1365-1417 */
static int scan_lt(const orc_bindex *bx, uint32_t *r, uint32_t c, int th) {
  side_t s;
  int e = locate(bx, c, &s);
  if (e == -1) {
    fill_words(r, bx->words, 0u, th);
    return 0;
  }
  if (e) return -1;
  copy_fv(bx, r, s.upper ? s.area - 1 : s.area, th);
  refine_side(bx, r, s.area, s.block, s.pos, s.upper, th);
  return 0;
}

/* bindex_scan_gt, bindex.cpp:1424-1476 (compare + 1 wraps, bindex.cpp:1426) */
static int scan_gt(const orc_bindex *bx, uint32_t *r, uint32_t c, int th) {
  c = c + 1u;
  side_t s;
  int e = locate(bx, c, &s);
  if (e == -1) {
    fill_words(r, bx->words, 0xFFFFFFFFu, th);
    return 0;
  }
  if (e) return -1;
  copy_fv_not(bx, r, s.upper ? s.area - 1 : s.area, th);
  refine_side(bx, r, s.area, s.block, s.pos, s.upper, th);
  return 0;
}

/* bindex_scan_bt, bindex.cpp:1483-1576 */
static int scan_bt(const orc_bindex *bx, uint32_t *r, uint32_t c1, uint32_t c2, int th) {
  if (!(c2 > c1)) return -1; /* assert(compare2 > compare1), bindex.cpp:1484 */
  c1 = c1 + 1u;
  side_t l, rr;
  int e = locate(bx, c1, &l);
  if (e == -1) { /* bindex.cpp:1493-1501: all ones, upper bound ignored */
    fill_words(r, bx->words, 0xFFFFFFFFu, th);
    return 0;
  }
  if (e) return -1;
  e = locate(bx, c2, &rr);
  if (e) return -1; /* assert(0), bindex.cpp:1521-1528 */
  copy_fv_bt(bx, r, l.upper ? l.area - 1 : l.area, rr.upper ? rr.area - 1 : rr.area, th);
  refine_side(bx, r, l.area, l.block, l.pos, l.upper, th);
  refine_side(bx, r, rr.area, rr.block, rr.pos, rr.upper, th);
  return 0;
}

/* bindex_scan_eq, bindex.cpp:1578-1691 */
static int scan_eq(const orc_bindex *bx, uint32_t *r, uint32_t c, int th) {
  int K = bx->K;
  int area = orc_in_which_area(bx, c);
  if (area < 0 || area > K - 1) return -1; /* assert, bindex.cpp:1583 */
  if (area != K - 1 && bx->start_values[area + 1] == c) {
    side_t a, b;
    if (locate(bx, c, &a)) return -1;
    uint32_t c1 = c + 1u;
    b.area = orc_in_which_area(bx, c1);
    if (b.area < 0) return -1;
    b.block = orc_in_which_block(bx, b.area, c1);
    if (b.block < 0) return -1;
    b.pos = orc_on_which_pos(bx, b.area, b.block, c1);
    /* bindex.cpp:1632-1641 */
    b.upper = b.area == 0 || (b.block < orc_area_block_num(bx, b.area) / 2 &&
                              bx->start_values[b.area - 1] != bx->start_values[b.area]);
    if (copy_fv_xor(bx, r, a.upper ? a.area - 1 : a.area, b.upper ? b.area - 1 : b.area, th))
      return -1;
    refine_side(bx, r, a.area, a.block, a.pos, a.upper, th);
    refine_side(bx, r, b.area, b.block, b.pos, b.upper, th);
    return 0;
  }
  /* bindex.cpp:1674-1689 with set_eq_bitmap_mt, bindex.cpp:1302-1328 */
  fill_words(r, bx->words, 0u, th);
  int block = orc_in_which_block(bx, area, c);
  if (block < 0) return -1;
  int nb = orc_area_block_num(bx, area);
  for (int i = block; i < nb; i++) {
    int f = block_first(bx, area, i), len = block_len(bx, area, i);
    if (bx->sorted_val[f] > c) break;
    for (int j = 0; j < len && bx->sorted_val[f + j] <= c; j++)
      if (bx->sorted_val[f + j] == c) {
        int p = bx->sorted_pos[f + j];
        r[p / BITSWIDTH] ^= row_bit(p);
      }
  }
  return 0;
}

int orc_bindex_scan(const orc_bindex *bx, uint32_t *result, int op, uint32_t t1, uint32_t t2,
                    int threads) {
  if (threads < 1) threads = 1;
  switch (op) {
    case ORC_LT: return scan_lt(bx, result, t1, threads);
    case ORC_LE: return scan_lt(bx, result, t1 + 1u, threads); /* bindex.cpp:1419-1422 */
    case ORC_GT: return scan_gt(bx, result, t1, threads);
    case ORC_GE: return scan_gt(bx, result, t1 - 1u, threads); /* bindex.cpp:1478-1481 */
    case ORC_BT: return scan_bt(bx, result, t1, t2, threads);
    case ORC_EQ: return scan_eq(bx, result, t1, threads);
    default: return -1;
  }
}

/* ------------------------------------------------------- GPU refine-area surface */

/* refineBlock, cuda_and.cu:74-99: the selected column is tested first, then the others;
 * every bound is exclusive. */
static void refine_block_rows(const orc_bindex *sb, const uint32_t *const *raws, int ncols,
                              const uint32_t *pred, int sel, int inverse, int s0, int s1,
                              uint32_t *bitmap) {
  for (int s = s0; s < s1; s++) {
    int p = sb->sorted_pos[s];
    uint32_t x = raws[sel][p];
    if (x <= pred[2 * sel] || x >= pred[2 * sel + 1]) continue;
    int j;
    for (j = 0; j < ncols; j++) {
      if (j == sel) continue;
      uint32_t y = raws[j][p];
      if (y <= pred[2 * j] || y >= pred[2 * j + 1]) break;
    }
    if (j < ncols) continue;
    if (inverse)
      bitmap[p / BITSWIDTH] &= ~row_bit(p);
    else
      bitmap[p / BITSWIDTH] |= row_bit(p);
  }
}

/* GPURefineAreaWithCuda, cuda_and.cu:510-640 (block choice cuda_and.cu:546-555, block
 * ranges cuda_and.cu:579-628) */
int orc_refine_area(const orc_bindex *const *bxs, const uint32_t *const *raws, int ncols,
                    const uint32_t *predicate, int sel, int inverse, uint32_t *bitmap) {
  const orc_bindex *b = bxs[sel];
  uint32_t compare = predicate[2 * sel + 1];
  int area = orc_in_which_area(b, compare);
  if (area < 0) return -1;
  if (b->start_values[area] == compare) {
    compare = predicate[2 * sel];
    area = orc_in_which_area(b, compare);
    if (area < 0) return -1;
  }
  int block = orc_in_which_block(b, area, compare);
  if (block < 0) return -1;
  int nb = orc_area_block_num(b, area);
  int b0 = inverse ? block : 0, b1 = inverse ? nb - 1 : block;
  if (b1 < b0) return 0;
  int s0 = block_first(b, area, b0), s1 = block_first(b, area, b1) + block_len(b, area, b1);
  refine_block_rows(b, raws, ncols, predicate, sel, inverse, s0, s1, bitmap);
  return 0;
}

/* GPURefineEqAreaWithCuda, cuda_and.cu:642-812 */
int orc_refine_eq_area(const orc_bindex *const *bxs, const uint32_t *const *raws, int ncols,
                       const uint32_t *predicate, int sel, int inverse, uint32_t *bitmap) {
  const orc_bindex *b = bxs[sel];
  uint32_t left = predicate[2 * sel], right = predicate[2 * sel + 1];
  if (left > right) {
    uint32_t t = left;
    left = right;
    right = t;
  }
  int al = orc_in_which_area(b, left), ar = orc_in_which_area(b, right);
  if (ar < 0) return -1;
  int br = orc_in_which_block(b, ar, right);
  if (br < 0) return -1;
  if (al == ar) { /* cuda_and.cu:701-721 */
    int bl = orc_in_which_block(b, al, left);
    for (int j = bl; j <= br; j++)
      refine_block_rows(b, raws, ncols, predicate, sel, inverse, block_first(b, al, j),
                        block_first(b, al, j) + block_len(b, al, j), bitmap);
    return 0;
  }
  if (al >= 0) { /* cuda_and.cu:725-750 */
    int bl = orc_in_which_block(b, al, left);
    for (int j = orc_area_block_num(b, al) - 1; j >= bl; j--)
      refine_block_rows(b, raws, ncols, predicate, sel, inverse, block_first(b, al, j),
                        block_first(b, al, j) + block_len(b, al, j), bitmap);
  }
  for (int i = al + 1; i <= ar - 1; i++) /* cuda_and.cu:751-776 */
    refine_block_rows(b, raws, ncols, predicate, sel, inverse, b->area_start[i],
                      b->area_start[i] + b->area_len[i], bitmap);
  for (int j = 0; j <= br; j++) /* cuda_and.cu:777-799 */
    refine_block_rows(b, raws, ncols, predicate, sel, inverse, block_first(b, ar, j),
                      block_first(b, ar, j) + block_len(b, ar, j), bitmap);
  return 0;
}

/* ------------------------------------------------------------------ remap surface */

int64_t orc_normal_encode(const uint32_t *data, int n, uint32_t encode_min, uint32_t encode_max,
                          uint32_t *codes, int32_t *sorted_pos, uint32_t *sorted_codes,
                          uint32_t *map_values, uint32_t *map_starts) {
  if (n <= 0) return 0;
  uint32_t *sval = (uint32_t *)malloc((size_t)n * sizeof(uint32_t));
  stable_argsort(data, n, sval, sorted_pos);
  /* remap.cpp:158: int intervalPerObject = (double(max) - double(min)) / double(n) */
  const int interval = (int)(((double)encode_max - (double)encode_min) / (double)n);
  /* remap.cpp:175-196 (setDataPerThread, moveDataByPosPerThread): the i-th sorted row gets
   * encode_min + i * interval, in 32-bit wrap-around arithmetic */
  for (int i = 0; i < n; i++) {
    uint32_t c = encode_min + (uint32_t)i * (uint32_t)interval;
    sorted_codes[i] = c;
    codes[sorted_pos[i]] = c;
  }
  /* remap.cpp:161-165: every distinct value starts where the previous run ended */
  int64_t m = 0;
  for (int i = 0; i < n; i++) {
    if (i == 0 || sval[i] != sval[i - 1]) {
      map_values[m] = sval[i];
      map_starts[m] = sorted_codes[i];
      m++;
    }
  }
  free(sval);
  return m;
}

int64_t orc_dispersed_map(const uint32_t *data, int n, uint32_t encode_min, uint32_t encode_max,
                          uint32_t *map_values, uint32_t *map_starts, uint32_t *map_steps) {
  if (n <= 0) return 0;
  uint32_t *sval = (uint32_t *)malloc((size_t)n * sizeof(uint32_t));
  int32_t *pos = (int32_t *)malloc((size_t)n * sizeof(int32_t));
  stable_argsort(data, n, sval, pos);
  /* remap.cpp:222: double intervalPerObject = (double(max) - double(min)) / double(n) */
  const double per_row = ((double)encode_max - (double)encode_min) / (double)n;
  /* remap.cpp:224-228: start begins at encode_min and grows by int(interval * count) */
  uint32_t start = encode_min;
  int64_t m = 0;
  int i = 0;
  while (i < n) {
    int j = i;
    while (j < n && sval[j] == sval[i]) j++;
    const uint32_t step = (uint32_t)(int64_t)(per_row * (double)(j - i));
    map_values[m] = sval[i];
    map_starts[m] = start;
    map_steps[m] = step;
    start += step;
    m++;
    i = j;
  }
  free(sval);
  free(pos);
  return m;
}

static int64_t lower_bound_u32(const uint32_t *a, int64_t m, uint32_t q) {
  int64_t lo = 0, hi = m;
  while (lo < hi) {
    int64_t mid = lo + (hi - lo) / 2;
    if (a[mid] < q) lo = mid + 1; else hi = mid;
  }
  return lo;
}

static int64_t upper_bound_u32(const uint32_t *a, int64_t m, uint32_t q) {
  int64_t lo = 0, hi = m;
  while (lo < hi) {
    int64_t mid = lo + (hi - lo) / 2;
    if (a[mid] <= q) lo = mid + 1; else hi = mid;
  }
  return lo;
}

int orc_encode_query(const uint32_t *map_values, const uint32_t *map_starts, int64_t m, uint32_t q,
                     int op, uint32_t sentinel, uint32_t *out) {
  int64_t i;
  switch (op) {
    case ORC_LT:
    case ORC_GE:
      i = lower_bound_u32(map_values, m, q);
      *out = i == m ? sentinel : map_starts[i];
      return 0;
    case ORC_LE:
      i = upper_bound_u32(map_values, m, q);
      *out = i == m ? sentinel : map_starts[i] - 1u;
      return 0;
    case ORC_GT:
      i = upper_bound_u32(map_values, m, q);
      *out = i == m ? sentinel : map_starts[i];
      return 0;
    default:
      return -1;
  }
}

uint32_t orc_find_key_by_value(const uint32_t *map_values, const uint32_t *map_starts, int64_t m,
                               uint32_t val) {
  uint32_t last = 0;
  for (int64_t i = 0; i < m; i++) {
    if (map_starts[i] > val) return last;
    last = map_values[i];
  }
  return 0;
}

/* ------------------------------------------------------------------ bitmap digest */

static uint64_t mix64(uint64_t x) {
  x ^= 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

/* The checker's conjunction of one multi-column query (bindex.cpp:2265-2289: raw_scan per
 * column into a zeroed bitmap, AND-merged), for closed ranges lo[c] <= x <= hi[c] (the
 * raw_scan predicates le / ge / bt of the query's targets), reduced to its popcount and
 * digest word by word so that a 1e9-row query needs no 125 MB bitmap.  The digest is
 * orc_bitmap_digest's. */
void orc_conjunction_digest(const uint32_t *const *cols, int ncols, int64_t n, const uint32_t *lo,
                            const uint32_t *hi, int threads, uint64_t *digest, uint64_t *count) {
  uint64_t d = 0, c = 0;
  const int64_t words = n <= 0 ? 0 : (n + 31) / 32;
#pragma omp parallel for num_threads(threads) schedule(static, 4096) reduction(+ : d, c)
  for (int64_t w = 0; w < words; w++) {
    const int64_t r0 = w * 32;
    const int rows = n - r0 < 32 ? (int)(n - r0) : 32;
    uint32_t acc = rows == 32 ? 0xFFFFFFFFu : ~(0xFFFFFFFFu >> rows);
    for (int k = 0; k < ncols && acc; k++) {
      const uint32_t *x = cols[k] + r0;
      uint32_t m = 0;
      for (int i = 0; i < rows; i++)
        if (x[i] >= lo[k] && x[i] <= hi[k]) m |= 0x80000000u >> i;
      acc &= m;
    }
    d += mix64(((uint64_t)w << 32) | acc);
    c += (uint64_t)__builtin_popcount(acc);
  }
  *digest = d;
  *count = c;
}

uint64_t orc_bitmap_digest(const uint32_t *bitmap, int64_t n) {
  if (n <= 0) return 0;
  const int64_t words = (n + 31) / 32;
  const uint32_t t = (uint32_t)(n & 31);
  const uint32_t last = t == 0 ? 0xFFFFFFFFu : ~(0xFFFFFFFFu >> t);
  uint64_t d = 0;
  for (int64_t i = 0; i < words; i++) {
    const uint32_t w = i == words - 1 ? bitmap[i] & last : bitmap[i];
    d += mix64(((uint64_t)i << 32) | w);
  }
  return d;
}
