/* Runs the oracle's C restatement (oracle/oracle.c, linked in directly) on small seeded
 * inputs under AddressSanitizer / UndefinedBehaviorSanitizer: raw_scan family, compare,
 * init_bindex, every bindex_scan operator, the refine-area surface, the digests and the
 * remap restatement.  Exits 0 when the sanitizers report nothing and the BinDex answers
 * equal raw_scan's on distinct values. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static uint64_t s = 88172645463325252ull;
static uint32_t rnd(void) {
  s ^= s << 13;
  s ^= s >> 7;
  s ^= s << 17;
  return (uint32_t)(s >> 16);
}

int main(void) {
  const int n = 600000;
  uint32_t *x = malloc(n * sizeof(uint32_t));
  for (int i = 0; i < n; i++) x[i] = rnd();
  const int W = orc_bits_num_needed(n);
  uint32_t *a = calloc(W, 4), *b = calloc(W, 4);
  int bad = 0;
  orc_bindex bx;
  if (orc_init_bindex(&bx, x, n, 128, 4) != 0) return 2;
  for (int q = 0; q < 40; q++) {
    uint32_t t1 = rnd(), t2 = rnd();
    if (q % 7 == 0) t1 = x[q];
    const int op = q % 6;
    if (op == ORC_BT && t1 > t2) {
      uint32_t t = t1;
      t1 = t2;
      t2 = t;
    }
    memset(a, 0, W * 4);
    orc_raw_scan_mt(x, n, a, t1, t2, op, 4);
    if (orc_bindex_scan(&bx, b, op, t1, t2, 4) == 0) {
      int total = 0, good = 0, first = 0;
      orc_compare_bitmap(a, b, n, &total, &good, &first);
      if (first != -1) bad++;
    }
    const uint32_t lo = t1 < t2 ? t1 : t2, hi = t1 < t2 ? t2 : t1;
    const uint32_t *cols[2] = {x, x};
    uint32_t los[2] = {lo, lo}, his[2] = {hi, hi};
    uint64_t d = 0, c = 0;
    orc_conjunction_digest(cols, 2, n, los, his, 3, &d, &c);
    (void)orc_bitmap_digest(a, n);
  }
  const orc_bindex *bxs[1] = {&bx};
  const uint32_t *raws[1] = {x};
  uint32_t pred[2] = {bx.start_values[5] - 1, bx.start_values[5] + 12345};
  memset(a, 0, W * 4);
  (void)orc_refine_area(bxs, raws, 1, pred, 0, 0, a);
  (void)orc_refine_eq_area(bxs, raws, 1, pred, 0, 1, a);
  const int m = 5000;
  uint32_t *codes = malloc(m * 4), *sc = malloc(m * 4), *mv = malloc(m * 4), *ms = malloc(m * 4), *st = malloc(m * 4);
  int32_t *sp = malloc(m * 4);
  const int64_t k = orc_normal_encode(x, m, 0, 0xFFFFFFFFu, codes, sp, sc, mv, ms);
  uint32_t out = 0;
  for (int q = 0; q < 100; q++) (void)orc_encode_query(mv, ms, k, rnd(), ORC_LT + q % 4, (uint32_t)m, &out);
  (void)orc_find_key_by_value(mv, ms, k, rnd());
  (void)orc_dispersed_map(x, m, 0, 0xFFFFFFFFu, mv, ms, st);
  orc_free_bindex(&bx);
  free(x), free(a), free(b), free(codes), free(sc), free(mv), free(ms), free(st), free(sp);
  if (bad) {
    fprintf(stderr, "%d BinDex answers differ from raw_scan\n", bad);
    return 1;
  }
  printf("oracle ok\n");
  return 0;
}
