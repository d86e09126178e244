// CPU test of the request loop's admission logic (rtscan_amd/csrc/admission.h) on a made-up
// pool budget: first fit with coalescing, arrival order per table, no overtaking, the
// memory bound, withdrawal, and many threads driving one Admission behind a mutex (run
// under AddressSanitizer / UndefinedBehaviorSanitizer and ThreadSanitizer builds).
#include <stdio.h>
#include <stdlib.h>

#include <mutex>
#include <random>
#include <thread>
#include <vector>

#include "admission.h"

using rts::Admission;
using rts::PoolAllocator;

static int failures = 0;
#define CHECK(c)                                                     \
  do {                                                               \
    if (!(c)) {                                                      \
      fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #c); \
      failures++;                                                    \
    }                                                                \
  } while (0)

static void pool_first_fit_and_coalescing() {
  PoolAllocator p(4096 + 100);  // rounds down to 16 granules
  CHECK(p.capacity() == 4096);
  const size_t a = p.alloc(1000), b = p.alloc(1), c = p.alloc(256);
  CHECK(a == 0 && b == 1024 && c == 1280);
  CHECK(p.in_use() == 1536);
  CHECK(p.alloc(4096) == SIZE_MAX);
  p.release(b, 1);
  CHECK(p.alloc(200) == 1024);  // first fit reuses the hole
  p.release(1024, 200);
  p.release(a, 1000);
  p.release(c, 256);
  CHECK(p.in_use() == 0 && p.free_ranges() == 1);  // every free coalesced back
  CHECK(p.alloc(4096) == 0);
  CHECK(p.peak() == 4096);
}

static void admission_order_and_memory_bound() {
  // table 0 bitmaps take 1000 bytes (1024 in the pool), table 1 bitmaps 3000 (3072)
  auto bytes_of = [](int t) { return t == 0 ? (size_t)1000 : (size_t)3000; };
  Admission adm(8192);
  for (uint64_t t = 1; t <= 6; t++) adm.enqueue(t, t == 3 ? 1 : 0);
  // oldest request is on table 0: its table-0 requests go in arrival order, table 1 waits
  auto got = adm.admit(3, bytes_of);
  CHECK(got.size() == 3 && got[0].ticket == 1 && got[1].ticket == 2 && got[2].ticket == 4);
  CHECK(adm.waiting() == 3 && adm.front_table() == 1);
  // next batch is table 1's: 3072 bytes fit in the 5120 left
  auto g2 = adm.admit(8, bytes_of);
  CHECK(g2.size() == 1 && g2[0].ticket == 3);
  // table 0 again: 2048 bytes left -> two fit, the third must wait for memory
  adm.enqueue(7, 0);
  auto g3 = adm.admit(8, bytes_of);
  CHECK(g3.size() == 2 && g3[0].ticket == 5 && g3[1].ticket == 6);
  CHECK(adm.admit(8, bytes_of).empty() && adm.waiting() == 1);
  adm.release(got[0].offset, 1000);
  auto g4 = adm.admit(8, bytes_of);
  CHECK(g4.size() == 1 && g4[0].ticket == 7 && g4[0].offset == got[0].offset);
  // no overtaking: a large request at the head blocks smaller later ones of its table
  Admission b(4096);
  auto big = [](int t) { return t == 0 ? (size_t)3000 : (size_t)100; };
  b.enqueue(1, 0);
  b.enqueue(2, 0);
  auto h = b.admit(8, big);
  CHECK(h.size() == 1 && h[0].ticket == 1);
  CHECK(b.admit(8, big).empty());
  CHECK(b.remove(2) && !b.remove(2) && b.waiting() == 0);
}

static void many_threads() {
  Admission adm(64 * 1024);
  std::mutex mu;
  auto bytes_of = [](int t) { return (size_t)(512 + 256 * t); };
  std::vector<std::thread> th;
  std::vector<uint64_t> admitted_total(8, 0);
  for (int k = 0; k < 8; k++) {
    th.emplace_back([&, k] {
      std::mt19937 rng(k);
      std::vector<Admission::Admitted> mine;
      std::vector<int> mine_table;
      for (int i = 0; i < 2000; i++) {
        std::lock_guard<std::mutex> lock(mu);
        const uint64_t ticket = (uint64_t)k << 32 | (uint64_t)i;
        adm.enqueue(ticket, (int)(rng() % 3));
        const int t = adm.front_table();
        auto got = adm.admit(4, bytes_of);
        for (auto &a : got) {
          mine.push_back(a);
          mine_table.push_back(t);
        }
        admitted_total[k] += got.size();
        while (mine.size() > 6) {
          adm.release(mine.front().offset, bytes_of(mine_table.front()));
          mine.erase(mine.begin());
          mine_table.erase(mine_table.begin());
        }
        CHECK(adm.pool().in_use() <= adm.pool().capacity());
      }
      std::lock_guard<std::mutex> lock(mu);
      for (size_t i = 0; i < mine.size(); i++) adm.release(mine[i].offset, bytes_of(mine_table[i]));
    });
  }
  for (auto &t : th) t.join();
  uint64_t total = 0;
  for (uint64_t v : admitted_total) total += v;
  CHECK(total + adm.waiting() == 8 * 2000);
  CHECK(adm.pool().in_use() == 0);
}

int main() {
  pool_first_fit_and_coalescing();
  admission_order_and_memory_bound();
  many_threads();
  if (failures) return 1;
  printf("admission ok\n");
  return 0;
}
