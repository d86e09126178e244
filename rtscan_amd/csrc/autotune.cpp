// autotune.cpp -- the batched launch's autotuner: times candidate launch shapes (tile and
// group sizes, workgroups, order, bin count, the planner's direct weight) on the plan's
// current queries and keeps the fastest; results cached in a text file keyed by the
// device, the table, the query count and the query-mix digest.
#include <hip/hip_runtime_api.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "rts_batch.h"

namespace {

std::string cache_key(int64_t n, int table_cols, int nq, int qcols, unsigned long long mix) {
  int dev = 0;
  char arch[256] = "unknown";
  int cus = 0;
  if (hipGetDevice(&dev) == hipSuccess) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess) {
      snprintf(arch, sizeof(arch), "%s", prop.gcnArchName);
      cus = prop.multiProcessorCount;
    }
  }
  for (char *c = arch; *c; c++)
    if (*c == ' ' || *c == ':') *c = '_';
  std::ostringstream k;
  char mixhex[32];
  snprintf(mixhex, sizeof(mixhex), "%016llx", mix);
  k << arch << "/cu" << cus << "/n" << n << "/c" << table_cols << "/q" << nq << "/qc" << qcols << "/mix" << mixhex;
  return k.str();
}

bool cache_lookup(const char *path, const std::string &key, rts_batch_config *cfg, float *ms) {
  std::ifstream in(path);
  std::string line;
  bool found = false;
  while (std::getline(in, line)) {
    std::istringstream ls(line);
    std::string k;
    rts_batch_config c;
    float m;
    if (ls >> k >> c.tile_words >> c.queries_per_group >> c.blocks_per_cu >> c.order >> c.threads >> c.bins >>
            c.direct_weight >> m &&
        k == key) {
      *cfg = c;  // the last entry for a key wins
      *ms = m;
      found = true;
    }
  }
  return found;
}

void cache_store(const char *path, const std::string &key, const rts_batch_config &c, float ms) {
  FILE *f = fopen(path, "a");
  if (!f) return;
  fprintf(f, "%s %d %d %d %d %d %d %d %.4f\n", key.c_str(), c.tile_words, c.queries_per_group, c.blocks_per_cu,
          c.order, c.threads, c.bins, c.direct_weight, ms);
  fclose(f);
}

}  // namespace

extern "C" hipError_t rts_batch_autotune(rts_batch_plan *plan, BITS *out, int64_t out_stride_words,
                                         const rts_batch_config *candidates, int ncandidates, int reps,
                                         const char *cache_path, rts_batch_config *chosen,
                                         float *best_ms, hipStream_t stream) {
  if (!plan || !out || !candidates || ncandidates < 1 || reps < 1) return hipErrorInvalidValue;
  int bcols = 0, dcols = 0;
  rts_batch_plan_stats(plan, &bcols, &dcols, nullptr, nullptr);
  long long n = 0;
  int table_cols = 0, nq = 0;
  rts_batch_plan_info(plan, &n, &table_cols, &nq);
  const std::string key = cache_key(n, table_cols, nq, bcols + dcols, rts_batch_plan_mix_digest(plan));
  rts_batch_config best{};
  float best_t = 0;
  if (cache_path && cache_lookup(cache_path, key, &best, &best_t)) {
    hipError_t e = rts_batch_plan_set_config(plan, &best, stream);
    if (e != hipSuccess) return e;
    if (chosen) *chosen = best;
    if (best_ms) *best_ms = best_t;
    return hipSuccess;
  }
  hipEvent_t a, b;
  hipError_t e = hipEventCreate(&a);
  if (e != hipSuccess) return e;
  e = hipEventCreate(&b);
  if (e != hipSuccess) {
    (void)hipEventDestroy(a);
    return e;
  }
  bool have = false;
  for (int i = 0; i < ncandidates && e == hipSuccess; i++) {
    e = rts_batch_plan_set_config(plan, &candidates[i], stream);
    if (e != hipSuccess) break;
    e = rts_batch_plan_launch(plan, out, out_stride_words, nullptr, stream);  // warm-up
    std::vector<float> t;
    for (int r = 0; r < reps && e == hipSuccess; r++) {
      e = hipEventRecord(a, stream);
      if (e == hipSuccess) e = rts_batch_plan_launch(plan, out, out_stride_words, nullptr, stream);
      if (e == hipSuccess) e = hipEventRecord(b, stream);
      if (e == hipSuccess) e = hipEventSynchronize(b);
      float ms = 0;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
      t.push_back(ms);
    }
    if (e != hipSuccess) break;
    std::sort(t.begin(), t.end());
    const float med = t[t.size() / 2];
    if (!have || med < best_t) {
      best = candidates[i];
      best_t = med;
      have = true;
    }
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (e != hipSuccess) return e;
  e = rts_batch_plan_set_config(plan, &best, stream);
  if (e != hipSuccess) return e;
  if (cache_path) cache_store(cache_path, key, best, best_t);
  if (chosen) *chosen = best;
  if (best_ms) *best_ms = best_t;
  return hipSuccess;
}
