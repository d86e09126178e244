// remap.hip -- value encoding of a column (the reference's remap surface, remap.cpp:110-291).
//
// Both encodings replace each value by a code that keeps the order of distinct values and
// gives every distinct value a code range proportional to its frequency:
//   rank form (remap.cpp:110-203): the row at sorted index i gets encode_min + i * I with
//     I = (int)((encode_max - encode_min) / n) in double; a value's map entry is the code of
//     its first sorted row.
//   dispersed form (remap.cpp:205-241): value v with c rows owns the codes
//     [start(v), start(v) + step(v) - 1) with step(v) = (int)(D * c), D = (max - min) / n in
//     double, start(v) = encode_min + sum of step(u) over u < v; each row draws an offset
//     inside its value's range (the reference uses rand(); here a counter hash of the row
//     index and a seed, so the result does not depend on thread timing).
// All code arithmetic is 32-bit wrap-around, as the reference's int/CODE mix is on x86.
//
// Device pipeline per column: stable radix sort of (value, row) -> run heads -> inclusive
// scan of the head flags (run index of every sorted row) -> codes scattered back to row
// order.  Every pass is a streaming pass over n words except the scatter (random 4-byte
// writes); the sort dominates.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <stdint.h>
#include <string.h>

#include "common.h"
#define RTS_REMAP_NO_CXX
#include "rts_remap.h"

namespace rts {
namespace {

constexpr int kRemapBlock = 256;

unsigned grid_for(int64_t items) { return (unsigned)((items + kRemapBlock - 1) / kRemapBlock); }

__global__ void k_remap_iota(uint32_t *p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = (uint32_t)i;
}

// head[i] = 1 where sorted row i starts a run of equal values.
__global__ void k_run_heads(const CODE *sval, int64_t n, uint32_t *head) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) head[i] = (i == 0 || sval[i] != sval[i - 1]) ? 1u : 0u;
}

// Rank form: codes in sorted order, scattered to row order, and the map entries at heads.
// run[i] is the inclusive scan of the head flags (run index + 1).
__global__ void k_rank_codes(const CODE *sval, const uint32_t *spos, const uint32_t *head,
                             const uint32_t *run, int64_t n, uint32_t encode_min, uint32_t interval,
                             CODE *codes, CODE *sorted_codes, CODE *map_values, CODE *map_starts) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t c = encode_min + (uint32_t)i * interval;
  if (sorted_codes) sorted_codes[i] = c;
  codes[spos[i]] = c;
  if (head[i]) {
    const uint32_t r = run[i] - 1;
    map_values[r] = sval[i];
    map_starts[r] = c;
  }
}

// Dispersed form, pass 1: run heads -> distinct values and first sorted index.
__global__ void k_run_firsts(const CODE *sval, const uint32_t *head, const uint32_t *run, int64_t n,
                             CODE *map_values, uint32_t *first) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && head[i]) {
    const uint32_t r = run[i] - 1;
    map_values[r] = sval[i];
    first[r] = (uint32_t)i;
  }
}

// Pass 2: step(v) = (int)(D * count(v)); kept as the 32-bit pattern of that int.
__global__ void k_run_steps(const uint32_t *first, int64_t m, int64_t n, double per_row,
                            uint32_t *step) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= m) return;
  const int64_t end = r + 1 < m ? (int64_t)first[r + 1] : n;
  const double w = per_row * (double)(end - (int64_t)first[r]);
  step[r] = (uint32_t)(int64_t)w;  // truncation toward zero, then the low 32 bits
}

// Pass 3: start(v) = encode_min + exclusive prefix sum of the steps (mod 2^32).
__global__ void k_run_starts(const uint32_t *excl, int64_t m, uint32_t encode_min, CODE *map_starts) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < m) map_starts[r] = encode_min + excl[r];
}

__device__ __forceinline__ uint32_t row_hash(uint64_t seed, uint32_t row) {
  uint64_t z = seed + 0x9E3779B97F4A7C15ull * ((uint64_t)row + 1);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return (uint32_t)(z ^ (z >> 31));
}

// Pass 4: code(row) = start(v) + (step(v) >= 2 ? hash(row) % (step(v) - 1) : 0).
__global__ void k_disperse(const uint32_t *spos, const uint32_t *run, int64_t n, const CODE *map_starts,
                           const uint32_t *step, uint64_t seed, CODE *codes) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t r = run[i] - 1;
  const uint32_t row = spos[i];
  const int32_t s = (int32_t)step[r];
  const uint32_t off = s >= 2 ? row_hash(seed, row) % (uint32_t)(s - 1) : 0u;
  codes[row] = map_starts[r] + off;
}

#define RTS_TRY(x)                      \
  do {                                  \
    err = (x);                          \
    if (err != hipSuccess) goto fail;   \
  } while (0)

// Shared front half: stable sort and run indices.  On success sval/spos/head/run are
// allocated (spos may be the caller's buffer) and *m holds the number of distinct values.
struct Runs {
  CODE *sval = nullptr;
  uint32_t *spos = nullptr;
  uint32_t *head = nullptr;
  uint32_t *run = nullptr;
  bool own_spos = false;
  int64_t m = 0;
  void release() {
    (void)hipFree(sval);
    if (own_spos) (void)hipFree(spos);
    (void)hipFree(head);
    (void)hipFree(run);
  }
};

hipError_t sorted_runs(const CODE *data, int64_t n, POSTYPE *caller_spos, Runs &R, hipStream_t s) {
  hipError_t err = hipSuccess;
  uint32_t *iota = nullptr;
  void *tmp = nullptr;
  size_t need = 0, have = 0;
  uint32_t last = 0;

  RTS_TRY(hipMalloc((void **)&R.sval, n * sizeof(CODE)));
  if (caller_spos) {
    R.spos = (uint32_t *)caller_spos;
  } else {
    RTS_TRY(hipMalloc((void **)&R.spos, n * sizeof(uint32_t)));
    R.own_spos = true;
  }
  RTS_TRY(hipMalloc((void **)&R.head, n * sizeof(uint32_t)));
  RTS_TRY(hipMalloc((void **)&R.run, n * sizeof(uint32_t)));
  RTS_TRY(hipMalloc((void **)&iota, n * sizeof(uint32_t)));
  hipLaunchKernelGGL(k_remap_iota, dim3(grid_for(n)), dim3(kRemapBlock), 0, s, iota, n);
  RTS_TRY(hipGetLastError());
  RTS_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, need, data, R.sval, iota, R.spos, (int)n, 0, 32, s));
  have = need;
  RTS_TRY(hipcub::DeviceScan::InclusiveSum(nullptr, need, R.head, R.run, (int)n, s));
  have = need > have ? need : have;
  RTS_TRY(hipMalloc(&tmp, have));
  need = have;
  RTS_TRY(hipcub::DeviceRadixSort::SortPairs(tmp, need, data, R.sval, iota, R.spos, (int)n, 0, 32, s));
  hipLaunchKernelGGL(k_run_heads, dim3(grid_for(n)), dim3(kRemapBlock), 0, s, R.sval, n, R.head);
  RTS_TRY(hipGetLastError());
  need = have;
  RTS_TRY(hipcub::DeviceScan::InclusiveSum(tmp, need, R.head, R.run, (int)n, s));
  RTS_TRY(hipMemcpyAsync(&last, R.run + (n - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  RTS_TRY(hipStreamSynchronize(s));
  R.m = last;

fail:
  (void)hipFree(iota);
  (void)hipFree(tmp);
  return err;
}

bool valid_column(const CODE *data, int64_t n, const CODE *codes, const CODE *mv, const CODE *ms) {
  return data && codes && mv && ms && n >= 1 && n <= 0x7FFFFFFFll;
}

}  // namespace

hipError_t encode_rank(const CODE *data, int64_t n, CODE encode_min, CODE encode_max, CODE *codes,
                       POSTYPE *sorted_pos, CODE *sorted_codes, CODE *map_values, CODE *map_starts,
                       int64_t *map_size, hipStream_t s) {
  if (!valid_column(data, n, codes, map_values, map_starts) || !map_size) return hipErrorInvalidValue;
  hipError_t err = hipSuccess;
  Runs R;
  // remap.cpp:158: int intervalPerObject = (double(max) - double(min)) / double(n)
  const double q = ((double)encode_max - (double)encode_min) / (double)n;
  const uint32_t interval = (uint32_t)(int64_t)q;
  RTS_TRY(sorted_runs(data, n, sorted_pos, R, s));
  hipLaunchKernelGGL(k_rank_codes, dim3(grid_for(n)), dim3(kRemapBlock), 0, s, R.sval, R.spos, R.head,
                     R.run, n, encode_min, interval, codes, sorted_codes, map_values, map_starts);
  RTS_TRY(hipGetLastError());
  *map_size = R.m;
fail:
  R.release();
  return err;
}

hipError_t encode_dispersed(const CODE *data, int64_t n, CODE encode_min, CODE encode_max, uint64_t seed,
                            CODE *codes, CODE *map_values, CODE *map_starts, int64_t *map_size,
                            hipStream_t s) {
  if (!valid_column(data, n, codes, map_values, map_starts) || !map_size) return hipErrorInvalidValue;
  hipError_t err = hipSuccess;
  Runs R;
  uint32_t *first = nullptr, *step = nullptr, *excl = nullptr;
  void *tmp = nullptr;
  size_t need = 0;
  // remap.cpp:222: double intervalPerObject = (double(max) - double(min)) / double(n)
  const double per_row = ((double)encode_max - (double)encode_min) / (double)n;
  RTS_TRY(sorted_runs(data, n, nullptr, R, s));
  RTS_TRY(hipMalloc((void **)&first, R.m * sizeof(uint32_t)));
  RTS_TRY(hipMalloc((void **)&step, R.m * sizeof(uint32_t)));
  RTS_TRY(hipMalloc((void **)&excl, R.m * sizeof(uint32_t)));
  hipLaunchKernelGGL(k_run_firsts, dim3(grid_for(n)), dim3(kRemapBlock), 0, s, R.sval, R.head, R.run, n,
                     map_values, first);
  RTS_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_run_steps, dim3(grid_for(R.m)), dim3(kRemapBlock), 0, s, first, R.m, n, per_row, step);
  RTS_TRY(hipGetLastError());
  RTS_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, need, step, excl, (int)R.m, s));
  RTS_TRY(hipMalloc(&tmp, need));
  RTS_TRY(hipcub::DeviceScan::ExclusiveSum(tmp, need, step, excl, (int)R.m, s));
  hipLaunchKernelGGL(k_run_starts, dim3(grid_for(R.m)), dim3(kRemapBlock), 0, s, excl, R.m, encode_min,
                     map_starts);
  RTS_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_disperse, dim3(grid_for(n)), dim3(kRemapBlock), 0, s, R.spos, R.run, n, map_starts,
                     step, seed, codes);
  RTS_TRY(hipGetLastError());
  RTS_TRY(hipStreamSynchronize(s));
  *map_size = R.m;
fail:
  (void)hipFree(first);
  (void)hipFree(step);
  (void)hipFree(excl);
  (void)hipFree(tmp);
  R.release();
  return err;
}

#undef RTS_TRY

}  // namespace rts

// ------------------------------------------------------------------------------- C ABI

extern "C" hipError_t rts_encode_column(const CODE *data, int64_t n, CODE encode_min, CODE encode_max,
                                        CODE *codes, POSTYPE *sorted_pos, CODE *sorted_codes,
                                        CODE *map_values, CODE *map_starts, int64_t *map_size,
                                        hipStream_t stream) {
  return rts::set_error(rts::encode_rank(data, n, encode_min, encode_max, codes, sorted_pos, sorted_codes,
                                         map_values, map_starts, map_size, stream),
                        "rts_encode_column");
}

extern "C" hipError_t rts_encode_column_dispersed(const CODE *data, int64_t n, CODE encode_min,
                                                  CODE encode_max, uint64_t seed, CODE *codes,
                                                  CODE *map_values, CODE *map_starts, int64_t *map_size,
                                                  hipStream_t stream) {
  return rts::set_error(rts::encode_dispersed(data, n, encode_min, encode_max, seed, codes, map_values,
                                              map_starts, map_size, stream),
                        "rts_encode_column_dispersed");
}

namespace {
int64_t first_not_below(const CODE *v, int64_t m, CODE q, bool strictly_above) {
  int64_t lo = 0, hi = m;
  while (lo < hi) {
    const int64_t mid = lo + (hi - lo) / 2;
    const bool go_right = strictly_above ? v[mid] <= q : v[mid] < q;
    if (go_right) lo = mid + 1; else hi = mid;
  }
  return lo;
}
}  // namespace

extern "C" hipError_t rts_encode_lookup(const CODE *map_values, const CODE *map_starts, int64_t m,
                                        CODE old_query, enum OPERATOR op, CODE sentinel, CODE *code) {
  if ((m > 0 && (!map_values || !map_starts)) || m < 0 || !code) return hipErrorInvalidValue;
  int64_t i;
  switch (op) {
    case LT:  // first value >= q (remap.cpp:244-249)
    case GE:  // remap.cpp:262-267
      i = first_not_below(map_values, m, old_query, false);
      *code = i == m ? sentinel : map_starts[i];
      return hipSuccess;
    case LE:  // first value > q, its start minus one (remap.cpp:250-255)
      i = first_not_below(map_values, m, old_query, true);
      *code = i == m ? sentinel : map_starts[i] - 1u;
      return hipSuccess;
    case GT:  // first value > q (remap.cpp:256-261)
      i = first_not_below(map_values, m, old_query, true);
      *code = i == m ? sentinel : map_starts[i];
      return hipSuccess;
    default:  // eq and bt are rejected (remap.cpp:268-271 exits)
      return hipErrorInvalidValue;
  }
}

extern "C" CODE rts_encode_key_of(const CODE *map_values, const CODE *map_starts, int64_t m, CODE code) {
  // findKeyByValue (remap.cpp:274-283): the key of the last entry whose start is <= code,
  // provided a later entry starts above code; 0 otherwise (also past the last entry).
  CODE last = 0;
  for (int64_t i = 0; i < m; i++) {
    if (map_starts[i] > code) return last;
    last = map_values[i];
  }
  return 0;
}
