// gen.cpp -- seeded synthetic columns for the benchmark and test workloads.
//
// The reference's generators are unseeded (std::random_device, tools/
// generate_uniform_data_n_col.cpp:21-22; np.random without a seed, tools/
// generate_zipf_data.py:12-15), so they cannot reproduce a dataset.  These are
// counter-based: row i of column c is a function of (seed, c, i) only, which lets any
// number of threads fill a column and gives the same table on every machine.
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <thread>
#include <vector>

#include "rtscan_hip.h"

namespace {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;

// splitmix64 finaliser
inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// murmur3 fmix32: a bijection on uint32 (xorshifts and odd multipliers are invertible).
inline uint32_t fmix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85EBCA6Bu;
  h ^= h >> 13;
  h *= 0xC2B2AE35u;
  h ^= h >> 16;
  return h;
}

inline uint64_t stream_base(uint64_t seed, int column) {
  return mix64(seed + kGolden * (uint64_t)(column + 1));
}

// k-th 64-bit draw of row i
inline uint64_t draw(uint64_t base, uint64_t i, int k) { return mix64(base + kGolden * (2 * i + k + 1)); }

// uniform double in (0, 1)
inline double unit_open(uint64_t v) { return ((double)(v >> 11) + 0.5) * (1.0 / 9007199254740992.0); }

template <class F>
void parallel_rows(uint64_t n, int threads, F f) {
  if (threads < 1) threads = 1;
  if ((uint64_t)threads > n / 4096 + 1) threads = (int)(n / 4096 + 1);
  std::vector<std::thread> pool;
  const uint64_t per = (n + threads - 1) / threads;
  for (int t = 0; t < threads; t++) {
    const uint64_t lo = std::min(n, per * t), hi = std::min(n, per * (t + 1));
    pool.emplace_back([=] { f(lo, hi); });
  }
  for (auto &th : pool) th.join();
}

}  // namespace

extern "C" int rts_generate_column(int dist, uint64_t seed, int column_id, uint64_t n, double p0,
                                   double p1, CODE *out, int threads) {
  if (!out && n) return -1;
  const uint64_t base = stream_base(seed, column_id);
  if (dist == RTS_DIST_UNIFORM) {
    parallel_rows(n, threads, [&](uint64_t lo, uint64_t hi) {
      for (uint64_t i = lo; i < hi; i++) out[i] = (CODE)(draw(base, i, 0) >> 32);
    });
    return 0;
  }
  if (dist == RTS_DIST_NORMAL) {
    const double mean = p0, sd = p1;
    if (!(sd >= 0)) return -1;
    parallel_rows(n, threads, [&](uint64_t lo, uint64_t hi) {
      for (uint64_t i = lo; i < hi; i++) {
        // Box-Muller, cosine branch
        const double u1 = unit_open(draw(base, i, 0)), u2 = unit_open(draw(base, i, 1));
        const double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
        double x = floor(mean + sd * z + 0.5);
        if (x < 0) x = 0;
        if (x > 4294967295.0) x = 4294967295.0;
        out[i] = (CODE)x;
      }
    });
    return 0;
  }
  if (dist == RTS_DIST_ZIPF) {
    const double s = p0;
    const uint64_t ranks = (uint64_t)p1;
    if (!(s > 0) || ranks < 1 || ranks > (1ull << 32)) return -1;
    // cdf[r-1] = P(rank <= r), P(rank = r) proportional to r^-s
    std::vector<double> cdf(ranks);
    double acc = 0;
    for (uint64_t r = 1; r <= ranks; r++) {
      acc += pow((double)r, -s);
      cdf[r - 1] = acc;
    }
    for (auto &c : cdf) c /= acc;
    cdf[ranks - 1] = 1.0;
    const uint32_t key_salt = (uint32_t)(mix64(base ^ 0xA5A5A5A5A5A5A5A5ull) >> 32);
    parallel_rows(n, threads, [&](uint64_t lo, uint64_t hi) {
      for (uint64_t i = lo; i < hi; i++) {
        const double u = unit_open(draw(base, i, 0));
        const uint64_t r = (uint64_t)(std::lower_bound(cdf.begin(), cdf.end(), u) - cdf.begin()) + 1;
        // distinct ranks -> distinct keys: fmix32 and xor with a constant are bijections
        out[i] = fmix32((uint32_t)r ^ key_salt);
      }
    });
    return 0;
  }
  return -1;
}
