// capi.hip -- extern "C" entry points: the reference-named cuda_and.cu / bindex.h
// wrappers and the additive rts_* scan, memory and timing entry points.
#include <stdio.h>
#include <string.h>

#include <atomic>
#include <map>
#include <mutex>
#include <thread>
#include <tuple>
#include <string>
#include <vector>

#include "common.h"
#include "rts_bindex.h"

namespace rts {

hipError_t scan_conjunction(const CODE *const *cols, const CODE *lo, const CODE *hi, int ncols,
                            uint64_t n, BITS *out, unsigned long long *count, int mode,
                            hipStream_t s, uint32_t *partials = nullptr);
hipError_t bitmap_and_popcount(const BITS *const *bms, int nb, uint64_t n, BITS *out,
                               unsigned long long *count, hipStream_t s);
hipError_t scan_ranges(const CODE *col, uint64_t n, const CODE *lo, const CODE *hi, int np, BITS *out,
                       uint64_t stride, unsigned long long *counts, hipStream_t s);
hipError_t bitop(int op, BITS *dst, const BITS *a, const BITS *b, uint64_t words, uint32_t value,
                 hipStream_t s);
hipError_t flip_rows(BITS *bm, uint64_t rows, hipStream_t s);

enum { kCopy = 0, kNot = 1, kAndNot = 2, kAnd = 3, kOr = 4, kFill = 5, kXorWith = 6 };

static thread_local std::string g_last_error;

hipError_t set_error(hipError_t e, const char *what) {
  if (e != hipSuccess) {
    char buf[256];
    snprintf(buf, sizeof(buf), "%s: %s (%d)", what, hipGetErrorString(e), (int)e);
    g_last_error = buf;
  }
  return e;
}

// Per-device caches of this file are fixed arrays (no allocation after load).
constexpr int kMaxDevices = 64;

int device_cu_count() {
  static std::atomic<int> cache[kMaxDevices];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 256;
  if (dev < 0 || dev >= kMaxDevices) return 256;
  int cus = cache[dev].load(std::memory_order_relaxed);
  if (cus == 0) {
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
      cus = 256;
    cache[dev].store(cus, std::memory_order_relaxed);
  }
  return cus;
}

// Scratch buffers of the calling host thread, keyed by (kind, device, stream).  Two threads
// enqueueing work on the same stream (the legacy null stream, shared by all threads) must
// not share scratch, since their launches can interleave between a kernel and the one that
// reads its partials; so the table is thread-local.  A thread keeps at most
// kScratchPerThread buffers (the oldest is freed first) and frees its buffers when it
// exits, so thread pools and stream churn do not grow device memory without bound.
// hipFree waits for the device's outstanding work, so a buffer is never freed under a
// kernel still using it.  The main thread's buffers are left to process teardown.
namespace {

constexpr size_t kScratchPerThread = 16;

struct ScratchEntry {
  int kind, dev;
  uintptr_t stream;
  void *dptr, *hptr;
};

const std::thread::id g_main_thread = std::this_thread::get_id();

struct ThreadScratch {
  std::vector<ScratchEntry> entries;  // oldest first
  static void release(ScratchEntry &e) {
    int cur = 0;
    const bool switched = hipGetDevice(&cur) == hipSuccess && cur != e.dev && hipSetDevice(e.dev) == hipSuccess;
    if (e.dptr) (void)hipFree(e.dptr);
    if (e.hptr) (void)hipHostFree(e.hptr);
    if (switched) (void)hipSetDevice(cur);
  }
  ~ThreadScratch() {
    if (std::this_thread::get_id() == g_main_thread) return;
    for (ScratchEntry &e : entries) release(e);
  }
};

thread_local ThreadScratch t_scratch;

}  // namespace

hipError_t thread_scratch(int kind, hipStream_t stream, size_t dev_bytes, size_t host_bytes, void **dptr,
                          void **hptr, bool *created) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (created) *created = false;
  std::vector<ScratchEntry> &tab = t_scratch.entries;
  for (const ScratchEntry &x : tab) {
    if (x.kind == kind && x.dev == dev && x.stream == (uintptr_t)stream) {
      *dptr = x.dptr;
      if (hptr) *hptr = x.hptr;
      return hipSuccess;
    }
  }
  if (created) *created = true;
  ScratchEntry x{kind, dev, (uintptr_t)stream, nullptr, nullptr};
  if (dev_bytes) e = hipMalloc(&x.dptr, dev_bytes);
  if (e == hipSuccess && host_bytes) e = hipHostMalloc(&x.hptr, host_bytes, 0);
  if (e != hipSuccess) {
    ThreadScratch::release(x);
    return e;
  }
  if (tab.size() >= kScratchPerThread) {
    ThreadScratch::release(tab.front());
    tab.erase(tab.begin());
  }
  tab.push_back(x);
  *dptr = x.dptr;
  if (hptr) *hptr = x.hptr;
  return hipSuccess;
}

size_t thread_scratch_entries() { return t_scratch.entries.size(); }

CountSink stream_count_sink(unsigned long long *out, hipStream_t stream) {
  void *p = nullptr;
  bool created = false;
  if (thread_scratch(kScratchCountAcc, stream, kCountAccBytes, 0, &p, nullptr, &created) != hipSuccess)
    return CountSink{out, nullptr, nullptr};
  // the counting kernels leave the accumulators at zero; a new set starts there too
  if (created && hipMemsetAsync(p, 0, kCountAccBytes, stream) != hipSuccess) return CountSink{out, nullptr, nullptr};
  return CountSink{out, nullptr, static_cast<unsigned long long *>(p)};
}

// One workgroup of 1024 threads; each thread keeps 8 independent loads in flight so the
// sum over tens of thousands of partials costs a few memory latencies, not one per slot.
constexpr int kFinalizeThreads = 1024;

__global__ __launch_bounds__(kFinalizeThreads) void k_count_finalize(const uint32_t *partials,
                                                                     unsigned nblocks,
                                                                     unsigned long long *out) {
  __shared__ unsigned long long part[kFinalizeThreads / kWave];
  unsigned long long s = 0;
  unsigned i = threadIdx.x;
  for (; i + 7 * kFinalizeThreads < nblocks; i += 8 * kFinalizeThreads) {
    uint32_t v[8];
#pragma unroll
    for (int k = 0; k < 8; k++) v[k] = partials[i + k * kFinalizeThreads];
#pragma unroll
    for (int k = 0; k < 8; k++) s += v[k];
  }
  for (; i < nblocks; i += kFinalizeThreads) s += partials[i];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, kWave);
  if ((threadIdx.x & (kWave - 1)) == 0) part[threadIdx.x / kWave] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int k = 0; k < kFinalizeThreads / kWave; k++) t += part[k];
    *out = t;
  }
}

hipError_t launch_count_finalize(const CountSink &sink, unsigned nblocks, hipStream_t stream) {
  if (!sink.out || sink.acc) return hipSuccess;
  hipLaunchKernelGGL(k_count_finalize, dim3(1), dim3(kFinalizeThreads), 0, stream, sink.partials,
                     nblocks, sink.out);
  return hipGetLastError();
}

// Reference wrappers: legacy null stream, synchronous (cuda_and.cu ends every wrapper with
// cudaDeviceSynchronize); launch errors are returned, never fatal.
static hipError_t finish_sync(hipError_t e, const char *what) {
  if (e != hipSuccess) return set_error(e, what);
  return set_error(hipStreamSynchronize(ref_stream()), what);
}

}  // namespace rts

using namespace rts;

extern "C" {

// ---------------------------------------------------------------- cuda_and.cu surface

hipError_t GPUbitAndWithCuda(BITS *a, BITS *b, unsigned int n) {
  return finish_sync(bitop(kAnd, a, b, nullptr, ref_words((int)n), 0, ref_stream()), "GPUbitAndWithCuda");
}

hipError_t GPUbitOrWithCuda(BITS *a, BITS *b, unsigned int n) {
  return finish_sync(bitop(kOr, a, b, nullptr, ref_words((int)n), 0, ref_stream()), "GPUbitOrWithCuda");
}

hipError_t GPUbitCopyWithCuda(BITS *a, BITS *b, unsigned int n) {
  return finish_sync(bitop(kCopy, a, b, nullptr, n, 0, ref_stream()), "GPUbitCopyWithCuda");
}

hipError_t GPUbitCopyNegationWithCuda(BITS *a, BITS *b, unsigned int n) {
  return finish_sync(bitop(kNot, a, b, nullptr, n, 0, ref_stream()), "GPUbitCopyNegationWithCuda");
}

// result = b & ~a (bitmap_copySIMDKernel, cuda_and.cu:54-61)
hipError_t GPUbitCopySIMDWithCuda(BITS *result, BITS *a, BITS *b, unsigned int bitnum) {
  return finish_sync(bitop(kAndNot, result, a, b, bitnum, 0, ref_stream()), "GPUbitCopySIMDWithCuda");
}

hipError_t GPUbitMemsetWithCuda(BITS *a, BITS value, unsigned int n) {
  return finish_sync(bitop(kFill, a, nullptr, nullptr, ref_words((int)n), value, ref_stream()),
                     "GPUbitMemsetWithCuda");
}

hipError_t CPUbitAndWithCuda(BITS *host_a, BITS *host_b, unsigned int size) {
  const size_t words = ref_words((int)size);
  const size_t bytes = words * sizeof(BITS);
  BITS *da = nullptr, *db = nullptr;
  hipError_t e = hipMalloc((void **)&da, bytes);
  if (e == hipSuccess) e = hipMalloc((void **)&db, bytes);
  if (e == hipSuccess) e = hipMemcpy(da, host_a, bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(db, host_b, bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = bitop(kAnd, da, db, nullptr, words, 0, ref_stream());
  if (e == hipSuccess) e = hipStreamSynchronize(ref_stream());
  if (e == hipSuccess) e = hipMemcpy(host_a, da, bytes, hipMemcpyDeviceToHost);
  if (da) (void)hipFree(da);
  if (db) (void)hipFree(db);
  return set_error(e, "CPUbitAndWithCuda");
}

hipError_t refineWithCuda(CODE *bitmap, unsigned int size) {
  const size_t bytes = (size_t)size * sizeof(BITS);
  if (size == 0) return hipSuccess;
  BITS *d = nullptr;
  hipError_t e = hipMalloc((void **)&d, bytes);
  if (e == hipSuccess) e = hipMemcpy(d, bitmap, bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = flip_rows(d, size, ref_stream());
  if (e == hipSuccess) e = hipStreamSynchronize(ref_stream());
  if (e == hipSuccess) e = hipMemcpy(bitmap, d, bytes, hipMemcpyDeviceToHost);
  if (d) (void)hipFree(d);
  return set_error(e, "refineWithCuda");
}

// ---------------------------------------------------------------------- predicates

int rts_predicate_interval(enum OPERATOR op, CODE t1, CODE t2, CODE *lo, CODE *hi) {
  CODE a = 0, b = 0xFFFFFFFFu;
  switch (op) {
    case LT:
      if (t1 == 0) return 0;
      b = t1 - 1;
      break;
    case LE: b = t1; break;
    case GT:
      if (t1 == 0xFFFFFFFFu) return 0;
      a = t1 + 1;
      break;
    case GE: a = t1; break;
    case EQ: a = b = t1; break;
    case BT: {
      CODE x = t1 < t2 ? t1 : t2, y = t1 < t2 ? t2 : t1;
      if ((uint64_t)y <= (uint64_t)x + 1) return 0;
      a = x + 1;
      b = y - 1;
      break;
    }
    default: return 0;
  }
  *lo = a;
  *hi = b;
  return 1;
}

// ------------------------------------------------------------------------ scans

hipError_t rts_scan_range(const CODE *column, uint64_t n, CODE lo, CODE hi, BITS *out,
                          unsigned long long *count, hipStream_t stream) {
  const CODE *cols[1] = {column};
  return set_error(scan_conjunction(cols, &lo, &hi, 1, n, out, count, kStoreSet, stream),
                   "rts_scan_range");
}

hipError_t rts_scan_ranges(const CODE *column, uint64_t n, const CODE *lo, const CODE *hi, int npred,
                           BITS *out, uint64_t out_stride_words, unsigned long long *counts,
                           hipStream_t stream) {
  return set_error(scan_ranges(column, n, lo, hi, npred, out, out_stride_words, counts, stream),
                   "rts_scan_ranges");
}

hipError_t rts_scan_range_xor(const CODE *column, uint64_t n, CODE lo, CODE hi, BITS *out,
                              hipStream_t stream) {
  const CODE *cols[1] = {column};
  return set_error(scan_conjunction(cols, &lo, &hi, 1, n, out, nullptr, kStoreXor, stream),
                   "rts_scan_range_xor");
}

hipError_t rts_conjunction_scan(const CODE *const *columns, const CODE *lo, const CODE *hi,
                                int ncols, uint64_t n, BITS *out, unsigned long long *count,
                                hipStream_t stream) {
  return set_error(scan_conjunction(columns, lo, hi, ncols, n, out, count, kStoreSet, stream),
                   "rts_conjunction_scan");
}

}  // extern "C"

// The allocation-free conjunction's scratch: the count partials of one counted launch
// (kMaxCountBlocks 32-bit slots, 256 KB), a file-scope device array, one copy per GPU held
// until the process ends.  Host code reaches it only through the address
// hipGetSymbolAddress gives for the current device (the array's name is a host-side
// stand-in, not a device address), fetched once per device on the first call there.
namespace {

__device__ __align__(256) uint32_t scratch[rts::kMaxCountBlocks];

std::atomic<uint32_t *> g_scratch_ptr[rts::kMaxDevices];
std::mutex g_scratch_mu;

hipError_t noalloc_scratch(uint32_t **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= rts::kMaxDevices) return hipErrorInvalidDevice;
  uint32_t *p = g_scratch_ptr[dev].load(std::memory_order_acquire);
  if (!p) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    p = g_scratch_ptr[dev].load(std::memory_order_acquire);
    if (!p) {
      void *q = nullptr;
      e = hipGetSymbolAddress(&q, HIP_SYMBOL(scratch));  // with `dev` current
      if (e != hipSuccess) return e;
      if (!q) return hipErrorInvalidSymbol;
      p = static_cast<uint32_t *>(q);
      g_scratch_ptr[dev].store(p, std::memory_order_release);
    }
  }
  *out = p;
  return hipSuccess;
}

}  // namespace

extern "C" {

hipError_t rts_conjunction_scan_noalloc(const CODE *const *columns, const CODE *lo, const CODE *hi,
                                        int ncols, uint64_t n, BITS *out, unsigned long long *count,
                                        hipStream_t stream) {
  if (n > RTS_NOALLOC_MAX_ROWS)
    return set_error(hipErrorInvalidValue, "rts_conjunction_scan_noalloc: more than RTS_NOALLOC_MAX_ROWS (2^40) rows");
  uint32_t *partials = nullptr;
  if (count) {
    // every workgroup of the counted launch writes its own slot before the finalize reads
    // it: nothing to clear between calls
    hipError_t e = noalloc_scratch(&partials);
    if (e != hipSuccess) return set_error(e, "rts_conjunction_scan_noalloc: scratch address");
  }
  return set_error(scan_conjunction(columns, lo, hi, ncols, n, out, count, kStoreSet, stream, partials),
                   "rts_conjunction_scan_noalloc");
}

hipError_t rts_bitmap_and_popcount(const BITS *const *bitmaps, int nbitmaps, uint64_t n,
                                   BITS *out, unsigned long long *count, hipStream_t stream) {
  return set_error(bitmap_and_popcount(bitmaps, nbitmaps, n, out, count, stream),
                   "rts_bitmap_and_popcount");
}

hipError_t rts_bitmap_popcount(const BITS *bitmap, uint64_t n, unsigned long long *count,
                               hipStream_t stream) {
  if (!count) return set_error(hipErrorInvalidValue, "rts_bitmap_popcount");
  return set_error(bitmap_and_popcount(&bitmap, 1, n, nullptr, count, stream),
                   "rts_bitmap_popcount");
}

// raw_scan (bindex.cpp:1851-1892) evaluated on the GPU: host column, host bitmap, every
// hit XOR-toggled into the bitmap.  With compare_bitmap given, the reference only reads
// (and its check is a no-op, bindex.cpp:1880-1886), so nothing is written.
void raw_scan(struct BinDex *bindex, BITS *bitmap, CODE t1, CODE t2, enum OPERATOR op,
              CODE *raw_data, BITS *compare_bitmap) {
  if (!bindex || !bitmap || !raw_data || compare_bitmap) return;
  const int64_t n = bindex->length;
  if (n <= 0) return;
  if (op == BT && t1 >= t2) return;  // a < x < b with b <= a: no row (no swap here)
  CODE lo, hi;
  if (!rts_predicate_interval(op, t1, t2, &lo, &hi)) return;
  const size_t words = words_for_rows((uint64_t)n);
  CODE *dcol = nullptr;
  BITS *dbm = nullptr;
  hipError_t e = hipMalloc((void **)&dcol, (size_t)n * sizeof(CODE));
  if (e == hipSuccess) e = hipMalloc((void **)&dbm, words * sizeof(BITS));
  if (e == hipSuccess) e = hipMemcpy(dcol, raw_data, (size_t)n * sizeof(CODE), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dbm, bitmap, words * sizeof(BITS), hipMemcpyHostToDevice);
  const CODE *cols[1] = {dcol};
  if (e == hipSuccess) e = scan_conjunction(cols, &lo, &hi, 1, (uint64_t)n, dbm, nullptr, kStoreXor, ref_stream());
  if (e == hipSuccess) e = hipStreamSynchronize(ref_stream());
  if (e == hipSuccess) e = hipMemcpy(bitmap, dbm, words * sizeof(BITS), hipMemcpyDeviceToHost);
  if (dcol) (void)hipFree(dcol);
  if (dbm) (void)hipFree(dbm);
  (void)set_error(e, "raw_scan");
}

// ------------------------------------------------------------------- helpers

hipError_t rts_malloc(void **ptr, size_t bytes) { return set_error(hipMalloc(ptr, bytes), "rts_malloc"); }
hipError_t rts_free(void *ptr) { return set_error(hipFree(ptr), "rts_free"); }
hipError_t rts_memcpy_htod(void *dst, const void *src, size_t bytes) {
  return set_error(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "rts_memcpy_htod");
}
hipError_t rts_memcpy_dtoh(void *dst, const void *src, size_t bytes) {
  return set_error(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "rts_memcpy_dtoh");
}
hipError_t rts_host_alloc(void **ptr, size_t bytes) {
  return set_error(hipHostMalloc(ptr, bytes, hipHostMallocDefault), "rts_host_alloc");
}
hipError_t rts_host_free(void *ptr) { return set_error(hipHostFree(ptr), "rts_host_free"); }
hipError_t rts_memcpy_dtoh_async(void *dst, const void *src, size_t bytes, hipStream_t stream) {
  return set_error(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream), "rts_memcpy_dtoh_async");
}
hipError_t rts_memcpy_htod_async(void *dst, const void *src, size_t bytes, hipStream_t stream) {
  return set_error(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream), "rts_memcpy_htod_async");
}
hipError_t rts_memcpy_dtod(void *dst, const void *src, size_t bytes) {
  hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice);
  if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
  return set_error(e, "rts_memcpy_dtod");
}
hipError_t rts_memset(void *dst, int v, size_t bytes) {
  // returns after the memory is set, so work the caller then enqueues on any stream
  // (including non-blocking ones) sees it
  hipError_t e = hipMemset(dst, v, bytes);
  if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
  return set_error(e, "rts_memset");
}
hipError_t rts_device_synchronize(void) {
  return set_error(hipDeviceSynchronize(), "rts_device_synchronize");
}
hipError_t rts_stream_create(hipStream_t *s) {
  return set_error(hipStreamCreateWithFlags(s, hipStreamNonBlocking), "rts_stream_create");
}
hipError_t rts_stream_destroy(hipStream_t s) {
  return set_error(hipStreamDestroy(s), "rts_stream_destroy");
}
hipError_t rts_stream_synchronize(hipStream_t s) {
  return set_error(hipStreamSynchronize(s), "rts_stream_synchronize");
}
hipError_t rts_event_create(hipEvent_t *e) { return set_error(hipEventCreate(e), "rts_event_create"); }
hipError_t rts_event_destroy(hipEvent_t e) { return set_error(hipEventDestroy(e), "rts_event_destroy"); }
hipError_t rts_event_record(hipEvent_t e, hipStream_t s) {
  return set_error(hipEventRecord(e, s), "rts_event_record");
}
hipError_t rts_event_synchronize(hipEvent_t e) {
  return set_error(hipEventSynchronize(e), "rts_event_synchronize");
}
hipError_t rts_event_elapsed_ms(float *ms, hipEvent_t a, hipEvent_t b) {
  return set_error(hipEventElapsedTime(ms, a, b), "rts_event_elapsed_ms");
}

int rts_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

hipError_t rts_device_info(char *name, size_t name_len, int *cu_count, size_t *total_mem) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return set_error(e, "rts_device_info");
  hipDeviceProp_t prop;
  e = hipGetDeviceProperties(&prop, dev);
  if (e != hipSuccess) return set_error(e, "rts_device_info");
  if (name && name_len) {
    strncpy(name, prop.gcnArchName, name_len - 1);
    name[name_len - 1] = 0;
  }
  if (cu_count) *cu_count = prop.multiProcessorCount;
  if (total_mem) *total_mem = prop.totalGlobalMem;
  return hipSuccess;
}

const char *rts_last_error(void) { return g_last_error.c_str(); }

int rts_thread_scratch_buffers(void) { return (int)thread_scratch_entries(); }

hipError_t rts_device_mem_info(size_t *free_bytes, size_t *total_bytes) {
  size_t f = 0, t = 0;
  hipError_t e = hipMemGetInfo(&f, &t);
  if (e != hipSuccess) return set_error(e, "rts_device_mem_info");
  if (free_bytes) *free_bytes = f;
  if (total_bytes) *total_bytes = t;
  return hipSuccess;
}

}  // extern "C"
