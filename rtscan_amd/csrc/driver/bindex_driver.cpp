// bindex_driver.cpp -- command-line query driver over the engine's C ABI, with the command
// line and the check output of the reference's GPU driver (bindex_cuda.cpp:1708-1995):
//
//   rts_bindex_driver [-f <data file>] [-b <columns>] [-p <query file>] [-n <rows>]
//                     [-s <seed>] [-m single|batch] [-e]
//
// Data: the data file (columns of n little-endian uint32, one after another), or seeded
// uniform columns when no file is given (the reference draws unseeded random data).  Its
// row count comes from -n, or from the file size.  Every column gets a BinDex
// (init_bindex_in_GPU).  Queries: the query file (default test/scan_cmd.txt, as there), one
// line per column.  Each query runs as one fused sieve + refine + AND launch
// (rts_bindex_query), or -m batch runs all of them in one batched launch (a query with a
// comma-separated list of targets on a line runs as XOR scans, see below); its bitmap is
// copied to page-locked host memory and checked against raw_scan of the same predicates
// with compare_bitmap, which prints "[CHECK]hit true/total" (and "[ERROR] ..." at the first
// wrong row).  -e encodes the columns first, as the reference's ENCODE build does
// (rtscan.cpp:1068-1074, 1016-1026): normalEncode to rank codes in [0, n), every target
// mapped through encodeQuery, the queries then run on the codes (eq and bt targets cannot
// be encoded, as there).  Exit status: 0 when every query matched, 1 on a mismatch, 2 on
// an error.
#include <hip/hip_runtime_api.h>

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

#include <chrono>
#include <string>
#include <vector>

#include "rts_batch.h"
#include "rts_bindex.h"
#include "rts_check.h"
#include "rts_query.h"
#include "rts_remap.h"

namespace {

int fail(const char *what) {
  fprintf(stderr, "[ERROR] %s: %s\n", what, rts_last_error());
  return 2;
}

double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

const char *op_name(int op) {
  static const char *names[] = {"eq", "lt", "gt", "le", "ge", "bt"};
  return op >= 0 && op <= 5 ? names[op] : "?";
}

}  // namespace

int main(int argc, char **argv) {
  std::string data_path, query_path = "test/scan_cmd.txt", mode = "single";
  bool encode = false;
  int ncols = 1;
  long long n = 0;
  unsigned long long seed = 1;
  int opt;
  while ((opt = getopt(argc, argv, "hf:b:p:n:s:m:e")) != -1) {
    switch (opt) {
      case 'h':
        printf("Usage: %s [-f <input-file>] [-b <columns>] [-p <scan-cmd-file>] [-n <rows>] [-s <seed>] "
               "[-m single|batch|ranges] [-e]\n", argv[0]);
        return 0;
      case 'f': data_path = optarg; break;
      case 'b': ncols = atoi(optarg); break;
      case 'p': query_path = optarg; break;
      case 'n': n = atoll(optarg); break;
      case 's': seed = strtoull(optarg, nullptr, 10); break;
      case 'm': mode = optarg; break;
      case 'e': encode = true; break;
      default:
        printf("Error: unknown option %c\n", (char)opt);
        return 2;
    }
  }
  if (ncols < 1 || ncols > RTS_BATCH_MAX_COLS || (mode != "single" && mode != "batch" && mode != "ranges")) {
    fprintf(stderr, "[ERROR] need 1 <= -b <= %d and -m single|batch|ranges\n", RTS_BATCH_MAX_COLS);
    return 2;
  }
  if (mode == "ranges" && ncols != 1) {  // rts_scan_ranges: many predicates over one column
    fprintf(stderr, "[ERROR] -m ranges runs one-column query files (-b 1)\n");
    return 2;
  }
  if (!data_path.empty() && n == 0) {
    struct stat st;
    if (stat(data_path.c_str(), &st) != 0) return fail("data file");
    n = (long long)(st.st_size / (4 * ncols));
  }
  if (n < RTS_VAREA_N || n > 0x7FFFFFFFll) {
    fprintf(stderr, "[ERROR] rows must be in [%d, 2^31): give -n or a data file\n", RTS_VAREA_N);
    return 2;
  }
  printf("N = %lld\n", n);

  // data (host), as the reference keeps it for its checker
  std::vector<std::vector<CODE>> data(ncols, std::vector<CODE>((size_t)n));
  std::vector<CODE *> cols(ncols);
  for (int c = 0; c < ncols; c++) cols[c] = data[c].data();
  if (!data_path.empty()) {
    if (rts_data_file_read(data_path.c_str(), ncols, n, cols.data()) != 0) return fail("rts_data_file_read");
  } else {
    printf("initing data by random (seed %llu)\n", seed);
    for (int c = 0; c < ncols; c++)
      if (rts_generate_column(0, seed, c, (uint64_t)n, 0.0, 0.0, cols[c], 0) != 0) return fail("generate");
  }

  if (encode) {  // rank codes in [0, n), in place (rtscan.cpp:1072)
    std::vector<POSTYPE *> sorted_pos(ncols, nullptr);
    std::vector<CODE *> sorted_data(ncols, nullptr);
    auto te = std::chrono::steady_clock::now();
    if (!normalEncode(cols.data(), ncols, 0, (CODE)n, (int)n, sorted_pos.data(), sorted_data.data()))
      return fail("normalEncode");
    printf("[Time] Uniform Encoding: %.3f ms\n", ms_since(te));
    for (int c = 0; c < ncols; c++) {
      printf("[+] column %d: %zu distinct values\n", c, encodeMap[c].size());
      free(sorted_pos[c]);
      free(sorted_data[c]);
    }
  }

  // indexes (bindex_cuda.cpp:1763-1780)
  std::vector<BinDex> bx(ncols);
  size_t free0 = 0, total = 0, free1 = 0, base_data = 0;
  (void)hipMemGetInfo(&free0, &total);
  auto t0 = std::chrono::steady_clock::now();
  for (int c = 0; c < ncols; c++) {
    printf("Build the bindex structure %d...\n", c);
    size_t before = base_data;
    init_bindex_in_GPU(&bx[c], cols[c], (POSTYPE)n, base_data);
    if (base_data == before) return fail("init_bindex_in_GPU");
  }
  printf("[Time] Index Build Time: %.3f ms\n", ms_since(t0));
  (void)hipMemGetInfo(&free1, &total);
  printf("[Mem] Device memory used(MB): %.1f\n", (double)(free0 - free1 - base_data) / (1 << 20));

  rts_query_file *qf = nullptr;
  int bad_line = 0;
  if (rts_query_file_load(query_path.c_str(), ncols, &qf, &bad_line) != 0) {
    fprintf(stderr, "[ERROR] query file %s: line %d\n", query_path.c_str(), bad_line);
    return 2;
  }
  const int nq = rts_query_file_count(qf);
  printf("BinDex scan... (%d queries, %s)\n", nq, mode.c_str());

  // target i of (query, column), mapped through encodeQuery with -e (rtscan.cpp:1022)
  auto target = [&](int q, int c, int which, int i = 0) -> CODE {
    const rts_column_predicate *p = rts_query_file_get(qf, q, c);
    const CODE t = which == 1 ? p->t1[i] : (p->t2 ? p->t2[i] : 0);
    if (!encode || which != 1) return t;
    std::string cmd = op_name(p->op);
    return encodeQuery(c, t, cmd);
  };
  // A line with several comma-separated targets selects, on that column, the rows hit by an
  // odd number of them: the reference's checker XORs each target's hits into the column's
  // bitmap (raw_scan_entry, bindex.cpp:1894-1930), the ground truth followed here.  Such
  // queries run as one XOR scan per target and an AND + popcount over the columns.
  std::vector<char> has_list(nq, 0);
  for (int q = 0; q < nq; q++)
    for (int c = 0; c < ncols; c++)
      if (rts_query_file_get(qf, q, c)->ntargets != 1) has_list[q] = 1;
  if (encode)
    for (int q = 0; q < nq; q++)
      for (int c = 0; c < ncols; c++) {
        const int op = rts_query_file_get(qf, q, c)->op;
        if (op == EQ || op == BT) {
          fprintf(stderr, "[ERROR] query %d column %d: eq and bt targets cannot be encoded\n", q, c);
          return 2;
        }
      }

  // closed ranges of every query
  std::vector<rts_query_spec> specs(nq);
  for (int q = 0; q < nq; q++) {
    rts_query_spec &sp = specs[q];
    memset(&sp, 0, sizeof(sp));
    sp.ncols = ncols;
    for (int c = 0; c < ncols; c++) {
      const rts_column_predicate *p = rts_query_file_get(qf, q, c);
      sp.column[c] = c;
      if (!rts_predicate_interval((enum OPERATOR)p->op, target(q, c, 1), target(q, c, 2), &sp.lo[c], &sp.hi[c])) {
        sp.lo[c] = 1;  // empty: lo > hi
        sp.hi[c] = 0;
      }
      sp.mode[c] = RTS_EVAL_BINDEX;
    }
  }

  const size_t words = (size_t)(n + 31) / 32, stride = (words + 3) / 4 * 4;
  BITS *dev_out = nullptr, *h_result = nullptr;
  unsigned long long *dev_count = nullptr;
  const bool many = mode == "batch" || mode == "ranges";  // all queries in one call
  const int nout = many ? (nq > 0 ? nq : 1) : 1;
  if (rts_malloc((void **)&dev_out, stride * 4 * nout) != hipSuccess ||
      rts_malloc((void **)&dev_count, 8 * nout) != hipSuccess ||
      rts_host_alloc((void **)&h_result, stride * 4) != hipSuccess)
    return fail("allocation");
  hipStream_t stream = nullptr;
  if (rts_stream_create(&stream) != hipSuccess) return fail("stream");

  // per-column bitmaps of the target-list queries
  std::vector<BITS *> colbm_dev(ncols, nullptr);
  bool any_list = false;
  for (int q = 0; q < nq; q++) any_list = any_list || has_list[q];
  for (int k = 0; any_list && k < ncols; k++)
    if (rts_malloc((void **)&colbm_dev[k], stride * 4) != hipSuccess) return fail("allocation");
  auto list_query = [&](int q, BITS *out, unsigned long long *count) -> hipError_t {
    for (int c = 0; c < ncols; c++) {
      const rts_column_predicate *p = rts_query_file_get(qf, q, c);
      // cleared on the query's stream (a hipMemset on the null stream is not ordered
      // before work on this non-blocking stream)
      hipError_t e = hipMemsetAsync(colbm_dev[c], 0, words * 4, stream);
      for (int i = 0; e == hipSuccess && i < p->ntargets; i++) {
        CODE lo, hi;
        if (rts_predicate_interval((enum OPERATOR)p->op, target(q, c, 1, i), target(q, c, 2, i), &lo, &hi))
          e = rts_scan_range_xor(bx[c].rawDataInGPU, (uint64_t)n, lo, hi, colbm_dev[c], stream);
      }
      if (e != hipSuccess) return e;
    }
    return rts_bitmap_and_popcount(colbm_dev.data(), ncols, (uint64_t)n, out, count, stream);
  };

  rts_batch_plan *plan = nullptr;
  double batch_ms = 0;
  std::vector<int> slot(nq, -1);  // the query's index in the batched plan
  std::vector<rts_query_spec> plan_specs;
  for (int q = 0; q < nq; q++)
    if (!has_list[q]) {
      slot[q] = (int)plan_specs.size();
      plan_specs.push_back(specs[q]);
    }
  if (mode == "batch" && !plan_specs.empty()) {
    std::vector<const CODE *> dcols(ncols);
    std::vector<const BinDex *> idx(ncols);
    for (int c = 0; c < ncols; c++) {
      dcols[c] = bx[c].rawDataInGPU;
      idx[c] = &bx[c];
    }
    const int np = (int)plan_specs.size();
    if (rts_batch_plan_create(&plan, dcols.data(), idx.data(), ncols, n, np, nullptr, stream) != hipSuccess ||
        rts_batch_plan_set_queries(plan, plan_specs.data(), np, stream) != hipSuccess)
      return fail("batch plan");
    auto tb = std::chrono::steady_clock::now();
    if (rts_batch_plan_launch(plan, dev_out, (int64_t)stride, dev_count, stream) != hipSuccess ||
        rts_stream_synchronize(stream) != hipSuccess)
      return fail("batch launch");
    batch_ms = ms_since(tb);
    printf("[Time] batched launch of %d queries: %.3f ms\n", np, batch_ms);
  }
  if (mode == "ranges" && !plan_specs.empty()) {
    // every query's range on the one column, the column read once per 64 of them
    const int np = (int)plan_specs.size();
    std::vector<CODE> lo(np), hi(np);
    for (int i = 0; i < np; i++) {
      lo[i] = plan_specs[i].lo[0];
      hi[i] = plan_specs[i].hi[0];
    }
    auto tb = std::chrono::steady_clock::now();
    if (rts_scan_ranges(bx[0].rawDataInGPU, (uint64_t)n, lo.data(), hi.data(), np, dev_out, (uint64_t)stride,
                        dev_count, stream) != hipSuccess ||
        rts_stream_synchronize(stream) != hipSuccess)
      return fail("rts_scan_ranges");
    batch_ms = ms_since(tb);
    printf("[Time] rts_scan_ranges of %d predicates: %.3f ms\n", np, batch_ms);
  }

  std::vector<BITS> check(words), colbm(words);
  int mismatches = 0;
  for (int q = 0; q < nq; q++) {
    for (int c = 0; c < ncols; c++) {
      const rts_column_predicate *p = rts_query_file_get(qf, q, c);
      std::string line = op_name(p->op), t1s, t2s;
      for (int i = 0; i < p->ntargets; i++) {
        t1s += (i ? "," : " ") + std::to_string(p->t1[i]);
        if (p->op == BT) t2s += (i ? "," : " ") + std::to_string(p->t2[i]);
      }
      if (encode && p->op != BT) {
        t2s = " [ENCODE]";
        for (int i = 0; i < p->ntargets; i++) t2s += (i ? "," : " ") + std::to_string(target(q, c, 1, i));
      }
      printf("%s%s%s\n", line.c_str(), t1s.c_str(), t2s.c_str());
    }
    BITS *res = dev_out;
    double scan_ms = 0;
    if (has_list[q]) {
      res = many ? dev_out + (size_t)(plan_specs.size() % nout) * stride : dev_out;
      auto ts = std::chrono::steady_clock::now();
      if (list_query(q, res, dev_count) != hipSuccess || rts_stream_synchronize(stream) != hipSuccess)
        return fail("target-list query");
      scan_ms = ms_since(ts);
    } else if (mode == "single") {
      std::vector<const BinDex *> idx(ncols);
      for (int c = 0; c < ncols; c++) idx[c] = &bx[c];
      auto ts = std::chrono::steady_clock::now();
      if (rts_bindex_query(idx.data(), specs[q].lo, specs[q].hi, ncols, dev_out, dev_count, stream) != hipSuccess ||
          rts_stream_synchronize(stream) != hipSuccess)
        return fail("rts_bindex_query");
      scan_ms = ms_since(ts);
    } else {
      res = dev_out + (size_t)slot[q] * stride;
    }
    auto tt = std::chrono::steady_clock::now();
    if (rts_memcpy_dtoh_async(h_result, res, words * 4, stream) != hipSuccess ||
        rts_stream_synchronize(stream) != hipSuccess)
      return fail("result transfer");
    const double xfer_ms = ms_since(tt);
    if (mode == "single" || has_list[q])
      printf("[Time] query %d: scan %.3f ms, transfer %.3f ms\n", q, scan_ms, xfer_ms);

    // ground truth: raw_scan of every column (bindex.cpp:1851-1892, the bt swap of
    // raw_scan_entry), ANDed across columns
    printf("[CHECK]check final result.\n");
    for (int c = 0; c < ncols; c++) {
      const rts_column_predicate *p = rts_query_file_get(qf, q, c);
      std::fill(colbm.begin(), colbm.end(), 0u);
      for (int i = 0; i < p->ntargets; i++) {  // raw_scan_entry: XOR of every target's hits
        CODE a = target(q, c, 1, i), b = target(q, c, 2, i);
        if (p->op == BT && a > b) std::swap(a, b);
        raw_scan(&bx[c], colbm.data(), a, b, (enum OPERATOR)p->op, cols[c], nullptr);
      }
      for (size_t w = 0; w < words; w++) check[w] = c ? check[w] & colbm[w] : colbm[w];
    }
    rts_compare_result r{};
    compare_bitmap(check.data(), h_result, (int)n, cols.data(), ncols);
    // the same comparison, counted, for the exit status
    BITS *dcheck = nullptr;
    if (rts_malloc((void **)&dcheck, words * 4) != hipSuccess ||
        rts_memcpy_htod(dcheck, check.data(), words * 4) != hipSuccess ||
        rts_bitmap_compare(dcheck, res, n, &r, stream) != hipSuccess)
      return fail("rts_bitmap_compare");
    (void)rts_free(dcheck);
    if (r.first_mismatch >= 0) mismatches++;
    printf("[CHECK]check final result done.\n\n");
    for (size_t w = 0; w < words; w++)
      if (h_result[w] != 0) {
        printf("[CHECK] not all 0!\n");
        break;
      }
  }
  printf("[CHECK] %d of %d queries matched raw_scan\n", nq - mismatches, nq);

  if (plan) rts_batch_plan_free(plan);
  for (BITS *b : colbm_dev) (void)rts_free(b);
  (void)rts_stream_destroy(stream);
  (void)rts_host_free(h_result);
  (void)rts_free(dev_out);
  (void)rts_free(dev_count);
  rts_query_file_free(qf);
  for (int c = 0; c < ncols; c++) rts_bindex_free(&bx[c]);
  return mismatches ? 1 : 0;
}
