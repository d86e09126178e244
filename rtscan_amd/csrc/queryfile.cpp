// queryfile.cpp -- the reference's query-file and data-file formats.
//
// Query file (parse loop bindex.cpp:2168-2189, bindex_cuda.cpp:1834-1858): one line per
// column, `ncols` lines per query, each `<op> <t1>[,<t1>...] [<t2>[,<t2>...]]` with op in
// lt le gt ge eq bt, fields split on single spaces (stringSplit, bindex.cpp:238-245);
// targets parsed as doubles and cast to uint32 (get_target_numbers, bindex.cpp:1828-1849).
// The first line whose first field is `exit` ends the file; the rest is ignored.
//
// Data file (getDataFromFile, bindex_cuda.cpp:1269-1286): `ncols` columns of n
// little-endian uint32 values, column after column.
#include <errno.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <fstream>
#include <string>
#include <vector>

#include "rts_query.h"

struct rts_query_file {
  int ncols = 0;
  std::vector<rts_column_predicate> preds;  // nqueries * ncols
  std::vector<CODE> targets;                // t1 / t2 storage
  std::vector<std::string> lines;           // source line of every predicate
};

namespace {

std::vector<std::string> split_spaces(const std::string &s) {
  // stringSplit keeps empty fields between consecutive delimiters
  std::vector<std::string> out;
  size_t start = 0;
  while (true) {
    size_t p = s.find(' ', start);
    out.push_back(s.substr(start, p == std::string::npos ? std::string::npos : p - start));
    if (p == std::string::npos) break;
    start = p + 1;
  }
  return out;
}

// (CODE)stod(value): decimal or floating text; values outside [0, 2^32) are rejected here
// instead of the reference's undefined double-to-unsigned conversion.
bool parse_targets(const std::string &s, std::vector<CODE> &out) {
  out.clear();
  size_t start = 0;
  while (start <= s.size()) {
    size_t p = s.find(',', start);
    std::string tok = s.substr(start, p == std::string::npos ? std::string::npos : p - start);
    if (tok.empty()) {
      if (p == std::string::npos) break;  // getline drops a trailing empty field
      return false;
    }
    errno = 0;
    char *end = nullptr;
    const double v = strtod(tok.c_str(), &end);
    while (end && (*end == ' ' || *end == '\t' || *end == '\r')) end++;
    if (errno || !end || *end != '\0' || !(v >= 0.0) || !(v < 4294967296.0)) return false;
    out.push_back((CODE)(uint64_t)v);
    if (p == std::string::npos) break;
    start = p + 1;
  }
  return !out.empty();
}

int op_code(const std::string &s) {
  if (s == "eq") return EQ;
  if (s == "lt") return LT;
  if (s == "gt") return GT;
  if (s == "le") return LE;
  if (s == "ge") return GE;
  if (s == "bt") return BT;
  return -1;
}

}  // namespace

extern "C" {

int rts_query_file_parse(const char *text, size_t len, int ncols, rts_query_file **out,
                         int *error_line) {
  if (error_line) *error_line = 0;
  if (!out || ncols < 1 || (!text && len)) return RTS_QF_EINVAL;
  *out = nullptr;
  auto *qf = new rts_query_file();
  qf->ncols = ncols;
  struct Pending {
    int op;
    size_t t1_off, t1_n, t2_off, t2_n;
    std::string line;
  };
  std::vector<Pending> pending, all;
  size_t pos = 0;
  int lineno = 0;
  bool ended = false;
  while (pos < len && !ended) {
    size_t nl = pos;
    while (nl < len && text[nl] != '\n') nl++;
    std::string line(text + pos, nl - pos);
    pos = nl + 1;
    lineno++;
    if (!line.empty() && line.back() == '\r') line.pop_back();
    std::vector<std::string> f = split_spaces(line);
    if (f[0] == "exit") {
      ended = true;
      break;
    }
    const int op = op_code(f[0]);
    std::vector<CODE> t1, t2;
    if (op < 0 || f.size() < 2 || f.size() > 3 || !parse_targets(f[1], t1) ||
        (f.size() == 3 && !parse_targets(f[2], t2)) || (op == BT && t2.size() != t1.size()) ||
        (op != BT && !t2.empty())) {
      if (error_line) *error_line = lineno;
      delete qf;
      return RTS_QF_ESYNTAX;
    }
    Pending p{op, qf->targets.size(), t1.size(), 0, t2.size(), line};
    qf->targets.insert(qf->targets.end(), t1.begin(), t1.end());
    p.t2_off = qf->targets.size();
    qf->targets.insert(qf->targets.end(), t2.begin(), t2.end());
    pending.push_back(p);
    if ((int)pending.size() == ncols) {
      all.insert(all.end(), pending.begin(), pending.end());
      pending.clear();
    }
  }
  if (!pending.empty()) {  // a query cut short by `exit` or end of file
    if (error_line) *error_line = lineno;
    delete qf;
    return RTS_QF_EINCOMPLETE;
  }
  for (const Pending &p : all) {
    rts_column_predicate cp;
    cp.op = p.op;
    cp.ntargets = (int)p.t1_n;
    cp.t1 = qf->targets.data() + p.t1_off;
    cp.t2 = p.t2_n ? qf->targets.data() + p.t2_off : nullptr;
    qf->preds.push_back(cp);
    qf->lines.push_back(p.line);
  }
  *out = qf;
  return 0;
}

int rts_query_file_load(const char *path, int ncols, rts_query_file **out, int *error_line) {
  if (error_line) *error_line = 0;
  std::ifstream in(path ? path : "", std::ios::binary);
  if (!in) return RTS_QF_EIO;
  std::string text((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
  return rts_query_file_parse(text.data(), text.size(), ncols, out, error_line);
}

int rts_query_file_count(const rts_query_file *qf) {
  return qf ? (int)(qf->preds.size() / qf->ncols) : 0;
}

int rts_query_file_columns(const rts_query_file *qf) { return qf ? qf->ncols : 0; }

const rts_column_predicate *rts_query_file_get(const rts_query_file *qf, int query, int column) {
  if (!qf || query < 0 || column < 0 || column >= qf->ncols ||
      (size_t)query * qf->ncols + column >= qf->preds.size())
    return nullptr;
  return &qf->preds[(size_t)query * qf->ncols + column];
}

void rts_query_file_free(rts_query_file *qf) { delete qf; }

int rts_data_file_read(const char *path, int ncols, int64_t n, CODE *const *columns) {
  if (!path || ncols < 1 || n < 0 || !columns) return RTS_QF_EINVAL;
  FILE *fp = fopen(path, "rb");
  if (!fp) return RTS_QF_EIO;
  for (int c = 0; c < ncols; c++) {
    if (fread(columns[c], sizeof(CODE), (size_t)n, fp) != (size_t)n) {
      fclose(fp);
      return RTS_QF_EIO;
    }
  }
  fclose(fp);
  return 0;
}

int rts_data_file_write(const char *path, int ncols, int64_t n, const CODE *const *columns) {
  if (!path || ncols < 1 || n < 0 || !columns) return RTS_QF_EINVAL;
  FILE *fp = fopen(path, "wb");
  if (!fp) return RTS_QF_EIO;
  for (int c = 0; c < ncols; c++) {
    if (fwrite(columns[c], sizeof(CODE), (size_t)n, fp) != (size_t)n) {
      fclose(fp);
      return RTS_QF_EIO;
    }
  }
  return fclose(fp) == 0 ? 0 : RTS_QF_EIO;
}

}  // extern "C"
