/*
 * rts_query.h -- the reference's query-file and data-file formats (host side).
 *
 * Query file: one line per column, ncols lines per query, `<op> <t1>[,...] [<t2>[,...]]`
 * with op in lt le gt ge eq bt; the first `exit` line ends the file (bindex.cpp:2168-2189,
 * bindex_cuda.cpp:1834-1858).  Several comma-separated targets on one line are applied
 * one after another; the reference's checker XORs their hits (raw_scan_entry,
 * bindex.cpp:1894-1930).  Data file: ncols columns of n little-endian uint32, column after
 * column (getDataFromFile, bindex_cuda.cpp:1269-1286).
 */
#ifndef RTS_QUERY_H_
#define RTS_QUERY_H_

#include "rtscan_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
  RTS_QF_EINVAL = -1,     /* bad argument */
  RTS_QF_EIO = -2,        /* file cannot be opened / read / written */
  RTS_QF_ESYNTAX = -3,    /* malformed line (unknown op, bad or out-of-range target) */
  RTS_QF_EINCOMPLETE = -4 /* the last query has fewer than ncols lines */
};

/* One column's predicate of one query. */
typedef struct rts_column_predicate {
  int op;          /* enum OPERATOR */
  int ntargets;    /* number of comma-separated targets */
  const CODE *t1;  /* ntargets values */
  const CODE *t2;  /* ntargets values for bt, else NULL */
} rts_column_predicate;

typedef struct rts_query_file rts_query_file;

/* Parse query-file text / a query file.  On a syntax error *error_line is the 1-based line. */
int rts_query_file_parse(const char *text, size_t len, int ncols, rts_query_file **out,
                         int *error_line);
int rts_query_file_load(const char *path, int ncols, rts_query_file **out, int *error_line);
int rts_query_file_count(const rts_query_file *qf);
int rts_query_file_columns(const rts_query_file *qf);
/* The predicate of (query, column); valid until rts_query_file_free. */
const rts_column_predicate *rts_query_file_get(const rts_query_file *qf, int query, int column);
void rts_query_file_free(rts_query_file *qf);

/* Read / write a data file of ncols columns x n rows into / from host arrays. */
int rts_data_file_read(const char *path, int ncols, int64_t n, CODE *const *columns);
int rts_data_file_write(const char *path, int ncols, int64_t n, const CODE *const *columns);

#ifdef __cplusplus
}
#endif
#endif /* RTS_QUERY_H_ */
