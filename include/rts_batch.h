/*
 * rts_batch.h -- batched evaluation of many range queries in one persistent launch
 * (additive: the reference runs one query at a time, bindex_cuda.cpp:1827-1995).
 *
 * A query is an AND of closed ranges over 1..RTS_BATCH_MAX_COLS columns of a table whose
 * columns all have n rows.  Each query-column is evaluated from the column's BinDex (filter
 * vectors + boundary-bin refine lists) or directly from its values; for RTS_EVAL_AUTO the
 * planner's cost model chooses (the bytes per row each form reads, the direct form weighted
 * by the config's direct_weight; a column without an index is evaluated directly).  Query q's bitmap is written at out + q * out_stride_words
 * (MSB-first words, bits past row n cleared) and its popcount at counts[q] (device).
 *
 * The launch is one persistent kernel over (row tile, query group) work items, handed out by
 * per-XCD counters so that one tile's work items run on one XCD back to back, plus one small
 * kernel summing the per-(query, tile) popcounts.  Launches of one plan must not overlap
 * (use one stream per plan, or order them): they share the plan's counters and partials.  rts_batch_plan_graph_launch captures both in a hipGraph and
 * replays it; queries set with rts_batch_plan_set_queries live in device memory at a fixed
 * address, so a replay evaluates the current set.
 */
#ifndef RTS_BATCH_H_
#define RTS_BATCH_H_

#include "rts_bindex.h"

#ifdef __cplusplus
extern "C" {
#endif

#define RTS_BATCH_MAX_COLS 4

enum rts_eval_mode { RTS_EVAL_AUTO = 0, RTS_EVAL_BINDEX = 1, RTS_EVAL_DIRECT = 2 };

typedef struct rts_query_spec {
  int ncols;                      /* 1..RTS_BATCH_MAX_COLS */
  int column[RTS_BATCH_MAX_COLS]; /* table column of each range */
  CODE lo[RTS_BATCH_MAX_COLS];    /* closed range lo <= x <= hi; lo > hi is empty */
  CODE hi[RTS_BATCH_MAX_COLS];
  int mode[RTS_BATCH_MAX_COLS];   /* enum rts_eval_mode */
} rts_query_spec;

/* Launch shape: tile_words output words per work item (512, 1024 or 2048, i.e. 16384,
 * 32768 or 65536 rows; a work item holds 16384 / tile_words query-columns),
 * queries_per_group queries per work item (1..64), blocks_per_cu persistent workgroups per
 * compute unit (1..16), order 0: group the queries in the caller's order, 1: in an order
 * that puts queries sharing filter vectors next to each other (results still land at the
 * caller's query index). */
typedef struct rts_batch_config {
  int tile_words;
  int queries_per_group;
  int blocks_per_cu;
  int order;
  int threads;  /* workgroup size: 512 (64 KB of LDS toggles) or 256 (32 KB, half the
                   column slots per work item, twice the work items per CU) */
  int bins;     /* engine bins of the index set the queries use (0: the set the plan was
                   created with; see rts_batch_plan_add_indexes) */
  int direct_weight; /* cost of evaluating an RTS_EVAL_AUTO query-column directly, per mille
                        of its column's 4 bytes per row, against the BinDex form's filter-
                        vector and list bytes per row; 0: never direct when indexed */
} rts_batch_config;

typedef struct rts_batch_plan rts_batch_plan;

void rts_batch_default_config(rts_batch_config *cfg);

/* columns: host array of ncols device column pointers; indexes: host array of ncols BinDex
 * pointers (entries or the whole array may be NULL: those columns are evaluated directly).
 * Allocates the device query table (max_queries) and per-(query, tile) count partials. */
hipError_t rts_batch_plan_create(rts_batch_plan **plan, const CODE *const *columns,
                                 const BinDex *const *indexes, int ncols, int64_t n,
                                 int max_queries, const rts_batch_config *cfg, hipStream_t stream);
/* Change the launch shape.  The current queries are planned again for it (a new tile size
 * or index set also rebuilds the list directories); a config naming bins no index set of
 * the plan has is refused. */
hipError_t rts_batch_plan_set_config(rts_batch_plan *plan, const rts_batch_config *cfg,
                                     hipStream_t stream);
/* Give the plan another index set of the same table: the same columns indexed with another
 * bin count (rts_bindex_build_bins; entries may be NULL), selected by a config whose bins
 * is that count.  A set with the bin count of an existing inactive set replaces it.
 * *bins (may be NULL) receives the set's bin count.  The indexes must outlive the plan. */
hipError_t rts_batch_plan_add_indexes(rts_batch_plan *plan, const BinDex *const *indexes, int *bins);
/* Plan nq <= max_queries queries (host array) and upload them; returns after the upload. */
hipError_t rts_batch_plan_set_queries(rts_batch_plan *plan, const rts_query_spec *queries,
                                      int nq, hipStream_t stream);
/* As rts_batch_plan_set_queries with each query's bitmap at outs[q] (device, 16-byte
 * aligned); launches then take out == NULL. */
hipError_t rts_batch_plan_set_queries_out(rts_batch_plan *plan, const rts_query_spec *queries,
                                          BITS *const *outs, int nq, hipStream_t stream);
/* Evaluate the current queries: out holds nq bitmaps of out_stride_words words each (or is
 * NULL when the queries carry their own outputs); counts (device, nq entries) may be NULL.
 * Asynchronous on stream. */
hipError_t rts_batch_plan_launch(rts_batch_plan *plan, BITS *out, int64_t out_stride_words,
                                 unsigned long long *counts, hipStream_t stream);
/* As rts_batch_plan_launch through a captured hipGraph (captured on first use and again
 * when out / counts / stride / query count change).  stream must not be the null stream. */
hipError_t rts_batch_plan_graph_launch(rts_batch_plan *plan, BITS *out, int64_t out_stride_words,
                                       unsigned long long *counts, hipStream_t stream);
/* Planner statistics of the current queries: query-columns on the BinDex path, on the direct
 * path, BinDex columns whose range lies inside one bin (no filter vector read); row tiles. */
void rts_batch_plan_stats(const rts_batch_plan *plan, int *bindex_cols, int *direct_cols,
                          int *narrow_cols, long long *ntiles);
/* RTS_EVAL_AUTO query-columns of the current queries that the cost model sent to the
 * direct form although their column has an index (-1 for a NULL plan). */
int rts_batch_plan_auto_direct(const rts_batch_plan *plan);
/* BinDex query-columns of the current queries in the list form: a range over one or two
 * bins whose refine lists hold all their rows, read from those lists alone with no filter
 * vector (-1 for a NULL plan).  No counterpart in the reference, whose lists leave out
 * every bin's start value. */
int rts_batch_plan_list_cols(const rts_batch_plan *plan);
/* 64-bit digest of the current query mix: mean columns per query (in quarters) and the
 * shares of query-columns per evaluation mode and per range width (empty, 1-8, 9-32 or more
 * of the column's 128 areas), in sixteenths.  Batches drawn from one distribution get the
 * same digest; the autotuner's cache is keyed by it. */
unsigned long long rts_batch_plan_mix_digest(const rts_batch_plan *plan);
/* Work-item groups the current queries were planned into, and (cfg non-NULL) the launch
 * shape they were planned for; -1 for a NULL plan. */
int rts_batch_plan_groups(const rts_batch_plan *plan, rts_batch_config *cfg);
/* Rows, table columns and current query count of a plan. */
void rts_batch_plan_info(const rts_batch_plan *plan, long long *n, int *ncols, int *nq);
void rts_batch_plan_free(rts_batch_plan *plan);

/* Autotuner: time every candidate launch shape on the plan's current queries (median of
 * `reps` launches each) and keep the fastest in the plan.  Candidates span tile size, group
 * size, workgroups per CU, order, workgroup size, bin count (the plan's index sets) and the
 * planner's direct weight.  With cache_path non-NULL the result is read from / appended to
 * that file, keyed by device, n, column count, query count and the query-mix digest, so
 * later runs skip the search.  best_ms (may be NULL) receives the winner's time. */
hipError_t rts_batch_autotune(rts_batch_plan *plan, BITS *out, int64_t out_stride_words,
                              const rts_batch_config *candidates, int ncandidates, int reps,
                              const char *cache_path, rts_batch_config *chosen, float *best_ms,
                              hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* RTS_BATCH_H_ */
