/*
 * rts_serve.h -- request loop over resident tables (additive; the reference runs one query
 * file at a time, bindex_cuda.cpp:1827-1995).
 *
 * Tables (device columns + BinDex) stay resident in HBM.  Requests are single queries
 * (rts_query_spec) on one table.  Admission: a request leaves the queue when the result
 * pool (pool_bytes of HBM reserved for result bitmaps) can hold its bitmap
 * (ceil(n/32) words) and a batch slot is free; its bitmap stays allocated until released.
 * Continuous batching: a loop thread launches every admitted request of a table as one
 * batched launch (rts_batch.h) on one of `streams` streams, so batches overlap on the
 * device and requests that arrive while others run are admitted as soon as memory and a
 * stream free up.  Latency = completion seen by the loop minus submission.
 *
 * The loop's streams share the process's hardware queues (GPU_MAX_HW_QUEUES, 4 by default)
 * with every other stream the process has used, the per-thread stream of a thread that
 * called a reference-named wrapper included, so a serving process should use few streams
 * besides the loop's (measured: 27.3k requests/s at saturation with the loop alone, 24.9k
 * after one wrapper call on the main thread).
 */
#ifndef RTS_SERVE_H_
#define RTS_SERVE_H_

#include "rts_batch.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rts_server_config {
  int max_batch;           /* queries per launched batch */
  int streams;             /* batches in flight at once */
  size_t pool_bytes;       /* HBM for result bitmaps (admission budget) */
  int queue_capacity;      /* requests waiting for admission, at most */
  rts_batch_config batch;  /* launch shape of the batches */
} rts_server_config;

typedef struct rts_server rts_server;

enum rts_request_state { RTS_REQ_QUEUED = 0, RTS_REQ_RUNNING = 1, RTS_REQ_DONE = 2, RTS_REQ_FAILED = 3 };

void rts_server_default_config(rts_server_config *cfg);
hipError_t rts_server_create(rts_server **server, const rts_server_config *cfg);
/* Register a resident table: device columns and their BinDex (entries may be NULL). */
hipError_t rts_server_add_table(rts_server *server, const CODE *const *columns,
                                const BinDex *const *indexes, int ncols, int64_t n, int *table_id);
hipError_t rts_server_start(rts_server *server);
/* Queue one query; thread-safe.  hipErrorNotReady when the queue is full,
 * hipErrorInvalidValue for a malformed query or one whose bitmap can never fit the pool. */
hipError_t rts_server_submit(rts_server *server, int table_id, const rts_query_spec *query,
                             uint64_t *ticket);
/* Wait up to timeout_ms (< 0: no limit) for a request.  Returns its rts_request_state; when
 * done, *bitmap is the device bitmap (valid until rts_server_release), *count its popcount,
 * *latency_ns submission-to-completion. */
int rts_server_wait(rts_server *server, uint64_t ticket, int timeout_ms, const BITS **bitmap,
                    unsigned long long *count, uint64_t *latency_ns);
/* Return a finished request's bitmap memory to the pool. */
void rts_server_release(rts_server *server, uint64_t ticket);
/* Counters since start: submitted, completed, batches launched, largest batch, pool bytes
 * in use now and at most. */
void rts_server_stats(const rts_server *server, uint64_t *submitted, uint64_t *completed,
                      uint64_t *batches, int *max_batch_seen, size_t *pool_in_use,
                      size_t *pool_peak);
hipError_t rts_server_stop(rts_server *server);
void rts_server_destroy(rts_server *server);

/* Replay an arrival trace: request i (table_ids[i], queries[i]) is submitted arrival_ns[i]
 * after the start by a submitter thread; each result is read (count) and released as soon
 * as it completes.  latency_ns[i] and counts[i] receive the outcome (counts may be NULL);
 * *elapsed_s the wall time from start to the last completion.  The server must be started. */
hipError_t rts_server_replay(rts_server *server, const int *table_ids, const rts_query_spec *queries,
                             const uint64_t *arrival_ns, int nq, uint64_t *latency_ns,
                             unsigned long long *counts, double *elapsed_s);

#ifdef __cplusplus
}
#endif
#endif /* RTS_SERVE_H_ */
