/*
 * rts_probe.h -- a streaming-read probe of the device (no reference counterpart): the
 * measured read ceiling the scan kernels are set against, next to the 8 TB/s data-sheet
 * figure.
 */
#ifndef RTS_PROBE_H_
#define RTS_PROBE_H_

#include "rtscan_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Read `bytes` (a multiple of 16, 16-byte aligned) of device memory once with non-temporal
 * 128-bit loads, the way kernel (a) reads a column, in a grid of blocks_per_cu 256-thread
 * workgroups per compute unit (1..64), each lane keeping `unroll` (1, 2, 4 or 8) loads in
 * flight; the XOR of all words goes to *sink (device, 16 bytes) so the loads cannot be
 * dropped.  Asynchronous on `stream`. */
hipError_t rts_probe_read(const void *buf, size_t bytes, int blocks_per_cu, int unroll,
                          unsigned long long *sink, hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* RTS_PROBE_H_ */
