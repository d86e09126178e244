// batch.h -- constants shared by the batched-query kernels and their host planner.
#pragma once

#include "common.h"
#include "rts_batch.h"

namespace rts {

constexpr int kBatchMaxCols = RTS_BATCH_MAX_COLS;  // columns per query
constexpr int kBatchMaxGroup = 64;                 // queries per work item, upper bound
constexpr int kBatchLdsWords = 16384;              // toggle words per workgroup (64 KB)
constexpr int kBatchMaxSlots = 32;                 // column slots per work item (512-word tiles)
constexpr int kXcds = 8;                           // MI355X: 8 XCDs, workgroup b -> XCD b % 8
constexpr int kSchedStride = 16;                   // 128 bytes between per-XCD item counters
constexpr int kSchedWords = (kXcds + 1) * kSchedStride;

}  // namespace rts
