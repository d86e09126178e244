// admission.h -- host-side admission of the request loop (no device calls): a first-fit
// allocator over the result pool and the FIFO of waiting requests.  The server owns one
// Admission behind its mutex; tests drive it on the CPU with a made-up budget.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <algorithm>
#include <deque>
#include <map>
#include <vector>

namespace rts {

// First-fit allocator over `capacity` bytes (offsets only; the server maps them onto one
// device allocation), 256-byte granules, coalescing frees.
class PoolAllocator {
 public:
  static constexpr size_t kGranule = 256;
  explicit PoolAllocator(size_t capacity = 0) { reset(capacity); }
  void reset(size_t capacity) {
    bytes_ = capacity / kGranule * kGranule;
    free_.clear();
    if (bytes_) free_[0] = bytes_;
    in_use_ = peak_ = 0;
  }
  static size_t rounded(size_t bytes) { return (bytes + kGranule - 1) / kGranule * kGranule; }
  // Offset of a free range of `bytes`, or SIZE_MAX.
  size_t alloc(size_t bytes) {
    bytes = rounded(bytes);
    if (bytes == 0) bytes = kGranule;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        const size_t off = it->first, rest = it->second - bytes;
        free_.erase(it);
        if (rest) free_[off + bytes] = rest;
        in_use_ += bytes;
        peak_ = std::max(peak_, in_use_);
        return off;
      }
    }
    return SIZE_MAX;
  }
  void release(size_t off, size_t bytes) {
    bytes = rounded(bytes);
    if (bytes == 0) bytes = kGranule;
    in_use_ -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto next = std::next(it);
    if (next != free_.end() && it->first + it->second == next->first) {
      it->second += next->second;
      free_.erase(next);
    }
    if (it != free_.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        free_.erase(it);
      }
    }
  }
  size_t capacity() const { return bytes_; }
  size_t in_use() const { return in_use_; }
  size_t peak() const { return peak_; }
  size_t free_ranges() const { return free_.size(); }

 private:
  size_t bytes_ = 0, in_use_ = 0, peak_ = 0;
  std::map<size_t, size_t> free_;  // offset -> size
};

// The queue of waiting requests (tickets) and the pool their bitmaps come from.  A batch
// takes, in arrival order, the waiting requests of the oldest request's table while each
// one's bitmap fits the pool and the batch has room; the first that does not fit stops the
// batch (no overtaking, so a large request is not starved by small ones behind it).
class Admission {
 public:
  struct Admitted {
    uint64_t ticket;
    size_t offset;
  };
  explicit Admission(size_t pool_bytes = 0) : pool_(pool_bytes) {}
  void reset(size_t pool_bytes) {
    pool_.reset(pool_bytes);
    queue_.clear();
  }
  void enqueue(uint64_t ticket, int table) { queue_.push_back({ticket, table}); }
  bool remove(uint64_t ticket) {  // a waiting request withdrawn by its caller
    for (auto it = queue_.begin(); it != queue_.end(); ++it)
      if (it->ticket == ticket) {
        queue_.erase(it);
        return true;
      }
    return false;
  }
  size_t waiting() const { return queue_.size(); }
  int front_table() const { return queue_.empty() ? -1 : queue_.front().table; }
  // bytes_of(table) = bitmap bytes of one request on that table.
  template <class BytesOf>
  std::vector<Admitted> admit(int max_batch, BytesOf bytes_of) {
    std::vector<Admitted> out;
    if (queue_.empty()) return out;
    const int table = queue_.front().table;
    const size_t bytes = bytes_of(table);
    for (auto it = queue_.begin(); it != queue_.end() && (int)out.size() < max_batch;) {
      if (it->table != table) {
        ++it;
        continue;
      }
      const size_t off = pool_.alloc(bytes);
      if (off == SIZE_MAX) break;  // memory: wait for releases, keep arrival order
      out.push_back({it->ticket, off});
      it = queue_.erase(it);
    }
    return out;
  }
  void release(size_t offset, size_t bytes) { pool_.release(offset, bytes); }
  const PoolAllocator &pool() const { return pool_; }

 private:
  struct Waiting {
    uint64_t ticket;
    int table;
  };
  PoolAllocator pool_;
  std::deque<Waiting> queue_;
};

}  // namespace rts
