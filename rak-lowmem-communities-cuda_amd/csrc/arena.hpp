// First-fit allocator of index ranges (vertex entries or edge positions of a batch arena), with
// coalescing of neighbouring free ranges on release.
#pragma once
#include <cstdint>
#include <iterator>
#include <map>

namespace rlc {

class RangeAlloc {
 public:
  void reset(uint64_t capacity) {
    free_.clear();
    if (capacity) free_[0] = capacity;
  }
  // [start, start + len) taken from the lowest free range that holds it; len 0 always succeeds.
  bool alloc(uint64_t len, uint64_t &start) {
    if (len == 0) {
      start = 0;
      return true;
    }
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < len) continue;
      start = it->first;
      const uint64_t rest = it->second - len;
      free_.erase(it);
      if (rest) free_[start + len] = rest;
      return true;
    }
    return false;
  }
  void release(uint64_t start, uint64_t len) {
    if (len == 0) return;
    auto next = free_.lower_bound(start);
    if (next != free_.end() && start + len == next->first) {
      len += next->second;
      next = free_.erase(next);
    }
    if (next != free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == start) {
        prev->second += len;
        return;
      }
    }
    free_[start] = len;
  }
  uint64_t largest() const {
    uint64_t r = 0;
    for (const auto &f : free_) r = f.second > r ? f.second : r;
    return r;
  }
  uint64_t total() const {
    uint64_t r = 0;
    for (const auto &f : free_) r += f.second;
    return r;
  }

 private:
  std::map<uint64_t, uint64_t> free_;  // start -> length
};

}  // namespace rlc
