// First-fit allocator over one pre-allocated device slab, used by the serving
// loop to admit requests by HBM budget.  Pure host bookkeeping on offsets: it
// never touches the device, so its policy is unit-testable without a GPU.
#pragma once

#include <cstddef>
#include <cstdint>
#include <iterator>
#include <map>

namespace nls {

class SlabPool {
public:
  explicit SlabPool(size_t capacity, size_t align = 256) : cap_(capacity), align_(align) {
    if (capacity) free_[0] = capacity;
  }
  // Returns the offset of a block of `bytes` (rounded up to the alignment), or
  // SIZE_MAX when no free extent is large enough.
  size_t alloc(size_t bytes) {
    bytes = round(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        const size_t off = it->first, len = it->second;
        free_.erase(it);
        if (len > bytes) free_[off + bytes] = len - bytes;
        used_[off] = bytes;
        in_use_ += bytes;
        return off;
      }
    }
    return SIZE_MAX;
  }
  void release(size_t off) {
    auto u = used_.find(off);
    if (u == used_.end()) return;
    size_t len = u->second;
    used_.erase(u);
    in_use_ -= len;
    auto next = free_.lower_bound(off);
    if (next != free_.end() && off + len == next->first) {
      len += next->second;
      next = free_.erase(next);
    }
    if (next != free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == off) {
        prev->second += len;
        return;
      }
    }
    free_[off] = len;
  }
  size_t in_use() const { return in_use_; }
  size_t capacity() const { return cap_; }
  size_t largest_free() const {
    size_t b = 0;
    for (auto &f : free_) b = f.second > b ? f.second : b;
    return b;
  }
  size_t round(size_t b) const { return (b + align_ - 1) / align_ * align_; }

private:
  size_t cap_, align_, in_use_ = 0;
  std::map<size_t, size_t> free_; // offset -> length
  std::map<size_t, size_t> used_;
};

} // namespace nls
