/*
 * dlp_pool.h -- ranges of one device allocation handed to batches (host
 * bookkeeping only, no device calls; internal).
 *
 * A multi-graph runner with a scratch pool does not give each graph its own
 * per-batch scratch: a batch leases one range of the pool when it is
 * admitted and returns it when its labels are final.  The allocator is
 * first fit over the free ranges sorted by offset, every range a multiple of
 * kPoolAlign bytes, and returned ranges merge with free neighbours, so a
 * pool that is empty again is one free range.  All members are safe to call
 * from several host threads (the hand-back lanes return leases).
 */
#ifndef DLP_POOL_H
#define DLP_POOL_H

#include <stdint.h>

#include <iterator>
#include <map>
#include <mutex>

constexpr int64_t kPoolAlign = 256;

class RangePool {
 public:
  explicit RangePool(int64_t capacity = 0) { reset(capacity); }

  /* forgets every lease; one free range of `capacity` (rounded down) */
  void reset(int64_t capacity) {
    std::lock_guard<std::mutex> lock(mu_);
    capacity_ = capacity > 0 ? capacity / kPoolAlign * kPoolAlign : 0;
    free_.clear();
    used_.clear();
    if (capacity_ > 0) free_[0] = capacity_;
    inUse_ = peak_ = leases_ = 0;
  }

  static int64_t rounded(int64_t bytes) {
    if (bytes < 1) bytes = 1;
    return (bytes + kPoolAlign - 1) / kPoolAlign * kPoolAlign;
  }

  /* a range of at least `bytes` at *offset (the lowest free range that
   * fits); false, and nothing changed, when none fits now */
  bool acquire(int64_t bytes, int64_t *offset) {
    const int64_t need = rounded(bytes);
    std::lock_guard<std::mutex> lock(mu_);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < need) continue;
      const int64_t off = it->first, left = it->second - need;
      free_.erase(it);
      if (left > 0) free_[off + need] = left;
      used_[off] = need;
      inUse_ += need;
      if (inUse_ > peak_) peak_ = inUse_;
      ++leases_;
      *offset = off;
      return true;
    }
    return false;
  }

  /* returns the range that starts at `offset`; false if there is none */
  bool release(int64_t offset) {
    std::lock_guard<std::mutex> lock(mu_);
    auto u = used_.find(offset);
    if (u == used_.end()) return false;
    int64_t off = u->first, len = u->second;
    used_.erase(u);
    inUse_ -= len;
    auto next = free_.lower_bound(off);
    if (next != free_.end() && next->first == off + len) {  // merge with the range after
      len += next->second;
      next = free_.erase(next);
    }
    if (next != free_.begin()) {  // merge with the range before
      auto prev = std::prev(next);
      if (prev->first + prev->second == off) {
        prev->second += len;
        return true;
      }
    }
    free_[off] = len;
    return true;
  }

  int64_t capacity() const { return capacity_; }
  int64_t inUse() {
    std::lock_guard<std::mutex> lock(mu_);
    return inUse_;
  }
  int64_t peak() {
    std::lock_guard<std::mutex> lock(mu_);
    return peak_;
  }
  int64_t leases() {
    std::lock_guard<std::mutex> lock(mu_);
    return leases_;
  }
  /* number of free ranges (1 when the pool is empty) */
  int freeRanges() {
    std::lock_guard<std::mutex> lock(mu_);
    return (int)free_.size();
  }
  /* largest free range */
  int64_t largestFree() {
    std::lock_guard<std::mutex> lock(mu_);
    int64_t m = 0;
    for (const auto &f : free_)
      if (f.second > m) m = f.second;
    return m;
  }

 private:
  std::mutex mu_;
  int64_t capacity_ = 0, inUse_ = 0, peak_ = 0, leases_ = 0;
  std::map<int64_t, int64_t> free_; /* offset -> bytes */
  std::map<int64_t, int64_t> used_; /* offset -> bytes */
};

#endif /* DLP_POOL_H */
