// first_fit.hpp -- first-fit bookkeeping of one region (offsets only, no
// device calls): the serving loop's operand arena keeps its free blocks
// here.  Sizes and offsets are whole 256-byte granules; a released block
// merges with free neighbours, so no two free blocks touch.
#pragma once

#include <algorithm>
#include <cstddef>
#include <iterator>
#include <map>

namespace dspmm {

class FirstFit {
  public:
    static size_t granule(size_t n) { return (std::max<size_t>(n, 1) + 255) / 256 * 256; }
    // A region of `size` bytes (a multiple of 256), all free.
    void reset(size_t size) {
        size_ = size;
        used_ = peak_ = 0;
        free_.clear();
        if (size_) free_[0] = size_;
    }
    // The lowest free block that holds n bytes (rounded up to a granule).
    bool alloc(size_t n, size_t *off) {
        n = granule(n);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            *off = it->first;
            const size_t rest = it->second - n, start = it->first + n;
            free_.erase(it);
            if (rest) free_[start] = rest;
            used_ += n;
            peak_ = std::max(peak_, used_);
            return true;
        }
        return false;
    }
    // Gives back a block alloc returned (same n).
    void release(size_t off, size_t n) {
        n = granule(n);
        used_ -= n;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    size_t size() const { return size_; }
    size_t used() const { return used_; }
    size_t peak() const { return peak_; }
    const std::map<size_t, size_t> &free_blocks() const { return free_; }

  private:
    size_t size_ = 0, used_ = 0, peak_ = 0;
    std::map<size_t, size_t> free_;   // offset -> bytes
};

}  // namespace dspmm
