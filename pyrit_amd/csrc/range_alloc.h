// range_alloc.h - first-fit range allocator with coalescing for the device
// pools (the reference's src/range_alloc.zig:10-59). Pure host bookkeeping:
// a free list sorted by offset, never overlapping and never adjacent.
#pragma once
#include <stdint.h>

#include <vector>

namespace pyr {

class RangeAlloc {
  public:
    struct Range {
        uint64_t offset, size;
    };

    explicit RangeAlloc(uint64_t capacity = 0) : capacity_(capacity) {
        if (capacity) free_.push_back({0, capacity});
    }

    // offset of `size` units, or -1 when no free range is large enough
    int64_t alloc(uint64_t size) {
        if (size == 0) return 0;
        for (size_t i = 0; i < free_.size(); i++) {
            Range& r = free_[i];
            if (r.size < size) continue;
            const uint64_t off = r.offset;
            r.offset += size;
            r.size -= size;
            if (r.size == 0) free_.erase(free_.begin() + (long)i);
            used_ += size;
            return (int64_t)off;
        }
        return -1;
    }

    void release(uint64_t offset, uint64_t size) {
        if (size == 0) return;
        used_ -= size;
        size_t i = 0;
        while (i < free_.size() && free_[i].offset < offset) i++;
        const bool mp = i > 0 && free_[i - 1].offset + free_[i - 1].size == offset;
        const bool mn = i < free_.size() && offset + size == free_[i].offset;
        if (mp && mn) {
            free_[i - 1].size += size + free_[i].size;
            free_.erase(free_.begin() + (long)i);
        } else if (mp) {
            free_[i - 1].size += size;
        } else if (mn) {
            free_[i].offset = offset;
            free_[i].size += size;
        } else {
            free_.insert(free_.begin() + (long)i, Range{offset, size});
        }
    }

    // reserve exactly [offset, offset+size) again (it must be free); used to
    // roll back a release
    bool take(uint64_t offset, uint64_t size) {
        if (size == 0) return true;
        for (size_t i = 0; i < free_.size(); i++) {
            Range r = free_[i];
            if (offset < r.offset || offset + size > r.offset + r.size) continue;
            free_.erase(free_.begin() + (long)i);
            if (offset + size < r.offset + r.size)
                free_.insert(free_.begin() + (long)i, Range{offset + size, r.offset + r.size - offset - size});
            if (offset > r.offset) free_.insert(free_.begin() + (long)i, Range{r.offset, offset - r.offset});
            used_ += size;
            return true;
        }
        return false;
    }

    uint64_t capacity() const { return capacity_; }
    uint64_t used() const { return used_; }
    size_t free_ranges() const { return free_.size(); }
    const std::vector<Range>& free_list() const { return free_; }

  private:
    uint64_t capacity_ = 0, used_ = 0;
    std::vector<Range> free_;
};

}  // namespace pyr
