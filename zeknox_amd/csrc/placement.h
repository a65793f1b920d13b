// First-fit placement of byte ranges in one region [0, size), host-only:
// free ranges kept by offset and coalesced on return.  Used by the commit
// server to place requests in its preallocated device region.
#pragma once
#include <stddef.h>

#include <iterator>
#include <map>

namespace zkx {

class Placement {
public:
    void reset(size_t size)
    {
        free_.clear();
        if (size) free_[0] = size;
    }

    // Lowest placement of n bytes that ends at or below `limit`.
    bool take_low(size_t n, size_t limit, size_t &off)
    {
        for (auto it = free_.begin(); it != free_.end() && it->first + n <= limit; ++it) {
            if (it->second < n) continue;
            off = it->first;
            const size_t rest = it->second - n;
            free_.erase(it);
            if (rest) free_[off + n] = rest;
            return true;
        }
        return false;
    }

    // Highest placement of n bytes.
    bool take_high(size_t n, size_t &off)
    {
        for (auto it = free_.rbegin(); it != free_.rend(); ++it) {
            if (it->second < n) continue;
            it->second -= n;
            off = it->first + it->second;
            if (!it->second) free_.erase(std::next(it).base());
            return true;
        }
        return false;
    }

    // Returns [off, off + n), which must have been taken.
    void give(size_t off, size_t n)
    {
        if (!n) return;
        auto next = free_.lower_bound(off);
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                off = prev->first;
                n += prev->second;
                free_.erase(prev);
            }
        }
        if (next != free_.end() && off + n == next->first) {
            n += next->second;
            free_.erase(next);
        }
        free_[off] = n;
    }

    const std::map<size_t, size_t> &free_ranges() const { return free_; }

private:
    std::map<size_t, size_t> free_;
};

}  // namespace zkx
