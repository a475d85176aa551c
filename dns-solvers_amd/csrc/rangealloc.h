// rangealloc.h -- first-fit bookkeeping of sub-ranges of one block
// (host-only: offsets, no device memory), used by the pair arenas
// (pairarena.hip).  Ranges are kept in granules; a release must name a live
// range exactly (offset and size), and freed ranges are merged with free
// neighbours.
#pragma once
#include <stddef.h>

#include <iterator>
#include <map>

class RangeAlloc {
public:
    RangeAlloc(size_t capacity, size_t granule) : cap_(capacity / granule * granule), gran_(granule)
    {
        if (cap_) free_[0] = cap_;
    }

    size_t capacity() const { return cap_; }
    size_t used() const { return used_; }
    bool empty() const { return used_ == 0; }
    size_t rounded(size_t n) const { return (n + gran_ - 1) / gran_ * gran_; }

    // First free range that holds n bytes (rounded up to the granule).
    bool alloc(size_t n, size_t* off)
    {
        const size_t r = rounded(n);
        if (r == 0) return false;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < r) continue;
            *off = it->first;
            const size_t rest = it->second - r;
            const size_t at = it->first + r;
            free_.erase(it);
            if (rest) free_[at] = rest;
            live_[*off] = r;
            used_ += r;
            return true;
        }
        return false;
    }

    // Returns a live range given by alloc (same offset and n); false, with
    // nothing changed, if it is not one.
    bool release(size_t off, size_t n)
    {
        const size_t r = rounded(n);
        auto lv = live_.find(off);
        if (lv == live_.end() || lv->second != r) return false;
        live_.erase(lv);
        auto next = free_.lower_bound(off);
        size_t start = off, len = r;
        if (next != free_.end() && next->first == off + r) {
            len += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                start = prev->first;
                len += prev->second;
                free_.erase(prev);
            }
        }
        free_[start] = len;
        used_ -= r;
        return true;
    }

    size_t free_ranges() const { return free_.size(); }

private:
    size_t cap_, gran_;
    size_t used_ = 0;
    std::map<size_t, size_t> free_;  // offset -> length
    std::map<size_t, size_t> live_;  // offset -> rounded length of each allocated range
};
