// sg_arena.hpp -- first-fit allocator of the serving loop's device region
// (host-only bookkeeping: offsets into one device allocation).  Free blocks
// are kept by offset; release coalesces with both neighbours.
#pragma once

#include <stddef.h>

#include <iterator>
#include <map>

namespace sg {

struct Arena {
    std::map<size_t, size_t> free_;   // offset -> bytes
    explicit Arena(size_t size = 0) {
        if (size) free_[0] = size;
    }
    bool alloc(size_t bytes, size_t* off) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                *off = it->first;
                const size_t rest = it->second - bytes, o = it->first + bytes;
                free_.erase(it);
                if (rest) free_[o] = rest;
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        auto it = free_.emplace(off, bytes).first;
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
};

}  // namespace sg
