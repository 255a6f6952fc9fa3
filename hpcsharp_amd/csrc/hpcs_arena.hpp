// hpcs_arena.hpp -- host-managed first-fit allocator over one device region
// (the serving loop's HBM arena).  Plain C++: the bookkeeping runs on the
// host and is tested there, with sanitizers, apart from any device.
#pragma once
#include <cstddef>
#include <cstdint>
#include <iterator>
#include <map>

namespace hpcs {

// Free blocks are kept by offset; allocation takes the first block that
// fits, release merges a block with free neighbours.
class Arena {
  public:
    void reset(char *base, size_t size)
    {
        base_ = base;
        free_.clear();
        if (size) free_[0] = size;
    }
    // Offset of a block of `bytes` (a multiple of 256), or SIZE_MAX.
    size_t alloc(size_t bytes)
    {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            const size_t off = it->first, left = it->second - bytes;
            free_.erase(it);
            if (left) free_[off + bytes] = left;
            return off;
        }
        return SIZE_MAX;
    }
    void release(size_t off, size_t bytes)
    {
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
    char *at(size_t off) const { return base_ + off; }

  private:
    char *base_ = nullptr;
    std::map<size_t, size_t> free_;
};

} // namespace hpcs
