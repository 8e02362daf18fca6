// rr_arena.hpp -- the request server's device-memory arena: offsets into one allocation.
#pragma once

#include <cstddef>
#include <iterator>
#include <map>

namespace rr
{

// First fit over [0, size) by address; a freed range is merged with its free neighbours,
// so the free list is always the maximal free runs in address order.
class arena
{
public:
    explicit arena(size_t size)
    {
        if(size > 0)
            free_[0] = size;
    }

    // Lowest-addressed free run of at least n bytes (n > 0); false when none.
    bool alloc(size_t n, size_t& start)
    {
        for(auto it = free_.begin(); it != free_.end(); ++it)
            if(it->second >= n)
            {
                start             = it->first;
                const size_t rest = it->second - n;
                free_.erase(it);
                if(rest > 0)
                    free_[start + n] = rest;
                return true;
            }
        return false;
    }

    // Returns [start, start + n), which alloc handed out, to the free list.
    void release(size_t start, size_t n)
    {
        auto next = free_.lower_bound(start);
        if(next != free_.end() && start + n == next->first)
        {
            n += next->second;
            next = free_.erase(next);
        }
        if(next != free_.begin())
        {
            auto prev = std::prev(next);
            if(prev->first + prev->second == start)
            {
                prev->second += n;
                return;
            }
        }
        free_[start] = n;
    }

    // Free runs in address order: start -> length.
    const std::map<size_t, size_t>& free_runs() const { return free_; }

private:
    std::map<size_t, size_t> free_;
};

} // namespace rr
