// CPU harness of the request server's arena (rocrand_amd/csrc/rr_arena.hpp) for
// tests/test_serving_host.py: the product's class behind a few C functions.
#include "rr_arena.hpp"

#include <cstddef>

extern "C" {

void* arena_new(size_t size) { return new rr::arena(size); }
void  arena_delete(void* a) { delete static_cast<rr::arena*>(a); }

// start of the allocated run, or -1
long long arena_alloc(void* a, size_t n)
{
    size_t s = 0;
    return static_cast<rr::arena*>(a)->alloc(n, s) ? (long long)s : -1;
}

void arena_release(void* a, size_t start, size_t n) { static_cast<rr::arena*>(a)->release(start, n); }

// writes up to cap (start, length) pairs; returns the number of free runs
size_t arena_free_runs(void* a, size_t* out, size_t cap)
{
    size_t k = 0;
    for(const auto& r : static_cast<rr::arena*>(a)->free_runs())
    {
        if(k < cap)
        {
            out[2 * k]     = r.first;
            out[2 * k + 1] = r.second;
        }
        k++;
    }
    return k;
}

} // extern "C"
