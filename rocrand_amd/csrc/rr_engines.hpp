// rr_engines.hpp -- the engine arithmetic (include/rocrand/rocrand_amd_engines.h, shared
// with the device API) in the library's namespace.
#pragma once

#include <rocrand/rocrand_amd_engines.h>

namespace rr
{
using namespace rocrand_amd_engine;
} // namespace rr
