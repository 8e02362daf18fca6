// rr_ext_stubs.cpp -- MT19937 is not built yet: its factory reports TYPE_ERROR.
#include "rr_generator.hpp"

namespace rr
{
rocrand_generator make_mt19937_generator() { return nullptr; }
} // namespace rr
