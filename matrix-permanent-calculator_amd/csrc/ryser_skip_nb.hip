// Skipping Ryser kernels for n in [33, 48] (see ryser_skip_impl.hpp).
#include "ryser_skip_impl.hpp"

PERM_SKIP_RANGE(b, 33, 48)
