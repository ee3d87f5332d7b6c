// Skipping Ryser kernels for n in [16, 32] (see ryser_skip_impl.hpp).
#include "ryser_skip_impl.hpp"

PERM_SKIP_RANGE(a, 16, 32)
