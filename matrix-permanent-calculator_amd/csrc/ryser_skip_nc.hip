// Skipping Ryser kernels for n in [49, 63] (see ryser_skip_impl.hpp).
#include "ryser_skip_impl.hpp"

PERM_SKIP_RANGE(c, 49, 63)
