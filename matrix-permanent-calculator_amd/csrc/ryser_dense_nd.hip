// Dense Ryser kernels instantiated for n = 31..33 (one of several
// translation units so the kernel family compiles in parallel).
#include "ryser_dense_impl.hpp"

PERM_DENSE_RANGE(d, 31, 33)
