// Dense Ryser kernels instantiated for n = 43..44 (one of several
// translation units so the kernel family compiles in parallel).
#include "ryser_dense_impl.hpp"

PERM_DENSE_RANGE(i, 43, 44)
