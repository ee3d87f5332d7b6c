// Dense Ryser kernels instantiated for n = 45..46 (one of several
// translation units so the kernel family compiles in parallel).
#include "ryser_dense_impl.hpp"

PERM_DENSE_RANGE(j, 45, 46)
