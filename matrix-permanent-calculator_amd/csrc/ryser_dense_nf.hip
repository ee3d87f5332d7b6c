// Dense Ryser kernels instantiated for n = 37..38 (one of several
// translation units so the kernel family compiles in parallel).
#include "ryser_dense_impl.hpp"

PERM_DENSE_RANGE(f, 37, 38)
