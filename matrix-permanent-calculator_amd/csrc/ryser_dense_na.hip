// Dense Ryser kernels instantiated for n = 1..20 (one of several
// translation units so the kernel family compiles in parallel).
#include "ryser_dense_impl.hpp"

PERM_DENSE_RANGE(a, 1, 20)
