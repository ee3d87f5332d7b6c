// Dense Ryser kernels instantiated for n = 34..36 (one of several
// translation units so the kernel family compiles in parallel).
#include "ryser_dense_impl.hpp"

PERM_DENSE_RANGE(e, 34, 36)
