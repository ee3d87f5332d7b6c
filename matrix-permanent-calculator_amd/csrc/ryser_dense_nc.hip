// Dense Ryser kernels instantiated for n = 27..30 (one of several
// translation units so the kernel family compiles in parallel).
#include "ryser_dense_impl.hpp"

PERM_DENSE_RANGE(c, 27, 30)
