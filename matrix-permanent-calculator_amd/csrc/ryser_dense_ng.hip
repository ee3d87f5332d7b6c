// Dense Ryser kernels instantiated for n = 39..40 (one of several
// translation units so the kernel family compiles in parallel).
#include "ryser_dense_impl.hpp"

PERM_DENSE_RANGE(g, 39, 40)
