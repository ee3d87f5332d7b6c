// Dense Ryser kernels instantiated for n = 21..26 (one of several
// translation units so the kernel family compiles in parallel).
#include "ryser_dense_impl.hpp"

PERM_DENSE_RANGE(b, 21, 26)
