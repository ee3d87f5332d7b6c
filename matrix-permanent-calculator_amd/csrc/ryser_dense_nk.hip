// Dense Ryser kernels instantiated for n = 47..48 (one of several
// translation units so the kernel family compiles in parallel).
#include "ryser_dense_impl.hpp"

PERM_DENSE_RANGE(k, 47, 48)
