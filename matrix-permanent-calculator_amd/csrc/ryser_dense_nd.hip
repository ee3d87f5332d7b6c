// Dense Ryser kernels instantiated for n = 37..42 (one of several
// translation units so the kernel family compiles in parallel).
#include "ryser_dense_impl.hpp"

PERM_DENSE_RANGE(d, 37, 42)
