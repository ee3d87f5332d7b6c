// Dense Ryser kernels instantiated for n = 41..42 (one of several
// translation units so the kernel family compiles in parallel).
#include "ryser_dense_impl.hpp"

PERM_DENSE_RANGE(h, 41, 42)
