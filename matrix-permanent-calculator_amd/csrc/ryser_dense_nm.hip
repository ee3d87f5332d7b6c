// Dense Ryser kernels instantiated for n = 54..58 (plain walk only: the
// nested shapes stop at n = 48).
#include "ryser_dense_impl.hpp"

PERM_DENSE_RANGE(m, 54, 58)
