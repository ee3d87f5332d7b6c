// Dense Ryser kernels instantiated for n = 49..53 (plain walk only: the
// nested shapes stop at n = 48).
#include "ryser_dense_impl.hpp"

PERM_DENSE_RANGE(l, 49, 53)
