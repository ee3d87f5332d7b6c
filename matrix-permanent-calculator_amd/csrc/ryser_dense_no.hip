// Dense Ryser kernels instantiated for n = 59..63 (plain walk only: the
// nested shapes stop at n = 48).
#include "ryser_dense_impl.hpp"

PERM_DENSE_RANGE(o, 59, 63)
