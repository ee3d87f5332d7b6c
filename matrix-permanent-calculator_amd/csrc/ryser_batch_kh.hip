// Batch kernel variants, range 7 of 8 (see ryser_batch_impl.hpp).
#include "ryser_batch_impl.hpp"

PERM_BATCH_RANGE(7)
