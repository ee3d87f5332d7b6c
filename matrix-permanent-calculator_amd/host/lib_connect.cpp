// lib_connect.cpp -- libConnect.so: the reference's Python/MATLAB connector
// library under its own exported names and argument lists, so its callers
// switch by changing only the library path.  The reference builds
// libConnect.so from py_connector.c (py_make) and binds
//   connect()                                    py_connector.c:59
//   read_calculate_return(filename, algorithm, nt, x, y, z)   :63
//   matlab_calculate_return(mat, algorithm, nt, x, y, z, nov, nnz)   :152
// (superpython.py:6-26 through ctypes, supermaTlab.m through loadlibrary
// with matlab_calculate_return.h).  Each forwards to the same-named perm_*
// entry point of libperman_hip.so: algorithms 4-8 (the exact ones) run on
// the GPU, 0-3 (the approximations, not part of this build) return NaN with
// the reason in perm_last_error().
#include "perman_hip.h"

extern "C" {

void connect() {
  perm_connect();
}

double read_calculate_return(char* filename, int algorithm, int nt, int x, int y, int z) {
  return perm_read_calculate_return(filename, algorithm, nt, x, y, z);
}

double matlab_calculate_return(int* mat, int algorithm, int nt, int x, int y, int z, int nov,
                               int nnz) {
  return perm_matlab_calculate_return(mat, algorithm, nt, x, y, z, nov, nnz);
}

}  // extern "C"
