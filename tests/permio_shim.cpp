// permio_shim.cpp -- C ABI over the `perman` CLI's matrix reader and
// -r1/-r2 orderings (matrix-permanent-calculator_amd/host/matrix_io.hpp),
// compiled by tests/test_capi.py so that its arrays can be compared byte for
// byte with the reference's ReadMatrix and matrix2compressed[_sortOrder|
// _skipOrder] (oracle/_ref/libref_algo.so).  Test infrastructure only.
#include <cstring>
#include <fstream>
#include <vector>

#include "../matrix-permanent-calculator_amd/host/matrix_io.hpp"

namespace {

template <class T>
int read_file(const char* path, int generic, T* mat, int cap) {
  std::ifstream in(path);
  permio::Header h;
  if (!permio::read_header(in, h) || h.n * h.n > cap) return -1;
  std::vector<T> m;
  permio::read_entries(in, h.n, generic == 0, m);
  std::memcpy(mat, m.data(), sizeof(T) * m.size());
  return h.n;
}

// order 0: compress only; 1: sort_order (-r1) then compress; 2: skip_order
// (-r2) then compress -- the CLI's sequence.  mat is rewritten in place.
template <class T>
int ordered(int order, T* mat, int n, int* cptrs, int* rows, T* cvals, int* rptrs, int* cols,
            T* rvals) {
  std::vector<T> m(mat, mat + (size_t)n * n);
  if (order == 1) permio::sort_order(m, n);
  if (order == 2) permio::skip_order(m, n);
  const permio::Compressed<T> c = permio::compress(m, n);
  std::memcpy(mat, m.data(), sizeof(T) * m.size());
  std::memcpy(cptrs, c.cptrs.data(), sizeof(int) * c.cptrs.size());
  std::memcpy(rptrs, c.rptrs.data(), sizeof(int) * c.rptrs.size());
  std::memcpy(rows, c.rows.data(), sizeof(int) * c.rows.size());
  std::memcpy(cols, c.cols.data(), sizeof(int) * c.cols.size());
  std::memcpy(cvals, c.cvals.data(), sizeof(T) * c.cvals.size());
  std::memcpy(rvals, c.rvals.data(), sizeof(T) * c.rvals.size());
  return (int)c.rows.size();
}

}  // namespace

extern "C" {
int shim_read_matrix_f64(const char* p, int g, double* m, int cap) { return read_file(p, g, m, cap); }
int shim_read_matrix_f32(const char* p, int g, float* m, int cap) { return read_file(p, g, m, cap); }
int shim_read_matrix_i32(const char* p, int g, int* m, int cap) { return read_file(p, g, m, cap); }
int shim_compressed_order_f64(int o, double* m, int n, int* cp, int* r, double* cv, int* rp,
                              int* c, double* rv) {
  return ordered(o, m, n, cp, r, cv, rp, c, rv);
}
int shim_compressed_order_f32(int o, float* m, int n, int* cp, int* r, float* cv, int* rp, int* c,
                              float* rv) {
  return ordered(o, m, n, cp, r, cv, rp, c, rv);
}
int shim_compressed_order_i32(int o, int* m, int n, int* cp, int* r, int* cv, int* rp, int* c,
                              int* rv) {
  return ordered(o, m, n, cp, r, cv, rp, c, rv);
}
}
