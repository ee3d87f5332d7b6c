// perman.hpp -- C++ drop-in for the reference's exact-permanent entry points.
//
// Re-creates, at global scope, the template names and signatures the
// reference's main.cu calls (RunAlgo, main.cu:20-248), forwarding to the C
// ABI of libperman_hip.so (perman_hip.h):
//   gpu_exact_dense.cu:402,467,525,583,641   gpu_perman64_<variant>(T*, int, int, int)
//   gpu_exact_sparse.cu:674,734,794,854      gpu_perman64_<variant>_sparse(...)
//   gpu_exact_sparse.cu:1124                 gpu_perman64_xshared_coalescing_mshared_skipper
//   gpu_exact_dense.cu:702,777,914           ..._multigpu, _multigpucpu_chunks,
//                                            _multigpu_manual_distribution
//   gpu_exact_sparse.cu:917,996,1193,1328    ..._multigpu_sparse, _multigpucpu_chunks_sparse,
//                                            _multigpucpu_chunks_skipper,
//                                            _multigpu_sparse_manual_distribution
//   algo.h:663, algo.h:569                   parallel_perman64 / _sparse
//   algo.h:749, algo.h:886                   parallel_skip_perman64_w / _w_balanced
// for T in {int, float, double}.  A caller that included the reference's .cu
// files replaces those includes with this header and links
// -lperman_hip.  Results are returned by value as double (NaN on error; see
// perm_last_error()).
#ifndef PERMAN_HPP
#define PERMAN_HPP

#include "perman_hip.h"

namespace perman_detail {
// Overload sets per element type; the templates below dispatch on T.
#define PERMAN_DENSE_OVERLOADS(NAME)                                                             \
  inline double NAME(int* m, int n, int g, int b) { return perm_##NAME##_dense_i32(m, n, g, b); } \
  inline double NAME(float* m, int n, int g, int b) {                                            \
    return perm_##NAME##_dense_f32(m, n, g, b);                                                  \
  }                                                                                              \
  inline double NAME(double* m, int n, int g, int b) {                                           \
    return perm_##NAME##_dense_f64(m, n, g, b);                                                  \
  }
PERMAN_DENSE_OVERLOADS(xglobal)
PERMAN_DENSE_OVERLOADS(xlocal)
PERMAN_DENSE_OVERLOADS(xshared)
PERMAN_DENSE_OVERLOADS(xshared_coalescing)
PERMAN_DENSE_OVERLOADS(xshared_coalescing_mshared)
#undef PERMAN_DENSE_OVERLOADS

#define PERMAN_SPARSE_OVERLOADS(NAME)                                                       \
  inline double NAME##_sparse(int* m, int* cp, int* r, int* cv, int n, int g, int b) {      \
    return perm_##NAME##_sparse_i32(m, cp, r, cv, n, g, b);                                 \
  }                                                                                         \
  inline double NAME##_sparse(float* m, int* cp, int* r, float* cv, int n, int g, int b) {  \
    return perm_##NAME##_sparse_f32(m, cp, r, cv, n, g, b);                                 \
  }                                                                                         \
  inline double NAME##_sparse(double* m, int* cp, int* r, double* cv, int n, int g, int b) { \
    return perm_##NAME##_sparse_f64(m, cp, r, cv, n, g, b);                                 \
  }
PERMAN_SPARSE_OVERLOADS(xlocal)
PERMAN_SPARSE_OVERLOADS(xshared)
PERMAN_SPARSE_OVERLOADS(xshared_coalescing)
PERMAN_SPARSE_OVERLOADS(xshared_coalescing_mshared)
#undef PERMAN_SPARSE_OVERLOADS

inline double cpu(int* m, int n, int t) { return perm_parallel_perman64_i32(m, n, t); }
inline double cpu(float* m, int n, int t) { return perm_parallel_perman64_f32(m, n, t); }
inline double cpu(double* m, int n, int t) { return perm_parallel_perman64_f64(m, n, t); }
inline double cpu_sparse(int* m, int* cp, int* r, int* cv, int n, int t) {
  return perm_parallel_perman64_sparse_i32(m, cp, r, cv, n, t);
}
inline double cpu_sparse(float* m, int* cp, int* r, float* cv, int n, int t) {
  return perm_parallel_perman64_sparse_f32(m, cp, r, cv, n, t);
}
inline double cpu_sparse(double* m, int* cp, int* r, double* cv, int n, int t) {
  return perm_parallel_perman64_sparse_f64(m, cp, r, cv, n, t);
}
inline double skipper(int* m, int* rp, int* c, int* cp, int* r, int* cv, int n, int g, int b) {
  return perm_xshared_coalescing_mshared_skipper_i32(m, rp, c, cp, r, cv, n, g, b);
}
inline double skipper(float* m, int* rp, int* c, int* cp, int* r, float* cv, int n, int g,
                      int b) {
  return perm_xshared_coalescing_mshared_skipper_f32(m, rp, c, cp, r, cv, n, g, b);
}
inline double skipper(double* m, int* rp, int* c, int* cp, int* r, double* cv, int n, int g,
                      int b) {
  return perm_xshared_coalescing_mshared_skipper_f64(m, rp, c, cp, r, cv, n, g, b);
}

#define PERMAN_TYPED_OVERLOADS(T, SFX)                                                        \
  inline double multigpu(T* m, int n, int g, int gd, int bd) {                                \
    return perm_xshared_coalescing_mshared_multigpu_dense_##SFX(m, n, g, gd, bd);             \
  }                                                                                           \
  inline double multigpu_manual(T* m, int n, int g, int gd, int bd) {                         \
    return perm_xshared_coalescing_mshared_multigpu_manual_distribution_dense_##SFX(m, n, g,  \
                                                                                   gd, bd);   \
  }                                                                                           \
  inline double chunks(T* m, int n, int g, bool cpu, int t, int gd, int bd) {                 \
    return perm_xshared_coalescing_mshared_multigpucpu_chunks_dense_##SFX(m, n, g, cpu, t,    \
                                                                         gd, bd);             \
  }                                                                                           \
  inline double multigpu_sparse(T* m, int* cp, int* r, T* cv, int n, int g, int gd, int bd) { \
    return perm_xshared_coalescing_mshared_multigpu_sparse_##SFX(m, cp, r, cv, n, g, gd, bd); \
  }                                                                                           \
  inline double multigpu_sparse_manual(T* m, int* cp, int* r, T* cv, int n, int g, int gd,    \
                                       int bd) {                                              \
    return perm_xshared_coalescing_mshared_multigpu_sparse_manual_distribution_##SFX(         \
        m, cp, r, cv, n, g, gd, bd);                                                          \
  }                                                                                           \
  inline double chunks_sparse(T* m, int* cp, int* r, T* cv, int n, int g, bool cpu, int t,    \
                              int gd, int bd) {                                               \
    return perm_xshared_coalescing_mshared_multigpucpu_chunks_sparse_##SFX(m, cp, r, cv, n,   \
                                                                          g, cpu, t, gd, bd); \
  }                                                                                           \
  inline double chunks_skipper(T* m, int* rp, int* c, int* cp, int* r, T* cv, int n, int g,  \
                               bool cpu, int t, int gd, int bd) {                             \
    return perm_xshared_coalescing_mshared_multigpucpu_chunks_skipper_##SFX(                  \
        m, rp, c, cp, r, cv, n, g, cpu, t, gd, bd);                                           \
  }                                                                                           \
  inline double cpu_skip(int* rp, int* c, T* rv, int* cp, int* r, T* cv, int n, int t) {      \
    return perm_parallel_skip_perman64_w_##SFX(rp, c, rv, cp, r, cv, n, t);                   \
  }                                                                                           \
  inline double cpu_skip_balanced(int* rp, int* c, T* rv, int* cp, int* r, T* cv, int n,      \
                                  int t) {                                                    \
    return perm_parallel_skip_perman64_w_balanced_##SFX(rp, c, rv, cp, r, cv, n, t);          \
  }
PERMAN_TYPED_OVERLOADS(int, i32)
PERMAN_TYPED_OVERLOADS(float, f32)
PERMAN_TYPED_OVERLOADS(double, f64)
#undef PERMAN_TYPED_OVERLOADS
}  // namespace perman_detail

// ---- dense, algorithm IDs 0-4 -------------------------------------------
template <class T>
double gpu_perman64_xglobal(T* mat, int nov, int grid_dim, int block_dim) {
  return perman_detail::xglobal(mat, nov, grid_dim, block_dim);
}
template <class T>
double gpu_perman64_xlocal(T* mat, int nov, int grid_dim, int block_dim) {
  return perman_detail::xlocal(mat, nov, grid_dim, block_dim);
}
template <class T>
double gpu_perman64_xshared(T* mat, int nov, int grid_dim, int block_dim) {
  return perman_detail::xshared(mat, nov, grid_dim, block_dim);
}
template <class T>
double gpu_perman64_xshared_coalescing(T* mat, int nov, int grid_dim, int block_dim) {
  return perman_detail::xshared_coalescing(mat, nov, grid_dim, block_dim);
}
template <class T>
double gpu_perman64_xshared_coalescing_mshared(T* mat, int nov, int grid_dim, int block_dim) {
  return perman_detail::xshared_coalescing_mshared(mat, nov, grid_dim, block_dim);
}

// ---- sparse (-s), algorithm IDs 1-4 ---------------------------------------
template <class T>
double gpu_perman64_xlocal_sparse(T* mat, int* cptrs, int* rows, T* cvals, int nov, int grid_dim,
                                  int block_dim) {
  return perman_detail::xlocal_sparse(mat, cptrs, rows, cvals, nov, grid_dim, block_dim);
}
template <class T>
double gpu_perman64_xshared_sparse(T* mat, int* cptrs, int* rows, T* cvals, int nov, int grid_dim,
                                   int block_dim) {
  return perman_detail::xshared_sparse(mat, cptrs, rows, cvals, nov, grid_dim, block_dim);
}
template <class T>
double gpu_perman64_xshared_coalescing_sparse(T* mat, int* cptrs, int* rows, T* cvals, int nov,
                                              int grid_dim, int block_dim) {
  return perman_detail::xshared_coalescing_sparse(mat, cptrs, rows, cvals, nov, grid_dim,
                                                  block_dim);
}
template <class T>
double gpu_perman64_xshared_coalescing_mshared_sparse(T* mat, int* cptrs, int* rows, T* cvals,
                                                      int nov, int grid_dim, int block_dim) {
  return perman_detail::xshared_coalescing_mshared_sparse(mat, cptrs, rows, cvals, nov, grid_dim,
                                                          block_dim);
}

// ---- sparse skipper, algorithm ID 7 (gpu_exact_sparse.cu:1124) -----------
template <class T>
double gpu_perman64_xshared_coalescing_mshared_skipper(T* mat, int* rptrs, int* cols, int* cptrs,
                                                       int* rows, T* cvals, int nov,
                                                       int grid_dim, int block_dim) {
  return perman_detail::skipper(mat, rptrs, cols, cptrs, rows, cvals, nov, grid_dim, block_dim);
}

// ---- multi-GPU / hybrid IDs 5, 6, 66 (dense) and 5, 6, 8, 66 (sparse) -----
// One GPU computes the whole range; `cpu` adds the CPU worker of the chunked
// scheduler (IDs 6).
template <class T>
double gpu_perman64_xshared_coalescing_mshared_multigpu(T* mat, int nov, int gpu_num,
                                                        int grid_dim, int block_dim) {
  return perman_detail::multigpu(mat, nov, gpu_num, grid_dim, block_dim);
}
template <class T>
double gpu_perman64_xshared_coalescing_mshared_multigpucpu_chunks(T* mat, int nov, int gpu_num,
                                                                  bool cpu, int threads,
                                                                  int grid_dim, int block_dim) {
  return perman_detail::chunks(mat, nov, gpu_num, cpu, threads, grid_dim, block_dim);
}
template <class T>
double gpu_perman64_xshared_coalescing_mshared_multigpu_manual_distribution(T* mat, int nov,
                                                                            int gpu_num,
                                                                            int grid_dim,
                                                                            int block_dim) {
  return perman_detail::multigpu_manual(mat, nov, gpu_num, grid_dim, block_dim);
}
template <class T>
double gpu_perman64_xshared_coalescing_mshared_multigpu_sparse(T* mat, int* cptrs, int* rows,
                                                               T* cvals, int nov, int gpu_num,
                                                               int grid_dim, int block_dim) {
  return perman_detail::multigpu_sparse(mat, cptrs, rows, cvals, nov, gpu_num, grid_dim,
                                        block_dim);
}
template <class T>
double gpu_perman64_xshared_coalescing_mshared_multigpucpu_chunks_sparse(
    T* mat, int* cptrs, int* rows, T* cvals, int nov, int gpu_num, bool cpu, int threads,
    int grid_dim, int block_dim) {
  return perman_detail::chunks_sparse(mat, cptrs, rows, cvals, nov, gpu_num, cpu, threads,
                                      grid_dim, block_dim);
}
template <class T>
double gpu_perman64_xshared_coalescing_mshared_multigpucpu_chunks_skipper(
    T* mat, int* rptrs, int* cols, int* cptrs, int* rows, T* cvals, int nov, int gpu_num,
    bool cpu, int threads, int grid_dim, int block_dim) {
  return perman_detail::chunks_skipper(mat, rptrs, cols, cptrs, rows, cvals, nov, gpu_num, cpu,
                                       threads, grid_dim, block_dim);
}
template <class T>
double gpu_perman64_xshared_coalescing_mshared_multigpu_sparse_manual_distribution(
    T* mat, int* cptrs, int* rows, T* cvals, int nov, int gpu_num, int grid_dim, int block_dim) {
  return perman_detail::multigpu_sparse_manual(mat, cptrs, rows, cvals, nov, gpu_num, grid_dim,
                                               block_dim);
}

// ---- CPU (`-c`) ------------------------------------------------------------
template <class T>
double parallel_perman64(T* mat, int nov, int threads) {
  return perman_detail::cpu(mat, nov, threads);
}
template <class T>
double parallel_perman64_sparse(T* mat, int* cptrs, int* rows, T* cvals, int nov, int threads) {
  return perman_detail::cpu_sparse(mat, cptrs, rows, cvals, nov, threads);
}
template <class T>
double parallel_skip_perman64_w(int* rptrs, int* cols, T* rvals, int* cptrs, int* rows,
                                T* cvals, int nov, int threads) {
  return perman_detail::cpu_skip(rptrs, cols, rvals, cptrs, rows, cvals, nov, threads);
}
template <class T>
double parallel_skip_perman64_w_balanced(int* rptrs, int* cols, T* rvals, int* cptrs, int* rows,
                                         T* cvals, int nov, int threads) {
  return perman_detail::cpu_skip_balanced(rptrs, cols, rvals, cptrs, rows, cvals, nov, threads);
}

#endif  // PERMAN_HPP
