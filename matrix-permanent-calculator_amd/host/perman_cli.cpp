// perman_cli.cpp -- the `perman` command line (reference: main.cu:315-590,
// RunAlgo main.cu:20-248), backed by libperman_hip.so.
//
// Flags are the reference's (short options "bsr:t:f:gd:cap:x:y:z:im:n:" and
// the same long names, main.cu:337-357) with the same defaults: GPU unless
// -c is given without -g (main.cu:472-474), -p 1, -t 16.  Exact paths:
//   GPU dense  -p 0..4   gpu_perman64_{xglobal,xlocal,xshared,
//                        xshared_coalescing,xshared_coalescing_mshared}
//   GPU sparse -s -p 1..4  the same names with the _sparse suffix
//   GPU sparse -s -p 7     gpu_perman64_xshared_coalescing_mshared_skipper
//   GPU dense  -p 5, 66 / sparse -s -p 5, 66   the multi-GPU IDs: this build
//                        drives one GPU, which computes the whole range
//   GPU dense  -p 6 / sparse -s -p 6   the chunk scheduler; -g -c adds the
//                        CPU worker with -t threads (main.cu:62-66)
//   GPU sparse -s -p 8   the chunked skipper: the skipper on the GPU
//   CPU dense  -c        parallel_perman64 (any -p, main.cu:188-192)
//   CPU sparse -c -s -p1 parallel_perman64_sparse
//   CPU sparse -c -s -p2, -p3  parallel_skip_perman64_w / _w_balanced
// The result line is the reference's, `Result: <name> <perm> in <seconds>`
// in its stream format (6 significant digits; dense -p 0 prints the xlocal
// name and -p 66 its printf format, as main.cu:38 and :72 do), followed by
// `Value: <perm>` with 17 significant digits and the stated error of the
// value.  Approximation (-a) and grid graphs (-i) are not part of this build
// and exit with status 2.
#include <getopt.h>
#include <omp.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "matrix_io.hpp"
#include "perman_hip.h"

namespace {

// -p 0 runs gpu_perman64_xglobal; the reference's result line names it
// gpu_perman64_xlocal (main.cu:38), and so does this one
const char* kDenseNames[] = {"gpu_perman64_xlocal", "gpu_perman64_xlocal", "gpu_perman64_xshared",
                             "gpu_perman64_xshared_coalescing",
                             "gpu_perman64_xshared_coalescing_mshared"};
const char* kSparseNames[] = {nullptr, "gpu_perman64_xlocal_sparse",
                              "gpu_perman64_xshared_sparse",
                              "gpu_perman64_xshared_coalescing_sparse",
                              "gpu_perman64_xshared_coalescing_mshared_sparse"};

struct Options {
  bool binary = false, dense = true, approximation = false, gpu = false, cpu = false;
  bool grid = false;
  int gpu_num = 2, threads = 16, algo = 1, preprocessing = 0;
  int times = 100000, scale_intervals = 4, scale_times = 5, gridm = 36, gridn = 36;
  std::string file;
};

int dtype_of(const std::string& t) {
  if (t == "int") return PERM_DTYPE_I32;
  if (t == "float") return PERM_DTYPE_F32;
  return PERM_DTYPE_F64;
}

// The reference's result line, as `cout << ... << perman << " in " << secs`
// prints it (default stream format: %g, 6 significant digits; main.cu:35-58),
// so scripts that parse it read the same line; then the value with every
// digit (results are held to 1e-10 relative, which 6 digits cannot show).
void print_result(const char* name, double value, double secs) {
  std::printf("Result: %s %g in %g\n", name, value, secs);
  std::printf("Value: %.17g\n", value);
}

template <class T>
int run(const Options& o, std::ifstream& in, const permio::Header& h) {
  std::vector<T> mat;
  permio::read_entries(in, h.n, o.binary, mat);
  if (o.preprocessing == 1) permio::sort_order(mat, h.n);
  else if (o.preprocessing == 2) permio::skip_order(mat, h.n);
  const int n = h.n;
  const int dt = dtype_of(h.type);
  permio::Compressed<T> c = permio::compress(mat, n);

  if (o.gpu) {
    perm_result r;
    int st;
    const char* name;
    const double t0 = omp_get_wtime();
    int cpu_chunks = 0;
    if (o.dense) {
      if (o.algo >= 0 && o.algo <= 4) {
        name = kDenseNames[o.algo];
        st = perm_dense(dt, mat.data(), n, o.algo, &r);
      } else if (o.algo == 5 || o.algo == 66) {  // one GPU computes the whole range
        name = o.algo == 5 ? "gpu_perman64_xshared_coalescing_mshared_multigpu"
                           : "gpu_perman64_xshared_coalescing_mshared_multigpu_manual_distribution";
        st = perm_dense(dt, mat.data(), n, 4, &r);
      } else if (o.algo == 6) {  // chunk scheduler, CPU worker with -c (main.cu:62-66)
        name = "gpu_perman64_xshared_coalescing_mshared_multigpucpu_chunks";
        st = perm_chunked(dt, mat.data(), nullptr, nullptr, nullptr, n, o.cpu, o.threads, &r,
                          &cpu_chunks);
      } else {
        std::printf("Unknown Algorithm ID\n");
        return 2;
      }
    } else if (o.algo == 7 || o.algo == 8) {  // the skipper (main.cu:138-147)
      name = o.algo == 7 ? "gpu_perman64_xshared_coalescing_mshared_skipper"
                         : "gpu_perman64_xshared_coalescing_mshared_multigpucpu_chunks_skipper";
      st = perm_skip(dt, mat.data(), c.rptrs.data(), c.cols.data(), c.cptrs.data(),
                     c.rows.data(), c.cvals.data(), n, &r);
    } else if (o.algo >= 1 && o.algo <= 4) {
      name = kSparseNames[o.algo];
      st = perm_sparse(dt, mat.data(), c.cptrs.data(), c.rows.data(), c.cvals.data(), n, o.algo,
                       &r);
    } else if (o.algo == 5 || o.algo == 66) {
      name = o.algo == 5 ? "gpu_perman64_xshared_coalescing_mshared_multigpu_sparse"
                         : "gpu_perman64_xshared_coalescing_mshared_multigpu_sparse_manual_"
                           "distribution";
      st = perm_sparse(dt, mat.data(), c.cptrs.data(), c.rows.data(), c.cvals.data(), n, 4, &r);
    } else if (o.algo == 6) {
      name = "gpu_perman64_xshared_coalescing_mshared_multigpucpu_chunks_sparse";
      st = perm_chunked(dt, mat.data(), c.cptrs.data(), c.rows.data(), c.cvals.data(), n, o.cpu,
                        o.threads, &r, &cpu_chunks);
    } else {
      std::printf("Unknown Algorithm ID\n");
      return 2;
    }
    const double t1 = omp_get_wtime();
    if (st != PERM_OK) {
      std::fprintf(stderr, "perman: %s\n", perm_last_error());
      return 1;
    }
    std::printf("kernel in %.6f\n", r.kernel_ms / 1e3);
    if (cpu_chunks > 0) std::printf("CPU worker: %d chunk(s)\n", cpu_chunks);
    if (o.dense && o.algo == 66) {  // main.cu:72 prints this one with printf
      std::printf("Result: %s %2lf in %lf\n", name, r.value, t1 - t0);
      std::printf("Value: %.17g\n", r.value);
    } else {
      print_result(name, r.value, t1 - t0);
    }
    std::printf("Error: bound %.3e estimate %.3e (sum of |terms| %.6e, mode %d, %s)\n",
                r.err_bound, r.err_est, r.abs_sum, r.mode,
                r.zero_pattern ? "no perfect matching: exactly 0"
                : r.certified  ? "certified to 1e-10"
                               : "not certified to 1e-10");
    return 0;
  }
  // CPU
  const double t0 = omp_get_wtime();
  double v;
  const char* name;
  if (o.dense) {
    name = "parallel_perman64";
    if constexpr (std::is_same<T, int>::value) v = perm_parallel_perman64_i32(mat.data(), n, o.threads);
    else if constexpr (std::is_same<T, float>::value)
      v = perm_parallel_perman64_f32(mat.data(), n, o.threads);
    else v = perm_parallel_perman64_f64(mat.data(), n, o.threads);
  } else {
    if (o.algo == 2 || o.algo == 3) {  // the skip walks (main.cu:215-224)
      name = o.algo == 2 ? "parallel_skip_perman64_w" : "parallel_skip_perman64_w_balanced";
      perm_result r;
      const int st = perm_cpu_skip(dt, c.rptrs.data(), c.cols.data(), c.rvals.data(),
                                   c.cptrs.data(), c.rows.data(), c.cvals.data(), n, o.threads,
                                   &r);
      const double t1 = omp_get_wtime();
      if (st != PERM_OK) {
        std::fprintf(stderr, "perman: %s\n", perm_last_error());
        return 1;
      }
      print_result(name, r.value, t1 - t0);
      std::printf("Error: bound %.3e estimate %.3e (terms evaluated %.6g of %.6g)\n",
                  r.err_bound, r.err_est, r.main_subsets, r.subsets);
      return 0;
    }
    if (o.algo != 1) {
      std::printf("Unknown Algorithm ID\n");
      return 2;
    }
    name = "parallel_perman64_sparse";
    if constexpr (std::is_same<T, int>::value)
      v = perm_parallel_perman64_sparse_i32(mat.data(), c.cptrs.data(), c.rows.data(),
                                            c.cvals.data(), n, o.threads);
    else if constexpr (std::is_same<T, float>::value)
      v = perm_parallel_perman64_sparse_f32(mat.data(), c.cptrs.data(), c.rows.data(),
                                            c.cvals.data(), n, o.threads);
    else
      v = perm_parallel_perman64_sparse_f64(mat.data(), c.cptrs.data(), c.rows.data(),
                                            c.cvals.data(), n, o.threads);
  }
  const double t1 = omp_get_wtime();
  if (std::isnan(v)) {
    std::fprintf(stderr, "perman: invalid matrix\n");
    return 1;
  }
  print_result(name, v, t1 - t0);
  return 0;
}

bool needs_arg(const char* opt) {
  if (optarg && optarg[0] == '-') {
    std::fprintf(stderr, "Option -%s requires an argument.\n", opt);
    return true;
  }
  return false;
}

}  // namespace

int main(int argc, char** argv) {
  Options o;
  const char* short_opts = "bsr:t:f:gd:cap:x:y:z:im:n:";
  const struct option long_opts[] = {
      {"binary", 0, nullptr, 'b'},         {"sparse", 0, nullptr, 's'},
      {"preprocessing", 1, nullptr, 'r'},  {"threads", 1, nullptr, 't'},
      {"file", 1, nullptr, 'f'},           {"gpu", 0, nullptr, 'g'},
      {"device", 1, nullptr, 'd'},         {"cpu", 0, nullptr, 'c'},
      {"approximation", 0, nullptr, 'a'},  {"perman", 1, nullptr, 'p'},
      {"numOfTimes", 1, nullptr, 'x'},     {"scaleIntervals", 1, nullptr, 'y'},
      {"scaleTimes", 1, nullptr, 'z'},     {"grid", 0, nullptr, 'i'},
      {"gridm", 1, nullptr, 'm'},          {"gridn", 1, nullptr, 'n'},
      {nullptr, 0, nullptr, 0}};
  int c;
  while ((c = getopt_long(argc, argv, short_opts, long_opts, nullptr)) != -1) {
    switch (c) {
      case 'b': o.binary = true; break;
      case 's': o.dense = false; break;
      case 'r': if (needs_arg("r")) return 1; o.preprocessing = std::atoi(optarg); break;
      case 't': if (needs_arg("t")) return 1; o.threads = std::atoi(optarg); break;
      case 'f': if (needs_arg("f")) return 1; o.file = optarg; break;
      case 'a': o.approximation = true; break;
      case 'g': o.gpu = true; break;
      case 'd': if (needs_arg("d")) return 1; o.gpu_num = std::atoi(optarg); break;
      case 'c': o.cpu = true; break;
      case 'p': if (needs_arg("p")) return 1; o.algo = std::atoi(optarg); break;
      case 'x': if (needs_arg("x")) return 1; o.times = std::atoi(optarg); break;
      case 'y': if (needs_arg("y")) return 1; o.scale_intervals = std::atoi(optarg); break;
      case 'z': if (needs_arg("z")) return 1; o.scale_times = std::atoi(optarg); break;
      case 'i': o.grid = true; break;
      case 'm': if (needs_arg("m")) return 1; o.gridm = std::atoi(optarg); break;
      case 'n': if (needs_arg("n")) return 1; o.gridn = std::atoi(optarg); break;
      default: return 1;
    }
  }
  if (!o.grid && o.file.empty()) {
    std::fprintf(stderr, "Option -f is a required argument.\n");
    return 1;
  }
  for (int i = optind; i < argc; ++i) std::printf("Non-option argument %s\n", argv[i]);
  if (!o.cpu && !o.gpu) o.gpu = true;
  // with -g, -c only enables the CPU worker of the chunked IDs (main.cu:31, :62)
  if (o.grid || o.approximation) {
    std::fprintf(stderr, "perman: %s is not part of this build (exact Ryser paths only)\n",
                 o.grid ? "grid-graph mode (-i)" : "approximation (-a)");
    return 2;
  }
  std::ifstream in(o.file);
  permio::Header h;
  if (!in || !permio::read_header(in, h) || h.n < 1) {
    std::fprintf(stderr, "perman: cannot read matrix header from %s\n", o.file.c_str());
    return 1;
  }
  if (h.type == "int") return run<int>(o, in, h);
  if (h.type == "float") return run<float>(o, in, h);
  if (h.type == "double") return run<double>(o, in, h);
  std::fprintf(stderr, "perman: unknown element type '%s'\n", h.type.c_str());
  return 1;
}
