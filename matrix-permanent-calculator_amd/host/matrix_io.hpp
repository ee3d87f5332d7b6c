// matrix_io.hpp -- the reference's matrix text format and sparse structures
// (host side of the `perman` CLI).
//
//  * read_matrix: header line `n nnz type` (main.cu:484-488), then lines
//    `i j value` parsed as the element type T; lines that do not parse are
//    skipped; `binary` (-b) stores 1 for every listed entry (util.h:343-358).
//  * compress: CCS (cptrs/rows/cvals) and CRS (rptrs/cols/rvals) keeping only
//    entries > 0, ascending indices (util.h:522-551).
//  * sort_order (-r1): columns reordered by ascending nonzero count, ties in
//    original order; the dense matrix is rewritten in the new column order
//    (util.h:553-619).
//  * skip_order (-r2): greedy row/column ordering -- repeatedly take the
//    remaining column of smallest live degree, append its unvisited rows,
//    and lower the degrees of the columns those rows touch; the matrix is
//    permuted accordingly (util.h:621-684).
// Row and column permutations leave the permanent unchanged.
#pragma once
#include <algorithm>
#include <climits>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

namespace permio {

struct Header {
  int n = 0;
  int nnz = 0;
  std::string type;
};

inline bool read_header(std::ifstream& in, Header& h) {
  std::string line;
  if (!std::getline(in, line)) return false;
  std::istringstream iss(line);
  return static_cast<bool>(iss >> h.n >> h.nnz >> h.type);
}

template <class T>
void read_entries(std::ifstream& in, int n, bool binary, std::vector<T>& mat) {
  mat.assign((size_t)n * n, T(0));
  std::string line;
  while (std::getline(in, line)) {
    std::istringstream iss(line);
    int i, j;
    T v;
    if (!(iss >> i >> j >> v)) continue;
    if (i < 0 || j < 0 || i >= n || j >= n) continue;
    mat[(size_t)i * n + j] = binary ? T(1) : v;
  }
}

template <class T>
struct Compressed {
  std::vector<int> cptrs, rows, rptrs, cols;
  std::vector<T> cvals, rvals;
};

template <class T>
Compressed<T> compress(const std::vector<T>& mat, int n) {
  Compressed<T> c;
  c.cptrs.assign(n + 1, 0);
  c.rptrs.assign(n + 1, 0);
  for (int k = 0; k < n; ++k) {
    c.cptrs[k] = (int)c.rows.size();
    for (int i = 0; i < n; ++i)
      if (mat[(size_t)i * n + k] > 0) {
        c.rows.push_back(i);
        c.cvals.push_back(mat[(size_t)i * n + k]);
      }
  }
  c.cptrs[n] = (int)c.rows.size();
  for (int i = 0; i < n; ++i) {
    c.rptrs[i] = (int)c.cols.size();
    for (int k = 0; k < n; ++k)
      if (mat[(size_t)i * n + k] > 0) {
        c.cols.push_back(k);
        c.rvals.push_back(mat[(size_t)i * n + k]);
      }
  }
  c.rptrs[n] = (int)c.cols.size();
  return c;
}

// Rewrite mat with rows rp[r] and columns cp[c] moved to positions r, c.
template <class T>
void permute(std::vector<T>& mat, int n, const std::vector<int>& rp, const std::vector<int>& cp) {
  std::vector<T> old = mat;
  for (int r = 0; r < n; ++r)
    for (int c = 0; c < n; ++c) mat[(size_t)r * n + c] = old[(size_t)rp[r] * n + cp[c]];
}

template <class T>
void sort_order(std::vector<T>& mat, int n) {
  std::vector<int> cnt(n, 0), order(n);
  for (int k = 0; k < n; ++k) {
    order[k] = k;
    for (int i = 0; i < n; ++i)
      if (mat[(size_t)i * n + k] > 0) cnt[k]++;
  }
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cnt[a] < cnt[b]; });
  std::vector<int> rp(n);
  for (int i = 0; i < n; ++i) rp[i] = i;
  // Entries <= 0 are dropped from the rewritten matrix (it is rebuilt from
  // the CCS of the positive entries).
  for (auto& v : mat)
    if (!(v > 0)) v = T(0);
  permute(mat, n, rp, order);
}

template <class T>
void skip_order(std::vector<T>& mat, int n) {
  std::vector<int> deg(n, 0), rp, cp;
  std::vector<char> seen(n, 0);
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < n; ++k)
      if (mat[(size_t)i * n + k] != 0) deg[k]++;
  const int kDone = INT8_MAX;  // marks a column already placed
  for (int step = 0; step < n; ++step) {
    int best = -1, bd = kDone;
    for (int k = 0; k < n; ++k)
      if (deg[k] < bd) {
        bd = deg[k];
        best = k;
      }
    if (best < 0) {  // every remaining degree reached the marker value
      for (int k = 0; k < n; ++k)
        if (std::find(cp.begin(), cp.end(), k) == cp.end()) {
          best = k;
          break;
        }
    }
    deg[best] = kDone;
    cp.push_back(best);
    for (int i = 0; i < n; ++i)
      if (mat[(size_t)i * n + best] != 0 && !seen[i]) {
        seen[i] = 1;
        rp.push_back(i);
        for (int k = 0; k < n; ++k)
          if (mat[(size_t)i * n + k] != 0 && deg[k] != kDone) deg[k]--;
      }
  }
  for (int i = 0; i < n; ++i)  // rows never reached (empty columns) keep their order
    if (!seen[i]) rp.push_back(i);
  permute(mat, n, rp, cp);
}

}  // namespace permio
