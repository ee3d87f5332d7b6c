// perman_batch.cpp -- batched permanents of mixed n on one GPU (C ABI in
// include/perman_hip.h, perm_batch_*).
//
// A batch holds many matrices resident on the device.  Each matrix is split
// into items of one wave (64 lanes) walking 2^w Gray subsets per lane, with w
// chosen from the batch's item size (2^item_log2 subsets per item) and n
// alone -- so a matrix's items, and therefore its result, do not depend on
// what else is in the batch.  Items are grouped by kernel variant (n and
// walk shape);
// one persistent kernel per variant drains its group, the groups running
// concurrently on several streams.  The whole launch sequence (counter reset,
// variant kernels, per-matrix reduction, result copy) is captured once in a
// hipGraph at the second run and replayed from then on (the first run after a
// build launches directly: a batch computed once pays no capture).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/perman_hip.h"
#include "perman_internal.hpp"
#include "ryser_batch.hpp"
#include "ryser_host.hpp"
#include "ryser_kernels.hpp"

namespace {

using perm::BatchItem;
using perm::BatchMatrix;
using perm::Partial;

constexpr int kMaxStreams = 8;

struct Entry {
  int n = 0, variant = -1, walk_bits = 0, high_bits = 0;
  perm::NestShape shape{0, 0, 0};
  bool integral = false;
  bool zero = false;  // no perfect matching: exactly 0, no items
  double product_bits = 1e9;
  size_t col_off = 0, x0_off = 0;  // offsets (doubles) into the data buffer
  uint32_t first_item = 0, item_count = 0;
};

}  // namespace

struct perm_batch {
  int device = 0;
  int item_log2 = 17;   // subsets per item (64 lanes x 2^(item_log2-6) steps)
  int streams = 8;
  int max_grid = 1024;  // workgroups per variant kernel
  std::vector<double> data;
  std::vector<Entry> ents;
  std::vector<BatchItem> items;
  struct Group {
    int vid;
    uint32_t begin, end;
    double work;
  };
  std::vector<Group> groups;
  // device
  void* dmem = nullptr;
  size_t dbytes = 0;
  double* d_data = nullptr;
  BatchMatrix* d_mats = nullptr;
  BatchItem* d_items = nullptr;
  Partial* d_partials = nullptr;
  Partial* d_results = nullptr;
  uint32_t* d_counters = nullptr;
  Partial* h_results = nullptr;  // pinned
  hipStream_t s[kMaxStreams] = {};
  hipEvent_t ev_fork = nullptr, ev_join[kMaxStreams] = {}, ev0 = nullptr, ev1 = nullptr;
  hipGraphExec_t graph = nullptr;
  bool built = false;
  bool ran = false;  // one direct run since the build: the next run captures the graph
  // sampled verification: copies of the chosen items and their records
  void* vmem = nullptr;
  size_t vbytes = 0;
  // caller[m]: the caller's index of entry m, when entries are not in the
  // caller's order (perm_batch_compute_dense); empty: the same order
  std::vector<uint32_t> caller;
};

namespace {

void release_device(perm_batch* B) {
  if (B->graph) (void)hipGraphExecDestroy(B->graph);
  B->graph = nullptr;
  B->ran = false;
  if (B->dmem) (void)hipFree(B->dmem);
  B->dmem = nullptr;
  if (B->h_results) (void)hipHostFree(B->h_results);
  B->h_results = nullptr;
  if (B->vmem) (void)hipFree(B->vmem);
  B->vmem = nullptr;
  B->vbytes = 0;
  B->built = false;
}


// Fill entry e from the prepared matrix and append its column store and
// start vector to B->data.
int fill_entry(perm_batch* B, Entry& e, int n, perm::Prepared& p) {
  if (n > perm::kMaxBatchN)
    return perm::fail(PERM_ERR_SIZE, "batch matrices are limited to n <= " +
                                         std::to_string(perm::kMaxBatchN));
  e = Entry{};
  e.n = n;
  e.integral = p.integral;
  e.zero = p.zero;
  e.product_bits = p.product_bits;
  // the cheapest batch kernel shape covering the prepared one (extra rows
  // add zeros, exactly)
  e.shape = perm::batch_shape(n, {p.nu > 0 ? p.nu : n, p.nu1 > 0 ? p.nu1 : n,
                                  p.nu2 > 0 ? p.nu2 : n});
  e.variant = perm::batch_variant(n, e.shape);
  if (e.variant < 0) return perm::fail(PERM_ERR_ARG, "no batch kernel for this n");
  e.col_off = B->data.size();
  B->data.insert(B->data.end(), p.cs.begin(), p.cs.end());
  e.x0_off = B->data.size();
  B->data.insert(B->data.end(), p.x0.begin(), p.x0.end());
  while (B->data.size() % 32) B->data.push_back(0.0);  // 256-B alignment of the next store
  return PERM_OK;
}

int add_prepared(perm_batch* B, int n, perm::Prepared& p, int* index) {
  Entry e;
  if (int st = fill_entry(B, e, n, p)) return st;
  // (after perm_batch_compute_dense the entries are in processing order;
  // a matrix added later keeps the caller's insertion index)
  if (!B->caller.empty()) B->caller.push_back((uint32_t)B->ents.size());
  B->ents.push_back(e);
  B->built = false;
  if (index) *index = (int)B->ents.size() - 1;
  return perm::ok();
}

#define HIP_TRYB(expr)                                                                      \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return perm::fail(PERM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));   \
  } while (0)

size_t align256(size_t v) { return (v + 255) / 256 * 256; }

// Enqueue one run of the batch on its streams: reset -> fork -> variant
// kernels over streams -> join -> reduce -> result copy.  Groups come in
// decreasing work; each goes to the stream with the least work so far
// (longest-processing-time first), so the streams end close together instead
// of one long tail of round-robin leftovers.
int enqueue_run(perm_batch* B) {
  hipStream_t s0 = B->s[0];
  const size_t nm = B->ents.size();
  const int ns = std::max(1, std::min(B->streams, kMaxStreams));
  std::vector<double> stream_work(ns, 0.0);
  HIP_TRYB(perm::launch_batch_reset(B->d_counters, (uint32_t)B->groups.size(), s0));
  HIP_TRYB(hipEventRecord(B->ev_fork, s0));
  for (int i = 1; i < ns; ++i) HIP_TRYB(hipStreamWaitEvent(B->s[i], B->ev_fork, 0));
  for (size_t gi = 0; gi < B->groups.size(); ++gi) {
    const auto& G = B->groups[gi];
    const uint32_t waves = G.end - G.begin;
    const uint32_t grid = std::max(1u, std::min<uint32_t>((waves + 3) / 4, B->max_grid));
    const int si = (int)(std::min_element(stream_work.begin(), stream_work.end()) -
                         stream_work.begin());
    stream_work[si] += G.work;
    HIP_TRYB(perm::launch_batch_variant(G.vid, B->d_mats, B->d_items, G.begin, G.end,
                                        B->d_counters + gi, B->d_partials, grid, B->s[si]));
  }
  for (int i = 1; i < ns; ++i) {
    HIP_TRYB(hipEventRecord(B->ev_join[i], B->s[i]));
    HIP_TRYB(hipStreamWaitEvent(s0, B->ev_join[i], 0));
  }
  HIP_TRYB(perm::launch_batch_reduce(B->d_mats, B->d_partials, B->d_results, (uint32_t)nm, s0));
  HIP_TRYB(hipMemcpyAsync(B->h_results, B->d_results, nm * sizeof(Partial), hipMemcpyDeviceToHost,
                          s0));
  return PERM_OK;
}

// Capture one run into B->graph.  The capture is ended on every path, so a
// failure while enqueueing leaves the stream usable for the next call.
int capture(perm_batch* B) {
  hipStream_t s0 = B->s[0];
  hipGraph_t g = nullptr;
  HIP_TRYB(hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal));
  const int st = enqueue_run(B);
  const hipError_t ce = hipStreamEndCapture(s0, &g);
  if (st != PERM_OK || ce != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    if (st != PERM_OK) return st;
    return perm::fail(PERM_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
  }
  const hipError_t ie = hipGraphInstantiate(&B->graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ie != hipSuccess)
    return perm::fail(PERM_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
  return PERM_OK;
}

// Plan items and upload.
// Append the items of entries `idx` to B->items, grouped by variant, the
// groups in decreasing work, each group a contiguous item range.
void plan_items(perm_batch* B, const std::vector<uint32_t>& idx) {
  std::vector<std::vector<uint32_t>> by_var;
  std::vector<int> var_ids;
  for (uint32_t m : idx) {
    Entry& e = B->ents[m];
    perm::batch_split(e.n, B->item_log2, e.walk_bits, e.high_bits);
    e.first_item = 0;
    e.item_count = 0;
    if (e.zero) continue;  // no walk: its result is exactly 0
    auto it = std::find(var_ids.begin(), var_ids.end(), e.variant);
    if (it == var_ids.end()) {
      var_ids.push_back(e.variant);
      by_var.emplace_back();
      it = var_ids.end() - 1;
    }
    by_var[it - var_ids.begin()].push_back(m);
  }
  std::vector<size_t> order(var_ids.size());
  std::vector<double> work(var_ids.size(), 0.0);
  for (size_t v = 0; v < var_ids.size(); ++v) {
    order[v] = v;
    for (uint32_t m : by_var[v]) {
      const Entry& e = B->ents[m];
      work[v] += std::ldexp((double)perm::nest_cost(e.n, e.shape), e.n - 4);
    }
  }
  std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return work[a] > work[b]; });
  for (size_t oi : order) {
    perm_batch::Group g{var_ids[oi], (uint32_t)B->items.size(), 0, work[oi]};
    for (uint32_t m : by_var[oi]) {
      Entry& e = B->ents[m];
      const uint64_t lanes = 1ull << e.high_bits;
      const uint64_t cnt = (lanes + perm::kItemLanes - 1) / perm::kItemLanes;
      e.first_item = (uint32_t)B->items.size();
      e.item_count = (uint32_t)cnt;
      for (uint64_t i = 0; i < cnt; ++i) B->items.push_back({m, 0u, i * perm::kItemLanes});
    }
    g.end = (uint32_t)B->items.size();
    B->groups.push_back(g);
  }
}

// Device memory for `data` doubles, nm matrices, `items` items and `groups`
// variant groups: data, matrices, items, item records, results, counters
// (one per group, then one per matrix for double-double passes), and the
// pinned result copy.
int alloc_device(perm_batch* B, size_t data, size_t nm, size_t items, size_t groups) {
  const size_t data_b = align256(data * sizeof(double));
  const size_t mats_b = align256(nm * sizeof(BatchMatrix));
  const size_t items_b = align256(items * sizeof(BatchItem));
  const size_t part_b = align256(items * sizeof(Partial));
  const size_t res_b = align256(nm * sizeof(Partial));
  const size_t ctr_b = align256((groups + nm) * sizeof(uint32_t));
  B->dbytes = data_b + mats_b + items_b + part_b + res_b + ctr_b;
  HIP_TRYB(hipMalloc(&B->dmem, B->dbytes));
  char* p = (char*)B->dmem;
  B->d_data = (double*)p;
  B->d_mats = (BatchMatrix*)(p + data_b);
  B->d_items = (BatchItem*)(p + data_b + mats_b);
  B->d_partials = (Partial*)(p + data_b + mats_b + items_b);
  B->d_results = (Partial*)(p + data_b + mats_b + items_b + part_b);
  B->d_counters = (uint32_t*)(p + data_b + mats_b + items_b + part_b + res_b);
  HIP_TRYB(hipHostMalloc((void**)&B->h_results, nm * sizeof(Partial), hipHostMallocDefault));
  return PERM_OK;
}

BatchMatrix device_matrix(const perm_batch* B, const Entry& e) {
  return {B->d_data + e.col_off, B->d_data + e.x0_off, 1ull << e.high_bits, e.walk_bits,
          e.high_bits, e.variant, e.first_item, e.item_count, 0u};
}

int build(perm_batch* B) {
  HIP_TRYB(hipSetDevice(B->device));
  release_device(B);
  const size_t nm = B->ents.size();
  if (nm == 0) return perm::fail(PERM_ERR_ARG, "empty batch");
  B->items.clear();
  B->groups.clear();
  {
    std::vector<uint32_t> all(nm);
    for (size_t m = 0; m < nm; ++m) all[m] = (uint32_t)m;
    plan_items(B, all);
  }
  if (B->items.size() > 0xffffffffull) return perm::fail(PERM_ERR_SIZE, "too many items");
  if (int st = alloc_device(B, B->data.size(), nm, B->items.size(), B->groups.size())) return st;
  std::vector<BatchMatrix> mats(nm);
  for (size_t m = 0; m < nm; ++m) mats[m] = device_matrix(B, B->ents[m]);
  hipStream_t s0 = B->s[0];
  HIP_TRYB(hipMemcpyAsync(B->d_data, B->data.data(), B->data.size() * sizeof(double),
                          hipMemcpyHostToDevice, s0));
  HIP_TRYB(hipMemcpyAsync(B->d_mats, mats.data(), nm * sizeof(BatchMatrix), hipMemcpyHostToDevice,
                          s0));
  HIP_TRYB(hipMemcpyAsync(B->d_items, B->items.data(), B->items.size() * sizeof(BatchItem),
                          hipMemcpyHostToDevice, s0));
  HIP_TRYB(hipStreamSynchronize(s0));
  B->built = true;  // the graph is captured at the second run (a batch run once needs none)
  return perm::ok();
}

// Full double-double X passes (mode 2) of matrices of at most
// kFineMaxItems items walk a 2^kFineBits times finer split: lanes of
// 2^(walk_bits - kFineBits) steps, 2^kFineBits times the items.  A coarse
// item lasts about a millisecond in that mode, so a pass of a few hundred
// items was one item per SIMD with most of the GPU idle; the fine items
// spread it over the whole GPU.  The result is the matrix's in a batch of
// 2^kFineBits times smaller items, with the error model of that split, and
// depends only on the matrix.  Fine passes are launched in chunks of at
// most kFineChunkItems items (bounded scratch memory).
constexpr int kFineBits = 4;
constexpr uint32_t kFineMaxItems = 4096;
constexpr size_t kFineChunkItems = size_t(1) << 20;

// One round of the escalation (finish_run): the sampled error measurement
// of the matrices `ms` and the passes in a more accurate mode of the
// matrices `esc` (disjoint sets), launched together and finished with one
// synchronisation (fine passes beyond the first chunk in further launch
// sets).  pass_ms: device time.
//
// Sampled error measurement (perm::pps_sample) of the FP64 results of the
// matrices `ms` (non-integer, at least kVerifyMinUnits items): for each, up
// to kVerifyUnits of its items, drawn in proportion to their modelled
// variance, are walked again in mode 2 (copies of the items in a separate
// list, records in a separate array), and err_est becomes the estimate from
// their measured errors (perm::sampled_error); values are unchanged.
// Escalation: for each (matrix, mode) a persistent kernel of that mode's
// variant over the matrix's items (double-double X runs the plain walk,
// finer for small matrices as above), spread over the streams; then the
// per-matrix reduction and the results of those matrices refilled in the
// new mode.
int run_round(perm_batch* B, const std::vector<size_t>& ms,
              const std::vector<std::pair<size_t, int>>& esc, perm_result* results,
              float& pass_ms) {
  pass_ms = 0.f;
  hipStream_t s0 = B->s[0];
  const int ns = std::max(1, std::min(B->streams, kMaxStreams));
  // the sample: the FP64 item records of the sampled matrices only (not the
  // batch's), the picks, and copies of the picked items
  std::vector<std::vector<Partial>> fp(ms.size());
  for (size_t k = 0; k < ms.size(); ++k) {
    const Entry& e = B->ents[ms[k]];
    fp[k].resize(e.item_count);
    HIP_TRYB(hipMemcpyAsync(fp[k].data(), B->d_partials + e.first_item,
                            e.item_count * sizeof(Partial), hipMemcpyDeviceToHost, s0));
  }
  if (!ms.empty()) HIP_TRYB(hipStreamSynchronize(s0));
  // A picked item is walked on the fine split when its matrix has one (the
  // item's 64 lanes of 2^w steps are fine items 16 i .. 16 i + 15, lanes of
  // 2^(w-4) steps): its double-double value is then the sum of those 16
  // records in order.
  struct Pick {
    std::vector<uint32_t> ids;
    std::vector<double> weight;
    uint32_t certain = 0, begin = 0;
    int fine_walk = 0;  // > 0: walked on the fine split of this many walk bits
  };
  constexpr uint32_t kSub = 1u << kFineBits;
  std::vector<Pick> picks(ms.size());
  std::vector<BatchItem> vitems;  // coarse picks
  std::vector<BatchMatrix> sfmats;  // fine picks: one fine matrix per sampled matrix
  std::vector<BatchItem> sfitems;
  for (size_t k = 0; k < ms.size(); ++k) {
    const Entry& e = B->ents[ms[k]];
    std::vector<double> z(e.item_count);
    for (uint32_t i = 0; i < e.item_count; ++i) {
      const Partial& p = fp[k][i];
      z[i] = perm::unit_weight(p.sq_sum, p.sq_run, e.walk_bits, e.high_bits);
    }
    perm::pps_sample(z, perm::kVerifyUnits, picks[k].ids, picks[k].weight, picks[k].certain);
    const int w = std::max(std::min(perm::dense_unroll_bits(e.n), e.walk_bits),
                           e.walk_bits - kFineBits);
    if (w == e.walk_bits - kFineBits && e.high_bits >= 6) {
      // (high_bits >= 6: every coarse item is 64 whole lanes)
      picks[k].fine_walk = w;
      picks[k].begin = (uint32_t)sfitems.size();
      BatchMatrix bm = device_matrix(B, e);
      bm.lanes = 1ull << (e.high_bits + kFineBits);
      bm.walk_bits = w;
      bm.high_bits = e.high_bits + kFineBits;
      bm.first_item = picks[k].begin;
      bm.item_count = (uint32_t)picks[k].ids.size() * kSub;
      const uint32_t fm = (uint32_t)sfmats.size();
      sfmats.push_back(bm);
      for (uint32_t i : picks[k].ids)
        for (uint32_t t = 0; t < kSub; ++t)
          sfitems.push_back({fm, 0u, ((uint64_t)i * kSub + t) * perm::kItemLanes});
    } else {
      picks[k].begin = (uint32_t)vitems.size();
      for (uint32_t i : picks[k].ids) vitems.push_back(B->items[e.first_item + i]);
    }
  }
  const size_t n_sfi = sfitems.size();
  // the passes: coarse (the matrix's own items) or fine (its own split)
  struct Fine {
    size_t i;  // index in esc
    int walk, high;
    uint32_t count;
  };
  std::vector<size_t> coarse;
  std::vector<Fine> fine;
  for (size_t i = 0; i < esc.size(); ++i) {
    const Entry& e = B->ents[esc[i].first];
    const int w = std::max(std::min(perm::dense_unroll_bits(e.n), e.walk_bits),
                           e.walk_bits - kFineBits);
    if (esc[i].second != 2 || e.item_count > kFineMaxItems || w >= e.walk_bits) {
      coarse.push_back(i);
      continue;
    }
    const int h = e.walk_bits + e.high_bits - w;
    const uint64_t cnt = ((1ull << h) + perm::kItemLanes - 1) / perm::kItemLanes;
    fine.push_back({i, w, h, (uint32_t)cnt});
  }
  // launch sets: the first holds the sample, the coarse passes and the first
  // chunk of fine passes; each further set one more chunk
  std::vector<std::pair<size_t, size_t>> chunks;  // [begin, end) in fine
  for (size_t f = 0; f < fine.size();) {
    size_t g = f, items = 0;
    while (g < fine.size() && (g == f || items + fine[g].count <= kFineChunkItems))
      items += fine[g++].count;
    chunks.push_back({f, g});
    f = g;
  }
  const size_t n_sets = std::max<size_t>(1, chunks.size());
  uint32_t* ctr = B->d_counters + B->groups.size();
  const uint32_t nm = (uint32_t)B->ents.size();
  std::vector<Partial> ref(vitems.size()), sfref(n_sfi);
  for (size_t set = 0; set < n_sets; ++set) {
    const bool first = set == 0;
    const size_t f0 = chunks.empty() ? 0 : chunks[set].first;
    const size_t f1 = chunks.empty() ? 0 : chunks[set].second;
    // this set's fine matrices (device copies of the entry with the fine
    // split) and items
    std::vector<BatchMatrix> fmats;
    std::vector<BatchItem> fitems;
    if (first) {  // the fine picks first: their fine matrices, then items [0, n_sfi)
      fmats = sfmats;
      fitems = sfitems;
    }
    const size_t fe0 = fmats.size();  // this set's fine passes: fine matrices [fe0, ...)
    for (size_t f = f0; f < f1; ++f) {
      const Entry& e = B->ents[esc[fine[f].i].first];
      BatchMatrix bm = device_matrix(B, e);
      bm.lanes = 1ull << fine[f].high;
      bm.walk_bits = fine[f].walk;
      bm.high_bits = fine[f].high;
      bm.first_item = (uint32_t)fitems.size();
      bm.item_count = fine[f].count;
      const uint32_t fm = (uint32_t)fmats.size();
      fmats.push_back(bm);
      for (uint32_t k = 0; k < fine[f].count; ++k)
        fitems.push_back({fm, 0u, (uint64_t)k * perm::kItemLanes});
    }
    const size_t nv = first ? vitems.size() : 0;
    const size_t n_coarse = first ? coarse.size() : 0;
    const size_t n_samp = first ? ms.size() : 0;
    const size_t n_pass = fmats.size() - fe0;
    if (nv == 0 && n_coarse == 0 && fmats.empty()) continue;
    // scratch: sample items and records, fine matrices, items, records and
    // results
    const size_t vi_b = align256(nv * sizeof(BatchItem));
    const size_t vp_b = align256(nv * sizeof(Partial));
    const size_t fm_b = align256(fmats.size() * sizeof(BatchMatrix));
    const size_t fi_b = align256(fitems.size() * sizeof(BatchItem));
    const size_t fp_b = align256(fitems.size() * sizeof(Partial));
    const size_t fr_b = align256(n_pass * sizeof(Partial));
    const size_t need = vi_b + vp_b + fm_b + fi_b + fp_b + fr_b;
    if (need > B->vbytes) {
      if (B->vmem) HIP_TRYB(hipFree(B->vmem));
      B->vmem = nullptr;
      B->vbytes = 0;
      HIP_TRYB(hipMalloc(&B->vmem, need));
      B->vbytes = need;
    }
    char* vm = (char*)B->vmem;
    BatchItem* d_vitems = (BatchItem*)vm;
    Partial* d_vparts = (Partial*)(vm + vi_b);
    BatchMatrix* d_fmats = (BatchMatrix*)(vm + vi_b + vp_b);
    BatchItem* d_fitems = (BatchItem*)(vm + vi_b + vp_b + fm_b);
    Partial* d_fparts = (Partial*)(vm + vi_b + vp_b + fm_b + fi_b);
    Partial* d_fres = (Partial*)(vm + vi_b + vp_b + fm_b + fi_b + fp_b);
    // one counter per launched kernel (every matrix is in at most one list)
    const size_t n_ctr = n_samp + n_coarse + n_pass;
    if (n_ctr > nm) return perm::fail(PERM_ERR_ARG, "escalation round counters");
    HIP_TRYB(hipEventRecord(B->ev0, s0));
    if (nv)
      HIP_TRYB(hipMemcpyAsync(d_vitems, vitems.data(), nv * sizeof(BatchItem),
                              hipMemcpyHostToDevice, s0));
    if (!fmats.empty()) {
      HIP_TRYB(hipMemcpyAsync(d_fmats, fmats.data(), fmats.size() * sizeof(BatchMatrix),
                              hipMemcpyHostToDevice, s0));
      HIP_TRYB(hipMemcpyAsync(d_fitems, fitems.data(), fitems.size() * sizeof(BatchItem),
                              hipMemcpyHostToDevice, s0));
    }
    HIP_TRYB(perm::launch_batch_reset(ctr, (uint32_t)n_ctr, s0));
    HIP_TRYB(hipEventRecord(B->ev_fork, s0));
    for (int i = 1; i < ns; ++i) HIP_TRYB(hipStreamWaitEvent(B->s[i], B->ev_fork, 0));
    size_t j = 0;  // kernel index: counter and stream
    for (size_t k = 0; k < n_samp; ++k, ++j) {
      const Entry& e = B->ents[ms[k]];
      const bool fine_k = picks[k].fine_walk > 0;
      const uint32_t cnt = (uint32_t)picks[k].ids.size() * (fine_k ? kSub : 1u);
      if (cnt == 0) continue;
      const int vid = perm::batch_variant(e.n, perm::NestShape{e.n, e.n, e.n}, 2);
      if (vid < 0) return perm::fail(PERM_ERR_ARG, "no double-double X batch kernel for this n");
      const uint32_t grid = std::max(1u, std::min<uint32_t>((cnt + 3) / 4, B->max_grid));
      if (fine_k)
        HIP_TRYB(perm::launch_batch_variant(vid, d_fmats, d_fitems, picks[k].begin,
                                            picks[k].begin + cnt, ctr + j, d_fparts, grid,
                                            B->s[j % ns]));
      else
        HIP_TRYB(perm::launch_batch_variant(vid, B->d_mats, d_vitems, picks[k].begin,
                                            picks[k].begin + cnt, ctr + j, d_vparts, grid,
                                            B->s[j % ns]));
    }
    for (size_t c = 0; c < n_coarse; ++c, ++j) {
      const Entry& e = B->ents[esc[coarse[c]].first];
      const int md = esc[coarse[c]].second;
      const perm::NestShape sh = md == 2 ? perm::NestShape{e.n, e.n, e.n} : e.shape;
      const int vid = perm::batch_variant(e.n, sh, md);
      if (vid < 0) return perm::fail(PERM_ERR_ARG, "no batch kernel of this mode for this n");
      const uint32_t grid = std::max(1u, std::min<uint32_t>((e.item_count + 3) / 4, B->max_grid));
      HIP_TRYB(perm::launch_batch_variant(vid, B->d_mats, B->d_items, e.first_item,
                                          e.first_item + e.item_count, ctr + j, B->d_partials,
                                          grid, B->s[j % ns]));
    }
    for (size_t fm = fe0; fm < fmats.size(); ++fm, ++j) {
      const Entry& e = B->ents[esc[fine[f0 + fm - fe0].i].first];
      const int vid = perm::batch_variant(e.n, perm::NestShape{e.n, e.n, e.n}, 2);
      if (vid < 0) return perm::fail(PERM_ERR_ARG, "no double-double X batch kernel for this n");
      const uint32_t cnt = fmats[fm].item_count;
      const uint32_t grid = std::max(1u, std::min<uint32_t>((cnt + 3) / 4, B->max_grid));
      HIP_TRYB(perm::launch_batch_variant(vid, d_fmats, d_fitems, fmats[fm].first_item,
                                          fmats[fm].first_item + cnt, ctr + j, d_fparts, grid,
                                          B->s[j % ns]));
    }
    for (int i = 1; i < ns; ++i) {
      HIP_TRYB(hipEventRecord(B->ev_join[i], B->s[i]));
      HIP_TRYB(hipStreamWaitEvent(s0, B->ev_join[i], 0));
    }
    if (n_coarse) {
      HIP_TRYB(perm::launch_batch_reduce(B->d_mats, B->d_partials, B->d_results, nm, s0));
      HIP_TRYB(hipMemcpyAsync(B->h_results, B->d_results, nm * sizeof(Partial),
                              hipMemcpyDeviceToHost, s0));
    }
    std::vector<Partial> fres(n_pass);
    if (n_pass) {
      HIP_TRYB(perm::launch_batch_reduce(d_fmats + fe0, d_fparts, d_fres, (uint32_t)n_pass, s0));
      HIP_TRYB(hipMemcpyAsync(fres.data(), d_fres, n_pass * sizeof(Partial),
                              hipMemcpyDeviceToHost, s0));
    }
    if (first && n_sfi)
      HIP_TRYB(hipMemcpyAsync(sfref.data(), d_fparts, n_sfi * sizeof(Partial),
                              hipMemcpyDeviceToHost, s0));
    if (nv)
      HIP_TRYB(hipMemcpyAsync(ref.data(), d_vparts, nv * sizeof(Partial), hipMemcpyDeviceToHost,
                              s0));
    HIP_TRYB(hipEventRecord(B->ev1, s0));
    HIP_TRYB(hipStreamSynchronize(s0));
    float set_ms = 0.f;
    HIP_TRYB(hipEventElapsedTime(&set_ms, B->ev0, B->ev1));
    pass_ms += set_ms;
    for (size_t c = 0; c < n_coarse; ++c) {
      const auto& em = esc[coarse[c]];
      const Entry& e = B->ents[em.first];
      const perm::NestShape sh = em.second == 2 ? perm::NestShape{e.n, e.n, e.n} : e.shape;
      perm::fill_item_result(e.n, e.walk_bits, e.integral, e.product_bits, sh, em.second,
                             B->h_results[em.first], results[em.first]);
    }
    for (size_t fm = 0; fm < n_pass; ++fm) {
      const Fine& f = fine[f0 + fm];
      const Entry& e = B->ents[esc[f.i].first];
      perm::fill_item_result(e.n, f.walk, e.integral, e.product_bits,
                             perm::NestShape{e.n, e.n, e.n}, 2, fres[fm], results[esc[f.i].first]);
    }
  }
  for (size_t k = 0; k < ms.size(); ++k) {
    const Entry& e = B->ents[ms[k]];
    const Pick& pk = picks[k];
    if (pk.ids.empty()) continue;
    std::vector<double> err(pk.ids.size());
    for (size_t j = 0; j < pk.ids.size(); ++j) {
      const Partial& f = fp[k][pk.ids[j]];
      perm::HostDD r{0.0, 0.0};
      if (pk.fine_walk > 0) {
        for (uint32_t t = 0; t < kSub; ++t) {
          const Partial& q = sfref[pk.begin + j * kSub + t];
          r = perm::host_dd_add(r, {q.hi, q.lo});
        }
      } else {
        r = {ref[pk.begin + j].hi, ref[pk.begin + j].lo};
      }
      err[j] = (f.hi - r.hi) + (f.lo - r.lo);
    }
    perm_result& o = results[ms[k]];
    o.err_est = perm::sampled_error(err, pk.weight, pk.certain, e.n, o.value);
    o.certified = perm::is_certified(o) ? 1 : 0;
  }
  return PERM_OK;
}

// Results of the main run (ms: its device time) in entry order, the
// escalation rounds, then the results in the caller's order.
int finish_run(perm_batch* B, perm_result* out, float ms, double* device_ms) {
  const size_t nm = B->ents.size();
  std::vector<perm_result> res(nm);
  perm_result* results = res.data();
  for (size_t m = 0; m < nm; ++m) {
    const Entry& e = B->ents[m];
    if (e.zero) {
      perm::fill_zero_result(e.n, e.integral, results[m]);
      continue;
    }
    perm::fill_item_result(e.n, e.walk_bits, e.integral, e.product_bits, e.shape, 0,
                           B->h_results[m], results[m]);
    results[m].kernel_ms = ms;
  }
  double total_ms = ms;
  // Results that miss the accuracy target (perm::escalation_mode) get a pass
  // in the more accurate mode over the same items: double-double products
  // for integer matrices (same walk shape), double-double X for the others
  // (plain walk); one persistent kernel per matrix on the batch's streams,
  // then the per-matrix reduction again.  At most three rounds: a sampled
  // measurement, then passes whose mode only increases.
  // A non-integer matrix of many items first has its FP64 error measured on
  // a sample of its items (run_round); it gets the full pass only if that
  // does not certify it.
  std::vector<char> sampled(nm, 0);
  for (int round = 0; round < 3; ++round) {
    std::vector<std::pair<size_t, int>> esc;
    std::vector<size_t> sample;
    for (size_t m = 0; m < nm; ++m) {
      const Entry& e = B->ents[m];
      if (e.zero) continue;
      const int md = perm::escalation_mode(results[m]);
      if (md < 0) continue;
      if (md == 2 && results[m].mode == 0 && !sampled[m] && e.item_count >= perm::kVerifyMinUnits) {
        sample.push_back(m);
        sampled[m] = 1;
      } else {
        esc.push_back({m, md});
      }
    }
    if (esc.empty() && sample.empty()) break;
    float pass_ms = 0.f;
    // PERMAN_BATCH_TRACE=1: each round's passes on stderr (diagnostics)
    static const bool trace = perm::env_int("PERMAN_BATCH_TRACE", 0) != 0;
    std::vector<double> est0;
    if (trace)
      for (size_t m : sample) est0.push_back(results[m].err_est / std::fabs(results[m].value));
    if (int s = run_round(B, sample, esc, results, pass_ms)) return s;
    total_ms += pass_ms;
    if (trace) {
      std::fprintf(stderr, "[batch] main %.3f ms; round %d: %zu sampled, %zu escalated, %.3f ms\n",
                   ms, round, sample.size(), esc.size(), pass_ms);
      for (size_t k = 0; k < sample.size(); ++k)
        std::fprintf(stderr, "[batch]   sampled matrix %zu (n %d): err_est/|value| %.3g -> %.3g\n",
                     sample[k], B->ents[sample[k]].n, est0[k],
                     results[sample[k]].err_est / std::fabs(results[sample[k]].value));
    }
  }
  if (device_ms) *device_ms = total_ms;
  for (size_t m = 0; m < nm; ++m) {
    results[m].kernel_ms = total_ms;
    out[B->caller.empty() ? m : B->caller[m]] = results[m];
  }
  return perm::ok();
}

}  // namespace

namespace perm {

void fill_item_result(int n, int walk_bits, bool integral, double product_bits, NestShape shape,
                      int mode, const Partial& r, perm_result& o) {
  const int high_bits = n - 1 - walk_bits;
  std::memset(&o, 0, sizeof(o));
  const double scale = (double)(4 * (n & 1) - 2);
  o.n = n;
  o.sum_hi = r.hi;
  o.sum_lo = r.lo;
  o.value = scale * (r.hi + r.lo);
  o.abs_sum = std::fabs(scale) * r.abs_sum;
  o.err_bound =
      error_statement(o.abs_sum, o.value, n, walk_bits, dense_unroll_bits(n), integral, mode);
  o.err_est = error_estimate(scale * scale * r.sq_sum, scale * scale * r.sq_run, o.value, n,
                             walk_bits, high_bits, mode, integral);
  o.subsets = std::ldexp(1.0, n - 1);
  o.integral = integral ? 1 : 0;
  o.nu = shape.k0;
  o.nu1 = shape.k1;
  o.nu2 = shape.k2;
  o.mode = mode;
  if (integer_sum_exact(integral, product_bits, o.abs_sum, scale, n))
    o.err_bound = o.err_est = 0.0;  // exact arithmetic throughout
  round_integral(integral, o.value, o.err_bound, o.err_est);
  o.status = PERM_OK;
  o.certified = is_certified(o) ? 1 : 0;
}

bool needs_dd_pass(const perm_result& o) {
  return o.integral && o.mode == 0 && o.err_bound >= 0.5 &&
         std::fabs(o.value) < 9007199254740992.0;
}

void batch_split(int n, int item_log2, int& walk, int& high) {
  const int B = n - 1;
  const int R = dense_unroll_bits(n);
  int w = std::min(item_log2 - 6, B - 6);
  if (w < R) w = std::min(R, B);
  walk = w;
  high = B - w;
}
}  // namespace perm

extern "C" {

int perm_batch_create(perm_batch** out) {
  if (!out) return perm::fail(PERM_ERR_ARG, "null pointer");
  if (int s = perm::select_device()) return s;
  auto* B = new perm_batch();
  B->device = perm::current_device();
  hipError_t e = hipSuccess;
  for (int i = 0; i < kMaxStreams && e == hipSuccess; ++i)
    e = hipStreamCreateWithFlags(&B->s[i], hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&B->ev_fork, hipEventDisableTiming);
  for (int i = 0; i < kMaxStreams && e == hipSuccess; ++i)
    e = hipEventCreateWithFlags(&B->ev_join[i], hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreate(&B->ev0);
  if (e == hipSuccess) e = hipEventCreate(&B->ev1);
  if (e != hipSuccess) {
    perm_batch_destroy(B);
    return perm::fail(PERM_ERR_HIP, std::string("batch streams: ") + hipGetErrorString(e));
  }
  *out = B;
  return perm::ok();
}

int perm_batch_set(perm_batch* B, int item_log2, int streams, int max_grid) {
  if (!B) return perm::fail(PERM_ERR_ARG, "null batch");
  if (item_log2 > 0) {
    if (item_log2 < 6 || item_log2 > 30) return perm::fail(PERM_ERR_ARG, "item_log2 in [6, 30]");
    B->item_log2 = item_log2;
  }
  if (streams > 0) B->streams = std::min(streams, kMaxStreams);
  if (max_grid > 0) B->max_grid = max_grid;
  B->built = false;
  return perm::ok();
}

int perm_batch_add_dense(perm_batch* B, int dtype, const void* mat, int nov, int* index) {
  if (!B) return perm::fail(PERM_ERR_ARG, "null batch");
  perm::Prepared p;
  if (int s = perm::prepare_dense(dtype, mat, nov, p, perm::kMenuBatch)) return s;
  return add_prepared(B, nov, p, index);
}

int perm_batch_add_sparse(perm_batch* B, int dtype, const void* mat, const int* cptrs,
                          const int* rows, const void* cvals, int nov, int* index) {
  if (!B) return perm::fail(PERM_ERR_ARG, "null batch");
  perm::Prepared p;
  if (int s = perm::prepare_sparse(dtype, mat, cptrs, rows, cvals, nov, p, perm::kMenuBatch)) return s;
  return add_prepared(B, nov, p, index);
}

int perm_batch_add_dense_many(perm_batch* B, int count, const int* dtypes,
                              const void* const* mats, const int* novs, int threads,
                              int* first_index) {
  if (!B || count < 0 || (count > 0 && (!dtypes || !mats || !novs)))
    return perm::fail(PERM_ERR_ARG, "null pointer");
  std::vector<perm::Prepared> prep(count);
  std::vector<int> status(count, PERM_OK);
  std::vector<std::string> msg(count);
  if (threads < 1) threads = 1;
  threads = std::min(threads, std::max(1, count));
  // host preparation on `threads` threads, matrices in fixed strides (each
  // thread's status lives in its own thread-local slot; collected below)
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t)
    pool.emplace_back([&, t] {
      for (int i = t; i < count; i += threads) {
        if (novs[i] > perm::kMaxBatchN) {
          status[i] = PERM_ERR_SIZE;
          msg[i] = "batch matrices are limited to n <= " + std::to_string(perm::kMaxBatchN);
          continue;
        }
        status[i] = perm::prepare_dense(dtypes[i], mats[i], novs[i], prep[i], perm::kMenuBatch);
        if (status[i] != PERM_OK) msg[i] = perm::last_error();
      }
    });
  for (auto& th : pool) th.join();
  for (int i = 0; i < count; ++i)
    if (status[i] != PERM_OK)
      return perm::fail(status[i], "matrix " + std::to_string(i) + ": " + msg[i]);
  if (first_index) *first_index = (int)B->ents.size();
  for (int i = 0; i < count; ++i)
    if (int s = add_prepared(B, novs[i], prep[i], nullptr)) return s;
  return perm::ok();
}

int perm_batch_size(const perm_batch* B) { return B ? (int)B->ents.size() : 0; }

int perm_batch_run(perm_batch* B, perm_result* results, double* device_ms) {
  if (!B || !results) return perm::fail(PERM_ERR_ARG, "null pointer");
  if (!B->built)
    if (int s = build(B)) return s;
  HIP_TRYB(hipSetDevice(B->device));
  if (B->ran && !B->graph)
    if (int s = capture(B)) return s;
  HIP_TRYB(hipEventRecord(B->ev0, B->s[0]));
  if (B->graph) {
    HIP_TRYB(hipGraphLaunch(B->graph, B->s[0]));
  } else {  // first run after a build: launched directly
    if (int s = enqueue_run(B)) return s;
    B->ran = true;
  }
  HIP_TRYB(hipEventRecord(B->ev1, B->s[0]));
  HIP_TRYB(hipStreamSynchronize(B->s[0]));
  float ms = 0.f;
  HIP_TRYB(hipEventElapsedTime(&ms, B->ev0, B->ev1));
  return finish_run(B, results, ms, device_ms);
}

int perm_batch_compute_dense(perm_batch* B, int count, const int* dtypes,
                             const void* const* mats, const int* novs, int threads,
                             perm_result* results, double* device_ms) {
  // PERMAN_BATCH_TRACE=1: the host time of each stage on stderr (diagnostics)
  static const bool trace = perm::env_int("PERMAN_BATCH_TRACE", 0) != 0;
  const auto t_start = std::chrono::steady_clock::now();
  auto mark = [&](const char* what) {
    if (trace)
      std::fprintf(stderr, "[batch %.3f ms] %s\n",
                   std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() -
                                                             t_start).count(),
                   what);
  };
  if (!B || count <= 0 || !dtypes || !mats || !novs || !results)
    return perm::fail(PERM_ERR_ARG, "null pointer or empty input");
  if (!B->ents.empty()) return perm::fail(PERM_ERR_ARG, "the batch must be empty");
  for (int i = 0; i < count; ++i)
    if (novs[i] < 1 || novs[i] > perm::kMaxBatchN)
      return perm::fail(PERM_ERR_SIZE, "matrix " + std::to_string(i) +
                                           ": batch matrices are limited to 1 <= n <= " +
                                           std::to_string(perm::kMaxBatchN));
  HIP_TRYB(hipSetDevice(B->device));
  release_device(B);
  threads = std::max(1, std::min(threads, count));
  // processing order: decreasing n; three phases, cut at 15 % and 45 % of
  // the work: the first is small, so the GPU starts early, and each later
  // one is prepared while the earlier ones run
  std::vector<uint32_t> order((size_t)count);
  for (int i = 0; i < count; ++i) order[(size_t)i] = (uint32_t)i;
  std::stable_sort(order.begin(), order.end(),
                   [&](uint32_t a, uint32_t b) { return novs[a] > novs[b]; });
  auto work_of = [&](uint32_t i) { return std::ldexp((double)novs[i], novs[i] - 1); };
  double total = 0.0;
  for (int i = 0; i < count; ++i) total += work_of((uint32_t)i);
  size_t cut[2] = {0, 0};
  {
    const double frac[2] = {0.15, 0.45};
    size_t k = 0;
    double acc = 0.0;
    for (int c = 0; c < 2; ++c) {
      while (k < (size_t)count && acc < frac[c] * total) acc += work_of(order[k++]);
      cut[c] = std::max<size_t>(k, 1);
    }
  }
  // device memory from the orders alone (upper bounds: column store, start
  // vector and padding per matrix; every matrix's items)
  size_t data_ub = 0, items_ub = 0;
  for (int i = 0; i < count; ++i) {
    const size_t n = (size_t)novs[i];
    data_ub += (2 * n - 1) * n + 2 * n + 32;
    int w = 0, h = 0;
    perm::batch_split(novs[i], B->item_log2, w, h);
    items_ub += ((1ull << h) + perm::kItemLanes - 1) / perm::kItemLanes;
  }
  if (items_ub > 0xffffffffull) return perm::fail(PERM_ERR_SIZE, "too many items");
  B->ents.assign((size_t)count, Entry{});
  B->caller.assign(order.begin(), order.end());
  B->data.clear();
  B->data.reserve(data_ub);  // no reallocation: the phase uploads read from it
  B->items.clear();
  B->items.reserve(items_ub);
  B->groups.clear();
  auto undo = [&](int st) {
    (void)hipDeviceSynchronize();
    release_device(B);
    B->ents.clear();
    B->caller.clear();
    B->data.clear();
    B->items.clear();
    B->groups.clear();
    return st;
  };
  if (int st = alloc_device(B, data_ub, (size_t)count, items_ub, 3 * (size_t)count))
    return undo(st);
  mark("allocated");
  hipStream_t s0 = B->s[0];
  // the variant kernels go to `streams` streams of their own, s[1..nk]
  // (s0 carries the uploads, so a phase's copy does not wait behind the
  // earlier phases' kernels); with one stream fewer for the kernels than a
  // replay has, the kernels' tails overlapped less (3 instead of 4 in
  // flight: 69 against 66 ms of device time at the default split)
  const int nk = std::max(1, std::min(B->streams, kMaxStreams - 1));
  const int k0 = 1, ke = nk + 1;
  std::vector<double> stream_work((size_t)ke, 0.0);
  HIP_TRYB(hipEventRecord(B->ev0, s0));
  HIP_TRYB(perm::launch_batch_reset(B->d_counters, (uint32_t)(3 * count), s0));
  std::vector<perm::Prepared> prep((size_t)count);
  std::vector<int> status((size_t)count, PERM_OK);
  std::vector<std::string> msg((size_t)count);
  std::vector<BatchMatrix> dm((size_t)count);
  const size_t phases[4] = {0, cut[0], std::max(cut[0], cut[1]), (size_t)count};
  for (int ph = 0; ph < 3; ++ph) {
    const size_t p0 = phases[ph], p1 = phases[ph + 1];
    if (p0 == p1) continue;
    // prepare the phase's matrices on `threads` host threads (the first
    // phase's kernels run meanwhile)
    {
      const int t_n = std::max(1, std::min<int>(threads, (int)(p1 - p0)));
      std::vector<std::thread> pool;
      for (int t = 0; t < t_n; ++t)
        pool.emplace_back([&, t] {
          for (size_t k = p0 + (size_t)t; k < p1; k += (size_t)t_n) {
            const uint32_t i = order[k];
            status[k] = perm::prepare_dense(dtypes[i], mats[i], novs[i], prep[k], perm::kMenuBatch);
            if (status[k] != PERM_OK) msg[k] = perm::last_error();
          }
        });
      for (auto& th : pool) th.join();
    }
    mark(ph == 0 ? "phase 0 prepared" : ph == 1 ? "phase 1 prepared" : "phase 2 prepared");
    for (size_t k = p0; k < p1; ++k)
      if (status[k] != PERM_OK)
        return undo(perm::fail(status[k], "matrix " + std::to_string(order[k]) + ": " + msg[k]));
    const size_t d0 = B->data.size(), i0 = B->items.size(), g0 = B->groups.size();
    std::vector<uint32_t> idx;
    for (size_t k = p0; k < p1; ++k) {
      if (int st = fill_entry(B, B->ents[k], novs[order[k]], prep[k])) return undo(st);
      prep[k] = perm::Prepared();
      idx.push_back((uint32_t)k);
    }
    if (B->data.size() > data_ub) return undo(perm::fail(PERM_ERR_SIZE, "batch data bound"));
    plan_items(B, idx);
    for (size_t k = p0; k < p1; ++k) dm[k] = device_matrix(B, B->ents[k]);
    mark("planned");
    // this phase's data, matrices and items, then its variant kernels
    HIP_TRYB(hipMemcpyAsync(B->d_data + d0, B->data.data() + d0,
                            (B->data.size() - d0) * sizeof(double), hipMemcpyHostToDevice, s0));
    HIP_TRYB(hipMemcpyAsync(B->d_mats + p0, dm.data() + p0, (p1 - p0) * sizeof(BatchMatrix),
                            hipMemcpyHostToDevice, s0));
    if (B->items.size() > i0)
      HIP_TRYB(hipMemcpyAsync(B->d_items + i0, B->items.data() + i0,
                              (B->items.size() - i0) * sizeof(BatchItem), hipMemcpyHostToDevice,
                              s0));
    HIP_TRYB(hipEventRecord(B->ev_fork, s0));
    for (int i = k0; i < ke; ++i) HIP_TRYB(hipStreamWaitEvent(B->s[i], B->ev_fork, 0));
    for (size_t gi = g0; gi < B->groups.size(); ++gi) {
      const auto& G = B->groups[gi];
      const uint32_t waves = G.end - G.begin;
      const uint32_t grid = std::max(1u, std::min<uint32_t>((waves + 3) / 4, B->max_grid));
      const int si = (int)(std::min_element(stream_work.begin() + k0, stream_work.end()) -
                           stream_work.begin());  // in [k0, ke)
      stream_work[(size_t)si] += G.work;
      HIP_TRYB(perm::launch_batch_variant(G.vid, B->d_mats, B->d_items, G.begin, G.end,
                                          B->d_counters + gi, B->d_partials, grid, B->s[si]));
    }
    mark("launched");
  }
  for (int i = k0; i < ke; ++i) {
    HIP_TRYB(hipEventRecord(B->ev_join[i], B->s[i]));
    HIP_TRYB(hipStreamWaitEvent(s0, B->ev_join[i], 0));
  }
  HIP_TRYB(perm::launch_batch_reduce(B->d_mats, B->d_partials, B->d_results, (uint32_t)count, s0));
  HIP_TRYB(hipMemcpyAsync(B->h_results, B->d_results, (size_t)count * sizeof(Partial),
                          hipMemcpyDeviceToHost, s0));
  HIP_TRYB(hipEventRecord(B->ev1, s0));
  HIP_TRYB(hipStreamSynchronize(s0));
  mark("device done");
  float ms = 0.f;
  HIP_TRYB(hipEventElapsedTime(&ms, B->ev0, B->ev1));
  // the batch is built: later runs replay it (a captured graph from the
  // second perm_batch_run on), results in the caller's order
  B->built = true;
  B->ran = true;
  const int st = finish_run(B, results, ms, device_ms);
  mark("finished");
  return st;
}

// Autotuner of the batch's work split: item size (Gray-code chunk length per
// lane, hence lanes per matrix), streams (variant kernels in flight) and the
// grid of each persistent variant kernel (waves per variant, hence matrices
// in flight per CU).  Coordinate search from the current setting: item_log2
// over {12, 14, 16, 18, 20}, then streams over {1, 2, 4, 8}, then max_grid
// over {512, 1024, 2048, 4096}; each candidate is rebuilt (plan, upload,
// graph capture) and timed over `reps` graph replays (device time, HIP
// events); the fastest setting is kept.
int perm_batch_autotune(perm_batch* B, int reps, perm_batch_tuning* best) {
  if (!B || !best) return perm::fail(PERM_ERR_ARG, "null pointer");
  if (B->ents.empty()) return perm::fail(PERM_ERR_ARG, "empty batch");
  if (reps < 1) reps = 1;
  std::vector<perm_result> res(B->ents.size());
  int trials = 0;
  auto measure = [&](int il, int ns, int grid, double& ms_out) -> int {
    B->item_log2 = il;
    B->streams = ns;
    B->max_grid = grid;
    B->built = false;
    double ms = 0.0, sum = 0.0;
    if (int s = perm_batch_run(B, res.data(), &ms)) return s;  // build + warm
    for (int r = 0; r < reps; ++r) {
      if (int s = perm_batch_run(B, res.data(), &ms)) return s;
      sum += ms;
    }
    ms_out = sum / reps;
    ++trials;
    return PERM_OK;
  };
  int bi = B->item_log2, bs = B->streams, bg = B->max_grid;
  double bms = 0.0;
  if (int s = measure(bi, bs, bg, bms)) return s;
  for (int il : {12, 14, 16, 18, 20}) {
    if (il == bi) continue;
    double ms;
    if (int s = measure(il, bs, bg, ms)) return s;
    if (ms < bms) bms = ms, bi = il;
  }
  const int i0 = bi;
  for (int ns : {1, 2, 4, 8}) {
    if (ns == bs) continue;
    double ms;
    if (int s = measure(i0, ns, bg, ms)) return s;
    if (ms < bms) bms = ms, bs = ns;
  }
  const int s0 = bs;
  for (int grid : {512, 1024, 2048, 4096}) {
    if (grid == bg) continue;
    double ms;
    if (int s = measure(i0, s0, grid, ms)) return s;
    if (ms < bms) bms = ms, bg = grid;
  }
  B->item_log2 = bi;
  B->streams = bs;
  B->max_grid = bg;
  B->built = false;
  best->item_log2 = bi;
  best->streams = bs;
  best->max_grid = bg;
  best->device_ms = bms;
  best->trials = trials;
  return perm::ok();
}

int perm_batch_get(const perm_batch* B, perm_batch_tuning* out) {
  if (!B || !out) return perm::fail(PERM_ERR_ARG, "null pointer");
  out->item_log2 = B->item_log2;
  out->streams = B->streams;
  out->max_grid = B->max_grid;
  out->device_ms = 0.0;
  out->trials = 0;
  return perm::ok();
}

void perm_batch_destroy(perm_batch* B) {
  if (!B) return;
  (void)hipSetDevice(B->device);
  for (int i = 0; i < kMaxStreams; ++i)
    if (B->s[i]) (void)hipStreamSynchronize(B->s[i]);
  release_device(B);
  for (int i = 0; i < kMaxStreams; ++i) {
    if (B->s[i]) (void)hipStreamDestroy(B->s[i]);
    if (B->ev_join[i]) (void)hipEventDestroy(B->ev_join[i]);
  }
  if (B->ev_fork) (void)hipEventDestroy(B->ev_fork);
  if (B->ev0) (void)hipEventDestroy(B->ev0);
  if (B->ev1) (void)hipEventDestroy(B->ev1);
  delete B;
}

}  // extern "C"
