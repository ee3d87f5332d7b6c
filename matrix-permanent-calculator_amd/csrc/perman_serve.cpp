// perman_serve.cpp -- a permanent server on one GPU: continuous batching of
// requests of mixed n (C ABI in include/perman_hip.h, perm_server_*).
//
// Callers submit matrices from any thread and wait on (or poll) a ticket.
// A scheduler thread owns the GPU work:
//
//  * submission copies the caller's matrix and returns a ticket; a pool of
//    prep threads prepares it (start vector, column store, walk shape);
//  * admission: a prepared request gets a device slot for its column store
//    and start vector (a fixed arena, so admission is bounded by device
//    memory) while the queued work stays under a cap;
//  * iterations: each iteration takes 64-lane items (the batch kernels'
//    items, ryser_batch.hip) from the admitted requests, shortest remaining
//    work first (requests older than max_wait_ms ahead of that, oldest
//    first), up to a work budget of a few milliseconds, and launches one
//    persistent kernel per variant (n and walk shape) present over several
//    streams;
//    up to `inflight` iterations are queued on the GPU at once, so the next
//    one is ready when the previous drains;
//  * completion: an iteration's item partials come back to pinned memory
//    and are stored per request by item index; a request whose items are
//    all done is summed on the host in item order (double-double) -- so its
//    result depends only on the matrix, bitwise the same alone or in any mix
//    and whatever iterations its items ran in;
//  * a result that misses the accuracy target goes back to the active list
//    for a pass in a double-double mode (integer matrices: double-double
//    products, exact below 2^53; others: double-double X), its matrix still
//    in its arena slot;
//  * a matrix whose nonzero pattern has no perfect matching is exactly 0 and
//    finishes at submission.
//
// Iterations change every time, so they are launched directly (no graph).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <queue>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
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
using Clock = std::chrono::steady_clock;

constexpr int kMaxStreams = 8;
constexpr int kMaxInflight = 4;
// Request mode of the sampled verification pass (mode 2 kernels on a sample
// of the FP64 pass's items).
constexpr int kVerifyPass = 3;
// FP64 VALU lane-instructions per second of one MI355X at full issue
// (256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz): converts the iteration time
// target into a work budget.
constexpr double kLaneInstrPerSec = 256.0 * 4 * 16 * 2.4e9;
// Slot of one resident matrix: column store (2(n-1)+1) n and x0 with its
// low parts (2n), for n <= kMaxBatchN, rounded up to 256 bytes.
constexpr size_t kSlotDoubles =
    ((size_t)(2 * (perm::kMaxBatchN - 1) + 1) * perm::kMaxBatchN + 2 * perm::kMaxBatchN + 31) /
    32 * 32;

// A submitted matrix as the caller gave it (copied), prepared by the pool.
struct RawMatrix {
  int dtype = PERM_DTYPE_F64;
  bool sparse = false;
  std::vector<char> mat, cvals;
  std::vector<int> cptrs, rows;
};

struct Request {
  uint64_t id = 0;
  RawMatrix raw;  // until prepared
  int n = 0, variant = -1, walk_bits = 0, high_bits = 0;
  int mode = 0;  // 0 FP64 items; 1, 2 a pass in a double-double mode (escalation);
                 // kVerifyPass: sampled items in mode 2 measuring the FP64 error
  perm::NestShape shape{0, 0, 0};
  bool integral = false;
  perm::Prepared prep;
  int slot = -1;
  uint32_t items = 0, issued = 0, done = 0;
  uint32_t walk_items = 0;  // items of the whole walk (items: of the current pass)
  double item_cost = 0.0;  // lane-instructions per item (scheduling only)
  std::vector<Partial> parts;
  // sampled verification: the FP64 pass's result and item records, the
  // sampled item indices, their weights and how many were certain
  bool sampled = false;
  perm_result fp_result{};
  std::vector<Partial> fp_parts;
  std::vector<uint32_t> vids;
  std::vector<double> vweight;
  uint32_t vcertain = 0;
  Clock::time_point t_submit, t_done;
  perm_result result{};
  int status = PERM_OK;
  std::string error;
  bool finished = false;
  double remaining_cost() const { return (double)(items - issued) * item_cost; }
};

struct Segment {
  Request* r;
  uint32_t first, count, pos;  // items [first, first+count) of r at iteration position pos
};

struct Iteration {
  BatchMatrix* h_mats = nullptr;
  Partial* h_parts = nullptr;
  BatchMatrix* d_mats = nullptr;
  Partial* d_parts = nullptr;
  uint32_t* d_counters = nullptr;
  void* dmem = nullptr;
  void* hmem = nullptr;
  hipEvent_t start = nullptr, done = nullptr;  // timed: the iteration's GPU span
  std::vector<Segment> segs;
  bool busy = false;
};

}  // namespace

struct perm_server {
  perm_server_config cfg{};
  int device = 0;
  size_t max_items = 0;
  double iter_cost = 0.0;
  std::thread worker;
  std::mutex mu;
  std::condition_variable cv_submit, cv_done;
  std::deque<Request*> to_prep;                  // submitted, not prepared (mu)
  std::deque<Request*> pending;                  // prepared, not admitted (mu)
  std::condition_variable cv_prep;
  std::vector<std::thread> preppers;
  int prepping = 0;                              // requests taken by a prep thread (mu)
  std::unordered_map<uint64_t, Request*> table;  // every uncollected request (mu)
  uint64_t next_id = 1;
  bool stop = false;
  int fatal = PERM_OK;  // scheduler error (mu): every later call fails with it
  std::string fatal_msg;
  perm_server_stats stats{};  // (mu)
  // scheduler-owned
  std::vector<Request*> active;
  // free arena slots, smallest first (requests admitted together get
  // neighbouring slots, uploaded by one copy)
  std::priority_queue<int, std::vector<int>, std::greater<int>> free_slots;
  double queued_cost = 0.0;
  // work of the admitted items not yet issued (the sum of remaining_cost()
  // over the active list), kept as items are admitted, requeued and issued
  double unissued_cost = 0.0;
  // scheduler time per phase (ms) and upload copies: written by the
  // scheduler only, read by perm_server_stats_get
  std::atomic<double> t_admit{0.0}, t_issue{0.0}, t_retire{0.0}, t_wait{0.0}, t_gpu{0.0};
  std::atomic<long long> n_uploads{0};
  std::atomic<long long> n_dd{0};  // requests given a double-double pass
  double* d_arena = nullptr;
  double* h_arena = nullptr;  // pinned mirror of the arena (upload staging)
  Iteration it[kMaxInflight];
  int head = 0, n_busy = 0;  // oldest busy iteration, number busy
  hipStream_t s0 = nullptr, s[kMaxStreams] = {};
  hipEvent_t ev_fork = nullptr, ev_join[kMaxStreams] = {};
};

namespace {

#define HIP_TRYS(expr)                                                                      \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return perm::fail(PERM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));   \
  } while (0)

size_t align256(size_t v) { return (v + 255) / 256 * 256; }


// Publish a retired iteration's finished requests under one lock, one wake-up.
void publish_all(perm_server* S, const std::vector<Request*>& rs) {
  if (rs.empty()) return;
  const Clock::time_point now = Clock::now();
  {
    std::lock_guard<std::mutex> lk(S->mu);
    for (Request* r : rs) {
      r->finished = true;
      r->t_done = now;
    }
    S->stats.completed += (long long)rs.size();
  }
  S->cv_done.notify_all();
}

// FP64 VALU lane-instructions per subset of a pass in a double-double mode
// (scheduling estimate only).  Mode 1 over walk shape s: the FP64 adds,
// three instructions per TwoProd multiply, about 8 for the TwoSum
// accumulation.  Mode 2 (plain walk): a TwoSum per column addition (7), a
// double-double multiply step per factor (4), the accumulation (8).
double dd_subset_cost(int n, perm::NestShape s, int mode) {
  if (mode == 2) return 11.0 * n + 8.0;
  const double half = (perm::nest_cost(n, s) / 8.0 - 2.0) / 2.0;  // adds = multiplies
  return half + 3.0 * half + 8.0;
}

// Queue r again for a pass in `mode` over `items` items (the verification's
// sample or the whole walk), its matrix still in its slot.
void requeue(perm_server* S, Request* r, int mode, uint32_t items) {
  r->mode = mode;
  const int kmode = mode == kVerifyPass ? 2 : mode;
  if (kmode == 2) r->shape = perm::NestShape{r->n, r->n, r->n};
  r->variant = perm::batch_variant(r->n, r->shape, kmode);
  r->items = items;
  r->issued = r->done = 0;
  r->parts.assign(items, Partial{0.0, 0.0, 0.0, 0.0, 0.0, 0.0});
  r->item_cost =
      std::ldexp((double)perm::kItemLanes, r->walk_bits) * dd_subset_cost(r->n, r->shape, kmode);
  S->queued_cost += (double)r->items * r->item_cost;
  S->unissued_cost += (double)r->items * r->item_cost;
  S->active.push_back(r);
  S->n_dd.store(S->n_dd.load(std::memory_order_relaxed) + 1, std::memory_order_relaxed);
}

// All items of r's current pass are done.  FP64 and double-double passes:
// sum its partials in item order and state the error.  A non-integer result
// of many items whose modelled estimate misses the accuracy target first has
// its error measured on a sample of its items (a kVerifyPass in mode 2, as
// perm::pps_sample / perm::sampled_error define it); a result that still
// misses the target, or misses it otherwise (perm::escalation_mode), goes
// back to the active list for a pass in the more accurate mode over the same
// slot -- double-double products (same walk shape) for an integer matrix,
// double-double X (plain walk) for any other -- and this returns false;
// otherwise the slot is freed and r is ready to publish (true).
bool finalize(perm_server* S, Request* r) {
  if (r->mode == kVerifyPass) {
    std::vector<double> e(r->vids.size());
    for (size_t k = 0; k < r->vids.size(); ++k) {
      const Partial& f = r->fp_parts[r->vids[k]];
      e[k] = (f.hi - r->parts[k].hi) + (f.lo - r->parts[k].lo);
    }
    r->result = r->fp_result;
    r->result.err_est = perm::sampled_error(e, r->vweight, r->vcertain, r->n, r->result.value);
    r->result.certified = perm::is_certified(r->result) ? 1 : 0;
    r->fp_parts.clear();
    r->fp_parts.shrink_to_fit();
    if (!r->result.certified) {
      requeue(S, r, 2, r->walk_items);
      return false;
    }
  } else {
    perm::HostDD s = {0.0, 0.0};
    Partial tot{0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (const Partial& p : r->parts) {
      s = perm::host_dd_add(s, perm::HostDD{p.hi, p.lo});
      tot.abs_sum += p.abs_sum;
      tot.sq_sum += p.sq_sum;
      tot.sq_run += p.sq_run;
    }
    tot.hi = s.hi;
    tot.lo = s.lo;
    perm::fill_item_result(r->n, r->walk_bits, r->integral, r->prep.product_bits, r->shape,
                           r->mode, tot, r->result);
    const int next = perm::escalation_mode(r->result);
    if (next == 2 && r->mode == 0 && !r->sampled && r->walk_items >= perm::kVerifyMinUnits) {
      std::vector<double> z(r->parts.size());
      for (size_t i = 0; i < z.size(); ++i)
        z[i] = perm::unit_weight(r->parts[i].sq_sum, r->parts[i].sq_run, r->walk_bits,
                                 r->high_bits);
      perm::pps_sample(z, perm::kVerifyUnits, r->vids, r->vweight, r->vcertain);
      if (!r->vids.empty()) {
        r->sampled = true;
        r->fp_result = r->result;
        r->fp_parts.swap(r->parts);
        requeue(S, r, kVerifyPass, (uint32_t)r->vids.size());
        return false;
      }
    }
    if (next >= 0) {
      requeue(S, r, next, r->walk_items);
      return false;
    }
  }
  r->parts.clear();
  r->parts.shrink_to_fit();
  S->free_slots.push(r->slot);
  r->slot = -1;
  r->prep = perm::Prepared();
  return true;
}

// Copy newly admitted requests into their arena slots: each into the pinned
// mirror, then one asynchronous copy per run of nearby slots on s0, ahead of
// any kernel that reads them.  A run may span up to kUploadGapSlots slots
// that are not fresh: a resident slot's mirror holds the bytes its device
// copy already has, a free slot's content is unused, so copying them over is
// harmless and cheaper than another copy call (a slot is ~26 KB).
constexpr int kUploadGapSlots = 4;

int upload(perm_server* S, std::vector<Request*>& fresh) {
  std::sort(fresh.begin(), fresh.end(),
            [](const Request* a, const Request* b) { return a->slot < b->slot; });
  for (size_t i = 0; i < fresh.size();) {
    size_t j = i;
    for (;;) {
      Request* r = fresh[j];
      double* h = S->h_arena + (size_t)r->slot * kSlotDoubles;
      std::memcpy(h, r->prep.cs.data(), r->prep.cs.size() * sizeof(double));
      std::memcpy(h + r->prep.cs.size(), r->prep.x0.data(), r->prep.x0.size() * sizeof(double));
      if (j + 1 == fresh.size() || fresh[j + 1]->slot > r->slot + 1 + kUploadGapSlots) break;
      ++j;
    }
    const Request* last = fresh[j];
    const size_t first = (size_t)fresh[i]->slot * kSlotDoubles;
    const size_t end = (size_t)last->slot * kSlotDoubles + last->prep.cs.size() + last->prep.x0.size();
    HIP_TRYS(hipMemcpyAsync(S->d_arena + first, S->h_arena + first, (end - first) * sizeof(double),
                            hipMemcpyHostToDevice, S->s0));
    S->n_uploads.store(S->n_uploads.load(std::memory_order_relaxed) + 1,
                       std::memory_order_relaxed);
    i = j + 1;
  }
  return PERM_OK;
}

// Admit pending requests in arrival order while slots are free and queued
// work stays under max_queued_work (the first one always, on an idle server);
// the pending queue is taken under one lock.
int admit(perm_server* S) {
  std::vector<Request*> fresh;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    size_t slots = S->free_slots.size();
    double queued = S->queued_cost;
    while (!S->pending.empty() && slots > 0) {
      Request* c = S->pending.front();
      const double cost = (double)c->items * c->item_cost;
      if (!(S->active.empty() && fresh.empty()) && queued + cost > S->cfg.max_queued_work) break;
      S->pending.pop_front();
      fresh.push_back(c);
      queued += cost;
      --slots;
    }
    S->stats.admitted += (long long)fresh.size();
  }
  if (fresh.empty()) return PERM_OK;
  for (Request* r : fresh) {
    r->slot = S->free_slots.top();
    S->free_slots.pop();
    r->parts.assign(r->items, Partial{0.0, 0.0, 0.0, 0.0, 0.0, 0.0});
    S->queued_cost += (double)r->items * r->item_cost;
    S->unissued_cost += (double)r->items * r->item_cost;
  }
  if (int st = upload(S, fresh)) return st;
  for (Request* r : fresh) S->active.push_back(r);
  return PERM_OK;
}

// Choose the next iteration's items and launch it on a free context.
int issue(perm_server* S, bool& issued) {
  issued = false;
  if (S->n_busy >= S->cfg.inflight) return PERM_OK;
  // continuous batching: while an iteration runs, let small work accumulate
  // into the next one instead of paying an iteration's launch cost for it
  // (checked on the running total, before any scan of the active list)
  if (S->n_busy > 0 && S->unissued_cost < S->iter_cost / 16.0) return PERM_OK;
  std::vector<Request*> cand;
  double work = 0.0;
  for (Request* r : S->active)
    if (r->issued < r->items) {
      cand.push_back(r);
      work += r->remaining_cost();
    }
  if (cand.empty()) {
    S->unissued_cost = 0.0;  // drop the rounding drift of the running total
    return PERM_OK;
  }
  S->unissued_cost = work;
  // shortest remaining work first, except that requests waiting longer than
  // max_wait_ms go ahead, oldest first (no large request starves behind a
  // stream of small ones)
  const Clock::time_point aged_before =
      Clock::now() - std::chrono::microseconds((int64_t)(S->cfg.max_wait_ms * 1e3));
  std::sort(cand.begin(), cand.end(), [&](const Request* a, const Request* b) {
    const bool oa = a->t_submit < aged_before, ob = b->t_submit < aged_before;
    if (oa != ob) return oa;
    if (oa) return a->id < b->id;
    const double ca = a->remaining_cost(), cb = b->remaining_cost();
    return ca != cb ? ca < cb : a->id < b->id;
  });
  Iteration& I = S->it[(S->head + S->n_busy) % S->cfg.inflight];
  I.segs.clear();
  double budget = S->iter_cost;
  size_t cap = S->max_items;
  for (Request* r : cand) {
    if (budget <= 0.0 || cap == 0) break;
    uint64_t take = r->items - r->issued;
    const uint64_t by_budget = (uint64_t)std::ceil(budget / r->item_cost);
    take = std::min<uint64_t>(take, std::max<uint64_t>(1, by_budget));
    take = std::min<uint64_t>(take, cap);
    I.segs.push_back({r, r->issued, (uint32_t)take, 0});
    r->issued += (uint32_t)take;
    budget -= (double)take * r->item_cost;
    S->unissued_cost -= (double)take * r->item_cost;
    cap -= take;
  }
  // group by variant: segments of one variant are contiguous in the item list
  std::stable_sort(I.segs.begin(), I.segs.end(), [](const Segment& a, const Segment& b) {
    return a.r->variant < b.r->variant;
  });
  struct Group {
    int vid;
    uint32_t begin, end;
  };
  std::vector<Group> groups;
  uint32_t pos = 0;
  // staging: matrix records, then (from the next 256-byte boundary) items
  const size_t mats_bytes = align256(I.segs.size() * sizeof(BatchMatrix));
  BatchItem* h_items = (BatchItem*)((char*)I.h_mats + mats_bytes);
  for (size_t m = 0; m < I.segs.size(); ++m) {
    Segment& g = I.segs[m];
    Request* r = g.r;
    g.pos = pos;
    I.h_mats[m] = {S->d_arena + (size_t)r->slot * kSlotDoubles,
                   S->d_arena + (size_t)r->slot * kSlotDoubles + r->prep.cs.size(),
                   1ull << r->high_bits, r->walk_bits, r->high_bits, r->variant, 0u, 0u, 0u};
    // the item's index in the walk (a verification pass runs the sampled ones)
    for (uint32_t i = 0; i < g.count; ++i) {
      const uint64_t item = r->mode == kVerifyPass ? r->vids[g.first + i] : g.first + i;
      h_items[pos + i] = {(uint32_t)m, 0u, item * perm::kItemLanes};
    }
    if (groups.empty() || groups.back().vid != r->variant)
      groups.push_back({r->variant, pos, pos + g.count});
    else
      groups.back().end = pos + g.count;
    pos += g.count;
  }
  const uint32_t n_items = pos;
  hipStream_t s0 = S->s0;
  // one upload: the matrix records, then the items
  BatchItem* d_items = (BatchItem*)((char*)I.d_mats + mats_bytes);
  HIP_TRYS(hipEventRecord(I.start, s0));
  HIP_TRYS(hipMemcpyAsync(I.d_mats, I.h_mats, mats_bytes + n_items * sizeof(BatchItem),
                          hipMemcpyHostToDevice, s0));
  HIP_TRYS(perm::launch_batch_reset(I.d_counters, (uint32_t)groups.size(), s0));
  // variants run concurrently on the streams they need (one group: on s0)
  const int ns = (int)std::min<size_t>(groups.size() > 1 ? groups.size() : 0,
                                       (size_t)std::max(1, std::min(S->cfg.streams, kMaxStreams)));
  if (ns > 0) {
    HIP_TRYS(hipEventRecord(S->ev_fork, s0));
    for (int i = 0; i < ns; ++i) HIP_TRYS(hipStreamWaitEvent(S->s[i], S->ev_fork, 0));
  }
  for (size_t gi = 0; gi < groups.size(); ++gi) {
    const Group& G = groups[gi];
    const uint32_t waves = G.end - G.begin;
    const uint32_t grid = std::max(1u, std::min<uint32_t>((waves + 3) / 4, 2048u));
    HIP_TRYS(perm::launch_batch_variant(G.vid, I.d_mats, d_items, G.begin, G.end,
                                        I.d_counters + gi, I.d_parts, grid,
                                        ns > 0 ? S->s[gi % ns] : s0));
  }
  for (int i = 0; i < ns; ++i) {
    HIP_TRYS(hipEventRecord(S->ev_join[i], S->s[i]));
    HIP_TRYS(hipStreamWaitEvent(s0, S->ev_join[i], 0));
  }
  HIP_TRYS(hipMemcpyAsync(I.h_parts, I.d_parts, n_items * sizeof(Partial),
                          hipMemcpyDeviceToHost, s0));
  HIP_TRYS(hipEventRecord(I.done, s0));
  I.busy = true;
  S->n_busy += 1;
  issued = true;
  std::lock_guard<std::mutex> lk(S->mu);
  S->stats.iterations += 1;
  S->stats.items += n_items;
  return PERM_OK;
}

// Retire the oldest iteration if it is done (wait = block until it is).
int retire(perm_server* S, bool wait, bool& retired) {
  retired = false;
  if (S->n_busy == 0) return PERM_OK;
  Iteration& I = S->it[S->head];
  if (wait) {
    HIP_TRYS(hipEventSynchronize(I.done));
  } else {
    const hipError_t q = hipEventQuery(I.done);
    if (q == hipErrorNotReady) return PERM_OK;
    if (q != hipSuccess)
      return perm::fail(PERM_ERR_HIP, std::string("hipEventQuery: ") + hipGetErrorString(q));
  }
  float gpu_ms = 0.f;
  if (hipEventElapsedTime(&gpu_ms, I.start, I.done) == hipSuccess)
    S->t_gpu.store(S->t_gpu.load(std::memory_order_relaxed) + gpu_ms, std::memory_order_relaxed);
  std::vector<Request*> completed, finished;
  for (const Segment& g : I.segs) {
    Request* r = g.r;
    std::copy(I.h_parts + g.pos, I.h_parts + g.pos + g.count, r->parts.begin() + g.first);
    r->done += g.count;
    S->queued_cost -= (double)g.count * r->item_cost;
    if (r->done == r->items) completed.push_back(r);
  }
  if (!completed.empty())  // drop them from the active list in one pass, first
    S->active.erase(std::remove_if(S->active.begin(), S->active.end(),
                                   [](const Request* r) { return r->done == r->items; }),
                    S->active.end());
  for (Request* r : completed)
    if (finalize(S, r)) finished.push_back(r);
  publish_all(S, finished);
  I.segs.clear();
  I.busy = false;
  S->head = (S->head + 1) % S->cfg.inflight;
  S->n_busy -= 1;
  retired = true;
  return PERM_OK;
}

void add_ms(std::atomic<double>& acc, Clock::time_point t0) {
  const double ms = std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
  acc.store(acc.load(std::memory_order_relaxed) + ms, std::memory_order_relaxed);
}

int scheduler_step(perm_server* S, bool& idle) {
  idle = false;
  Clock::time_point t0 = Clock::now();
  if (int st = admit(S)) return st;
  add_ms(S->t_admit, t0);
  t0 = Clock::now();
  bool issued = false, retired = false;
  if (int st = issue(S, issued)) return st;
  add_ms(S->t_issue, t0);
  t0 = Clock::now();
  if (int st = retire(S, false, retired)) return st;
  add_ms(S->t_retire, t0);
  if (issued || retired) return PERM_OK;
  if (S->n_busy >= S->cfg.inflight) {  // GPU queue full
    t0 = Clock::now();
    const int st = retire(S, true, retired);
    add_ms(S->t_wait, t0);
    return st;
  }
  if (S->n_busy > 0) {
    // nothing to issue: wait briefly for the GPU or a new request
    std::unique_lock<std::mutex> lk(S->mu);
    S->cv_submit.wait_for(lk, std::chrono::microseconds(100),
                          [&] { return S->stop || !S->pending.empty(); });
    return PERM_OK;
  }
  idle = true;
  return PERM_OK;
}

void scheduler_loop(perm_server* S) {
  for (;;) {
    bool idle = false;
    const int st = scheduler_step(S, idle);
    if (st != PERM_OK) {
      // a HIP failure leaves the device state unknown: fail everything
      std::lock_guard<std::mutex> lk(S->mu);
      S->fatal = st;
      S->fatal_msg = perm::last_error();
      for (auto& kv : S->table)
        if (!kv.second->finished) {
          kv.second->status = st;
          kv.second->error = S->fatal_msg;
          kv.second->finished = true;
        }
      S->cv_done.notify_all();
      return;
    }
    if (idle) {
      std::unique_lock<std::mutex> lk(S->mu);
      if (S->stop && S->to_prep.empty() && S->prepping == 0 && S->pending.empty() &&
          S->active.empty())
        return;
      S->cv_submit.wait(lk, [&] { return S->stop || !S->pending.empty(); });
    }
  }
}

void scheduler_main(perm_server* S) {
  perm::set_device(S->device);
  (void)hipSetDevice(S->device);
  scheduler_loop(S);
}

size_t dtype_size(int dtype) { return dtype == PERM_DTYPE_F64 ? 8 : 4; }

// Prepare r from its raw input (a prep thread, no lock held): start vector,
// column store, split, walk shape, variant.  Returns PERM_OK or an error
// status (r->error set).
int prepare_request(perm_server* S, Request* r) {
  const RawMatrix& in = r->raw;
  const int n = r->n;
  perm::Prepared p;
  int st = in.sparse ? perm::prepare_sparse(in.dtype, in.mat.data(), in.cptrs.data(),
                                            in.rows.data(), in.cvals.data(), n, p,
                                            perm::kMenuBatch)
                     : perm::prepare_dense(in.dtype, in.mat.data(), n, p, perm::kMenuBatch);
  if (st != PERM_OK) {
    r->error = perm::last_error();
    return st;
  }
  r->integral = p.integral;
  perm::batch_split(n, S->cfg.item_log2, r->walk_bits, r->high_bits);
  const uint64_t lanes = 1ull << r->high_bits;
  r->items = (uint32_t)((lanes + perm::kItemLanes - 1) / perm::kItemLanes);
  r->walk_items = r->items;
  // pair / nested kernels for requests with real work; tiny ones use the
  // plain variant (fewer kernels per iteration)
  r->shape = r->items >= 8 ? perm::batch_shape(n, {p.nu > 0 ? p.nu : n, p.nu1 > 0 ? p.nu1 : n,
                                                   p.nu2 > 0 ? p.nu2 : n})
                           : perm::NestShape{n, n, n};
  r->variant = perm::batch_variant(n, r->shape);
  if (r->variant < 0 || perm::batch_variant(n, r->shape, 1) < 0 ||
      perm::batch_variant(n, {n, n, n}, 2) < 0) {
    r->error = "no batch kernel for this n";
    return PERM_ERR_ARG;
  }
  r->item_cost = std::ldexp((double)perm::kItemLanes, r->walk_bits) *
                 (double)perm::nest_cost(n, r->shape) / 8.0;
  if (p.zero) perm::fill_zero_result(n, r->integral, r->result);
  r->prep = std::move(p);
  r->raw = RawMatrix();
  return PERM_OK;
}

// Prep thread: take submitted requests, prepare them, hand them to the
// scheduler (or finish them: exactly 0, or a preparation error).
void prep_loop(perm_server* S) {
  perm::set_device(S->device);
  for (;;) {
    Request* r;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv_prep.wait(lk, [&] { return S->stop || !S->to_prep.empty(); });
      if (S->to_prep.empty()) return;  // stopping and drained
      r = S->to_prep.front();
      S->to_prep.pop_front();
      S->prepping += 1;
    }
    const int st = prepare_request(S, r);
    bool done_now = false;
    {
      std::lock_guard<std::mutex> lk(S->mu);
      S->prepping -= 1;
      if (st != PERM_OK || r->prep.zero) {  // finished at once
        if (st != PERM_OK) r->status = st;
        r->prep = perm::Prepared();
        r->t_done = Clock::now();
        r->finished = true;
        S->stats.completed += 1;
        done_now = true;
      } else {
        S->pending.push_back(r);
      }
    }
    if (done_now) S->cv_done.notify_all();
    else S->cv_submit.notify_one();
  }
}

// Copy a caller's matrix into a new request (no preparation yet).
int make_request(int dtype, const void* mat, const int* cptrs, const int* rows,
                 const void* cvals, int nov, Request*& out) {
  if (!mat) return perm::fail(PERM_ERR_ARG, "null pointer");
  if (dtype < PERM_DTYPE_I32 || dtype > PERM_DTYPE_F64) return perm::fail(PERM_ERR_ARG, "bad dtype");
  if (nov < 1) return perm::fail(PERM_ERR_ARG, "nov must be >= 1");
  if (nov > perm::kMaxBatchN)
    return perm::fail(PERM_ERR_SIZE, "server requests are limited to n <= " +
                                         std::to_string(perm::kMaxBatchN));
  auto* r = new Request();
  r->n = nov;
  RawMatrix& m = r->raw;
  m.dtype = dtype;
  const size_t es = dtype_size(dtype);
  m.mat.assign((const char*)mat, (const char*)mat + (size_t)nov * nov * es);
  if (cptrs) {
    if (cptrs[0] != 0 || cptrs[nov] < 0 || (cptrs[nov] > 0 && (!rows || !cvals))) {
      delete r;
      return perm::fail(PERM_ERR_ARG, "malformed cptrs");
    }
    m.sparse = true;
    const int nnz = cptrs[nov];
    m.cptrs.assign(cptrs, cptrs + nov + 1);
    m.rows.assign(rows, rows + nnz);
    m.cvals.assign((const char*)cvals, (const char*)cvals + (size_t)nnz * es);
  }
  out = r;
  return PERM_OK;
}

// Register requests (one lock) and wake the prep pool.
int enqueue(perm_server* S, Request* const* rs, int count, unsigned long long* ids) {
  {
    std::lock_guard<std::mutex> lk(S->mu);
    const char* why = S->fatal != PERM_OK ? "server failed: " : S->stop ? "server is shutting down" : nullptr;
    if (why) {
      for (int i = 0; i < count; ++i) delete rs[i];
      return perm::fail(S->fatal != PERM_OK ? S->fatal : PERM_ERR_ARG,
                        std::string(why) + (S->fatal != PERM_OK ? S->fatal_msg : ""));
    }
    const Clock::time_point now = Clock::now();
    for (int i = 0; i < count; ++i) {
      Request* r = rs[i];
      r->t_submit = now;
      r->id = S->next_id++;
      S->table[r->id] = r;
      S->to_prep.push_back(r);
      ids[i] = r->id;
    }
    S->stats.submitted += count;
  }
  if (count == 1) S->cv_prep.notify_one();
  else S->cv_prep.notify_all();
  return perm::ok();
}

// Collect a finished request (mu held): copy out, erase, free.
int collect(perm_server* S, std::unordered_map<uint64_t, Request*>::iterator it,
            perm_result* out, double* latency_ms) {
  Request* r = it->second;
  const int st = r->status;
  if (out) *out = r->result;
  const double lat = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
  if (latency_ms) *latency_ms = lat;
  if (out && st == PERM_OK) out->kernel_ms = lat;
  const std::string err = r->error;
  S->table.erase(it);
  delete r;
  S->cv_done.notify_all();  // other waiters on this ticket see it gone
  if (st != PERM_OK) return perm::fail(st, err);
  return perm::ok();
}

// Wait (mu held through lk) until request `id` is finished and collect it.
// The ticket is looked up again after every wake-up: another thread may
// have collected it (wait, poll or wait_many on the same id), and then this
// call returns PERM_ERR_ARG with out->status = PERM_ERR_ARG instead of
// touching the freed request.
int wait_collect(perm_server* S, std::unique_lock<std::mutex>& lk, uint64_t id, perm_result* out,
                 double* latency_ms) {
  S->cv_done.wait(lk, [&] {
    const auto it = S->table.find(id);
    return it == S->table.end() || it->second->finished;
  });
  const auto it = S->table.find(id);
  if (it == S->table.end()) {
    if (out) {
      std::memset(out, 0, sizeof(perm_result));
      out->status = PERM_ERR_ARG;
    }
    return perm::fail(PERM_ERR_ARG, "request collected by another caller");
  }
  return collect(S, it, out, latency_ms);
}

void release(perm_server* S) {
  for (auto& I : S->it) {
    if (I.dmem) (void)hipFree(I.dmem);
    if (I.hmem) (void)hipHostFree(I.hmem);
    if (I.done) (void)hipEventDestroy(I.done);
    if (I.start) (void)hipEventDestroy(I.start);
  }
  if (S->d_arena) (void)hipFree(S->d_arena);
  if (S->h_arena) (void)hipHostFree(S->h_arena);
  for (int i = 0; i < kMaxStreams; ++i) {
    if (S->s[i]) (void)hipStreamDestroy(S->s[i]);
    if (S->ev_join[i]) (void)hipEventDestroy(S->ev_join[i]);
  }
  if (S->ev_fork) (void)hipEventDestroy(S->ev_fork);
  if (S->s0) (void)hipStreamDestroy(S->s0);
  for (auto& kv : S->table) delete kv.second;
  delete S;
}

int setup(perm_server* S) {
  perm_server_config& c = S->cfg;
  const size_t slots = (size_t)c.max_resident;
  HIP_TRYS(hipMalloc((void**)&S->d_arena, slots * kSlotDoubles * sizeof(double)));
  HIP_TRYS(hipHostMalloc((void**)&S->h_arena, slots * kSlotDoubles * sizeof(double),
                         hipHostMallocDefault));
  for (size_t i = 0; i < slots; ++i) S->free_slots.push((int)i);
  HIP_TRYS(hipStreamCreateWithFlags(&S->s0, hipStreamNonBlocking));
  for (int i = 0; i < kMaxStreams; ++i) {
    HIP_TRYS(hipStreamCreateWithFlags(&S->s[i], hipStreamNonBlocking));
    HIP_TRYS(hipEventCreateWithFlags(&S->ev_join[i], hipEventDisableTiming));
  }
  HIP_TRYS(hipEventCreateWithFlags(&S->ev_fork, hipEventDisableTiming));
  const size_t mi = S->max_items;
  // [matrix records | items] are staged and uploaded together
  const size_t mats_b = align256(mi * sizeof(BatchMatrix)) + align256(mi * sizeof(BatchItem));
  const size_t parts_b = align256(mi * sizeof(Partial));
  const size_t ctr_b = align256(mi * sizeof(uint32_t));
  for (int k = 0; k < c.inflight; ++k) {
    Iteration& I = S->it[k];
    HIP_TRYS(hipMalloc(&I.dmem, mats_b + parts_b + ctr_b));
    HIP_TRYS(hipHostMalloc(&I.hmem, mats_b + parts_b, hipHostMallocDefault));
    char* d = (char*)I.dmem;
    char* h = (char*)I.hmem;
    I.d_mats = (BatchMatrix*)d;
    I.d_parts = (Partial*)(d + mats_b);
    I.d_counters = (uint32_t*)(d + mats_b + parts_b);
    I.h_mats = (BatchMatrix*)h;
    I.h_parts = (Partial*)(h + mats_b);
    HIP_TRYS(hipEventCreate(&I.done));
    HIP_TRYS(hipEventCreate(&I.start));
  }
  return PERM_OK;
}

}  // namespace

extern "C" {

void perm_server_default_config(perm_server_config* c) {
  if (!c) return;
  c->item_log2 = 18;
  c->streams = 4;
  c->inflight = 2;
  c->max_items = 32768;
  c->iteration_us = 8000.0;
  c->max_resident = 2048;
  c->max_queued_work = 4.0e13;  // lane-instructions (about one second of GPU work)
  c->prep_threads = 4;
  c->max_wait_ms = 50.0;
}

int perm_server_create(const perm_server_config* cfg, perm_server** out) {
  if (!out) return perm::fail(PERM_ERR_ARG, "null pointer");
  perm_server_config c;
  perm_server_default_config(&c);
  if (cfg) {
    if (cfg->item_log2 > 0) c.item_log2 = cfg->item_log2;
    if (cfg->streams > 0) c.streams = cfg->streams;
    if (cfg->inflight > 0) c.inflight = cfg->inflight;
    if (cfg->max_items > 0) c.max_items = cfg->max_items;
    if (cfg->iteration_us > 0) c.iteration_us = cfg->iteration_us;
    if (cfg->max_resident > 0) c.max_resident = cfg->max_resident;
    if (cfg->max_queued_work > 0) c.max_queued_work = cfg->max_queued_work;
    if (cfg->prep_threads > 0) c.prep_threads = std::min(cfg->prep_threads, 64);
    if (cfg->max_wait_ms > 0) c.max_wait_ms = cfg->max_wait_ms;
  }
  if (c.item_log2 < 6 || c.item_log2 > 30) return perm::fail(PERM_ERR_ARG, "item_log2 in [6, 30]");
  if (c.streams > kMaxStreams) c.streams = kMaxStreams;
  if (c.inflight > kMaxInflight) c.inflight = kMaxInflight;
  if (int s = perm::select_device()) return s;
  auto* S = new perm_server();
  S->cfg = c;
  S->device = perm::current_device();
  S->max_items = (size_t)c.max_items;
  S->iter_cost = c.iteration_us * 1e-6 * kLaneInstrPerSec;
  if (int st = setup(S)) {
    const std::string msg = perm::last_error();
    release(S);
    return perm::fail(st, msg);
  }
  S->worker = std::thread(scheduler_main, S);
  for (int t = 0; t < c.prep_threads; ++t) S->preppers.emplace_back(prep_loop, S);
  *out = S;
  return perm::ok();
}

int perm_server_submit_dense(perm_server* S, int dtype, const void* mat, int nov,
                             unsigned long long* id) {
  if (!S || !id) return perm::fail(PERM_ERR_ARG, "null pointer");
  Request* r = nullptr;
  if (int s = make_request(dtype, mat, nullptr, nullptr, nullptr, nov, r)) return s;
  return enqueue(S, &r, 1, id);
}

int perm_server_submit_many(perm_server* S, int count, const int* dtypes,
                            const void* const* mats, const int* const* cptrs,
                            const int* const* rows, const void* const* cvals, const int* novs,
                            unsigned long long* ids) {
  if (!S || count < 0 || (count > 0 && (!dtypes || !mats || !novs || !ids)))
    return perm::fail(PERM_ERR_ARG, "null pointer");
  if (cptrs && (!rows || !cvals)) return perm::fail(PERM_ERR_ARG, "null rows/cvals arrays");
  std::vector<Request*> rs;
  rs.reserve(count);
  int st = PERM_OK;
  for (int i = 0; i < count && st == PERM_OK; ++i) {
    Request* r = nullptr;
    const bool sp = cptrs && cptrs[i];
    st = make_request(dtypes[i], mats[i], sp ? cptrs[i] : nullptr, sp ? rows[i] : nullptr,
                      sp ? cvals[i] : nullptr, novs[i], r);
    if (st == PERM_OK) rs.push_back(r);
  }
  const std::string err = st != PERM_OK ? perm::last_error() : "";
  if (!rs.empty())
    if (int s = enqueue(S, rs.data(), (int)rs.size(), ids)) return s;
  if (st != PERM_OK)
    return perm::fail(st, "matrix " + std::to_string(rs.size()) + ": " + err);
  return perm::ok();
}

int perm_server_submit_sparse(perm_server* S, int dtype, const void* mat, const int* cptrs,
                              const int* rows, const void* cvals, int nov,
                              unsigned long long* id) {
  if (!S || !id || !cptrs) return perm::fail(PERM_ERR_ARG, "null pointer");
  Request* r = nullptr;
  if (int s = make_request(dtype, mat, cptrs, rows, cvals, nov, r)) return s;
  return enqueue(S, &r, 1, id);
}

int perm_server_wait(perm_server* S, unsigned long long id, perm_result* out,
                     double* latency_ms) {
  if (!S) return perm::fail(PERM_ERR_ARG, "null server");
  std::unique_lock<std::mutex> lk(S->mu);
  if (S->table.find(id) == S->table.end())
    return perm::fail(PERM_ERR_ARG, "unknown request id");
  return wait_collect(S, lk, id, out, latency_ms);
}

int perm_server_wait_many(perm_server* S, const unsigned long long* ids, int count,
                          perm_result* out, double* latency_ms) {
  if (!S || (count > 0 && (!ids || !out))) return perm::fail(PERM_ERR_ARG, "null pointer");
  std::unique_lock<std::mutex> lk(S->mu);
  for (int i = 0; i < count; ++i)
    if (S->table.find(ids[i]) == S->table.end())
      return perm::fail(PERM_ERR_ARG, "unknown request id");
  int first_error = PERM_OK;
  for (int i = 0; i < count; ++i) {
    auto it = S->table.find(ids[i]);
    if (it == S->table.end()) {  // listed twice: collected already
      std::memset(out + i, 0, sizeof(perm_result));
      out[i].status = PERM_ERR_ARG;
      if (first_error == PERM_OK) first_error = perm::fail(PERM_ERR_ARG, "request id listed twice");
      continue;
    }
    // (a ticket collected meanwhile by another caller comes back as PERM_ERR_ARG)
    const int st = wait_collect(S, lk, ids[i], out + i, latency_ms ? latency_ms + i : nullptr);
    if (st != PERM_OK && first_error == PERM_OK) first_error = st;
  }
  return first_error == PERM_OK ? perm::ok() : first_error;
}

int perm_server_poll(perm_server* S, unsigned long long id, perm_result* out,
                     double* latency_ms) {
  if (!S) return perm::fail(PERM_ERR_ARG, "null server");
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->table.find(id);
  if (it == S->table.end()) return perm::fail(PERM_ERR_ARG, "unknown request id");
  if (!it->second->finished) return PERM_PENDING;
  return collect(S, it, out, latency_ms);
}

int perm_server_stats_get(perm_server* S, perm_server_stats* out) {
  if (!S || !out) return perm::fail(PERM_ERR_ARG, "null pointer");
  std::lock_guard<std::mutex> lk(S->mu);
  *out = S->stats;
  out->pending = (long long)S->pending.size();
  // scheduler-owned counters, read without its lock (statistics only)
  out->uploads = S->n_uploads.load(std::memory_order_relaxed);
  out->admit_ms = S->t_admit.load(std::memory_order_relaxed);
  out->issue_ms = S->t_issue.load(std::memory_order_relaxed);
  out->retire_ms = S->t_retire.load(std::memory_order_relaxed);
  out->gpu_wait_ms = S->t_wait.load(std::memory_order_relaxed);
  out->gpu_ms = S->t_gpu.load(std::memory_order_relaxed);
  out->dd_passes = S->n_dd.load(std::memory_order_relaxed);
  return perm::ok();
}

void perm_server_destroy(perm_server* S) {
  if (!S) return;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
  }
  S->cv_prep.notify_all();
  for (auto& t : S->preppers)
    if (t.joinable()) t.join();  // every submitted request is prepared (or failed)
  S->cv_submit.notify_all();
  if (S->worker.joinable()) S->worker.join();  // drains admitted and pending work
  (void)hipSetDevice(S->device);
  if (S->s0) (void)hipStreamSynchronize(S->s0);
  release(S);
}

}  // extern "C"
