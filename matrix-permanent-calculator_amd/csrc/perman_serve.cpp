// perman_serve.cpp -- a permanent server on one GPU: continuous batching of
// requests of mixed n (C ABI in include/perman_hip.h, perm_server_*) on one
// persistent kernel (ryser_serve.hip) launched as a captured graph.
//
// Callers submit matrices from any thread and wait on (or poll) a ticket.
//
//  * Submission copies the caller's matrix; a pool of prep threads prepares
//    it: start vector, column store and walk shape (kMenuBatch), padded to
//    the served order N = serve_padded_n(n) (rows of x equal to 1 that no
//    Gray column touches; n < 8 embedded as A (+) I), the walk split and the
//    item split (about 2^item_log2 subsets per item, at most max_items items
//    per pass, several 64-lane groups per item above that).
//  * Admission (scheduler thread): a prepared request gets one block of the
//    device arena for its whole life -- a header slot for each of its (at
//    most three) passes, its column store and start vector, room for
//    sampled item ids, one record per item (reused by every pass) -- while
//    resident requests stay under max_resident and queued work under
//    max_queued_work.  Blocks are taken from a ring in order and freed in
//    order; since a later pass never needs new space, the oldest request
//    always finishes and frees the ring's front.  Blocks are written into a
//    pinned mirror of the arena at the same offsets.
//  * Publication: each pass's header address and the mirror range it needs
//    copied (its header; on the first pass also the matrix; on a
//    verification pass also the sampled ids) go into the ring of the pinned
//    control block and `tail` advances.  The kernel's polling wave copies
//    the ranges into the arena and publishes the passes to the other waves:
//    the host issues no copy and waits on no event.  The waves claim items
//    in publication order, so no request waits behind later ones.  If no
//    launch of the pass's kind is live, the scheduler replays its graph.
//  * Completion: the wave that finishes a pass's last item appends the
//    pass's sum (its records summed in item order, so a result depends only
//    on its matrix: bitwise the same alone or in any mix) to a completion
//    queue in pinned memory, which the scheduler reads in order (no scan of
//    the passes in flight).  The scheduler turns it into a result with the
//    stated error; a result that misses the accuracy target gets another
//    pass over the same matrix (double-double products for an integer
//    matrix; for other matrices first a sampled measurement of the FP64
//    error -- up to 256 items walked again in double-double X -- then, if
//    needed, the double-double X pass), else the request is published to its
//    waiters.
//  * Cache rule of the persistent kernel: a launch reads headers and
//    matrices with plain loads, through caches that may still hold what an
//    earlier tenant of the same arena space left there if this launch read
//    it, and its waves may scan any pass from the oldest one in flight when
//    it began.  Each range is copied once per use (a header slot, the
//    matrix, the ids are never rewritten while their request lives).  So arena space is reused only if it
//    was freed before the live launch began; otherwise the live launch is
//    stopped first (about once per lap of the arena ring) and the next one
//    starts with clean caches.
//  * A matrix whose nonzero pattern has no perfect matching is exactly 0 and
//    finishes at preparation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <deque>
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
#include "ryser_serve.hpp"

namespace {

using perm::Partial;
using perm::ServeControl;
using perm::ServeDone;
using perm::ServeHeader;
using perm::ServeState;
using Clock = std::chrono::steady_clock;

// Pass kinds: 0 FP64, 1 double-double products, 2 double-double X;
// kVerifyPass: sampled items of the FP64 pass walked in mode 2;
// kExportPass: the FP64 pass's records copied to the host mirror (no walk),
// before the sampling that picks the verification's items.
constexpr int kVerifyPass = 3;
constexpr int kExportPass = 4;
// Request block: header slot 0, the column store and start vector, header
// slots 1 and 2, the sampled ids, header slot 3, the item records -- so each
// pass's copy is one contiguous range: slot 0 with the matrix, slot 2 with
// the ids (the verification pass follows the FP64 pass and its export),
// slots 1 and 3 alone.
constexpr size_t kHeaderBytes = 256;
constexpr int kPassSlots = 4;  // FP64; then double-double products, or export, verification, X
constexpr size_t kIdsBytes = (perm::kVerifyUnits * sizeof(uint32_t) + 255) / 256 * 256;
static_assert(sizeof(perm::ServeHeader) <= kHeaderBytes, "header slot");

size_t align256(size_t v) { return (v + 255) / 256 * 256; }

// Offsets in a request's block, from its column store + start vector size.
struct BlockLayout {
  size_t data_bytes;
  size_t slot(int i) const {
    return i == 0 ? 0
         : i == 1 ? kHeaderBytes + data_bytes
         : i == 2 ? 2 * kHeaderBytes + data_bytes
                  : 3 * kHeaderBytes + data_bytes + kIdsBytes;
  }
  size_t data() const { return kHeaderBytes; }
  size_t ids() const { return slot(2) + kHeaderBytes; }
  size_t records() const { return slot(3) + kHeaderBytes; }
};

// A submitted matrix as the caller gave it (copied), prepared by the pool.
struct RawMatrix {
  int dtype = PERM_DTYPE_F64;
  bool sparse = false;
  std::vector<char> mat, cvals;
  std::vector<int> cptrs, rows;
};

struct Request;

// One pass of a request: a header (and records) in one arena block.
struct Pass {
  Request* r = nullptr;
  int kind = 0;          // 0, 1, 2, kVerifyPass, kExportPass
  uint32_t items = 0;
  uint64_t seq = 0;      // publication sequence
  bool first = false;    // the request's first pass (its copy carries the matrix)
  int slot = 0;          // header slot in the request's block
  size_t up_off = 0;     // arena offset and size of the mirror range the kernel copies
  size_t up_bytes = 0;
  uint64_t up2_src = 0;  // a second range (the first pass's matrix): source address,
  size_t up2_off = 0;    // arena offset and size (0: none)
  size_t up2_bytes = 0;
};

struct Request {
  uint64_t id = 0;
  RawMatrix raw;   // until prepared
  int n = 0;       // caller's order
  int n_eff = 0;   // order walked: n, or 8 for an embedded n < 8
  int n_pad = 0;   // served kernel order
  int walk_bits = 0, high_bits = 0;
  uint32_t walk_items = 0, groups = 0;
  perm::NestShape shape{0, 0, 0};    // walk shape, padded-order terms
  perm::NestShape shape_n{0, 0, 0};  // the same shape in terms of n_eff (reported)
  bool integral = false;
  double product_bits = 1e9;
  double cost = 0.0;     // FP64 lane-instructions of the FP64 pass (admission)
  // padded column store then start vector (2N doubles), written by the prep
  // thread into pinned staging memory that the kernel copies from
  char* staged = nullptr;
  uint64_t staged_dev = 0;  // its device address
  size_t staged_bytes = 0;  // a multiple of 256
  uint64_t vblock = 0;   // virtual arena position of the request's block
  size_t block_bytes = 0;
  size_t block = 0;      // block index in allocation order
  size_t cs_doubles = 0;
  BlockLayout lay{0};
  int passes = 0;        // passes staged so far
  bool resident = false;
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
};

// One persistent kernel instance: its stream, graph, the events around its
// last launch, and where the arena's live data began when it started.
struct KernelSlot {
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool launched = false;  // a launch was made and not yet accounted
  uint64_t epoch = 0;     // launch number (the kernel reports the one that ends)
  uint64_t launch_vfront = 0;
};

// Arena block: virtual positions grow forever; the physical offset is the
// position modulo the arena size (a block never straddles the end).
struct Block {
  uint64_t vstart = 0, size = 0;
  bool live = true;
};

// Pinned, coherent host memory for prepared matrices, read by the kernel's
// copying waves: blocks of one size per served order (free lists by size),
// cut from 8 MiB slabs allocated on demand and kept until the server ends.
struct StagingPool {
  static constexpr size_t kSlab = 8u << 20;
  struct Slab {
    char* host;
    uint64_t dev;
  };
  std::mutex mu;
  std::vector<Slab> slabs;
  std::unordered_map<size_t, std::vector<char*>> free;
  size_t used = 0;  // bytes handed out of the current slab

  // A block of `bytes` (<= kSlab): host pointer, device address; nullptr on a
  // HIP failure.
  char* get(size_t bytes, uint64_t& dev) {
    std::lock_guard<std::mutex> lk(mu);
    auto& fl = free[bytes];
    char* p = nullptr;
    if (!fl.empty()) {
      p = fl.back();
      fl.pop_back();
    } else {
      if (slabs.empty() || used + bytes > kSlab) {
        Slab sl{nullptr, 0};
        if (hipHostMalloc((void**)&sl.host, kSlab, hipHostMallocCoherent) != hipSuccess)
          return nullptr;
        void* d = nullptr;
        if (hipHostGetDevicePointer(&d, sl.host, 0) != hipSuccess) {
          (void)hipHostFree(sl.host);
          return nullptr;
        }
        sl.dev = (uint64_t)(uintptr_t)d;
        slabs.push_back(sl);
        used = 0;
      }
      p = slabs.back().host + used;
      used += bytes;
    }
    dev = 0;
    for (const Slab& sl : slabs)
      if (p >= sl.host && p < sl.host + kSlab) dev = sl.dev + (uint64_t)(p - sl.host);
    return p;
  }
  void put(char* p, size_t bytes) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(mu);
    free[bytes].push_back(p);
  }
  void release() {
    for (const Slab& sl : slabs) (void)hipHostFree(sl.host);
    slabs.clear();
    free.clear();
  }
};

}  // namespace

struct perm_server {
  perm_server_config cfg{};
  int device = 0;
  std::thread worker;
  std::mutex mu;
  std::condition_variable cv_done;               // (mu)
  // the queues between submitters, prep threads and the scheduler have a
  // lock of their own, so that the scheduler and the prep threads do not
  // contend with waiters on mu
  std::mutex mu_q;
  std::condition_variable cv_submit, cv_prep;    // (mu_q)
  std::deque<Request*> to_prep;                  // submitted, not prepared (mu_q)
  std::deque<Request*> pending;                  // prepared, not admitted (mu_q)
  std::atomic<int> pending_n{0};                 // pending.size(), read without mu_q
  std::atomic<bool> stopping{false};             // set under mu and mu_q
  int prep_ahead = 0;  // prep threads wait while this many requests are prepared and pending
  StagingPool staging;
  std::vector<std::thread> preppers;
  size_t preppers_n = 0;
  int prepping = 0;                              // requests taken by a prep thread (mu_q)
  std::unordered_map<uint64_t, Request*> table;  // every uncollected request (mu)
  uint64_t next_id = 1;
  int fatal = PERM_OK;  // scheduler error (mu): every later call fails with it
  std::string fatal_msg;
  perm_server_stats stats{};  // (mu)
  // ---- scheduler-owned ----
  char* d_arena = nullptr;
  char* h_arena = nullptr;     // pinned mirror (the kernel copies from it, same offsets)
  char* h_arena_dev = nullptr; // the mirror's device address
  uint64_t arena_bytes = 0;
  std::deque<Block> blocks;    // blocks in allocation order, from index block_base
  size_t block_base = 0;
  uint64_t vhead = 0;          // virtual position of the next block
  ServeControl* ctl = nullptr;  // pinned, coherent
  ServeDone* done = nullptr;    // pinned, coherent: the completion queue (kServeRing)
  ServeState* d_state = nullptr;
  uint64_t next_seq = 0;        // publication sequence of the next pass
  uint64_t lo_seq = 0;          // every pass before it has completed
  uint64_t consumed = 0;        // completion records read
  int resident = 0;
  double queued_cost = 0.0;
  std::deque<Pass*> to_publish;  // written to the mirror, not published yet
  std::vector<Pass*> by_seq;     // passes in flight at seq mod kServeRing
  int n_inflight[2] = {0, 0};    // passes in flight per kernel kind
  bool dd_waiting = false;       // double-double X passes wait for a main-kernel stop
  Clock::time_point dd_wait_since;
  KernelSlot K[2];  // [0] the main kernel, [1] the double-double X kernel
  std::atomic<double> t_admit{0.0}, t_publish{0.0}, t_retire{0.0}, t_stop{0.0}, t_gpu{0.0};
  std::atomic<double> t_prep{0.0}, t_submit{0.0};  // (added to under mu)
  std::atomic<long long> n_publish{0}, n_dd{0}, n_launches{0}, n_items{0}, n_stops{0};
  std::atomic<long long> n_admitted{0};
  std::atomic<uint64_t> vhead_seen{0};  // vhead, for the statistics
};

namespace {

#define HIP_TRYS(expr)                                                                      \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return perm::fail(PERM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));   \
  } while (0)

// PERMAN_SERVE_TRACE=1: one stderr line per scheduler event (diagnostics).
bool trace_on() {
  static const bool on = perm::env_int("PERMAN_SERVE_TRACE", 0) != 0;
  return on;
}
#define SERVE_TRACE(...)                                                                   \
  do {                                                                                     \
    if (trace_on()) {                                                                      \
      std::fprintf(stderr, "[serve %.3f ms] ",                                             \
                   std::chrono::duration<double, std::milli>(                              \
                       Clock::now().time_since_epoch()).count());                          \
      std::fprintf(stderr, __VA_ARGS__);                                                   \
      std::fprintf(stderr, "\n");                                                          \
    }                                                                                      \
  } while (0)

void add_ms(std::atomic<double>& acc, Clock::time_point t0) {
  const double ms = std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
  double cur = acc.load(std::memory_order_relaxed);
  while (!acc.compare_exchange_weak(cur, cur + ms, std::memory_order_relaxed)) {
  }
}
void bump(std::atomic<long long>& c, long long v = 1) {
  c.store(c.load(std::memory_order_relaxed) + v, std::memory_order_relaxed);
}

// FP64 VALU lane-instructions per subset of a pass (admission estimate).
double subset_cost(int n_pad, perm::NestShape s, int kind) {
  if (kind == 2 || kind == kVerifyPass) return 11.0 * n_pad + 8.0;
  const double fp64 = (double)perm::nest_cost(n_pad, s) / 8.0;
  if (kind == 0) return fp64;
  const double half = (fp64 - 2.0) / 2.0;  // adds = multiplies
  return half + 3.0 * half + 8.0;
}

double pass_cost(const Request* r, int kind, uint32_t items) {
  if (kind == kExportPass) return 0.0;
  return (double)items * r->groups * std::ldexp((double)perm::kWave, r->walk_bits) *
         subset_cost(r->n_pad, r->shape, kind);
}

// ---- the persistent kernels ---------------------------------------------

bool is_dd(int kind) { return kind == 2 || kind == kVerifyPass; }

// The last launch of kind k may still run (exact: its end event).
bool launch_live(perm_server* S, int k) {
  return S->K[k].launched && hipEventQuery(S->K[k].e1) == hipErrorNotReady;
}

// The last launch of kind k has decided to end (its waves leave between
// items): a new launch may be queued behind it.  A read of host memory.
bool launch_ending(perm_server* S, int k) {
  return !S->K[k].launched || __atomic_load_n(&S->ctl->ended[k], __ATOMIC_ACQUIRE) == S->K[k].epoch;
}

// Account an ended launch's GPU time (dropped if it has not ended yet).
void account_launch(perm_server* S, int k) {
  KernelSlot& K = S->K[k];
  if (!K.launched) return;
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, K.e0, K.e1) == hipSuccess)
    S->t_gpu.store(S->t_gpu.load(std::memory_order_relaxed) + ms, std::memory_order_relaxed);
  K.launched = false;
}

uint64_t vfront(const perm_server* S) {
  return S->blocks.empty() ? S->vhead : S->blocks.front().vstart;
}

// Replay the graph of kernel k.  Its waves start scanning at the oldest
// pass of its kind in flight (written before the launch).
int launch(perm_server* S, int k) {
  KernelSlot& K = S->K[k];
  account_launch(S, k);
  uint64_t head = S->next_seq;
  for (uint64_t q = S->lo_seq; q < S->next_seq; ++q) {
    const Pass* p = S->by_seq[q & (perm::kServeRing - 1)];
    if (p && is_dd(p->kind) == (k == 1)) {
      head = q;
      break;
    }
  }
  __atomic_store_n(&S->ctl->head[k], head, __ATOMIC_RELEASE);
  K.epoch += 1;
  __atomic_store_n(&S->ctl->epoch[k], K.epoch, __ATOMIC_RELEASE);
  // the arena positions this launch may have read begin at the live front;
  // a previous launch still finishing its items keeps its older front (the
  // new launch starts after it on the stream, so its events cover both)
  if (!launch_live(S, k)) K.launch_vfront = vfront(S);
  HIP_TRYS(hipEventRecord(K.e0, K.s));
  // the device copy of the stop flag starts clear (ordered after the
  // previous launch of the kind, which may still be finishing its items,
  // and before this one; outside the captured graph)
  HIP_TRYS(hipMemsetAsync(&S->d_state->stop[k], 0, sizeof(uint64_t), K.s));
  HIP_TRYS(hipGraphLaunch(K.exec, K.s));
  HIP_TRYS(hipEventRecord(K.e1, K.s));
  K.launched = true;
  bump(S->n_launches);
  SERVE_TRACE("launch kernel %d head %llu (previous ended: %llu, why %llu, activity %llu)", k,
              (unsigned long long)head, (unsigned long long)S->ctl->ended[k],
              (unsigned long long)(S->ctl->why[k] & 0xff),
              (unsigned long long)(S->ctl->why[k] >> 8));
  return PERM_OK;
}

// Stop the live launches (waves finish their current item, then leave).
int stop_kernels(perm_server* S) {
  if (!S->K[0].launched && !S->K[1].launched) return PERM_OK;
  const Clock::time_point t0 = Clock::now();
  __atomic_store_n(&S->ctl->stop, 1ull, __ATOMIC_RELEASE);
  hipError_t e = hipSuccess;
  for (int k = 0; k < 2; ++k)
    if (S->K[k].launched && e == hipSuccess) e = hipEventSynchronize(S->K[k].e1);
  __atomic_store_n(&S->ctl->stop, 0ull, __ATOMIC_RELEASE);
  if (e != hipSuccess)
    return perm::fail(PERM_ERR_HIP, std::string("hipEventSynchronize: ") + hipGetErrorString(e));
  account_launch(S, 0);
  account_launch(S, 1);
  bump(S->n_stops);
  SERVE_TRACE("stopped the kernels");
  add_ms(S->t_stop, t0);
  return PERM_OK;
}

// Launch the kernels the passes in flight need.  The double-double X kernel
// (1 wave per SIMD) can only become resident where the main one (4 waves
// per SIMD take every register file) leaves room: when double-double X
// passes have waited kDDWaitMs behind a live main launch, that launch is
// stopped (its waves finish their items), the double-double X kernel
// launched, and the main one launched again after it.  The wait lets one
// stop serve the double-double passes that arrive close together.
constexpr double kDDWaitMs = 5.0;
int ensure_kernels(perm_server* S) {
  const bool main_work = S->n_inflight[0] > 0, dd_work = S->n_inflight[1] > 0;
  if (dd_work && launch_ending(S, 1)) {
    bool go = true;
    if (launch_live(S, 0)) {
      const Clock::time_point now = Clock::now();
      if (!S->dd_waiting) {
        S->dd_waiting = true;
        S->dd_wait_since = now;
      }
      go = std::chrono::duration<double, std::milli>(now - S->dd_wait_since).count() >= kDDWaitMs;
      if (go)
        if (int s = stop_kernels(S)) return s;
    }
    if (go) {
      S->dd_waiting = false;
      if (int s = launch(S, 1)) return s;
    }
  }
  if (main_work && launch_ending(S, 0))
    if (int s = launch(S, 0)) return s;
  return PERM_OK;
}

// ---- arena ring -----------------------------------------------------------

// Reserve a block of `size` bytes (a multiple of 256): its virtual position,
// or UINT64_MAX when the ring has no room.  Space that the live launch may
// have read is reused only after stopping that launch.
int arena_alloc(perm_server* S, uint64_t size, uint64_t& vpos) {
  vpos = UINT64_MAX;
  const uint64_t A = S->arena_bytes;
  if (size > A / 2) return PERM_OK;
  uint64_t v = S->vhead;
  if (v % A + size > A) v = (v / A + 1) * A;  // blocks never straddle the end
  // live blocks span at most half the arena: a launch stopped so that space
  // it may have read can be reused (the cache rule below) is then followed by
  // half a lap of allocations before the next stop, not one per allocation
  if (!S->blocks.empty() && v + size - vfront(S) > A / 2) return PERM_OK;
  // the block reuses space last used at positions [v - A, v + size - A): a
  // live launch that began before that space was freed may have read it
  if (v + size > A)
    for (int k = 0; k < 2; ++k)
      if (v + size - A > S->K[k].launch_vfront && launch_live(S, k)) {
        if (int s = stop_kernels(S)) return s;
        break;
      }
  S->vhead = v + size;
  S->vhead_seen.store(S->vhead, std::memory_order_relaxed);
  S->blocks.push_back({v, size, true});
  vpos = v;
  return PERM_OK;
}

void arena_free(perm_server* S, size_t block) {
  S->blocks[block - S->block_base].live = false;
  while (!S->blocks.empty() && !S->blocks.front().live) {
    S->blocks.pop_front();
    ++S->block_base;
  }
}

// ---- passes ---------------------------------------------------------------

// Stage pass `kind` of r.  The first pass reserves the request's block and
// writes header 0, the column store and the start vector to the mirror;
// a later pass writes header slot r->passes (and the sampled ids of a
// verification pass).  Returns nullptr (st = PERM_OK) when the ring has no
// room for a first pass now.
Pass* make_pass(perm_server* S, Request* r, int kind, int& st) {
  st = PERM_OK;
  const bool first = !r->resident;
  const uint32_t items =
      kind == kVerifyPass ? (uint32_t)r->vids.size()
      : kind == kExportPass
          ? (r->walk_items + perm::kServeExportChunk - 1) / perm::kServeExportChunk
          : r->walk_items;
  const int vmode = kind == kVerifyPass ? 2 : kind;
  const perm::NestShape vshape =
      vmode == 2 ? perm::NestShape{r->n_pad, r->n_pad, r->n_pad} : r->shape;
  const int variant = kind == kExportPass ? 0 : perm::serve_variant(r->n_pad, vshape, vmode);
  const int slot = r->passes;
  if (variant < 0 || slot >= kPassSlots || items > r->walk_items ||
      (kind == kVerifyPass && (items > perm::kVerifyUnits || slot != 2)) ||
      (kind == kExportPass && slot != 1)) {  // not expected
    st = perm::fail(PERM_ERR_ARG, "no served kernel variant or pass slot for this request");
    return nullptr;
  }
  if (first) {
    r->lay.data_bytes = r->staged_bytes;
    const size_t bytes = r->lay.records() + align256((size_t)r->walk_items * sizeof(Partial));
    uint64_t v = 0;
    if ((st = arena_alloc(S, bytes, v)) != PERM_OK || v == UINT64_MAX) return nullptr;
    r->vblock = v;
    r->block_bytes = bytes;
    r->block = S->block_base + S->blocks.size() - 1;
    r->resident = true;
  }
  const size_t off = (size_t)(r->vblock % S->arena_bytes);
  char* h = S->h_arena + off;
  char* d = S->d_arena + off;
  const BlockLayout& L = r->lay;
  size_t ids_bytes = 0;
  if (kind == kVerifyPass) {
    std::memcpy(h + L.ids(), r->vids.data(), items * sizeof(uint32_t));
    ids_bytes = align256(items * sizeof(uint32_t));
  }
  r->passes += 1;
  auto* p = new Pass();
  p->r = r;
  p->kind = kind;
  p->items = items;
  p->first = first;
  p->slot = slot;
  p->up_off = off + L.slot(slot);
  p->up_bytes = kHeaderBytes + ids_bytes;  // the header slot (and the ids)
  if (first) {  // and the matrix, from its staging block
    p->up2_src = r->staged_dev;
    p->up2_off = off + L.data();
    p->up2_bytes = L.data_bytes;
  }
  ServeHeader hd;
  std::memset(&hd, 0, sizeof(hd));
  hd.colpm = (const double*)(d + L.data());
  hd.x0 = hd.colpm + r->cs_doubles;
  hd.parts = (Partial*)(d + L.records());
  // the host reads the records of the verification pass, and those of the
  // FP64 pass through its export pass
  if (kind == kVerifyPass || kind == kExportPass)
    hd.host_parts = (Partial*)(S->h_arena_dev + off + L.records());
  hd.ids = kind == kVerifyPass ? (const uint32_t*)(d + L.ids()) : nullptr;
  hd.lanes = kind == kExportPass ? r->walk_items : 1ull << r->high_bits;
  hd.walk_bits = r->walk_bits;
  hd.high_bits = r->high_bits;
  hd.variant = kind == kExportPass ? (int32_t)perm::kServeExportFlag
                                   : variant | (vmode == 2 ? (int32_t)perm::kServeDDFlag : 0);
  hd.items = items;
  hd.groups = r->groups;
  std::memcpy(h + L.slot(slot), &hd, sizeof(hd));  // seq: at publication
  S->queued_cost += pass_cost(r, kind, items);
  S->to_publish.push_back(p);
  return p;
}

// Publish the staged passes in order while the ring has room: sequence
// number, header address and copy range into the control block's ring, then
// `tail`; launch the kernels the passes in flight need.
int publish(perm_server* S, bool& progressed) {
  size_t count = 0;
  while (!S->to_publish.empty() && S->next_seq - S->lo_seq < perm::kServeRing) {
    Pass* p = S->to_publish.front();
    S->to_publish.pop_front();
    const uint64_t seq = S->next_seq++;
    p->seq = seq;
    std::memcpy(S->h_arena + p->up_off + offsetof(ServeHeader, seq), &seq, sizeof(seq));
    const uint64_t q = seq & (perm::kServeRing - 1);
    S->ctl->ring[q] = (uint64_t)(uintptr_t)(S->d_arena + p->up_off);
    S->ctl->up[q] = ((uint64_t)(p->up_off / 256) << 32) | (uint64_t)(p->up_bytes / 256);
    S->ctl->up2_src[q] = p->up2_src;
    S->ctl->up2[q] = ((uint64_t)(p->up2_off / 256) << 32) | (uint64_t)(p->up2_bytes / 256);
    S->by_seq[q] = p;
    S->n_inflight[is_dd(p->kind) ? 1 : 0] += 1;
    bump(S->n_items, p->items);
    ++count;
  }
  if (count > 0) {
    // the mirror and the ring entries are visible before the tail that covers them
    __atomic_store_n(&S->ctl->tail, S->next_seq, __ATOMIC_RELEASE);
    bump(S->n_publish);
    progressed = true;
    SERVE_TRACE("published %zu passes, tail %llu", count, (unsigned long long)S->next_seq);
  }
  return ensure_kernels(S);
}

// ---- results ----------------------------------------------------------------

void publish_finished(perm_server* S, const std::vector<Request*>& rs) {
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

// The first `count` item records of r (sampled error measurement), from the
// host mirror: the kernel stored them there (an export or verification
// pass) before that pass's completion record.
int read_parts(perm_server* S, const Request* r, uint32_t count, std::vector<Partial>& out) {
  const Partial* parts =
      (const Partial*)(S->h_arena + (size_t)(r->vblock % S->arena_bytes) + r->lay.records());
  out.resize(count);
  for (uint32_t i = 0; i < count; ++i) {
    const volatile Partial* q = parts + i;
    out[i] = Partial{q->hi, q->lo, q->abs_sum, q->sq_sum, q->sq_run, 0.0};
  }
  return PERM_OK;
}

// Pass p of r is complete, its sum in rec.  Returns the kind of the next
// pass r needs, or -1 when r is finished (r->result filled).
int finalize(perm_server* S, Request* r, const Pass* p, const Partial& rec, int& st) {
  st = PERM_OK;
  const int n = r->n_eff;
  if (p->kind == kVerifyPass) {
    std::vector<Partial> ver;
    if ((st = read_parts(S, r, p->items, ver)) != PERM_OK) return -1;
    std::vector<double> e(r->vids.size());
    for (size_t k = 0; k < r->vids.size(); ++k) {
      const Partial& f = r->fp_parts[r->vids[k]];
      e[k] = (f.hi - ver[k].hi) + (f.lo - ver[k].lo);
    }
    r->result = r->fp_result;
    r->result.err_est = perm::sampled_error(e, r->vweight, r->vcertain, n, r->result.value);
    r->result.certified = perm::is_certified(r->result) ? 1 : 0;
    std::vector<Partial>().swap(r->fp_parts);
    return r->result.certified ? -1 : 2;
  }
  if (p->kind == kExportPass) {
    // the FP64 pass's records are in the mirror: sample the items to walk
    // again in double-double X (r->result still holds the FP64 result)
    std::vector<Partial> parts;
    if ((st = read_parts(S, r, r->walk_items, parts)) != PERM_OK) return -1;
    std::vector<double> z(parts.size());
    for (size_t i = 0; i < z.size(); ++i)
      z[i] = perm::unit_weight(parts[i].sq_sum, parts[i].sq_run, r->walk_bits, r->high_bits);
    perm::pps_sample(z, perm::kVerifyUnits, r->vids, r->vweight, r->vcertain);
    r->sampled = true;
    if (r->vids.empty()) return 2;
    r->fp_result = r->result;
    r->fp_parts.swap(parts);
    return kVerifyPass;
  }
  perm::fill_item_result(n, r->walk_bits, r->integral, r->product_bits, r->shape_n, p->kind, rec,
                         r->result);
  const int next = perm::escalation_mode(r->result);
  if (next == 2 && p->kind == 0 && !r->sampled && r->walk_items >= perm::kVerifyMinUnits)
    return kExportPass;
  return next;
}

// Most completions retired, and most requests admitted, per scheduler round:
// the GPU gets new work while the host still has completions to read
// (retiring thousands before admitting any leaves the kernel idle).
constexpr int kRoundBatch = 256;

// Completed passes, in completion order: results, further passes,
// publication to waiters.
int retire(perm_server* S, bool& progressed) {
  std::vector<Request*> finished;
  int st = PERM_OK;
  for (int taken = 0; taken < kRoundBatch; ++taken) {
    ServeDone& d = S->done[S->consumed & (perm::kServeRing - 1)];
    if (__atomic_load_n(&d.tag, __ATOMIC_ACQUIRE) != S->consumed + 1) break;
    Partial rec = d.sum;
    rec.pad = 0.0;
    const uint64_t seq = d.seq;
    S->consumed += 1;
    Pass*& slot = S->by_seq[seq & (perm::kServeRing - 1)];
    Pass* p = slot;
    if (!p || p->seq != seq) {
      st = perm::fail(PERM_ERR_HIP, "served kernel completed an unknown pass");
      break;
    }
    slot = nullptr;
    S->n_inflight[is_dd(p->kind) ? 1 : 0] -= 1;
    Request* r = p->r;
    SERVE_TRACE("pass %llu (request %llu, kind %d) complete", (unsigned long long)p->seq,
                (unsigned long long)r->id, p->kind);
    S->queued_cost -= pass_cost(r, p->kind, p->items);
    if (p->first) {  // the kernel has copied the matrix in: its staging block is free
      S->staging.put(r->staged, r->staged_bytes);
      r->staged = nullptr;
    }
    const int next = finalize(S, r, p, rec, st);
    delete p;
    if (st != PERM_OK) break;
    if (next >= 0) {
      // the next pass lives in the request's block: it never waits for space
      if (next != kExportPass) bump(S->n_dd);
      make_pass(S, r, next, st);
      if (st != PERM_OK) break;
    } else {
      r->result.n = r->n;
      r->result.subsets = std::ldexp(1.0, r->n - 1);
      arena_free(S, r->block);
      r->resident = false;
      S->resident -= 1;
      std::vector<uint32_t>().swap(r->vids);
      std::vector<double>().swap(r->vweight);
      finished.push_back(r);
    }
    progressed = true;
  }
  while (S->lo_seq < S->next_seq && !S->by_seq[S->lo_seq & (perm::kServeRing - 1)]) ++S->lo_seq;
  publish_finished(S, finished);
  return st;
}

// Admit pending requests (arrival order) while resident requests, queued
// work and arena space allow.  Only the scheduler takes from `pending`: the
// front requests are staged first, then popped together.
int admit(perm_server* S, bool& progressed) {
  size_t k = 0;
  while (S->pending_n.load(std::memory_order_acquire) > (int)k &&
         S->resident < S->cfg.max_resident && k < (size_t)kRoundBatch) {
    Request* c;
    {
      std::lock_guard<std::mutex> lk(S->mu_q);
      if (k >= S->pending.size()) break;
      c = S->pending[k];
    }
    if (S->resident > 0 && S->queued_cost + c->cost > S->cfg.max_queued_work) break;
    int st = PERM_OK;
    Pass* p = make_pass(S, c, 0, st);
    if (st != PERM_OK) return st;
    if (!p) break;  // no arena room now
    S->resident += 1;
    ++k;
  }
  if (k > 0) {
    {
      std::lock_guard<std::mutex> lk(S->mu_q);
      for (size_t i = 0; i < k; ++i) S->pending.pop_front();
      S->pending_n.store((int)S->pending.size(), std::memory_order_relaxed);
    }
    S->cv_prep.notify_all();  // prep threads may wait on the look-ahead bound
    progressed = true;
    bump(S->n_admitted, (long long)k);
  }
  return PERM_OK;
}

// One round of the scheduler: completions first (they free arena space and
// may stage further passes, and a launch that ended after its last pass is
// not launched again for it), then admission, then publication (which
// launches the kernels the passes in flight need).
int scheduler_step(perm_server* S, bool& progressed) {
  Clock::time_point t0 = Clock::now();
  if (int st = retire(S, progressed)) return st;
  add_ms(S->t_retire, t0);
  t0 = Clock::now();
  if (int st = admit(S, progressed)) return st;
  add_ms(S->t_admit, t0);
  t0 = Clock::now();
  if (int st = publish(S, progressed)) return st;
  add_ms(S->t_publish, t0);
  return PERM_OK;
}

// Diagnostics (PERMAN_SERVE_TRACE): every 5 ms while passes are in flight,
// the host-side queue positions.
void trace_state(perm_server* S) {
  static Clock::time_point last;
  const Clock::time_point now = Clock::now();
  if (now - last < std::chrono::milliseconds(5)) return;
  last = now;
  // completion records written and not yet read
  uint64_t ready = 0;
  while (ready < perm::kServeRing &&
         __atomic_load_n(&S->done[(S->consumed + ready) & (perm::kServeRing - 1)].tag,
                         __ATOMIC_ACQUIRE) == S->consumed + ready + 1)
    ++ready;
  SERVE_TRACE("host: tail %llu device pub %llu lo %llu consumed %llu ready %llu in flight %d/%d"
              " | live %d%d | pending %d resident %d admitted %lld",
              (unsigned long long)S->next_seq, (unsigned long long)S->ctl->dev_pub,
              (unsigned long long)S->lo_seq, (unsigned long long)S->consumed,
              (unsigned long long)ready, S->n_inflight[0], S->n_inflight[1],
              (int)!launch_ending(S, 0), (int)!launch_ending(S, 1), S->pending_n.load(),
              S->resident, S->n_admitted.load());
}

bool drained(perm_server* S) {
  return S->n_inflight[0] == 0 && S->n_inflight[1] == 0 && S->to_publish.empty();
}

void fail_all(perm_server* S, int st) {
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
}

void scheduler_loop(perm_server* S) {
  for (;;) {
    bool progressed = false;
    const int st = scheduler_step(S, progressed);
    if (st != PERM_OK) {
      // a HIP failure leaves the device state unknown: fail everything
      fail_all(S, st);
      return;
    }
    if (progressed) continue;
    if (trace_on() && !drained(S)) trace_state(S);
    std::unique_lock<std::mutex> lk(S->mu_q);
    if (!drained(S)) {
      // work on the GPU: look again shortly (or at once on a new request)
      S->cv_submit.wait_for(lk, std::chrono::microseconds(20),
                            [&] { return S->stopping.load() || !S->pending.empty(); });
      continue;
    }
    if (S->stopping.load() && S->to_prep.empty() && S->prepping == 0 && S->pending.empty())
      return;
    S->cv_submit.wait(lk, [&] { return S->stopping.load() || !S->pending.empty(); });
  }
}

void scheduler_main(perm_server* S) {
  perm::set_device(S->device);
  (void)hipSetDevice(S->device);
  scheduler_loop(S);
  // every request is finished: end the live launches
  (void)stop_kernels(S);
}

size_t dtype_size(int dtype) { return dtype == PERM_DTYPE_F64 ? 8 : 4; }

// A raw matrix of order n < 8 embedded as A (+) I of order 8 (same
// permanent), dense and (if sparse) CCS, in the raw dtype.
void embed_identity(RawMatrix& m, int n) {
  constexpr int E = perm::kServeMinN;
  const size_t es = dtype_size(m.dtype);
  std::vector<char> mat((size_t)E * E * es, 0);
  auto put_one = [&](char* dst) {
    if (m.dtype == PERM_DTYPE_F64) {
      const double one = 1.0;
      std::memcpy(dst, &one, 8);
    } else if (m.dtype == PERM_DTYPE_F32) {
      const float one = 1.0f;
      std::memcpy(dst, &one, 4);
    } else {
      const int one = 1;
      std::memcpy(dst, &one, 4);
    }
  };
  for (int i = 0; i < n; ++i)
    std::memcpy(mat.data() + (size_t)i * E * es, m.mat.data() + (size_t)i * n * es, n * es);
  for (int i = n; i < E; ++i) put_one(mat.data() + ((size_t)i * E + i) * es);
  m.mat.swap(mat);
  if (m.sparse) {
    std::vector<int> cptrs(E + 1), rows;
    std::vector<char> cvals;
    for (int c = 0; c < E; ++c) {
      cptrs[c] = (int)rows.size();
      if (c < n) {
        for (int e = m.cptrs[c]; e < m.cptrs[c + 1]; ++e) {
          rows.push_back(m.rows[e]);
          cvals.insert(cvals.end(), m.cvals.begin() + (size_t)e * es,
                       m.cvals.begin() + (size_t)(e + 1) * es);
        }
      } else {
        rows.push_back(c);
        cvals.resize(cvals.size() + es);
        put_one(cvals.data() + cvals.size() - es);
      }
    }
    cptrs[E] = (int)rows.size();
    m.cptrs.swap(cptrs);
    m.rows.swap(rows);
    m.cvals.swap(cvals);
  }
}

// Prepare r from its raw input (a prep thread, no lock held): start vector,
// column store and walk shape, padded to the served order; walk and item
// split.  Returns PERM_OK or an error status (r->error set).
int prepare_request(perm_server* S, Request* r) {
  RawMatrix& in = r->raw;
  int n = r->n;
  if (n < perm::kServeMinN) {
    embed_identity(in, n);
    n = perm::kServeMinN;
  }
  r->n_eff = n;
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
  r->product_bits = p.product_bits;
  if (p.zero) {
    perm::fill_zero_result(r->n, r->integral, r->result);
    r->raw = RawMatrix();
    return PERM_OK;
  }
  const int N = perm::serve_padded_n(n);
  r->n_pad = N;
  // walk shape: column l < 3 touches rows [0, k_l); rows from n on are the
  // padding (untouched, like the matrix's own rows beyond k2)
  const int k0 = p.nu > 0 && p.nu < n ? p.nu : n, k1 = p.nu1 > 0 && p.nu1 < n ? p.nu1 : n,
            k2 = p.nu2 > 0 && p.nu2 < n ? p.nu2 : n;
  r->shape_n = {k0, k1, k2};
  r->shape = {k0 < n ? k0 : N, k1 < n ? k1 : N, k2 < n ? k2 : N};
  if (perm::serve_variant(N, r->shape, 0) < 0) r->shape = {N, N, N};  // (not expected)
  perm::batch_split(n, S->cfg.item_log2, r->walk_bits, r->high_bits);
  const uint64_t lanes = 1ull << r->high_bits;
  const uint64_t groups64 = (lanes + perm::kWave - 1) / perm::kWave;
  const uint64_t max_items = (uint64_t)std::max(1, S->cfg.max_items);
  const uint64_t items0 = std::min<uint64_t>(groups64, max_items);
  r->groups = (uint32_t)((groups64 + items0 - 1) / items0);
  r->walk_items = (uint32_t)((groups64 + r->groups - 1) / r->groups);
  // padded column store (block b of n values -> block b of N values), then
  // the start vector and its low parts, written straight into a pinned
  // staging block the kernel copies from
  const size_t blocks_n = (size_t)(2 * (n - 1));
  r->cs_doubles = (size_t)(2 * (N - 1) + 1) * N;
  r->staged_bytes = align256((r->cs_doubles + 2 * (size_t)N) * sizeof(double));
  r->staged = S->staging.get(r->staged_bytes, r->staged_dev);
  if (!r->staged) {
    r->error = "hipHostMalloc of a staging slab failed";
    return PERM_ERR_HIP;
  }
  double* cs = (double*)r->staged;
  double* x0 = cs + r->cs_doubles;
  std::memset(r->staged, 0, r->staged_bytes);
  for (size_t b = 0; b < blocks_n; ++b)
    std::memcpy(cs + b * N, p.cs.data() + b * n, (size_t)n * sizeof(double));
  for (int j = 0; j < N; ++j) x0[j] = j < n ? p.x0[j] : 1.0;
  for (int j = 0; j < n; ++j) x0[N + j] = p.x0[n + j];
  r->cost = pass_cost(r, 0, r->walk_items);
  r->raw = RawMatrix();
  return PERM_OK;
}

// Prep thread: take submitted requests (a share of the queue at a time),
// prepare them, hand them to the scheduler (or finish them: exactly 0, or a
// preparation error).
void prep_loop(perm_server* S) {
  perm::set_device(S->device);
  const size_t threads = std::max<size_t>(1, S->preppers_n);
  std::vector<Request*> got, ready, done_now;
  for (;;) {
    got.clear();
    {
      std::unique_lock<std::mutex> lk(S->mu_q);
      // (at most prep_ahead prepared requests wait for admission: their
      // staging blocks are pinned memory)
      S->cv_prep.wait(lk, [&] {
        return S->stopping.load() ||
               (!S->to_prep.empty() && S->pending_n.load() < S->prep_ahead);
      });
      if (S->to_prep.empty()) return;  // stopping and drained
      const size_t take = std::min<size_t>(16, std::max<size_t>(1, S->to_prep.size() / threads));
      for (size_t i = 0; i < take; ++i) {
        got.push_back(S->to_prep.front());
        S->to_prep.pop_front();
      }
      S->prepping += (int)got.size();
    }
    ready.clear();
    done_now.clear();
    const Clock::time_point t0 = Clock::now();
    for (Request* r : got) {
      const int st = prepare_request(S, r);
      if (st != PERM_OK || r->n_pad == 0) {  // finished at once (error, or exactly 0)
        if (st != PERM_OK) r->status = st;
        done_now.push_back(r);
      } else {
        ready.push_back(r);
      }
    }
    add_ms(S->t_prep, t0);
    if (!done_now.empty()) {
      {
        std::lock_guard<std::mutex> lk(S->mu);
        const Clock::time_point now = Clock::now();
        for (Request* r : done_now) {
          r->t_done = now;
          r->finished = true;
        }
        S->stats.completed += (long long)done_now.size();
      }
      S->cv_done.notify_all();
    }
    {
      std::lock_guard<std::mutex> lk(S->mu_q);
      for (Request* r : ready) S->pending.push_back(r);
      S->prepping -= (int)got.size();
      S->pending_n.store((int)S->pending.size(), std::memory_order_release);
    }
    S->cv_submit.notify_one();
  }
}

// Copy a caller's matrix into a new request (no preparation yet).
int make_request(int dtype, const void* mat, const int* cptrs, const int* rows,
                 const void* cvals, int nov, Request*& out) {
  if (!mat) return perm::fail(PERM_ERR_ARG, "null pointer");
  if (dtype < PERM_DTYPE_I32 || dtype > PERM_DTYPE_F64) return perm::fail(PERM_ERR_ARG, "bad dtype");
  if (nov < 1) return perm::fail(PERM_ERR_ARG, "nov must be >= 1");
  if (nov > perm::kServeMaxN)
    return perm::fail(PERM_ERR_SIZE, "server requests are limited to n <= " +
                                         std::to_string(perm::kServeMaxN));
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

// Register requests (one lock) and wake the prep pool.  t0: when the
// submission call began (statistics).
int enqueue(perm_server* S, Request* const* rs, int count, unsigned long long* ids,
            Clock::time_point t0, const Clock::time_point* arrival = nullptr) {
  {
    std::lock_guard<std::mutex> lk(S->mu);
    const char* why = S->fatal != PERM_OK ? "server failed: " : S->stopping.load() ? "server is shutting down" : nullptr;
    if (why) {
      for (int i = 0; i < count; ++i) delete rs[i];
      return perm::fail(S->fatal != PERM_OK ? S->fatal : PERM_ERR_ARG,
                        std::string(why) + (S->fatal != PERM_OK ? S->fatal_msg : ""));
    }
    const Clock::time_point now = Clock::now();
    for (int i = 0; i < count; ++i) {
      Request* r = rs[i];
      r->t_submit = arrival ? arrival[i] : now;
      r->id = S->next_id++;
      S->table[r->id] = r;
      ids[i] = r->id;
    }
    S->stats.submitted += count;
  }
  {
    std::lock_guard<std::mutex> lk(S->mu_q);
    for (int i = 0; i < count; ++i) S->to_prep.push_back(rs[i]);
  }
  add_ms(S->t_submit, t0);
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
  for (Pass* p : S->to_publish) delete p;
  for (Pass* p : S->by_seq) delete p;
  for (KernelSlot& K : S->K) {
    if (K.exec) (void)hipGraphExecDestroy(K.exec);
    if (K.graph) (void)hipGraphDestroy(K.graph);
    if (K.e0) (void)hipEventDestroy(K.e0);
    if (K.e1) (void)hipEventDestroy(K.e1);
    if (K.s) (void)hipStreamDestroy(K.s);
  }
  if (S->d_state) (void)hipFree(S->d_state);
  if (S->d_arena) (void)hipFree(S->d_arena);
  if (S->h_arena) (void)hipHostFree(S->h_arena);
  S->staging.release();
  if (S->ctl) (void)hipHostFree(S->ctl);
  if (S->done) (void)hipHostFree(S->done);
  for (auto& kv : S->table) delete kv.second;
  delete S;
}

int setup(perm_server* S) {
  const perm_server_config& c = S->cfg;
  S->arena_bytes = (uint64_t)(c.arena_mb * 1048576.0) / 256 * 256;
  HIP_TRYS(hipMalloc((void**)&S->d_arena, S->arena_bytes));
  HIP_TRYS(hipHostMalloc((void**)&S->h_arena, S->arena_bytes, hipHostMallocCoherent));
  HIP_TRYS(hipHostGetDevicePointer((void**)&S->h_arena_dev, S->h_arena, 0));
  S->by_seq.assign(perm::kServeRing, nullptr);
  HIP_TRYS(hipHostMalloc((void**)&S->ctl, sizeof(ServeControl), hipHostMallocCoherent));
  std::memset(S->ctl, 0, sizeof(ServeControl));
  HIP_TRYS(hipHostMalloc((void**)&S->done, perm::kServeRing * sizeof(ServeDone),
                         hipHostMallocCoherent));
  std::memset(S->done, 0, perm::kServeRing * sizeof(ServeDone));
  HIP_TRYS(hipMalloc((void**)&S->d_state, sizeof(ServeState)));
  HIP_TRYS(hipMemset(S->d_state, 0, sizeof(ServeState)));
  HIP_TRYS(hipDeviceSynchronize());
  // the two persistent kernels, each captured once and replayed per launch;
  // the double-double X one on 1/16 of the main one's workgroups (each of its
  // workgroups holds a whole CU's register files)
  // (100 MHz counter; the double-double X kernel waits twice as long: it is
  // launched only after stopping the main one)
  const uint64_t idle_ticks = (uint64_t)std::max(1.0, c.idle_us * 100.0);
  for (int k = 0; k < 2; ++k) {
    KernelSlot& K = S->K[k];
    HIP_TRYS(hipStreamCreateWithFlags(&K.s, hipStreamNonBlocking));
    HIP_TRYS(hipEventCreate(&K.e0));
    HIP_TRYS(hipEventCreate(&K.e1));
    const uint32_t grid = k == 0 ? (uint32_t)c.grid : (uint32_t)std::max(16, c.grid / 16);
    HIP_TRYS(hipStreamBeginCapture(K.s, hipStreamCaptureModeThreadLocal));
    const hipError_t le =
        perm::launch_serve(k == 1, S->ctl, S->d_state, S->done, S->h_arena_dev, S->d_arena,
                           k == 1 ? 2 * idle_ticks : idle_ticks, grid, K.s);
    const hipError_t ce = hipStreamEndCapture(K.s, &K.graph);
    HIP_TRYS(le);
    HIP_TRYS(ce);
    HIP_TRYS(hipGraphInstantiate(&K.exec, K.graph, nullptr, nullptr, 0));
  }
  return PERM_OK;
}

}  // namespace

extern "C" {

void perm_server_default_config(perm_server_config* c) {
  if (!c) return;
  c->item_log2 = 16;
  c->grid = 1024;
  c->max_items = 4096;
  c->idle_us = 2000.0;
  c->max_resident = 8192;
  c->max_queued_work = 4.0e13;  // lane-instructions (about one second of GPU work)
  c->prep_threads = 4;
  c->arena_mb = 256.0;
}

int perm_server_create(const perm_server_config* cfg, perm_server** out) {
  if (!out) return perm::fail(PERM_ERR_ARG, "null pointer");
  perm_server_config c;
  perm_server_default_config(&c);
  if (cfg) {
    if (cfg->item_log2 > 0) c.item_log2 = cfg->item_log2;
    if (cfg->grid > 0) c.grid = cfg->grid;
    if (cfg->max_items > 0) c.max_items = cfg->max_items;
    if (cfg->idle_us > 0) c.idle_us = cfg->idle_us;
    if (cfg->max_resident > 0) c.max_resident = cfg->max_resident;
    if (cfg->max_queued_work > 0) c.max_queued_work = cfg->max_queued_work;
    if (cfg->prep_threads > 0) c.prep_threads = std::min(cfg->prep_threads, 64);
    if (cfg->arena_mb > 0) c.arena_mb = cfg->arena_mb;
  }
  if (c.item_log2 < 6 || c.item_log2 > 30) return perm::fail(PERM_ERR_ARG, "item_log2 in [6, 30]");
  if (c.max_items > (int)perm::kServeMaxItems) c.max_items = (int)perm::kServeMaxItems;
  if (c.grid > 65536) c.grid = 65536;
  // passes in flight never exceed the completion ring (one per resident request)
  if (c.max_resident > (int)perm::kServeRing / 2) c.max_resident = (int)perm::kServeRing / 2;
  if (c.arena_mb < 16.0) c.arena_mb = 16.0;
  if (int s = perm::select_device()) return s;
  auto* S = new perm_server();
  S->cfg = c;
  S->device = perm::current_device();
  if (int st = setup(S)) {
    const std::string msg = perm::last_error();
    release(S);
    return perm::fail(st, msg);
  }
  S->worker = std::thread(scheduler_main, S);
  S->preppers_n = (size_t)c.prep_threads;
  S->prep_ahead = std::max(1024, 2 * c.max_resident);
  for (int t = 0; t < c.prep_threads; ++t) S->preppers.emplace_back(prep_loop, S);
  *out = S;
  return perm::ok();
}

int perm_server_submit_dense(perm_server* S, int dtype, const void* mat, int nov,
                             unsigned long long* id) {
  if (!S || !id) return perm::fail(PERM_ERR_ARG, "null pointer");
  const Clock::time_point t0 = Clock::now();
  Request* r = nullptr;
  if (int s = make_request(dtype, mat, nullptr, nullptr, nullptr, nov, r)) return s;
  return enqueue(S, &r, 1, id, t0);
}

int perm_server_submit_many(perm_server* S, int count, const int* dtypes,
                            const void* const* mats, const int* const* cptrs,
                            const int* const* rows, const void* const* cvals, const int* novs,
                            unsigned long long* ids) {
  if (!S || count < 0 || (count > 0 && (!dtypes || !mats || !novs || !ids)))
    return perm::fail(PERM_ERR_ARG, "null pointer");
  if (cptrs && (!rows || !cvals)) return perm::fail(PERM_ERR_ARG, "null rows/cvals arrays");
  const Clock::time_point t0 = Clock::now();
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
    if (int s = enqueue(S, rs.data(), (int)rs.size(), ids, t0)) return s;
  if (st != PERM_OK)
    return perm::fail(st, "matrix " + std::to_string(rs.size()) + ": " + err);
  return perm::ok();
}

int perm_server_submit_sparse(perm_server* S, int dtype, const void* mat, const int* cptrs,
                              const int* rows, const void* cvals, int nov,
                              unsigned long long* id) {
  if (!S || !id || !cptrs) return perm::fail(PERM_ERR_ARG, "null pointer");
  const Clock::time_point t0 = Clock::now();
  Request* r = nullptr;
  if (int s = make_request(dtype, mat, cptrs, rows, cvals, nov, r)) return s;
  return enqueue(S, &r, 1, id, t0);
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
  out->admitted = S->n_admitted.load(std::memory_order_relaxed);
  out->pending = S->pending_n.load(std::memory_order_relaxed);
  // scheduler-owned counters, read without its lock (statistics only)
  out->launches = S->n_launches.load(std::memory_order_relaxed);
  out->items = S->n_items.load(std::memory_order_relaxed);
  out->uploads = S->n_publish.load(std::memory_order_relaxed);
  out->stops = S->n_stops.load(std::memory_order_relaxed);
  out->laps = (long long)(S->vhead_seen.load(std::memory_order_relaxed) / std::max<uint64_t>(1, S->arena_bytes));
  out->admit_ms = S->t_admit.load(std::memory_order_relaxed);
  out->publish_ms = S->t_publish.load(std::memory_order_relaxed);
  out->retire_ms = S->t_retire.load(std::memory_order_relaxed);
  out->stop_ms = S->t_stop.load(std::memory_order_relaxed);
  out->gpu_ms = S->t_gpu.load(std::memory_order_relaxed);
  out->dd_passes = S->n_dd.load(std::memory_order_relaxed);
  out->prep_ms = S->t_prep.load(std::memory_order_relaxed);
  out->submit_ms = S->t_submit.load(std::memory_order_relaxed);
  return perm::ok();
}

void perm_server_destroy(perm_server* S) {
  if (!S) return;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    std::lock_guard<std::mutex> lq(S->mu_q);
    S->stopping.store(true);
  }
  S->cv_prep.notify_all();
  for (auto& t : S->preppers)
    if (t.joinable()) t.join();  // every submitted request is prepared (or failed)
  S->cv_submit.notify_all();
  if (S->worker.joinable()) S->worker.join();  // drains admitted and pending work
  (void)hipSetDevice(S->device);
  for (KernelSlot& K : S->K)
    if (K.s) (void)hipStreamSynchronize(K.s);
  release(S);
}

int perm_server_drive(perm_server* S, int count, const int* dtypes, const void* const* mats,
                      const int* const* cptrs, const int* const* rows,
                      const void* const* cvals, const int* novs, const double* arrival_s,
                      perm_result* out, double* latency_ms, double* wall_s) {
  if (!S || count < 0 || (count > 0 && (!dtypes || !mats || !novs || !arrival_s || !out)))
    return perm::fail(PERM_ERR_ARG, "null pointer");
  std::vector<unsigned long long> ids(count);
  std::atomic<int> submitted{0};
  std::atomic<int> sub_status{PERM_OK};
  std::string sub_error;
  const Clock::time_point t0 = Clock::now();
  auto due = [&](int i) {
    return t0 + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(arrival_s[i]));
  };
  // Every request is handed over at (or after) its arrival time, and its
  // latency counts from the arrival time: a submitter that falls behind
  // shows in the latencies.  Requests already due are handed over in one
  // call (up to kDriveBatch), as an ingress hands over simultaneous arrivals.
  constexpr int kDriveBatch = 256;
  std::thread submitter([&] {
    std::vector<Request*> rs;
    std::vector<Clock::time_point> arr;
    for (int i = 0; i < count;) {
      const Clock::time_point d = due(i);
      for (;;) {
        const Clock::time_point now = Clock::now();
        if (now >= d) break;
        if (d - now > std::chrono::microseconds(200))
          std::this_thread::sleep_for(d - now - std::chrono::microseconds(100));
      }
      const Clock::time_point ts = Clock::now();
      rs.clear();
      arr.clear();
      int st = PERM_OK;
      for (int j = i; j < count && j - i < kDriveBatch && (j == i || due(j) <= ts); ++j) {
        const bool sp = cptrs && cptrs[j];
        Request* r = nullptr;
        st = make_request(dtypes[j], mats[j], sp ? cptrs[j] : nullptr, sp ? rows[j] : nullptr,
                          sp ? cvals[j] : nullptr, novs[j], r);
        if (st != PERM_OK) break;
        rs.push_back(r);
        arr.push_back(due(j));
      }
      if (st != PERM_OK) {
        for (Request* r : rs) delete r;
      } else {
        st = enqueue(S, rs.data(), (int)rs.size(), &ids[i], ts, arr.data());
      }
      if (st != PERM_OK) {
        sub_error = perm::last_error();
        sub_status.store(st);
        submitted.store(-1);
        return;
      }
      i += (int)rs.size();
      submitted.store(i, std::memory_order_release);
    }
  });
  int first_error = PERM_OK;
  for (int i = 0; i < count; ++i) {
    int k;
    while ((k = submitted.load(std::memory_order_acquire)) >= 0 && k <= i)
      std::this_thread::sleep_for(std::chrono::microseconds(20));
    if (k < 0) break;
    const int st = perm_server_wait(S, ids[i], out + i, latency_ms ? latency_ms + i : nullptr);
    if (st != PERM_OK && first_error == PERM_OK) first_error = st;
  }
  submitter.join();
  if (wall_s) *wall_s = std::chrono::duration<double>(Clock::now() - t0).count();
  if (sub_status.load() != PERM_OK) return perm::fail(sub_status.load(), sub_error);
  return first_error == PERM_OK ? perm::ok() : first_error;
}

// Closed-loop throughput of one configuration on the sample (requests per
// second), or a negative status.
static double serve_trial(const perm_server_config& c, int count, const int* dtypes,
                          const void* const* mats, const int* const* cptrs,
                          const int* const* rows, const void* const* cvals, const int* novs,
                          int& st) {
  perm_server* S = nullptr;
  if ((st = perm_server_create(&c, &S)) != PERM_OK) return -1.0;
  std::vector<unsigned long long> ids(count);
  std::vector<perm_result> res(count);
  const Clock::time_point t0 = Clock::now();
  st = perm_server_submit_many(S, count, dtypes, mats, cptrs, rows, cvals, novs, ids.data());
  if (st == PERM_OK) st = perm_server_wait_many(S, ids.data(), count, res.data(), nullptr);
  const double s = std::chrono::duration<double>(Clock::now() - t0).count();
  perm_server_destroy(S);
  return st == PERM_OK ? count / s : -1.0;
}

int perm_server_autotune(const perm_server_config* base, int count, const int* dtypes,
                         const void* const* mats, const int* const* cptrs, const int* const* rows,
                         const void* const* cvals, const int* novs, perm_server_config* best,
                         double* best_req_per_s) {
  if (!best || count <= 0 || !dtypes || !mats || !novs)
    return perm::fail(PERM_ERR_ARG, "null pointer or empty sample");
  perm_server_config c;
  perm_server_default_config(&c);
  if (base) c = *base;
  int st = PERM_OK;
  // warm the kernel and the host paths once
  (void)serve_trial(c, std::min(count, 256), dtypes, mats, cptrs, rows, cvals, novs, st);
  if (st != PERM_OK) return st;
  double best_rate = serve_trial(c, count, dtypes, mats, cptrs, rows, cvals, novs, st);
  if (st != PERM_OK) return st;
  // coordinate descent: item size, then the kernel's grid, then item size again
  const int item_opts[] = {12, 14, 16, 18, 20};
  const int grid_opts[] = {256, 512, 1024, 2048};
  for (int round = 0; round < 2; ++round) {
    for (int il : item_opts) {
      if (il == c.item_log2) continue;
      perm_server_config t = c;
      t.item_log2 = il;
      const double rate = serve_trial(t, count, dtypes, mats, cptrs, rows, cvals, novs, st);
      if (st != PERM_OK) return st;
      if (rate > best_rate) {
        best_rate = rate;
        c = t;
      }
    }
    if (round == 1) break;
    for (int g : grid_opts) {
      if (g == c.grid) continue;
      perm_server_config t = c;
      t.grid = g;
      const double rate = serve_trial(t, count, dtypes, mats, cptrs, rows, cvals, novs, st);
      if (st != PERM_OK) return st;
      if (rate > best_rate) {
        best_rate = rate;
        c = t;
      }
    }
  }
  *best = c;
  if (best_req_per_s) *best_req_per_s = best_rate;
  return perm::ok();
}

}  // extern "C"
