// ryser_serve.hip -- the persistent serving kernel (see ryser_serve.hpp for
// the protocol between the host and the kernel).
//
// Each wave of the kernel loops:
//   1. claim: first from its workgroup's current pass (kept in LDS, so the
//      waves of a CU tend to run one variant and the instruction cache holds
//      few walk loops), else from its cursor (>= the device head of its kind)
//      over the published passes of its kind, one atomic increment of a
//      pass's item counter per pass tried; the first pass with an unclaimed
//      item gives the wave that item and becomes the workgroup's pass;
//   2. walk it: the item's 64-lane groups with the lane walk of the
//      single-matrix kernel (ryser_dense_impl.hpp) for the pass's variant,
//      called through one non-inlined function per variant, so each body is
//      register-allocated on its own (under its kernel's waves per SIMD: two
//      instances, FP64 / double-double products at 4, double-double X at 1);
//   3. store the item's record (also in the host mirror when the header asks
//      for it); the wave whose increment of the pass's done counter is the
//      last sums the pass's records in item order (lane i takes items i,
//      i+64, ... in order, then the fixed butterfly) and appends the
//      completion record to the queue in host memory: the sum depends only
//      on the pass, never on which waves ran which items;
//   4. with nothing to claim: one wave at a time (at most once per 0.5 us)
//      reads the host's tail into device memory; idle waves take up to 16
//      published passes at a time, copy their ranges from the host mirror
//      into the arena (many waves in parallel: each GPU read of host memory
//      is a PCIe round trip), mark them copied and advance `pub` over the
//      passes copied so far, in publication order.  All waves of a launch
//      leave together (a device stop flag of their kind) on the host's stop
//      (read by each wave at most once per ms), or when the kind's activity
//      count (claims, publications) has not moved for the idle time on the
//      deciding wave's own clock and nothing published waits to be copied.
// Every wait is bounded by the idle time measured on the 100 MHz real-time
// counter, so the kernel always ends once work stops arriving.
#include "ryser_dense_impl.hpp"
#include "ryser_serve.hpp"

namespace perm {

namespace {

__device__ __forceinline__ uint64_t uni64(uint64_t v) {
  return ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32)) << 32) |
         (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
}
__device__ __forceinline__ uint32_t uni32(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}

// Cross-wave data (counters, ring copies, item records) is read and written
// with relaxed atomics of device (agent) or system scope only: each such
// access carries its own coherence bits (the per-XCD L2 is bypassed or
// written through), so no access needs a whole-cache write-back or
// invalidation.  Acquire/release orderings at device scope would emit
// those (buffer_wbl2 / buffer_inv of the XCD's L2) on every item, and at
// hundreds of thousands of items per second they serialise the L2s.
// Ordering between a wave's own accesses comes from wait_mem().
__device__ __forceinline__ uint64_t load_sys(const uint64_t* p) {
  return __hip_atomic_load(const_cast<uint64_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ uint64_t load_dev(const uint64_t* p) {
  return __hip_atomic_load(const_cast<uint64_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double load_dev_d(const double* p) {
  return __builtin_bit_cast(double, load_dev((const uint64_t*)p));
}
__device__ __forceinline__ void store_sys(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ void store_dev(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_dev_d(double* p, double v) {
  store_dev((uint64_t*)p, __builtin_bit_cast(uint64_t, v));
}
// The wave's earlier vector memory accesses have completed (stores are
// acknowledged at the scope they were issued with) before any later one
// issues; the compiler moves no memory access across it.
__device__ __forceinline__ void wait_mem() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// A 64-bit word of a pass header, wave-uniform.  Header fields are read
// with device-scope loads: a launch's waves scan past completed passes'
// headers, whose arena space may be reused (and the new header copied in)
// while the launch lives, so a plain load could return a stale line of an
// XCD's L2.
__device__ __forceinline__ uint64_t hword(const ServeHeader* h, size_t off) {
  return uni64(load_dev((const uint64_t*)((const char*)h + off)));
}
#define HDR(h, field) hword((h), offsetof(ServeHeader, field))

struct ItemSum {
  DD s;
  Moments m;
};

// One item of a pass of variant V: `groups` consecutive 64-lane groups of
// the walk, summed per lane in double-double, then over the wave (lane 0
// holds the item's sum).
template <int V>
__device__ __noinline__ ItemSum serve_item(const ServeHeader* h, uint32_t walk_item) {
  constexpr ServeVariant v = kServeVariants[V];
  constexpr int N = v.n;
  constexpr int R = unroll_bits_for<N>();
  const double* colpm = (const double*)HDR(h, colpm);
  const double* x0 = (const double*)HDR(h, x0);
  const uint64_t lanes = HDR(h, lanes);
  const uint64_t bits = HDR(h, walk_bits);  // walk_bits, high_bits
  const int wb = (int)(uint32_t)bits;
  const int hb = (int)(uint32_t)(bits >> 32);
  const uint32_t groups = (uint32_t)HDR(h, groups);
  ItemSum r = {{0.0, 0.0}, {0.0, 0.0, 0.0}};
  const uint64_t base = (uint64_t)walk_item * groups * kWave + (threadIdx.x & (kWave - 1));
  for (uint32_t g = 0; g < groups; ++g) {
    const uint64_t lane = base + (uint64_t)g * kWave;
    DD s;
    Moments m;
    // (lanes past the walk's last occur only when it has fewer than 64: their
    // pattern 0 keeps the high bits wave-uniform, and their sums are dropped)
    lane_walk<N, R, v.mode, Nest<v.k0, v.k1, v.k2>>(colpm, x0, lane < lanes ? lane : 0ull, wb,
                                                    hb, s, m);
    if (lane < lanes) {
      r.s = dd_add(r.s, s);
      r.m.a += m.a;
      r.m.q += m.q;
      r.m.r += m.r;
    }
  }
  wave_reduce(r.s, r.m);
  return r;
}

// Variant I is compiled into the kernel of its kind only: the double-double
// X variants into the DD kernel, the others into the main one.
template <bool DD, int I>
__device__ __forceinline__ void try_variant(int vid, const ServeHeader* h, uint32_t walk_item,
                                            ItemSum& r) {
  if constexpr ((kServeVariants[I].mode == 2) == DD) {
    if (vid == I) r = serve_item<I>(h, walk_item);
  }
}

template <bool DD, int... I>
__device__ __forceinline__ ItemSum dispatch(int vid, const ServeHeader* h, uint32_t walk_item,
                                            std::integer_sequence<int, I...>) {
  ItemSum r = {{0.0, 0.0}, {0.0, 0.0, 0.0}};
  (try_variant<DD, I>(vid, h, walk_item, r), ...);
  return r;
}

// The last item of pass H is done: sum its records in item order and append
// the completion record (lane 0) at the next queue position, tag last.
__device__ void finish_pass(const ServeHeader* H, ServeState* st, ServeDone* done, uint32_t lane) {
  const uint64_t vi = HDR(H, variant);  // variant, items
  // (an export pass's sum is zero: its records are the walk's, not its own)
  const uint32_t items = ((uint32_t)vi & kServeExportFlag) ? 0u : (uint32_t)(vi >> 32);
  const Partial* parts = (const Partial*)HDR(H, parts);
  const uint64_t seq = HDR(H, seq);
  DD s = {0.0, 0.0};
  Moments m = {0.0, 0.0, 0.0};
  for (uint32_t i = lane; i < items; i += kWave) {
    const Partial* p = parts + i;
    s = dd_add(s, DD{load_dev_d(&p->hi), load_dev_d(&p->lo)});
    m.a += load_dev_d(&p->abs_sum);
    m.q += load_dev_d(&p->sq_sum);
    m.r += load_dev_d(&p->sq_run);
  }
  wave_reduce(s, m);
  if (lane == 0) {
    const uint64_t slot = atomicAdd((unsigned long long*)&st->done_count, 1ull);
    ServeDone* d = done + (slot & (kServeRing - 1));
    store_sys((uint64_t*)&d->sum.hi, __builtin_bit_cast(uint64_t, s.hi));
    store_sys((uint64_t*)&d->sum.lo, __builtin_bit_cast(uint64_t, s.lo));
    store_sys((uint64_t*)&d->sum.abs_sum, __builtin_bit_cast(uint64_t, m.a));
    store_sys((uint64_t*)&d->sum.sq_sum, __builtin_bit_cast(uint64_t, m.q));
    store_sys((uint64_t*)&d->sum.sq_run, __builtin_bit_cast(uint64_t, m.r));
    store_sys(&d->seq, seq);
    wait_mem();  // the record is in host memory before its tag
    store_sys(&d->tag, slot + 1);
  }
}

// Copy up to 64 ranges (lane i: `words` 8-byte words from `src` to `dst`;
// 0 words for none) as one flat run of words spread over the lanes:
// kCopyDepth loads in flight per lane, so the copy of a batch of small
// passes takes a few PCIe round trips, not two per pass.  pre/from/to: this
// wave's LDS scratch.
constexpr uint32_t kCopyDepth = 12;
__device__ void copy_ranges(const char* src, char* dst, uint32_t words, uint32_t lane,
                            uint32_t* pre, uint64_t* from, uint64_t* to) {
  uint32_t incl = words;  // inclusive prefix sum over the lanes
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const uint32_t y = (uint32_t)__shfl_up((int)incl, (unsigned)d, kWave);
    if ((int)lane >= d) incl += y;
  }
  pre[lane] = incl - words;
  from[lane] = (uint64_t)(uintptr_t)src;
  to[lane] = (uint64_t)(uintptr_t)dst;
  if (lane == kWave - 1) pre[kWave] = incl;
  __builtin_amdgcn_wave_barrier();
  asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
  const uint32_t total = pre[kWave];
  uint32_t q = 0;  // this lane's range
  for (uint32_t w0 = 0; w0 < total; w0 += kWave * kCopyDepth) {
    uint64_t v[kCopyDepth], a[kCopyDepth];
#pragma unroll
    for (uint32_t j = 0; j < kCopyDepth; ++j) {
      const uint32_t w = w0 + j * kWave + lane;
      a[j] = 0;
      v[j] = 0;
      if (w < total) {
        while (pre[q + 1] <= w) ++q;
        const uint64_t k = (uint64_t)(w - pre[q]) * 8;
        a[j] = to[q] + k;
        v[j] = load_sys((const uint64_t*)(uintptr_t)(from[q] + k));
      }
    }
#pragma unroll
    for (uint32_t j = 0; j < kCopyDepth; ++j)
      if (w0 + j * kWave + lane < total) store_dev((uint64_t*)(uintptr_t)a[j], v[j]);
  }
}

// Item `item` of an export pass: its chunk of the walk's records, device
// arena -> host mirror (8 words in flight per lane).
__device__ void export_item(const ServeHeader* H, uint32_t item, uint32_t lane) {
  const uint64_t count = HDR(H, lanes);
  const uint64_t r0 = (uint64_t)item * kServeExportChunk;
  const uint64_t r1 = r0 + kServeExportChunk < count ? r0 + kServeExportChunk : count;
  if (r1 <= r0) return;
  constexpr uint32_t kWords = sizeof(Partial) / 8;
  const uint64_t* src = (const uint64_t*)((const Partial*)HDR(H, parts) + r0);
  uint64_t* dst = (uint64_t*)((Partial*)HDR(H, host_parts) + r0);
  const uint32_t words = (uint32_t)(r1 - r0) * kWords;
  for (uint32_t w0 = 0; w0 < words; w0 += kWave * 8) {
    uint64_t v[8];
#pragma unroll
    for (uint32_t j = 0; j < 8; ++j) {
      const uint32_t w = w0 + j * kWave + lane;
      v[j] = w < words ? load_dev(src + w) : 0ull;
    }
#pragma unroll
    for (uint32_t j = 0; j < 8; ++j) {
      const uint32_t w = w0 + j * kWave + lane;
      if (w < words) store_sys(dst + w, v[j]);
    }
  }
}

// Passes one wave tries per claim round.
constexpr uint64_t kScan = 64;
// Least real-time ticks (10 ns) between two reads of the host's tail.
constexpr uint64_t kPollTicks = 50;
// Most passes one idle wave takes to copy in at a time (many waves copy in
// parallel: a GPU read of host memory is a PCIe round trip).
constexpr uint64_t kCopyBatch = 16;

// Advance `pub` over the passes whose copies are complete (64 flags per
// round, one per lane); a wave that advanced it bumps the kind's activity.
__device__ void advance_pub(ServeControl* ctl, ServeState* st, int k, uint32_t lane) {
  const uint64_t p0 = uni64(load_dev(&st->pub));
  uint64_t p = p0;
  for (;;) {
    const uint64_t q = p + lane;
    const bool ok = load_dev(&st->copied[q & (kServeRing - 1)]) == q + 1;
    const uint64_t miss = ~__builtin_amdgcn_ballot_w64(ok);
    const uint64_t run = miss ? (uint64_t)__builtin_ctzll(miss) : (uint64_t)kWave;
    p += run;
    if (run < (uint64_t)kWave) break;
  }
  if (p > p0 && lane == 0) {
    atomicMax((unsigned long long*)&st->pub, (unsigned long long)p);
    atomicAdd((unsigned long long*)&st->activity[k], 1ull);
    store_sys(&ctl->dev_pub, p);
  }
}
// Least real-time ticks between two reads of the host's stop flag by one wave
// (1 ms: a few million reads per second over the whole launch).
constexpr uint64_t kStopTicks = 100000;
// Most real-time ticks between two increments of the activity count by a
// wave that keeps claiming (and at most a quarter of the idle time).
constexpr uint64_t kActivityTicks = 20000;

// DD = false: the passes of FP64 and double-double-product variants, 4 waves
// per SIMD (128 VGPRs).  DD = true: the double-double X passes (x and its
// low parts take 4N VGPRs), 1 wave per SIMD.  Each kind skips the other's
// passes and keeps its own head, stop flag and activity count; either
// kind's idle waves poll the host and copy new passes for both.
template <bool DD>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(DD ? 1 : 4)))
void serve_kernel(ServeControl* ctl, ServeState* st, ServeDone* done, const char* h_arena,
                  char* d_arena, uint64_t idle_ticks) {
  constexpr int k = DD ? 1 : 0;
  // The workgroup's current pass: its waves claim from it first, so the
  // waves of a CU tend to run one variant at a time and the instruction
  // cache holds few walk loops (several per CU thrash it).
  __shared__ unsigned long long wg_pass;
  __shared__ uint32_t copy_pre[256 / kWave][kWave + 1];  // each wave's copy plan
  __shared__ uint64_t copy_from[256 / kWave][kWave];
  __shared__ uint64_t copy_to[256 / kWave][kWave];
  if (threadIdx.x == 0) wg_pass = 0ull;
  __syncthreads();
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint32_t wave = threadIdx.x / kWave;
  uint64_t cursor = uni64(load_sys(&ctl->head[k]));
  const uint64_t epoch = uni64(load_sys(&ctl->epoch[k]));
  // idle: the kind's activity count has not changed for idle_ticks of this
  // wave's own clock
  uint64_t idle_seen = ~0ull, idle_since = 0, last_activity = 0, last_stop = 0;
  uint32_t idle_rounds = 0;  // rounds in a row without work (longer sleeps after a few)
  const uint64_t activity_ticks = idle_ticks / 4 < kActivityTicks ? idle_ticks / 4 : kActivityTicks;
  for (;;) {
    const uint64_t now = __builtin_amdgcn_s_memrealtime();
    if (uni64(load_dev(&st->stop[k])) != 0) break;  // the host's stop, or the launch went idle
    if (now - last_stop >= kStopTicks) {
      last_stop = now;
      if (uni64(load_sys(&ctl->stop)) != 0) {
        if (lane == 0) {
          atomicMax((unsigned long long*)&st->stop[k], 1ull);
          store_sys(&ctl->why[k], 2ull);
          store_sys(&ctl->ended[k], epoch);
        }
        break;
      }
    }
    ServeHeader* H = nullptr;
    uint32_t item = 0, items = 0;
    // first the workgroup's current pass
    {
      ServeHeader* h = (ServeHeader*)uni64(__hip_atomic_load(&wg_pass, __ATOMIC_RELAXED,
                                                             __HIP_MEMORY_SCOPE_WORKGROUP));
      if (h) {
        const uint32_t n_items = (uint32_t)(HDR(h, variant) >> 32);
        uint32_t t = 0;
        if (lane == 0) t = atomicAdd(&h->next_item, 1u);
        t = uni32(t);
        if (t < n_items) {
          H = h;
          item = t;
          items = n_items;
        } else if (lane == 0) {
          // every item is claimed: the workgroup's waves stop trying it (an
          // atomic on a hot line per idle round slows every claimer)
          unsigned long long cur = (unsigned long long)(uintptr_t)h;
          __hip_atomic_compare_exchange_strong(&wg_pass, &cur, 0ull, __ATOMIC_RELAXED,
                                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
      }
    }
    if (!H) {
      const uint64_t pub = uni64(load_dev(&st->pub));
      const uint64_t head = uni64(load_dev(&st->head[k]));
      if (cursor < head) cursor = head;
      const uint64_t scan_end = pub < cursor + kScan ? pub : cursor + kScan;
      for (uint64_t q = cursor; q < scan_end; ++q) {
        ServeHeader* h = (ServeHeader*)uni64(load_dev(&st->dring[q & (kServeRing - 1)]));
        const uint64_t vi = HDR(h, variant);  // variant, items
        if ((((uint32_t)vi & kServeDDFlag) != 0) != DD) {
          if (q == cursor) ++cursor;  // the other kernel's pass
          continue;
        }
        const uint32_t n_items = (uint32_t)(vi >> 32);
        uint32_t t = 0;
        if (lane == 0) t = atomicAdd(&h->next_item, 1u);
        t = uni32(t);
        if (t < n_items) {
          H = h;
          item = t;
          items = n_items;
          break;
        }
        if (q == cursor) ++cursor;  // every item of pass q is claimed
      }
      if (cursor > head) atomicMax((unsigned long long*)&st->head[k], (unsigned long long)cursor);
      if (H && lane == 0)
        __hip_atomic_store(&wg_pass, (unsigned long long)(uintptr_t)H, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    if (H) {
      idle_seen = ~0ull;
      idle_rounds = 0;
      // the kind's activity count moves at least every activity_ticks while
      // this wave claims items (not on every claim: one counter per kind)
      if (now - last_activity >= activity_ticks) {
        last_activity = now;
        if (lane == 0) atomicAdd((unsigned long long*)&st->activity[k], 1ull);
      }
      const uint32_t vflags = (uint32_t)HDR(H, variant);
      if (!DD && (vflags & kServeExportFlag)) {
        export_item(H, item, lane);
        wait_mem();  // the chunk is in host memory before the item counts as done
        uint32_t d = 0;
        if (lane == 0) d = atomicAdd(&H->done_items, 1u);
        d = uni32(d);
        if (d + 1 == items) finish_pass(H, st, done, lane);
        continue;
      }
      uint32_t* ids = (uint32_t*)HDR(H, ids);
      const uint32_t walk_item =
          ids ? uni32(__hip_atomic_load(ids + item, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
              : item;
      const int vid = (int)(vflags & ~kServeDDFlag);
      const ItemSum r = dispatch<DD>(vid, H, walk_item,
                                     std::make_integer_sequence<int, kNumServeVariants>{});
      if (lane == 0) {
        const Partial v = make_partial(r.s, r.m);
        Partial* p = (Partial*)HDR(H, parts) + item;
        store_dev_d(&p->hi, v.hi);
        store_dev_d(&p->lo, v.lo);
        store_dev_d(&p->abs_sum, v.abs_sum);
        store_dev_d(&p->sq_sum, v.sq_sum);
        store_dev_d(&p->sq_run, v.sq_run);
        Partial* hp = (Partial*)HDR(H, host_parts);
        if (hp) {  // the host reads this pass's records
          hp += item;
          store_sys((uint64_t*)&hp->hi, __builtin_bit_cast(uint64_t, v.hi));
          store_sys((uint64_t*)&hp->lo, __builtin_bit_cast(uint64_t, v.lo));
          store_sys((uint64_t*)&hp->abs_sum, __builtin_bit_cast(uint64_t, v.abs_sum));
          store_sys((uint64_t*)&hp->sq_sum, __builtin_bit_cast(uint64_t, v.sq_sum));
          store_sys((uint64_t*)&hp->sq_run, __builtin_bit_cast(uint64_t, v.sq_run));
        }
      }
      wait_mem();  // the record is stored before the item counts as done
      uint32_t d = 0;
      if (lane == 0) d = atomicAdd(&H->done_items, 1u);
      d = uni32(d);
      if (d + 1 == items) {
        wait_mem();  // every record is read after the last count
        finish_pass(H, st, done, lane);
      }
      continue;
    }
    // nothing to claim: (a) one wave per kPollTicks reads the host's tail
    // into device memory; (b) idle waves take up to kCopyBatch of the passes
    // the host published, copy their ranges from the host mirror into the
    // arena and their ring entries into dring, mark them copied, and
    // advance `pub` over the passes copied so far (in order)
    {
      const uint64_t last = uni64(load_dev(&st->poll));
      if (now - last >= kPollTicks) {
        unsigned long long old = 0;
        if (lane == 0)
          old = atomicCAS((unsigned long long*)&st->poll, (unsigned long long)last,
                          (unsigned long long)now);
        if (uni64(old) == last) {
          const uint64_t tail = uni64(load_sys(&ctl->tail));
          if (lane == 0) atomicMax((unsigned long long*)&st->host_tail, (unsigned long long)tail);
        }
      }
    }
    {
      const uint64_t ht = uni64(load_dev(&st->host_tail));
      const uint64_t c = uni64(load_dev(&st->copy_claim));
      if (c < ht) {
        const uint64_t n = ht - c < kCopyBatch ? ht - c : kCopyBatch;
        unsigned long long old = ~0ull;
        if (lane == 0)
          old = atomicCAS((unsigned long long*)&st->copy_claim, (unsigned long long)c,
                          (unsigned long long)(c + n));
        if (uni64(old) == c) {
          // (the host wrote the ring entries before the tail this wave or
          // another read: they are read after it)
          // lane i < m: pass c+i's mirror range; lane m+i: its second range
          // (the first pass's matrix, from its pinned staging block)
          const uint32_t m = (uint32_t)n;  // <= kCopyBatch = 16
          const uint32_t pi = lane < m ? lane : lane - m;
          const bool mine = lane < 2 * m;
          const uint64_t q = (c + pi) & (kServeRing - 1);
          const uint64_t entry = lane < m ? load_sys(&ctl->ring[q]) : 0ull;
          const uint64_t up = mine ? load_sys(lane < m ? &ctl->up[q] : &ctl->up2[q]) : 0ull;
          const uint64_t src2 = mine && lane >= m ? load_sys(&ctl->up2_src[q]) : 0ull;
          const uint64_t off = (up >> 32) * 256;
          const uint32_t words = (lane >= m && src2 == 0) ? 0u : (uint32_t)(up & 0xffffffffu) * 32u;
          copy_ranges(lane < m ? h_arena + off : (const char*)(uintptr_t)src2, d_arena + off,
                      mine ? words : 0u, lane, copy_pre[wave], copy_from[wave], copy_to[wave]);
          if (lane < m) store_dev(&st->dring[q], entry);
          wait_mem();  // the copies and entries land before they are marked
          if (lane < m) store_dev(&st->copied[q], c + lane + 1);
          wait_mem();
          advance_pub(ctl, st, k, lane);
          idle_rounds = 0;
          continue;
        }
      }
    }
    // the launch ends when the kind's activity count stood still for
    // idle_ticks: the wave that sees it stops every wave of the launch, so a
    // launch never lingers with a few waves (the host launches again for
    // anything published meanwhile)
    const uint64_t act = uni64(load_dev(&st->activity[k]));
    if (act != idle_seen) {
      idle_seen = act;
      idle_since = now;
    } else if (now - idle_since > idle_ticks &&
               // never while the host has published passes not yet copied in
               // (a slow first touch of the host mirror must not end the
               // launch)
               uni64(load_sys(&ctl->tail)) > uni64(load_dev(&st->pub))) {
      idle_since = now;
    } else if (now - idle_since > idle_ticks) {
      if (lane == 0) {
        atomicMax((unsigned long long*)&st->stop[k], 1ull);
        store_sys(&ctl->why[k], 1ull | (act << 8));
        store_sys(&ctl->ended[k], epoch);  // the host may launch again (after this one)
      }
      break;
    }
    // back off: every idle round reads a few hot device-scope lines, and
    // thousands of idle waves polling them slow the claims of the busy ones
    ++idle_rounds;
    if (idle_rounds < 4) {
      __builtin_amdgcn_s_sleep(16);  // about 0.4 us
    } else {
      const uint32_t naps = idle_rounds < 12 ? 1u : idle_rounds < 20 ? 2u : 8u;  // 3.4 .. 27 us
      for (uint32_t i = 0; i < naps; ++i) __builtin_amdgcn_s_sleep(127);
    }
  }
}

}  // namespace

int serve_variant(int n_pad, NestShape s, int mode) {
  for (int i = 0; i < kNumServeVariants; ++i) {
    const ServeVariant& v = kServeVariants[i];
    if (v.n == n_pad && v.k0 == s.k0 && v.k1 == s.k1 && v.k2 == s.k2 && v.mode == mode) return i;
  }
  return -1;
}

hipError_t launch_serve(bool dd, ServeControl* ctl, ServeState* st, ServeDone* done,
                        const char* h_arena, char* d_arena, uint64_t idle_ticks, uint32_t grid,
                        hipStream_t stream) {
  if (grid == 0) return hipErrorInvalidValue;
  if (dd)
    serve_kernel<true><<<dim3(grid), dim3(256), 0, stream>>>(ctl, st, done, h_arena, d_arena,
                                                             idle_ticks);
  else
    serve_kernel<false><<<dim3(grid), dim3(256), 0, stream>>>(ctl, st, done, h_arena, d_arena,
                                                              idle_ticks);
  return hipGetLastError();
}

}  // namespace perm
