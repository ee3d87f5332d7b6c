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
//   3. store the item's record; the wave whose increment of the pass's done
//      counter is the last sums the pass's records in item order (lane i
//      takes items i, i+64, ... in order, then the fixed butterfly) and
//      writes the completion record to host memory: the sum depends only on
//      the pass, never on which waves ran which items;
//   4. with nothing to claim: one wave at a time (at most once per 0.5 us)
//      reads the host's stop flag and tail and copies new ring entries into
//      device memory; all waves leave together (a device stop flag) on the
//      host's stop, or when the launch's activity count (claims, new passes)
//      has not moved for the idle time on the deciding wave's own clock.
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
  const double* colpm = (const double*)uni64((uint64_t)h->colpm);
  const double* x0 = (const double*)uni64((uint64_t)h->x0);
  const uint64_t lanes = uni64(h->lanes);
  const int wb = (int)uni32((uint32_t)h->walk_bits);
  const int hb = (int)uni32((uint32_t)h->high_bits);
  const uint32_t groups = uni32(h->groups);
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

// The last item of pass H is done: sum its records in item order and write
// the completion record (lane 0), tag last.
__device__ void finish_pass(const ServeHeader* H, ServeDone* done, uint32_t lane) {
  const uint32_t items = uni32(H->items);
  const Partial* parts = (const Partial*)uni64((uint64_t)H->parts);
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
    ServeDone* d = done + H->done_slot;
    store_sys((uint64_t*)&d->sum.hi, __builtin_bit_cast(uint64_t, s.hi));
    store_sys((uint64_t*)&d->sum.lo, __builtin_bit_cast(uint64_t, s.lo));
    store_sys((uint64_t*)&d->sum.abs_sum, __builtin_bit_cast(uint64_t, m.a));
    store_sys((uint64_t*)&d->sum.sq_sum, __builtin_bit_cast(uint64_t, m.q));
    store_sys((uint64_t*)&d->sum.sq_run, __builtin_bit_cast(uint64_t, m.r));
    wait_mem();  // the sum is in host memory before its tag
    store_sys(&d->tag, H->tag + 1);
  }
}

// Passes one wave tries per claim round.
constexpr uint64_t kScan = 64;
// Least real-time ticks (10 ns) between two reads of the host's tail.
constexpr uint64_t kPollTicks = 50;
// Most real-time ticks between two increments of the activity count by a
// wave that keeps claiming (and at most a quarter of the idle time).
constexpr uint64_t kActivityTicks = 20000;

// DD = false: the passes of FP64 and double-double-product variants, 4 waves
// per SIMD (128 VGPRs).  DD = true: the double-double X passes (x and its
// low parts take 4N VGPRs), 1 wave per SIMD.  Each kind skips the other's
// passes and keeps its own head.
template <bool DD>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(DD ? 1 : 4)))
void serve_kernel(ServeControl* ctl, ServeState* st, ServeDone* done, uint64_t idle_ticks) {
  // The workgroup's current pass: its waves claim from it first, so the
  // waves of a CU tend to run one variant at a time and the instruction
  // cache holds few walk loops (several per CU thrash it).
  __shared__ unsigned long long wg_pass;
  if (threadIdx.x == 0) wg_pass = 0ull;
  __syncthreads();
  const uint32_t lane = threadIdx.x & (kWave - 1);
  uint64_t* const head_p = DD ? &st->head_dd : &st->head;
  uint64_t cursor = uni64(load_sys(DD ? &ctl->head_dd : &ctl->head));
  // idle: the launch-wide activity count has not changed for idle_ticks of
  // this wave's own clock
  uint64_t idle_seen = ~0ull, idle_since = 0, last_activity = 0;
  const uint64_t activity_ticks = idle_ticks / 4 < kActivityTicks ? idle_ticks / 4 : kActivityTicks;
  for (;;) {
    if (uni64(load_dev(&st->stop)) != 0) break;  // the host's stop, or the launch went idle
    ServeHeader* H = nullptr;
    uint32_t item = 0, items = 0;
    // first the workgroup's current pass
    {
      ServeHeader* h = (ServeHeader*)uni64(__hip_atomic_load(&wg_pass, __ATOMIC_RELAXED,
                                                             __HIP_MEMORY_SCOPE_WORKGROUP));
      if (h) {
        const uint32_t n_items = uni32(__hip_atomic_load(&h->items, __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT));
        // one increment per wave (lane 0 adds 1, the others 0)
        uint32_t t = 0;
        if (lane == 0) t = atomicAdd(&h->next_item, 1u);
        t = uni32(t);
        if (t < n_items) {
          H = h;
          item = t;
          items = n_items;
        }
      }
    }
    const uint64_t pub = uni64(load_dev(&st->pub));
    if (!H) {
      const uint64_t head = uni64(load_dev(head_p));
      if (cursor < head) cursor = head;
      const uint64_t scan_end = pub < cursor + kScan ? pub : cursor + kScan;
      for (uint64_t q = cursor; q < scan_end; ++q) {
        ServeHeader* h = (ServeHeader*)uni64(load_dev(&st->dring[q & (kServeRing - 1)]));
        if (((uni32((uint32_t)h->variant) & kServeDDFlag) != 0) != DD) {
          if (q == cursor) ++cursor;  // the other kernel's pass
          continue;
        }
        const uint32_t n_items = uni32(__hip_atomic_load(&h->items, __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT));
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
      if (cursor > head) atomicMax((unsigned long long*)head_p, (unsigned long long)cursor);
      if (H && lane == 0)
        __hip_atomic_store(&wg_pass, (unsigned long long)(uintptr_t)H, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    const uint64_t now = __builtin_amdgcn_s_memrealtime();
    if (H) {
      idle_seen = ~0ull;
      // the launch's activity count moves at least every activity_ticks
      // while this wave claims items (not on every claim: one counter for
      // the whole launch)
      if (now - last_activity >= activity_ticks) {
        last_activity = now;
        if (lane == 0) atomicAdd((unsigned long long*)&st->activity, 1ull);
      }
      const uint32_t* ids = (const uint32_t*)uni64((uint64_t)H->ids);
      const uint32_t walk_item = ids ? uni32(ids[item]) : item;
      const int vid = (int)(uni32((uint32_t)H->variant) & ~kServeDDFlag);
      const ItemSum r = dispatch<DD>(vid, H, walk_item,
                                     std::make_integer_sequence<int, kNumServeVariants>{});
      if (lane == 0) {
        Partial* p = (Partial*)uni64((uint64_t)H->parts) + item;
        const Partial v = make_partial(r.s, r.m);
        store_dev_d(&p->hi, v.hi);
        store_dev_d(&p->lo, v.lo);
        store_dev_d(&p->abs_sum, v.abs_sum);
        store_dev_d(&p->sq_sum, v.sq_sum);
        store_dev_d(&p->sq_run, v.sq_run);
      }
      wait_mem();  // the record is stored before the item counts as done
      uint32_t d = 0;
      if (lane == 0) d = atomicAdd(&H->done_items, 1u);
      d = uni32(d);
      if (d + 1 == items) {
        wait_mem();  // every record is read after the last count
        finish_pass(H, done, lane);
      }
      continue;
    }
    // nothing to claim: at most one wave at a time reads the host's control
    // block (stop flag, tail) and copies what it finds into device memory
    const uint64_t last = uni64(load_dev(&st->poll));
    bool poll = false;
    if (now - last >= kPollTicks) {
      unsigned long long old = 0;
      if (lane == 0)
        old = atomicCAS((unsigned long long*)&st->poll, (unsigned long long)last,
                        (unsigned long long)now);
      poll = uni64(old) == last;
    }
    if (poll) {
      if (uni64(load_sys(&ctl->stop)) != 0) {
        if (lane == 0) atomicMax((unsigned long long*)&st->stop, 1ull);
        break;
      }
      const uint64_t tail = uni64(load_sys(&ctl->tail));
      wait_mem();  // the ring entries are read after the tail
      if (tail > pub) {
        for (uint64_t q = pub + lane; q < tail; q += kWave)
          store_dev(&st->dring[q & (kServeRing - 1)], load_sys(&ctl->ring[q & (kServeRing - 1)]));
        wait_mem();  // the entries are stored before they are published
        if (lane == 0) {
          atomicMax((unsigned long long*)&st->pub, (unsigned long long)tail);
          atomicAdd((unsigned long long*)&st->activity, 1ull);
        }
        continue;
      }
    }
    // the launch ends when the activity count stood still for idle_ticks:
    // the wave that sees it stops every wave, so a launch never lingers with
    // a few waves (the host launches again for anything published meanwhile)
    const uint64_t act = uni64(load_dev(&st->activity));
    if (act != idle_seen) {
      idle_seen = act;
      idle_since = now;
    } else if (now - idle_since > idle_ticks) {
      if (lane == 0) atomicMax((unsigned long long*)&st->stop, 1ull);
      break;
    }
    __builtin_amdgcn_s_sleep(16);
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
                        uint64_t idle_ticks, uint32_t grid, hipStream_t stream) {
  if (grid == 0) return hipErrorInvalidValue;
  if (dd) serve_kernel<true><<<dim3(grid), dim3(256), 0, stream>>>(ctl, st, done, idle_ticks);
  else serve_kernel<false><<<dim3(grid), dim3(256), 0, stream>>>(ctl, st, done, idle_ticks);
  return hipGetLastError();
}

}  // namespace perm
