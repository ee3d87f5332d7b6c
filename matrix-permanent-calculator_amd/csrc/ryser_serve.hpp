// ryser_serve.hpp -- the persistent serving kernel: its descriptors, the
// host-written control block it polls, and its variant table.
//
// The server (perman_serve.cpp) keeps every resident matrix and every
// partial sum in one device arena, mirrored at the same offsets in pinned
// host memory.  Per pass of a request (its FP64 walk, or a double-double
// pass) the host writes a ServeHeader (and, on the first pass, the matrix's
// column store and start vector) into the mirror, then publishes it in a
// ring inside pinned, coherent host memory (ServeControl): the header's
// device address and the mirror range to copy.  One wave at a time of the
// persistent kernel reads `tail`, copies the new ranges from the mirror into
// the device arena, and publishes them to the other waves (`pub`): no copy
// is issued by the host.  The waves claim the published passes' items
// through the headers' atomic counters (publication order), walk them,
// store each item's record in the pass's record array (and in the host
// mirror when the host asks for a pass's records), and the wave that
// finishes a pass's last item sums the records in item order and appends
// the pass's sum to a completion queue in pinned host memory (ServeDone),
// which the host reads in order.  A launch ends when the host sets `stop`,
// or when no wave of its kind has claimed an item or seen a new pass for
// `idle` ticks: its waves leave together, never one by one while work may
// still come.  The host launches it again (one captured graph) when it
// publishes work and no launch of that kind is live.
//
// Orders are served by kernels of padded orders: a request of order n runs
// the kernel of N = serve_padded_n(n) >= n, its start vector padded with
// rows equal to 1 that no Gray column touches (every product is unchanged);
// n < 8 is embedded as A (+) I (the same permanent) and walked at N = 8.
#pragma once
#include <array>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nest_shape.hpp"
#include "ryser_device.hpp"

namespace perm {

constexpr int kServeMinN = 8;
constexpr int kServeMaxN = 40;
constexpr int serve_padded_n(int n) { return n < kServeMinN ? kServeMinN : (n + 3) / 4 * 4; }
// Passes published and not yet completed are kept in a ring of this size
// (the host never has more in flight).
constexpr uint32_t kServeRing = 1u << 16;
// Most items of one pass (the per-pass record array); a pass with more
// 64-lane groups gives each item several groups.
constexpr uint32_t kServeMaxItems = 16384;
// Records one item of an export pass copies to the host mirror.
constexpr uint32_t kServeExportChunk = 512;

// One pass of one request, in the device arena (copied from the host mirror
// by the polling wave, which thereby zeroes the counters).  192 bytes.
struct ServeHeader {
  const double* colpm;   // column store of the padded order N: 2(N-1)+1 blocks of N
  const double* x0;      // start vector (N doubles), then its low parts (N more)
  Partial* parts;        // one record per item of this pass
  Partial* host_parts;   // nullptr, or the same records in the host mirror
  const uint32_t* ids;   // walk item index of each item (nullptr: item i is walk item i)
  uint64_t lanes;        // lanes of the walk, 2^high_bits
  uint64_t seq;          // publication sequence (echoed in the completion record)
  int32_t walk_bits, high_bits;
  int32_t variant;       // served kernel variant (padded order, walk shape, mode)
  uint32_t items;        // items of this pass
  uint32_t groups;       // 64-lane groups per walk item
  // third 64-byte line: the counters the kernel updates (zero in the mirror)
  alignas(64) uint32_t next_item;  // items claimed
  uint32_t done_items;             // items finished
};
static_assert(sizeof(ServeHeader) == 192, "ServeHeader layout");

// Completion record of one pass (pinned host memory, a queue in completion
// order): the pass's sum and moments, its sequence, and `tag` = the record's
// queue position + 1, stored after the rest has landed.
struct ServeDone {
  Partial sum;
  uint64_t tag;
  uint64_t seq;
};
static_assert(sizeof(ServeDone) == 64, "ServeDone layout");

// Host-written control block (pinned, coherent host memory), polled by the
// kernel with system-scope loads.
struct ServeControl {
  uint64_t tail;     // passes published: ring/up[0, tail mod kServeRing) are written
  uint64_t stop;     // nonzero: waves leave after their current item
  uint64_t head[2];  // oldest pass of each kernel kind in flight when its launch began
  uint64_t epoch[2];  // launch number of each kind (written before the launch)
  uint64_t ended[2];  // written by the kernel: the launch number that decided to end
  uint64_t why[2];    // and why (diagnostics): 1 idle, 2 the host's stop; activity << 8
  uint64_t dev_pub;   // the device's publication count after each poll (diagnostics)
  uint64_t pad[5];
  uint64_t ring[kServeRing];  // device address of the header of pass q at q mod kServeRing
  uint64_t up[kServeRing];    // mirror range of pass q: (offset / 256) << 32 | bytes / 256
  // a second range of pass q (its matrix, on the first pass): the device
  // address of its source in pinned memory (0: none), and the arena range
  // as in `up`
  uint64_t up2_src[kServeRing];
  uint64_t up2[kServeRing];
};

// Device-side state, kept across launches.  Index k: 0 the main kernel, 1
// the double-double X kernel.
struct ServeState {
  uint64_t pub;          // passes copied in and in dring, in order (claimable below it)
  uint64_t poll;         // real-time tick of the last read of the host's tail
  uint64_t host_tail;    // the host's tail as last read
  uint64_t copy_claim;   // passes below it are taken by a copying wave
  uint64_t done_count;   // completion records written
  uint64_t head[2];      // passes of kind k before head[k] have every item claimed
  uint64_t stop[2];      // all waves of kind k leave (zeroed before each launch)
  uint64_t activity[2];  // kind k's claims and publications (idle: it stands still)
  uint64_t pad[5];
  uint64_t dring[kServeRing];   // device copy of ServeControl::ring
  uint64_t copied[kServeRing];  // q + 1 once pass q's range and ring entry are in
};

// Served kernel variants: padded order N in {8, 12, ..., 40}; FP64 and
// double-double products over the plain walk and the kMenuBatch shapes of
// N; double-double X over the plain walk.
struct ServeVariant {
  int n, k0, k1, k2, mode;
};

constexpr int count_serve_variants() {
  int c = 0;
  for (int n = kServeMinN; n <= kServeMaxN; n += 4) {
    c += 3;
    for (int k0 = 4; k0 <= n - 4; k0 += 4)
      for (int j = 0; j < kNestShapesPerK0; ++j) c += nest_menu_has(kMenuBatch, n, k0, j) ? 2 : 0;
  }
  return c;
}
constexpr int kNumServeVariants = count_serve_variants();

constexpr std::array<ServeVariant, kNumServeVariants> make_serve_variants() {
  std::array<ServeVariant, kNumServeVariants> v{};
  int c = 0;
  for (int n = kServeMinN; n <= kServeMaxN; n += 4) {
    v[c++] = {n, n, n, n, 0};
    v[c++] = {n, n, n, n, 1};
    v[c++] = {n, n, n, n, 2};
    for (int k0 = 4; k0 <= n - 4; k0 += 4)
      for (int j = 0; j < kNestShapesPerK0; ++j)
        if (nest_menu_has(kMenuBatch, n, k0, j)) {
          const NestShape s = nest_menu_shape(n, k0, j);
          v[c++] = {n, s.k0, s.k1, s.k2, 0};
          v[c++] = {n, s.k0, s.k1, s.k2, 1};
        }
  }
  return v;
}
constexpr std::array<ServeVariant, kNumServeVariants> kServeVariants = make_serve_variants();

// Variant id of (padded order N, shape s, mode), or -1.
int serve_variant(int n_pad, NestShape s, int mode);
// Flag in ServeHeader::variant: a double-double X pass (run by the DD kernel).
constexpr uint32_t kServeDDFlag = 1u << 16;
// Flag in ServeHeader::variant: an export pass (main kernel) -- no walk; item
// i copies records [i * kServeExportChunk, ...) of `parts` (`lanes` records
// in all) to `host_parts`, for the host's sampled error measurement.
constexpr uint32_t kServeExportFlag = 1u << 17;

// Launch the persistent kernel of one kind (dd: the double-double X passes;
// else all others): `grid` workgroups of 256 lanes; it returns after
// `idle_ticks` (100 MHz ticks) without claimable work of its kind, or when
// ctl->stop is set.  h_arena: the device address of the host mirror of
// d_arena.
hipError_t launch_serve(bool dd, ServeControl* ctl, ServeState* st, ServeDone* done,
                        const char* h_arena, char* d_arena, uint64_t idle_ticks, uint32_t grid,
                        hipStream_t stream);

}  // namespace perm
