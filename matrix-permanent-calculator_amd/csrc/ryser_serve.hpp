// ryser_serve.hpp -- the persistent serving kernel: its descriptors, the
// host-written control block it polls, and its variant table.
//
// The server (perman_serve.cpp) keeps every resident matrix and every
// partial sum in one device arena.  Per pass of a request (its FP64 walk, or
// a double-double pass) the host uploads a ServeHeader followed by the
// matrix's column store and start vector, then publishes the header's
// device address in a ring inside pinned, coherent host memory
// (ServeControl) and advances `tail`.  One persistent kernel polls `tail`:
// its waves claim the published passes' items through the headers' atomic
// counters (publication order), walk them, store each item's record in the
// pass's record array, and the wave that finishes a pass's last item sums
// the records in item order and writes the pass's sum into a completion
// record in pinned host memory (ServeDone).  Only one wave at a time reads
// the host's control block (at most once per 0.5 us) and copies what it
// finds into device memory.  A launch ends when the host sets `stop`, or
// when no wave has claimed an item or seen a new pass for `idle` ticks: its
// waves leave together, never one by one while work may still come.  The host launches it again (one captured graph) when it
// publishes work and no launch is live.
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
constexpr uint32_t kServeMaxItems = 4096;

// One pass of one request, in the device arena (uploaded; the kernel only
// writes the two counters, which the upload zeroes).  128 bytes.
struct ServeHeader {
  const double* colpm;   // column store of the padded order N: 2(N-1)+1 blocks of N
  const double* x0;      // start vector (N doubles), then its low parts (N more)
  Partial* parts;        // one record per item of this pass
  const uint32_t* ids;   // walk item index of each item (nullptr: item i is walk item i)
  uint64_t lanes;        // lanes of the walk, 2^high_bits
  int32_t walk_bits, high_bits;
  int32_t variant;       // served kernel variant (padded order, walk shape, mode)
  uint32_t items;        // items of this pass
  uint32_t groups;       // 64-lane groups per walk item
  uint32_t done_slot;    // completion record index
  // second 64-byte line: the counters the kernel updates (zero at upload)
  alignas(64) uint32_t next_item;  // items claimed
  uint32_t done_items;             // items finished
  uint64_t tag;                    // echoed in the completion record
};
static_assert(sizeof(ServeHeader) == 128, "ServeHeader layout");

// Completion record of one pass (pinned host memory): the pass's sum and
// moments; `tag` = header tag + 1 once written (stored after the sum has landed).
struct ServeDone {
  Partial sum;
  uint64_t tag;
  uint64_t pad;
};
static_assert(sizeof(ServeDone) == 64, "ServeDone layout");

// Host-written control block (pinned, coherent host memory), polled by the
// kernel with system-scope loads.
struct ServeControl {
  uint64_t tail;     // passes published: ring[0, tail mod kServeRing) hold header addresses
  uint64_t stop;     // nonzero: waves leave after their current item
  uint64_t head;     // oldest main-kernel pass in flight when its launch began
  uint64_t head_dd;  // the same for the double-double X kernel
  uint64_t pad[4];
  uint64_t ring[kServeRing];  // device address of the header of pass q at q mod kServeRing
};

// Device-side state, kept across launches.
struct ServeState {
  uint64_t head;      // main-kernel passes before it have every item claimed
  uint64_t head_dd;   // the same for the double-double X kernel
  uint64_t pub;       // passes whose header addresses are in dring
  uint64_t poll;      // real-time tick of the last read of the host's control block
  uint64_t activity;  // count of claims and new passes (idle: it stands still)
  uint64_t stop;      // all waves leave: the host's stop, or the launch went idle
                      // (zeroed before each launch)
  uint64_t pad[2];
  uint64_t dring[kServeRing];  // device copy of ServeControl::ring
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

// Launch the persistent kernel of one kind (dd: the double-double X passes;
// else all others): `grid` workgroups of 256 lanes; it returns after
// `idle_ticks` (100 MHz ticks) without claimable work of its kind, or when
// ctl->stop is set.
hipError_t launch_serve(bool dd, ServeControl* ctl, ServeState* st, ServeDone* done,
                        uint64_t idle_ticks, uint32_t grid, hipStream_t stream);

}  // namespace perm
