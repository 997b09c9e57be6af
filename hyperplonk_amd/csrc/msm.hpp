// Pippenger G1 MSM against a device-resident prepared SRS layer.
#pragma once
#include <vector>

#include "curve.hpp"
#include "runtime.hpp"

namespace hp {
namespace msm {

// Window plan.  Windows use signed digits in [-2^(c-1), 2^(c-1)] of scalars
// folded below (r - 1) / 2 < 2^254 (s > (r - 1) / 2 is taken as r - s against
// the negated point), so W = ceil(255 / c) windows end with no carry.  The prepared layer stores 2^(c w) P_i for every window w, so all W
// windows share one set of B = 2^(c-1) buckets and the bucket reduction runs
// once per MSM rather than once per window.
struct Plan {
  uint32_t c = 0;        // window bits
  uint32_t windows = 0;  // W
  uint32_t buckets = 0;  // B
};

Plan choose_plan(size_t n, uint32_t override_c = 0);
// The plan for a layer of n points on this context: override, tuned, or rule.
Plan plan_for(const rt::Context& c, size_t n, uint32_t override_c);
// Time candidate window sizes and lane task caps for a layer of 2^log_n
// random points (HIP events, median of 3 MSMs each); installs and returns the
// fastest in the context's tuning table.
void autotune(rt::Context& c, size_t log_n, uint32_t* best_c, uint32_t* best_cap, float* best_ms);

struct Layer {
  size_t n = 0;               // points per window copy
  Plan plan;
  rt::BufferRef bases;        // W * n affine points, window-major, packed Fq29 coordinates,
                              // 128 bytes per point (96 used)
};

// Validate 97-byte packets on the host (tags, canonical coordinates) and
// upload them; on-curve checks run on the device.  Returns false on a bad packet.
bool upload_layer(rt::Context& c, Layer& layer, const uint8_t* packets, size_t n,
                  uint32_t override_c = 0);
// Build a layer from device-resident affine points (window 0 copy).
void prepare_layer_from_device(rt::Context& c, Layer& layer, const G1Affine* points, size_t n,
                               uint32_t override_c = 0);

// Stage boundaries recorded on the context stream when profiling.
struct StageEvents {
  hipEvent_t marks[7];  // before digits, sort, tasks, accumulate, buckets, reduce, end
  uint32_t* entries_out = nullptr;  // device address of the non-zero digit count
};

// A batch of MSMs (layer, Montgomery scalar table of layer.n entries) run
// together: every stage is one launch over all of them in one global bucket
// space, so many small MSMs cost about as much as one.  enqueue() leaves
// result_slots() XYZZ values in `result` (device); after they are copied to
// the host, finish_packets() writes every MSM's point.  The layers must outlive
// the launch.
class Batch {
 public:
  void add(const Layer& layer, const Fr* scalars);
  size_t size() const { return entries_.size(); }
  void set_cap(uint32_t cap) { cap_override_ = cap; }
  size_t result_slots() const;
  unsigned red_group() const;  // buckets per level-1 reduction group
  // `hold`, when given, receives every device buffer the launches use, for
  // callers whose stream is not the context's (the buffers must outlive it)
  void enqueue(rt::Context& c, G1Xyzz* result, StageEvents* events = nullptr,
               std::vector<rt::BufferRef>* hold = nullptr) const;
  // After the result slots are copied to the host: one 97-byte packet per MSM.
  void finish_packets(const G1Xyzz* host_slots, uint8_t* out97) const;

 private:
  struct Entry {
    const Layer* layer;
    const Fr* scalars;
    uint32_t bucket_base, chunks;
    size_t cc_base;
  };
  std::vector<Entry> entries_;
  uint32_t gbuckets_ = 0;
  uint32_t cap_override_ = 0;
  size_t largest_n_ = 0;
  size_t cc_total_ = 0, m_total_ = 0;
};

// Host: XYZZ -> 97-byte packet (unique affine representative).
void to_packet(const G1Xyzz& p, uint8_t* out97);
void affine_to_packet(const G1Affine& p, uint8_t* out97);

// Copy layer points (window 0) back as packets.
void download_layer(rt::Context& c, const Layer& layer, uint8_t* packets);

// Compressed G1 encoding of SRS files (48 bytes per point: x big-endian, flag
// bits 0x80 compressed, 0x40 infinity, 0x20 y > -y on the first byte).
// upload_layer_compressed decompresses on the device (square root of
// x^3 + 4) and returns false for a missing compression flag, inconsistent
// infinity bits, x >= q or x off the curve; no subgroup check.
bool upload_layer_compressed(rt::Context& c, Layer& layer, const uint8_t* compressed, size_t n,
                             uint32_t override_c = 0);
void download_layer_compressed(rt::Context& c, const Layer& layer, uint8_t* compressed);

// Byte table d * 2^(8 w) * g (w < 32, d < 256) of the standard generator.
rt::BufferRef make_fixed_table(rt::Context& c);
// Device fixed-base generation of g * s_i for a Montgomery scalar table.
void generate_points(rt::Context& c, const Fr* scalars, size_t n, G1Affine* out,
                     const G1Affine* table);

}  // namespace msm
}  // namespace hp
