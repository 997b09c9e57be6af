// Pippenger G1 MSM against a device-resident prepared SRS layer.
#pragma once
#include <vector>

#include "curve.hpp"
#include "runtime.hpp"

namespace hp {
namespace msm {

// Window plan.  Windows use signed digits in [-2^(c-1), 2^(c-1)], so
// W = ceil(256 / c) windows cover every canonical Fr scalar with no final
// carry.  The prepared layer stores 2^(c w) P_i for every window w, so all W
// windows share one set of B = 2^(c-1) buckets and the bucket reduction runs
// once per MSM rather than once per window.
struct Plan {
  uint32_t c = 0;        // window bits
  uint32_t windows = 0;  // W
  uint32_t buckets = 0;  // B
  uint32_t task_cap = 0; // entries per accumulation task (one wavefront)
};

Plan choose_plan(size_t n, uint32_t override_c = 0);

struct Layer {
  size_t n = 0;               // points per window copy
  Plan plan;
  rt::BufferRef bases;        // W * n affine Montgomery points, window-major
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

// Queue the device part of sum_i scalar[i] * base_i (scalars: Montgomery Fr
// table of layer.n entries).  It leaves result_slots(layer) XYZZ values in
// `result` (device); finish() combines them on the host after they are copied.
size_t result_slots(const Layer& layer);
void enqueue(rt::Context& c, const Layer& layer, const Fr* scalars, G1Xyzz* result,
             StageEvents* events = nullptr);
G1Xyzz finish(const Layer& layer, const G1Xyzz* host_slots);

// Host: XYZZ -> 97-byte packet (unique affine representative).
void to_packet(const G1Xyzz& p, uint8_t* out97);
void affine_to_packet(const G1Affine& p, uint8_t* out97);

// Copy layer points (window 0) back as packets.
void download_layer(rt::Context& c, const Layer& layer, uint8_t* packets);

// Device fixed-base generation of g * s_i for a Montgomery scalar table.
void generate_points(rt::Context& c, const Fr* scalars, size_t n, G1Affine* out);

}  // namespace msm
}  // namespace hp
