// Shared device helpers: vectorised Fr/Fq loads and stores, the
// canonical <-> Montgomery conversion kernels used at the host boundary, and
// host helpers that move canonical tables in and out of HBM.
#pragma once
#include "curve.hpp"
#include "runtime.hpp"

namespace hp {

// Fr occupies 32 bytes (two 16-byte loads), Fq 48 bytes (three).
__device__ __forceinline__ Fr load_fr(const Fr* base, size_t i) {
  const uint4* p = reinterpret_cast<const uint4*>(base + i);
  uint4 a = p[0], b = p[1];
  Fr r;
  r.v[0] = a.x; r.v[1] = a.y; r.v[2] = a.z; r.v[3] = a.w;
  r.v[4] = b.x; r.v[5] = b.y; r.v[6] = b.z; r.v[7] = b.w;
  return r;
}

__device__ __forceinline__ void store_fr(Fr* base, size_t i, const Fr& v) {
  uint4* p = reinterpret_cast<uint4*>(base + i);
  p[0] = make_uint4(v.v[0], v.v[1], v.v[2], v.v[3]);
  p[1] = make_uint4(v.v[4], v.v[5], v.v[6], v.v[7]);
}

__device__ __forceinline__ Fq load_fq(const Fq* p) {
  const uint4* q = reinterpret_cast<const uint4*>(p);
  uint4 a = q[0], b = q[1], c = q[2];
  Fq r;
  r.v[0] = a.x; r.v[1] = a.y; r.v[2] = a.z; r.v[3] = a.w;
  r.v[4] = b.x; r.v[5] = b.y; r.v[6] = b.z; r.v[7] = b.w;
  r.v[8] = c.x; r.v[9] = c.y; r.v[10] = c.z; r.v[11] = c.w;
  return r;
}

__device__ __forceinline__ void store_fq(Fq* p, const Fq& v) {
  uint4* q = reinterpret_cast<uint4*>(p);
  q[0] = make_uint4(v.v[0], v.v[1], v.v[2], v.v[3]);
  q[1] = make_uint4(v.v[4], v.v[5], v.v[6], v.v[7]);
  q[2] = make_uint4(v.v[8], v.v[9], v.v[10], v.v[11]);
}

__device__ __forceinline__ G1Affine load_affine(const G1Affine* base, size_t i) {
  return {load_fq(&base[i].x), load_fq(&base[i].y)};
}

__device__ __forceinline__ void store_affine(G1Affine* base, size_t i, const G1Affine& a) {
  store_fq(&base[i].x, a.x);
  store_fq(&base[i].y, a.y);
}

__device__ __forceinline__ G1Xyzz load_xyzz(const G1Xyzz* base, size_t i) {
  const G1Xyzz* p = base + i;
  return {load_fq(&p->X), load_fq(&p->Y), load_fq(&p->ZZ), load_fq(&p->ZZZ)};
}

__device__ __forceinline__ void store_xyzz(G1Xyzz* base, size_t i, const G1Xyzz& v) {
  G1Xyzz* p = base + i;
  store_fq(&p->X, v.X);
  store_fq(&p->Y, v.Y);
  store_fq(&p->ZZ, v.ZZ);
  store_fq(&p->ZZZ, v.ZZZ);
}

namespace rt {

// Upload `count` canonical Fr values (validated) into a new Montgomery table.
BufferRef upload_fr_table(Context& c, const uint8_t* host, size_t count, size_t num_vars);
// Upload `count` canonical Fr values into an existing device array (Montgomery).
void upload_fr_values(Context& c, Fr* dst, const uint8_t* host, size_t count);
// Copy `count` Montgomery Fr values back to host as canonical bytes (synchronises).
void download_fr(Context& c, const Fr* src, size_t count, uint8_t* host);
// Enqueues the canonical form of src[0..count) into dst (device or pinned
// host memory) on the context stream; no synchronisation.
void enqueue_from_mont(Context& c, const Fr* src, size_t count, Fr* dst);
// Copy raw device bytes to host (synchronises).
void download_bytes(Context& c, const void* src, size_t bytes, uint8_t* host);

}  // namespace rt
}  // namespace hp
