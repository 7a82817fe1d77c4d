// Data records of the fused render pipeline.
#pragma once

#include "common.h"

namespace nsvf {

// One ray-voxel hit: entry and exit depth and the voxel (leaf) id.  16 bytes,
// so a ray's hit row is read with one 128-bit load per hit.
struct HitRec {
  float lo;
  float hi;
  int32_t idx;
  int32_t pad;
};
static_assert(sizeof(HitRec) == 16, "HitRec must be 16 bytes");

// One sample of a sample stream: voxel id (-1 marks an unused slot), midpoint
// depth along the ray, interval length, and the ray id in the frame.  16
// bytes, written and read with one 128-bit access.
struct SampleRec {
  int32_t idx;
  float depth;
  float dist;
  int32_t ray;
};
static_assert(sizeof(SampleRec) == 16, "SampleRec must be 16 bytes");

}  // namespace nsvf
