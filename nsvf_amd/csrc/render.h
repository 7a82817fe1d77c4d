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

// A compacted sample stream (structure of arrays, one entry per sample).
struct SampleStream {
  int32_t* idx;    // voxel id
  float* depth;    // midpoint depth along the ray
  float* dist;     // interval length
  int32_t* ray;    // ray id in the frame
};

}  // namespace nsvf
