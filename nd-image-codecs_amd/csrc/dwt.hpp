// 5/3 DWT device passes (definitions in dwt.hip).
#pragma once

#include "common.hpp"

namespace ndic {

enum DwtPass { kDwtInvH = 0, kDwtInvV = 1, kDwtFwdV = 2, kDwtFwdH = 3, kDwtInvLevel = 4 };

// One pass of one level over every plane of one chunk (all planes of a
// chunk share their geometry).  Planes are plane_elems apart in each buffer.
struct DwtJob {
  const int32_t* src;     // pass input
  const int32_t* src_ll;  // inverse horizontal: LL rows (y < hl) come from here
  int32_t* dst;
  uint64_t stride;        // plane row pitch (= plane width)
  uint64_t plane_elems;
  uint32_t w, h;          // level region
  uint32_t hl;            // rows of the LL half (inverse horizontal)
  uint32_t tiles_x;       // fused inverse level: 64x64 output tiles per tile row
  uint64_t tiles;         // 64-wide tiles per line (fused inverse level: tiles per plane)
  uint64_t items, item_begin;
};

void launch_dwt_pass(int pass, const DwtJob* d_jobs, uint32_t njobs, uint64_t items, hipStream_t stream);

}  // namespace ndic
