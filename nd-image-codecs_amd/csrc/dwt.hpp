// 5/3 DWT device passes (definitions in dwt.hip).
#pragma once

#include "common.hpp"

namespace ndic {

enum DwtPass { kDwtInvH = 0, kDwtInvV = 1, kDwtFwdV = 2, kDwtFwdH = 3, kDwtInvLevel = 4, kDwt97InvH = 5, kDwt97InvV = 6 };

struct IrrevPlane;

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
  // planes whose codestream is 9/7 (IrrevPlane::on): the 5/3 passes skip
  // them, the 9/7 passes (f32 bits in the int32 buffers) take only them;
  // null: every plane is 5/3
  const IrrevPlane* irrev;
};

void launch_dwt_pass(int pass, const DwtJob* d_jobs, uint32_t njobs, uint64_t items, hipStream_t stream);

struct DevPlane;

// The finest synthesis level of one chunk's planes fused with its inverse
// nd_lift step along a plane axis and the narrowing store (dwt_lift_kernel).
// Plane p of the chunk is ((o * len + z) * R + r) for outer slab o, lift-axis
// position z and remaining plane index r.  Work items: outer x segs x R x
// tiles (64 x 32 output tiles of a plane).
struct DwtLiftJob {
  const int32_t* A;         // Mallat coefficients (planes w * h apart, rows w apart)
  const int32_t* LL;        // reconstruction of the next coarser level (same layout)
  const DevPlane* planes;   // the chunk's planes: DC un-shift of unsigned ones
  uint8_t* dst;             // dtype samples in shape order (null: range only)
  int32_t* flag;            // set when a sample does not fit the dtype
  int64_t* minmax;          // range mode (dst null): lifted sample min / max
  uint64_t plane_elems;     // w * h
  uint32_t w, h;
  uint32_t tiles_x, tiles;  // tiles per tile row / per plane
  uint64_t outer, len, R, group, segs;
  uint32_t kind, levels, dtype, dsize;
  uint64_t items, item_begin;
};
#ifndef NDIC_DWT_LIFT_TILE_W
#define NDIC_DWT_LIFT_TILE_W 64
#endif
#ifndef NDIC_DWT_LIFT_ROW_PAIRS
#define NDIC_DWT_LIFT_ROW_PAIRS 2
#endif
// the fused pass's output tile: 256 threads, each one column and
// kDwtLiftRowPairs row pairs of it
constexpr uint32_t kDwtLiftTileW = NDIC_DWT_LIFT_TILE_W, kDwtLiftRowPairs = NDIC_DWT_LIFT_ROW_PAIRS;
constexpr uint32_t kDwtLiftTileH = 2 * kDwtLiftRowPairs * (256 / kDwtLiftTileW);
void launch_dwt_lift(const DwtLiftJob* d_jobs, uint32_t njobs, uint64_t items, hipStream_t stream);

}  // namespace ndic
