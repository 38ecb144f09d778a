// nd_lift device launches (definitions in lift.hip).
#pragma once

#include "common.hpp"

namespace ndic {

// One chunk x one step, as a list of work items: outer slabs x group
// segments x 64-column tiles (crates/ndic-lift/src/chunk.rs:110-150).
struct LiftJob {
  void* plane;          // the chunk's int32 / int64 coefficient plane
  void* tmp;            // chunk-sized temporary (global path only)
  const uint8_t* src;   // forward: widen from these dtype samples on load (or null)
  uint8_t* dst;         // inverse: narrow into these dtype samples on store (or null)
  int32_t* flag;        // set when a narrowed value does not fit the dtype
  uint64_t outer, len, inner, group;
  uint64_t tiles, segs, items, item_begin;
  uint32_t kind;        // 0 delta, 1 haar, 2 lift53, 3 none (convert only)
  uint32_t levels;
  uint32_t dtype;
  uint32_t dsize;
};

// Largest segment length staged in LDS (longer ones run from global memory).
size_t lift_lds_limit(bool wide);
// Longest segment the register-resident inverse kernel takes (int32 planes,
// at most two synthesis levels per segment).
uint32_t lift_col_limit();
// col: the register-resident inverse kernel (every job of the launch must be
// an inverse int32 job whose segments have at most two bands levels and at
// most lift_col_limit() rows).
void launch_lift_step(const LiftJob* d_jobs, uint32_t njobs, uint64_t items, bool fwd, bool wide, uint32_t max_n,
                      bool col, hipStream_t stream);
// Elementwise widen/narrow with optional per-call min/max (2 x int64) and
// range flag.
void launch_convert(const uint8_t* src, uint32_t src_dtype, void* plane, bool wide, uint8_t* dst, uint32_t dst_dtype,
                    uint64_t n, int64_t* minmax, int32_t* flag, hipStream_t stream);

}  // namespace ndic
