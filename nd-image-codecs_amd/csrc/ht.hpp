// HT (FBCOT) block coder on the device (definitions in ht.hip).
#pragma once

#include "common.hpp"

namespace ndic {

// One code-block to decode: the packet parser fills these.
struct HtDecTask {
  const uint8_t* seg;   // codeword segment (cleanup then refinement bytes)
  void* out;            // top-left output sample
  int32_t* err;         // error slot of the owning chunk (first error wins)
  uint32_t stride;      // output row pitch in samples
  uint32_t len1, len2;  // cleanup / refinement lengths
  uint16_t w, h;
  uint8_t mmsb, passes, kmax;
  uint8_t flags;        // kTaskIncluded (seg is set) | kTaskSignMag | kTaskCausal | kTaskHeader
  uint32_t err_offset;  // byte offset reported with an error
};

// One code-block to encode.
struct HtEncTask {
  const int32_t* in;    // top-left coefficient in the Mallat plane
  uint8_t* slot;        // [ms_cap bytes MagSgn][4096 MEL][4096 VLC, production order]
  uint32_t* lens;       // out: {ms_len, mel_len, vlc_len, included}
  int32_t* err;
  uint32_t stride;
  uint32_t ms_cap;
  uint16_t w, h;
  uint8_t kmax, flags;  // kTaskSignMag: `in` holds sign-magnitude words
  uint16_t pad;
};

// kTaskCausal: the COD style's vertically causal bit (SigProp ignores the
// stripe below).  kTaskHeader: the packet header included the block; the
// parser turns it into kTaskIncluded only when it also sets the segment
// pointer from the packet body, so a packet that fails part-way leaves no
// included task without a segment.
enum : uint8_t { kTaskIncluded = 1, kTaskSignMag = 2, kTaskCausal = 4, kTaskHeader = 8 };
constexpr uint32_t kMelCap = 4096, kVlcCap = 4096;

// Upper bound of the MagSgn bytes of a w x h block at K_max.
__host__ __device__ inline uint32_t ht_ms_cap(uint32_t w, uint32_t h, uint32_t kmax) {
  return (uint32_t)((uint64_t)w * h * (kmax + 1) / 7 + 16);
}

void ht_init_tables();
// Orders the tasks heaviest first by work class (the bit length of the
// cleanup segment; blocks left out come last): a counting sort on the
// device into d_order (n entries), using d_counts (kHtClasses + 1 words).
constexpr uint32_t kHtClasses = 34;
void launch_ht_order(const HtDecTask* d_tasks, uint64_t n, uint32_t* d_counts, uint32_t* d_order, hipStream_t stream);
// d_order: the task indices to decode, in order (nullptr: 0..n-1)
void launch_ht_decode(const HtDecTask* d_tasks, const uint32_t* d_order, uint64_t n, hipStream_t stream);
// One code-block per wavefront with the lanes working on its columns
// together (wave-parallel stream unstuffing and MagSgn decode): the lowest
// latency per block.  A persistent launch fed by the work queue at `queue`
// (two zeroed counters the calling context owns; left zeroed).
void launch_ht_decode_wave(const HtDecTask* d_tasks, uint64_t n, uint64_t* queue, hipStream_t stream);
// The same decode in two passes: the VLC / MEL decode one lane per block,
// writing per-quad records (kHtRecBytes per task) to d_rec, then the MagSgn
// decode one wavefront per block from those records.
constexpr uint64_t kHtRecBytes = 32 * 32 * 4;
void launch_ht_decode_split(const HtDecTask* d_tasks, uint64_t n, uint32_t* d_rec, hipStream_t stream);
// its two passes, for callers that time them apart
void launch_ht_vlc(const HtDecTask* d_tasks, uint64_t n, uint32_t* d_rec, hipStream_t stream);
void launch_ht_magsgn(const HtDecTask* d_tasks, uint64_t n, const uint32_t* d_rec, hipStream_t stream);
// SigProp/MagRef of the tasks with more than one coding pass (after launch_ht_decode)
void launch_ht_refine(const HtDecTask* d_tasks, uint64_t n, hipStream_t stream);
void launch_ht_encode(const HtEncTask* d_tasks, uint64_t n, hipStream_t stream);
// The same encode with one code-block per wavefront (lanes over columns, bit
// offsets from wave prefix sums): the lowest latency per block.
void launch_ht_encode_wave(const HtEncTask* d_tasks, uint64_t n, uint64_t* queue, hipStream_t stream);
// Byte k of an encoded segment from its three pieces (see HtEncTask::slot).
__host__ __device__ inline uint8_t ht_seg_byte(const uint8_t* slot, uint32_t ms_cap, const uint32_t* lens, uint32_t k) {
  if (k < lens[0]) return slot[k];
  k -= lens[0];
  if (k < lens[1]) return slot[ms_cap + k];
  k -= lens[1];
  return slot[ms_cap + kMelCap + (lens[2] - 1 - k)];
}

}  // namespace ndic
