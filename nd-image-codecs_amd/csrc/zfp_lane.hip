// Dispatch of the lane-per-block ZFP coders (zfp_lane_kernels.hpp) by field
// class: dimensionality 1-3 and scalar type.
#include "zfp.hpp"

namespace ndic {

template <class S> void zfp_lane_encode_launch(uint32_t dims, const ZfpChunk* d, uint32_t n, uint64_t total,
                                               hipStream_t s);
template <class S> void zfp_lane_decode_launch(uint32_t dims, const ZfpChunk* d, uint32_t n, uint64_t total,
                                               const ZfpBrick* br, uint32_t words, hipStream_t s);

namespace {

// LDS stage per wave: the stretch of 64 fixed-rate blocks, else 32 KiB
uint32_t stage_words_for(uint32_t fixed_maxbits) {
  const uint64_t want = fixed_maxbits ? (64ull * fixed_maxbits + 63) / 64 + 4 : 4096;
  return (uint32_t)(want < 4096 ? want : 4096);
}

}  // namespace

bool launch_zfp_lane_encode(const ZfpChunk* d_chunks, uint32_t nchunks, uint64_t total, uint32_t dims,
                            uint32_t scalar, hipStream_t stream) {
  if (dims < 1 || dims > 4) return false;
  if (total == 0) return true;
  switch (scalar) {
    case kZfpInt32: zfp_lane_encode_launch<int32_t>(dims, d_chunks, nchunks, total, stream); break;
    case kZfpInt64: zfp_lane_encode_launch<int64_t>(dims, d_chunks, nchunks, total, stream); break;
    case kZfpFloat: zfp_lane_encode_launch<float>(dims, d_chunks, nchunks, total, stream); break;
    case kZfpDouble: zfp_lane_encode_launch<double>(dims, d_chunks, nchunks, total, stream); break;
    default: return false;
  }
  NDIC_HIP(hipGetLastError());
  return true;
}

bool launch_zfp_lane_decode(const ZfpChunk* d_chunks, uint32_t nchunks, uint64_t total, uint32_t dims,
                            uint32_t scalar, const ZfpBrick* d_bricks, uint32_t fixed_maxbits, hipStream_t stream) {
  if (dims < 1 || dims > 4) return false;
  if (total == 0) return true;
  const uint32_t words = stage_words_for(fixed_maxbits);
  switch (scalar) {
    case kZfpInt32: zfp_lane_decode_launch<int32_t>(dims, d_chunks, nchunks, total, d_bricks, words, stream); break;
    case kZfpInt64: zfp_lane_decode_launch<int64_t>(dims, d_chunks, nchunks, total, d_bricks, words, stream); break;
    case kZfpFloat: zfp_lane_decode_launch<float>(dims, d_chunks, nchunks, total, d_bricks, words, stream); break;
    case kZfpDouble: zfp_lane_decode_launch<double>(dims, d_chunks, nchunks, total, d_bricks, words, stream); break;
    default: return false;
  }
  NDIC_HIP(hipGetLastError());
  return true;
}

}  // namespace ndic
