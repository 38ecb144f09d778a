// Lane-per-block ZFP coders of int64_t fields (1D-3D).
#include "zfp_lane_kernels.hpp"

namespace ndic {

template void zfp_lane_encode_launch<int64_t>(uint32_t, const ZfpChunk*, uint32_t, uint64_t, hipStream_t);
template void zfp_lane_decode_launch<int64_t>(uint32_t, const ZfpChunk*, uint32_t, uint64_t, const ZfpBrick*, uint32_t,
                                          hipStream_t);

}  // namespace ndic
