// Lane-per-block ZFP coders of double fields (1D-3D).
#include "zfp_lane_kernels.hpp"

namespace ndic {

template void zfp_lane_encode_launch<double>(uint32_t, const ZfpChunk*, uint32_t, uint64_t, hipStream_t);
template void zfp_lane_decode_launch<double>(uint32_t, const ZfpChunk*, uint32_t, uint64_t, const ZfpBrick*, uint32_t,
                                          hipStream_t);

}  // namespace ndic
