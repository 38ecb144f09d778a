/*
 * ndic_hip_dev.h — device-kernel test hooks of libndic_hip.so.
 *
 * Not part of the codec boundary: these run one stage of the path on the GPU
 * over host buffers so the parity tests can check a stage in isolation
 * against the oracle.  No reference interface corresponds to them.
 */
#ifndef NDIC_HIP_DEV_H
#define NDIC_HIP_DEV_H

#include "ndic_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* In-place 5/3 DWT (forward != 0) or its inverse over a width x height
 * int32 plane, `levels` decompositions (crates/ndic-htj2k/src/dwt/mod.rs). */
ndic_status ndic_dev_dwt53(int32_t forward, int32_t* plane, size_t width, size_t height, uint32_t levels,
                           ndic_error* err);

/* HT cleanup-pass encode of one code-block of sign-magnitude samples
 * (crates/ndic-htj2k/src/block/encode.rs:113).  *out_len = 0 when the block
 * has nothing significant. */
ndic_status ndic_dev_ht_encode_block(const uint32_t* samples, size_t width, size_t height, uint32_t missing_msbs,
                                     ndic_buffer* out, ndic_error* err);

/* HT decode of one code-block segment into sign-magnitude samples
 * (crates/ndic-htj2k/src/block/decode.rs:54). */
ndic_status ndic_dev_ht_decode_block(const uint8_t* coded, size_t len, uint32_t* samples, size_t width,
                                     size_t height, uint32_t missing_msbs, uint32_t num_passes, size_t len_cleanup,
                                     size_t len_refinement, ndic_error* err);

/* Stage timing of the decode pipeline, for the roofline report: enable (and
 * reset) with ndic_dev_profile(1); ndic_dev_stage_times fills, per stage,
 * the summed device milliseconds (HIP events on the library's stream) and
 * the algorithmic bytes moved, and returns the number of batch windows timed.
 * Stages: 0 container/codestream/packet parse, 1 HT block decode, 2 inverse
 * DWT, 3 DC un-shift, 4 inverse nd_lift with dtype narrowing. */
void ndic_dev_profile(int32_t enable);
size_t ndic_dev_stage_times(double* ms, double* bytes, size_t n);

#ifdef __cplusplus
}
#endif

#endif /* NDIC_HIP_DEV_H */
