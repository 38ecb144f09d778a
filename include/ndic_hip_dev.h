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

/* HT decode of one code-block segment into sign-magnitude samples: the
 * cleanup pass, then SigProp and MagRef when num_passes is 2 or 3
 * (crates/ndic-htj2k/src/block/decode.rs:54, the stripe_causal argument as
 * there). */
ndic_status ndic_dev_ht_decode_block(const uint8_t* coded, size_t len, uint32_t* samples, size_t width,
                                     size_t height, uint32_t missing_msbs, uint32_t num_passes, size_t len_cleanup,
                                     size_t len_refinement, int stripe_causal, ndic_error* err);

/* Stage timing of the decode pipeline, for the roofline report: enable (and
 * reset) with ndic_dev_profile(1); ndic_dev_stage_times fills, per stage,
 * the summed device milliseconds (HIP events on the library's stream) and
 * the algorithmic bytes moved, and returns the number of batch windows timed.
 * Stages: 0 container/codestream/packet parse, 1 HT block decode, 2 inverse
 * DWT, 3 DC un-shift, 4 inverse nd_lift with dtype narrowing; when the HT
 * stage runs as two passes, entries 5, 6 and 7 split stage 1 into the VLC/MEL
 * pass, the MagSgn pass and the SigProp/MagRef passes (times only, bytes 0).
 * n is at most 8. */
void ndic_dev_profile(int32_t enable);
size_t ndic_dev_stage_times(double* ms, double* bytes, size_t n);

/* Device time of the ZFP coders' kernels while the profiler is on (HIP
 * events around their launches on the call's stream): per group, the summed
 * milliseconds and the number of timed spans.  Groups: 0 block encode
 * (zfp_encode_kernel), 1 placement of variable-size blocks and their index
 * (zfp_place_kernel, zfp_index_write_kernel), 2 block decode
 * (zfp_decode_kernel), 3 serial stream walk (zfp_index_kernel).  Returns the
 * number of groups filled (at most n). */
size_t ndic_dev_zfp_times(double* ms, uint64_t* spans, size_t n);

/* Device memory on the library's own HIP runtime and stream, for callers
 * (tests, the benchmark) that hold no other GPU framework.  kind of
 * ndic_dev_memcpy: 0 host->device, 1 device->host, 2 device->device; every
 * helper returns after the stream has drained. */
ndic_status ndic_dev_malloc(size_t bytes, void** out, ndic_error* err);
void ndic_dev_free(void* p);
ndic_status ndic_dev_memcpy(void* dst, const void* src, size_t n, int32_t kind, ndic_error* err);
ndic_status ndic_dev_memset(void* dst, int32_t value, size_t n, ndic_error* err);
ndic_status ndic_dev_synchronize(ndic_error* err);

/* Synthetic z,y,x volumes on the device: all-zero, uniform random over the
 * dtype, or 1000 + a sum of Gaussian blobs with Poisson-like noise (normal
 * approximation), clipped to 12 bits and scaled to the dtype (u8 >> 4,
 * u16 << 4, i16 (<< 4) - 32768).  Deterministic in (shape, blobs, seed). */
#define NDIC_SYNTH_ZERO 0
#define NDIC_SYNTH_RANDOM 1
#define NDIC_SYNTH_BLOBS 2
#define NDIC_SYNTH_MAX_BLOBS 8
typedef struct ndic_synth_blobs {
  int32_t count;
  float cz[NDIC_SYNTH_MAX_BLOBS], cy[NDIC_SYNTH_MAX_BLOBS], cx[NDIC_SYNTH_MAX_BLOBS];
  float amp[NDIC_SYNTH_MAX_BLOBS], sigma[NDIC_SYNTH_MAX_BLOBS];
} ndic_synth_blobs;
ndic_status ndic_dev_synth_chunk(void* d_out, uint32_t z, uint32_t y, uint32_t x, int32_t dtype, int32_t content,
                                 const ndic_synth_blobs* blobs, uint64_t seed, ndic_error* err);

/* A plain vectorised device-to-device copy kernel (16-byte aligned, n a
 * multiple of 16): the known-size reference for HBM traffic counters. */
ndic_status ndic_dev_copy(void* dst, const void* src, size_t n, ndic_error* err);

/* Number of differing bytes of two 16-byte aligned device buffers. */
ndic_status ndic_dev_count_mismatch(const void* a, const void* b, size_t n, uint64_t* count, ndic_error* err);

/* The same count added to a device counter (*d_count, 8-byte aligned) on the
 * calling thread's stream, without waiting: ordered after the thread's
 * earlier library calls, read back by the caller when it synchronises. */
ndic_status ndic_dev_count_mismatch_async(const void* a, const void* b, size_t n, uint64_t* d_count,
                                          ndic_error* err);

/* Content hash of n device bytes added to *d_slot (8-byte aligned) on the
 * calling thread's stream, without waiting: the sum mod 2^64 over the 64-bit
 * little-endian words w_i (the last zero-padded) of mix64(w_i ^ i*K) plus
 * mix64(n ^ K), K = 0x9e3779b97f4a7c15, mix64 the splitmix64 finaliser.
 * The serving benchmark checks responses against the hashes of expected
 * outputs computed on the host. */
ndic_status ndic_dev_hash_async(const void* p, size_t n, uint64_t* d_slot, ndic_error* err);

#ifdef __cplusplus
}
#endif

#endif /* NDIC_HIP_DEV_H */
