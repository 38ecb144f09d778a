/*
 * ndic_hip.h — C ABI of the MI355X-native nd-lift-ht / nd-zfp codec path.
 *
 * Every entry point takes plain pointers and sizes.  The single-chunk calls
 * mirror the reference's feature-free chunk cores, which every ecosystem
 * (zarrs plugins, PyO3, WASM) calls; a binding swaps the Rust core for these
 * functions and keeps its codec metadata, stream formats and errors.
 *
 * Errors mirror ndic_core::Error (crates/ndic-core/src/error.rs:14-43): the
 * status is the variant, ndic_error carries the byte offset (Codestream) and
 * the message.  Output buffers are allocated by the library and released with
 * ndic_buffer_free.  All calls are safe to make from several host threads; the
 * library serialises device work onto its own HIP stream.
 */
#ifndef NDIC_HIP_H
#define NDIC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  NDIC_OK = 0,
  NDIC_INVALID_ARGUMENT = 1, /* Error::InvalidArgument */
  NDIC_CODESTREAM = 2,       /* Error::Codestream { offset } */
  NDIC_UNSUPPORTED = 3,      /* Error::Unsupported */
  NDIC_IO = 4,               /* Error::Io (device/runtime failures) */
} ndic_status;

typedef struct {
  int32_t status;
  uint64_t offset;
  char message[256];
} ndic_error;

typedef struct {
  uint8_t* data;
  size_t len;
} ndic_buffer;

/* Element types: SampleType (crates/ndic-core/src/dtype.rs:12) and LiftDtype
 * (crates/ndic-zarr/src/lift.rs:18-35) share these codes. */
typedef enum {
  NDIC_U8 = 0, NDIC_I8 = 1, NDIC_U16 = 2, NDIC_I16 = 3,
  NDIC_U32 = 4, NDIC_I32 = 5, NDIC_U64 = 6, NDIC_I64 = 7,
  NDIC_F32 = 8, NDIC_F64 = 9,
} ndic_dtype;

/* One nd_lift step: AxisTransform (crates/ndic-lift/src/lib.rs:82-101).
 * kind: 0 delta, 1 haar, 2 lift53.  group 0 = whole extent. */
typedef struct {
  uint32_t dimension;
  uint32_t kind;
  uint32_t levels;
  uint32_t group;
} ndic_lift_step;

/* Htj2kConfig (crates/ndic-zarr/src/htj2k.rs:22-38).  progression: 0 LRCP,
 * 1 RLCP, 2 RPCL, 3 PCRL, 4 CPRL. */
typedef struct {
  uint32_t xy_levels;
  int32_t reversible;
  int32_t progression;
  int32_t index;
} ndic_htj2k_config;

void ndic_buffer_free(ndic_buffer* buf);
/* Library version string. */
const char* ndic_version(void);

/* Replaces ndic_zarr::lift::forward_chunk (crates/ndic-zarr/src/lift.rs:323):
 * little-endian dtype elements in C order -> int32 (int64 for 64-bit dtypes)
 * coefficient plane, with the encode-time overflow budget. */
ndic_status ndic_lift_forward_chunk(const uint8_t* bytes, size_t len, const size_t* shape, size_t ndim,
                                    ndic_dtype dtype, const char* version, const ndic_lift_step* steps,
                                    size_t nsteps, ndic_buffer* out, ndic_error* err);

/* Replaces ndic_zarr::lift::inverse_chunk (crates/ndic-zarr/src/lift.rs:339):
 * coefficient plane -> dtype elements, refusing coefficients that do not
 * narrow back. */
ndic_status ndic_lift_inverse_chunk(const uint8_t* bytes, size_t len, const size_t* shape, size_t ndim,
                                    ndic_dtype dtype, const char* version, const ndic_lift_step* steps,
                                    size_t nsteps, ndic_buffer* out, ndic_error* err);

/* Replaces ndic_zarr::htj2k::encode_chunk (crates/ndic-zarr/src/htj2k.rs:286):
 * chunk bytes -> ndht container of per-plane Part 1/15 codestreams. */
ndic_status ndic_htj2k_encode_chunk(const uint8_t* bytes, size_t len, const size_t* shape, size_t ndim,
                                    ndic_dtype dtype, const ndic_htj2k_config* config, ndic_buffer* out,
                                    ndic_error* err);

/* Replaces ndic_zarr::htj2k::decode_chunk (crates/ndic-zarr/src/htj2k.rs:417). */
ndic_status ndic_htj2k_decode_chunk(const uint8_t* chunk, size_t len, const size_t* shape, size_t ndim,
                                    ndic_dtype dtype, ndic_buffer* out, ndic_error* err);

/* Replaces Codestream::parse / parse_prefix + decode_to_resolution
 * (crates/ndic-codestream/src/reader.rs:103-117, 478): decodes resolutions
 * 0..=max_res (max_res < 0: all) of one plane codestream, or of a byte-range
 * prefix of it (prefix != 0), to int32 samples (width*height of the reduced
 * image, DC level shift undone for unsigned planes).  The thumbnail path of
 * the ndht byte-range index. */
ndic_status ndic_htj2k_decode_plane(const uint8_t* codestream, size_t len, int32_t max_res, int32_t prefix,
                                    ndic_buffer* out, size_t* width, size_t* height, ndic_error* err);

/* ---- fused nd-lift-ht chain over device-resident batches --------------- */

/* One chunk of a batch: the codec chain transpose(done by the caller) ->
 * nd_lift -> htj2k over the int32 coefficient plane, as zarrs runs it for
 * codec_series output (crates/ndic-zarr/src/series.rs:203-404). */
#define NDIC_MAX_NDIM 8
#define NDIC_MAX_STEPS 4
typedef struct {
  uint64_t in_offset;  /* decode: compressed chunk offset in the input arena;
                          encode: sample offset (bytes) in the input arena */
  uint64_t in_len;
  uint64_t out_offset; /* decode: decoded samples (bytes) in the output arena;
                          encode: filled by the library */
  uint64_t out_len;    /* encode: filled by the library */
  uint32_t ndim;
  uint32_t shape[NDIC_MAX_NDIM];
  int32_t dtype;
  uint32_t xy_levels;
  uint32_t nsteps;
  ndic_lift_step steps[NDIC_MAX_STEPS];
  int32_t status;       /* filled by the library: ndic_status of this chunk */
} ndic_chunk_desc;

#define NDIC_BATCH_PLAIN 1 /* one chunk per launch sequence (the comparison bar) */

/* Decodes every chunk of a batch on the device: d_in holds the compressed
 * chunks, d_out receives the samples.  stream is a hipStream_t (NULL: the
 * library's stream).  Returns after the work is complete; per-chunk status in
 * descs[i].status, the first failure in err. */
ndic_status ndic_lift_ht_decode_batch(const uint8_t* d_in, uint8_t* d_out, ndic_chunk_desc* descs, size_t n,
                                      uint32_t flags, void* stream, ndic_error* err);

/* Encodes every chunk of a batch: d_in holds the samples, the codestreams are
 * written packed into d_out (capacity bytes); descs[i].out_offset/out_len
 * report where. */
ndic_status ndic_lift_ht_encode_batch(const uint8_t* d_in, uint8_t* d_out, size_t capacity, ndic_chunk_desc* descs,
                                      size_t n, uint32_t flags, void* stream, ndic_error* err);

/* One plane codestream (or a byte-range prefix of it, e.g. up to the ndht
 * index's prefix[r]) of a device-resident batch, decoded to resolution
 * max_res: the serving form of ndic_htj2k_decode_plane.  width/height/levels
 * are the plane's (known from the chunk it belongs to); out_width/out_height
 * are filled with the reduced size, whose int32 samples land at out_offset. */
typedef struct {
  uint64_t in_offset, in_len, out_offset;
  uint32_t width, height, levels;
  int32_t max_res;      /* < 0: all resolutions */
  int32_t prefix;       /* the bytes may stop after the last needed packet */
  uint32_t out_width, out_height;
  int32_t status;
} ndic_plane_desc;
ndic_status ndic_htj2k_decode_plane_batch(const uint8_t* d_in, uint8_t* d_out, ndic_plane_desc* descs, size_t n,
                                          void* stream, ndic_error* err);

/* The decode-only lossy path: as ndic_htj2k_decode_plane_batch, but a
 * codestream coded with the irreversible 9/7 wavelet and an expounded QCD
 * (style 2) is decoded instead of refused as Unsupported (reader.rs:326-330,
 * the check every other decode entry point keeps).  It assembles the pieces
 * the reference ships for that path -- the unscaled 9/7 lifting
 * (crates/ndic-htj2k/src/dwt/mod.rs:243-325), Quant::irrev_delta
 * (crates/ndic-codestream/src/quant.rs:204-221) and the block decoder's
 * mid-point reconstruction (block/decode.rs:5-6) -- as T.800 Annex E/F define
 * the decode: each coefficient is sign x magnitude x (delta x 2^depth x the K
 * factor the unscaled lifting leaves in its band) in f32, synthesised in f32,
 * rounded to the nearest integer (ties to even), clamped to the sample range
 * and level shifted.  Scope: one component without a component transform,
 * at most 30-bit samples; 5/3 codestreams in the same batch decode as usual.
 * Samples are int32 at out_offset, as for the reversible batch. */
ndic_status ndic_htj2k_decode_plane_batch_irreversible(const uint8_t* d_in, uint8_t* d_out, ndic_plane_desc* descs,
                                                       size_t n, void* stream, ndic_error* err);

/* ---- byte-range plans and plane-selective decodes ---------------------- */

/* Replaces RangeIndex::from_chunk_prefix + thumbnail / plane / thumbnail_3d
 * (crates/ndic-codestream/src/range.rs:141-369): what to fetch from an ndht
 * chunk, planned from its first `len` bytes (at least header + index).
 * target: 0 thumbnail (arg = max_px, plane 0), 1 plane (arg = z), 2
 * thumbnail_3d (arg = max_px; steps = the chunk's nd_lift steps: the
 * low-pass planes of every lifted leading axis).  `ranges` holds (start, end)
 * uint64 pairs, end inclusive, coalesced and ascending; `planes` the chunk
 * plane indices in fetch order (uint64); decoded_size the result's shape,
 * outermost first.  Both buffers are released with ndic_buffer_free. */
typedef struct {
  int32_t target;
  uint32_t max_res;
  uint32_t ndims;
  uint64_t decoded_size[34];
  uint64_t total_bytes;
  uint64_t header_len;   /* bytes of [header | index] (the first range) */
  ndic_buffer planes;
  ndic_buffer ranges;
} ndic_range_plan;
ndic_status ndic_range_plan_chunk(const uint8_t* prefix, size_t len, int32_t target, uint64_t arg,
                                  const ndic_lift_step* steps, size_t nsteps, ndic_range_plan* out, ndic_error* err);

/* Plane-selective decodes of device-resident ndht chunks, driven from each
 * chunk's own index in device memory (only the selected planes' codestreams
 * are read and decoded).  target 0 replaces the zarrs partial decoder
 * (crates/ndic-zarr/src/htj2k_codec.rs:319-411): the planes an axis-aligned
 * subset (start/count per axis) touches, decoded in full, their (y, x)
 * window written in C order as `dtype` samples (a sample that does not fit
 * is an error, as in narrow_samples).  target 1 decodes what thumbnail_3d
 * plans: the low-pass planes under `steps` at the largest resolution whose
 * long side fits max_px, as int32 samples.  The library fills max_res,
 * out_ndim / out_shape and out_bytes; with d_out NULL it only fills them. */
typedef struct {
  uint64_t in_offset, in_len, out_offset;
  uint32_t ndim;
  uint32_t shape[NDIC_MAX_NDIM];
  int32_t dtype;         /* target 0: the htj2k codec's dtype */
  uint32_t xy_levels;
  int32_t target;        /* 0: subset, 1: thumbnail_3d */
  uint32_t start[NDIC_MAX_NDIM], count[NDIC_MAX_NDIM];
  uint32_t max_px;
  uint32_t nsteps;
  ndic_lift_step steps[NDIC_MAX_STEPS];
  uint32_t max_res;      /* filled by the library */
  uint32_t out_ndim;
  uint64_t out_shape[NDIC_MAX_NDIM];
  uint64_t out_bytes;
  int32_t status;
} ndic_select_desc;
ndic_status ndic_htj2k_decode_select_batch(const uint8_t* d_in, uint8_t* d_out, ndic_select_desc* descs, size_t n,
                                           void* stream, ndic_error* err);

/* ---- nd_zfp / zfp (LLNL ZFP streams, codec version 5) ------------------ */

/* ZfpMode (crates/ndic-zfp/src/lib.rs:77-86) as carried by NdZfpConfig
 * (crates/ndic-zfp/src/chunk.rs:101-128): the mode, its one parameter, and
 * the legacy `dims` member (0 = absent: the registered `zfp` mapping, the
 * chunk shape is the field shape; 1..4 = squeeze singletons, left-pad). */
#define NDIC_ZFP_FIXED_RATE 0
#define NDIC_ZFP_FIXED_ACCURACY 1
#define NDIC_ZFP_FIXED_PRECISION 2
#define NDIC_ZFP_REVERSIBLE 3
typedef struct {
  int32_t mode;
  uint32_t precision; /* fixed_precision: bit planes, 1..=64 */
  double rate;        /* fixed_rate: bits per value, positive and finite */
  double tolerance;   /* fixed_accuracy: absolute error bound, >= 0 and finite */
  int32_t dims;       /* legacy nd_zfp dims (0: none) */
} ndic_zfp_config;

/* Replaces ndic_zfp::encode_chunk (crates/ndic-zfp/src/chunk.rs:358): chunk
 * bytes (u8/i8/u16/i16 promoted to int32, i32, i64, f32, f64) -> one full
 * ZFP stream (header + blocks, padded to a 64-bit word). */
ndic_status ndic_zfp_encode_chunk(const uint8_t* bytes, size_t len, const size_t* shape, size_t ndim,
                                  ndic_dtype dtype, const ndic_zfp_config* config, ndic_buffer* out,
                                  ndic_error* err);

/* Replaces ndic_zfp::decode_chunk (crates/ndic-zfp/src/chunk.rs:383): the
 * stream header must declare exactly this field, type and mode; fixed-rate
 * streams must have the exact computed length. */
ndic_status ndic_zfp_decode_chunk(const uint8_t* stream, size_t len, const size_t* shape, size_t ndim,
                                  ndic_dtype dtype, const ndic_zfp_config* config, ndic_buffer* out,
                                  ndic_error* err);

/* Replaces ndic_zfp::decode_chunk_brick / NdZfpBrickDecoder::decode_brick
 * (crates/ndic-zfp/src/chunk.rs:446-540): one 4^d brick of a fixed-rate
 * stream at per-axis brick coordinates over the effective shape; brick_shape
 * (nbrick entries) receives the clipped shape of the returned samples. */
ndic_status ndic_zfp_decode_brick(const uint8_t* stream, size_t len, const size_t* shape, size_t ndim,
                                  ndic_dtype dtype, const ndic_zfp_config* config, const size_t* brick,
                                  size_t nbrick, ndic_buffer* out, size_t* brick_shape, ndic_error* err);

/* One chunk of a device-resident ZFP batch. */
typedef struct {
  uint64_t in_offset;  /* encode: samples (element-aligned); decode: stream (8-byte aligned) */
  uint64_t in_len;
  uint64_t out_offset; /* decode: samples; encode: filled by the library */
  uint64_t out_len;    /* encode: filled by the library */
  uint32_t ndim;
  uint32_t shape[NDIC_MAX_NDIM];
  int32_t dtype;
  ndic_zfp_config config;
  int32_t status;      /* filled by the library */
  uint64_t index_offset; /* block index of a variable-size stream in the index arena (the *_indexed calls) */
  uint64_t index_len;    /* its bytes: 4 * (blocks + 1); 0: none */
} ndic_zfp_desc;

/* Batch forms of ndic_zfp_encode_chunk / ndic_zfp_decode_chunk over device
 * memory (the zarrs codec calls them one chunk at a time,
 * crates/ndic-zarr/src/zfp_codec.rs:174-520; the caller batches). */
ndic_status ndic_zfp_encode_batch(const uint8_t* d_in, uint8_t* d_out, size_t capacity, ndic_zfp_desc* descs,
                                  size_t n, void* stream, ndic_error* err);
ndic_status ndic_zfp_decode_batch(const uint8_t* d_in, uint8_t* d_out, ndic_zfp_desc* descs, size_t n,
                                  void* stream, ndic_error* err);

/* Block index of a variable-size stream (reversible, fixed-accuracy,
 * fixed-precision).  No counterpart in the reference, whose docs list an
 * explicit table for these modes as future work (docs/architecture/zfp.md,
 * "Brick index"): the stream bytes stay exactly the ZFP stream, and beside
 * it the index holds blocks + 1 little-endian uint32 bit offsets from the
 * stream's first block bit (the bit after the header), the last one the end
 * of the last block; block k occupies stream bits [header_bits + idx[k],
 * header_bits + idx[k + 1]).  Streams of 2^32 block bits or more get none.
 * ndic_zfp_encode_batch_indexed writes each variable-size stream's index
 * packed into d_index (8-byte aligned entries; index_capacity bytes) and
 * reports it in descs[i].index_offset / index_len (0 for fixed-rate
 * streams, which need none).  ndic_zfp_decode_batch_indexed decodes every
 * block of an indexed stream in parallel instead of walking the stream to
 * find them; a block whose bits disagree with the index fails its chunk with
 * a Codestream error.  Descriptors with index_len 0 decode as in
 * ndic_zfp_decode_batch. */
ndic_status ndic_zfp_encode_batch_indexed(const uint8_t* d_in, uint8_t* d_out, size_t capacity, uint8_t* d_index,
                                          size_t index_capacity, ndic_zfp_desc* descs, size_t n, void* stream,
                                          ndic_error* err);
ndic_status ndic_zfp_decode_batch_indexed(const uint8_t* d_in, const uint8_t* d_index, uint8_t* d_out,
                                          ndic_zfp_desc* descs, size_t n, void* stream, ndic_error* err);

/* One 4^d brick of a device-resident fixed-rate stream (the serving form of
 * NdZfpBrickDecoder::decode_brick, chunk.rs:484): shape is the effective
 * field shape, brick the per-axis brick coordinates; the clipped brick's
 * samples land at out_offset in C order. */
typedef struct {
  uint64_t in_offset;  /* 8-byte aligned */
  uint64_t in_len, out_offset;
  uint32_t ndim;
  uint32_t shape[4];
  uint32_t brick[4];
  int32_t dtype;
  double rate;
  int32_t status;
} ndic_zfp_brick_desc;
ndic_status ndic_zfp_decode_bricks(const uint8_t* d_in, uint8_t* d_out, ndic_zfp_brick_desc* descs, size_t n,
                                   void* stream, ndic_error* err);

/* The allocation-free form of ndic_zfp_decode_bricks, with the same
 * arguments, results and errors.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  It calls no
 * hipMalloc, hipFree or other allocating routine: its device tables live in
 * a fixed library-owned array on each GPU (under 1 MiB, held until the
 * process ends), so a call takes at most NDIC_ZFP_NOALLOC_MAX_BRICKS bricks
 * (more is an InvalidArgument error).  Calls on one GPU share that array and
 * the library serialises them; each returns after its kernels complete. */
#define NDIC_ZFP_NOALLOC_MAX_BRICKS 4096
ndic_status ndic_zfp_decode_bricks_noalloc(const uint8_t* d_in, uint8_t* d_out, ndic_zfp_brick_desc* descs,
                                           size_t n, void* stream, ndic_error* err);

/* Free and total memory of the device the library runs on (hipMemGetInfo).
 * No counterpart in the reference: a serving loop admits work by it. */
ndic_status ndic_device_memory(size_t* free_bytes, size_t* total_bytes, ndic_error* err);

/* ---- launch-parameter knobs (no counterpart in the reference) ----------
 * The autotuner (nd-image-codecs_amd/autotune.py) measures the kernel
 * variants per shape class and stores its choices here; tests pin them.
 * Knobs: "ht_decode_wave_max_blocks" -- batch windows of at most this many
 * code-blocks decode one block per wavefront, larger ones one block per lane;
 * "ht_encode_wave_max_blocks" -- the same choice for the encoder;
 * "parse_wave_max_planes" -- windows of at most this many planes parse one
 * plane per wavefront (bytes and tag trees staged in LDS), larger ones one
 * plane per thread; "ht_wave_grain" -- code-blocks a wavefront of the
 * persistent wave-per-block coders takes from its work queue at a time;
 * "fuse_dwt_lift" -- a lift-ht decode window whose chunks have a single lift
 * step along a plane axis synthesises the finest DWT level, lifts and
 * narrows in one pass when it holds at least this many fused work items
 * (64 x 32 tiles x lift segments; default 4096); 0: three passes (same
 * output); "pipeline_windows" --
 * 1 (default): the windows of a large decode batch alternate between two
 * streams so that their passes overlap; 2, 3: three or four windows in
 * flight on as many streams; 0: one after the other;
 * "window_elems" -- samples per batch window (0: NDIC_WINDOW_ELEMS, else 2^31);
 * "ht_lane_waves" -- waves per workgroup of the lane-per-block HT decoder
 * (0: four on windows with at least four waves per CU, else one; 1-3: one;
 * 4: four); "graph_windows" -- 1: each decode window's launch sequence is
 * captured into a HIP graph (cached per thread context, updated in place)
 * and launched as one; 0 (default): launched directly.
 * ndic_tune_set returns 0, -1 for an unknown key, -2 for a value out of
 * range; ndic_tune_get returns the value in force (-1: unknown key);
 * ndic_tune_reset restores the built-in default; ndic_tune_name lists the
 * keys (NULL past the last). */
int32_t ndic_tune_set(const char* key, int64_t value);
int64_t ndic_tune_get(const char* key);
int32_t ndic_tune_reset(const char* key);
const char* ndic_tune_name(int32_t index);

#ifdef __cplusplus
}
#endif

#endif /* NDIC_HIP_H */
