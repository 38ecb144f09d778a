// ZFP (LLNL ZFP stream format, codec version 5) on the device: the block
// coders (zfp_lane_kernels.hpp: one lane per 4^d block for 1D-3D fields,
// four lanes per block for 4D fields), the stream walk, placement and block
// index (zfp.hip); orchestration in engine_zfp.cpp.
#pragma once

#include <vector>

#include "common.hpp"

namespace ndic {

// zfp_type codes (header field value + 1)
enum ZfpScalar : uint32_t { kZfpInt32 = 1, kZfpInt64 = 2, kZfpFloat = 3, kZfpDouble = 4 };
constexpr uint32_t kZfpMaxBits = 16657;  // ZFP_MAX_BITS of the C library 0.5.5
constexpr int32_t kZfpMinExp = -1074;    // ZFP_MIN_EXP

// Coding parameters of one stream (zfp_stream minbits/maxbits/maxprec/minexp).
struct ZfpParams {
  uint32_t minbits, maxbits, maxprec;
  int32_t minexp;
  bool reversible() const { return minexp < kZfpMinExp; }
};

// One field (chunk) of a ZFP batch.
struct ZfpChunk {
  uint32_t dims;          // 1..4
  uint32_t n[4];          // extents, x fastest (x = last axis of the row-major shape)
  uint32_t nb[4];         // blocks per axis
  uint64_t nblocks;
  uint64_t block_begin;   // global index of this field's block 0 within the launch
  uint32_t scalar;        // ZfpScalar
  int32_t dtype;          // element dtype of the samples (u8/i8/u16/i16 ride int32)
  uint32_t minbits, maxbits, maxprec;
  int32_t minexp;
  uint32_t fixed;         // fixed-rate layout: block k at header_bits + k * maxbits
  uint32_t slot_words;    // encode, variable modes: scratch words per block
  uint64_t header_bits;
  const uint8_t* src;     // encode: samples; decode: stream (8-byte aligned)
  uint64_t src_len;       // decode: stream bytes
  uint8_t* dst;           // encode: stream words (zeroed); decode: samples
  uint64_t* offs;         // variable modes: per-block bit lengths (encode) / bit offsets (decode)
  uint64_t* slots;        // encode, variable modes: per-block scratch
  // Block index of a variable-size stream (nblocks + 1 bit offsets from the
  // first block bit; the last entry is the end of the last block): encode
  // writes it, decode seeks with it instead of walking the stream.
  uint32_t* index;
  uint32_t* bad;          // decode with an index: set when a block's bits disagree with it
};

// One brick (block) decoded into its own clipped, C-ordered output.
struct ZfpBrick {
  uint32_t chunk;
  uint32_t pad;
  uint64_t block;         // block number within the field
  uint8_t* out;
};

// Stream header words (the first header_bits of a stream) for one chunk.
struct ZfpHeader {
  uint64_t* dst;
  uint64_t w[3];
  uint32_t nwords;
  uint32_t pad;
};
// Where the first bytes of a stream are, for the host-side header checks.
struct ZfpHead {
  const uint8_t* src;
  uint64_t len;
};
constexpr uint32_t kZfpHeadBytes = 24;  // 148 header bits at most

size_t zfp_scan_temp_bytes(uint64_t n);
void zfp_exclusive_scan(void* temp, size_t temp_bytes, const uint64_t* d_in, uint64_t* d_out, uint64_t n,
                        hipStream_t stream);
void launch_zfp_headers(const ZfpHeader* d_headers, uint32_t n, hipStream_t stream);
// totals[i] = bits of chunk i's blocks (variable-size modes), from the scan
void launch_zfp_totals(const ZfpChunk* d_chunks, uint32_t nchunks, const uint64_t* d_scan, uint64_t* d_totals,
                       hipStream_t stream);
void launch_zfp_heads(const ZfpHead* d_heads, uint32_t n, uint8_t* d_out, hipStream_t stream);
// Launches over `total` blocks of chunks of one (dims, scalar) class.
void launch_zfp_encode(const ZfpChunk* d_chunks, uint32_t nchunks, uint64_t total, uint32_t dims, uint32_t scalar,
                       hipStream_t stream);
void launch_zfp_place(const ZfpChunk* d_chunks, uint32_t nchunks, uint64_t total, const uint64_t* d_scan,
                      hipStream_t stream);
void launch_zfp_index(const ZfpChunk* d_chunks, uint32_t nchunks, uint32_t dims, uint32_t scalar,
                      hipStream_t stream);
// fixed_maxbits: the largest block size when every chunk is fixed-rate (it
// sizes the lane decoder's LDS stage), else 0.
void launch_zfp_decode(const ZfpChunk* d_chunks, uint32_t nchunks, uint64_t total, uint32_t dims, uint32_t scalar,
                       const ZfpBrick* d_bricks, uint32_t fixed_maxbits, hipStream_t stream);
// One lane per block for the 1D-3D classes (zfp_lane.hip); false for 4D.
bool launch_zfp_lane_encode(const ZfpChunk* d_chunks, uint32_t nchunks, uint64_t total, uint32_t dims,
                            uint32_t scalar, hipStream_t stream);
bool launch_zfp_lane_decode(const ZfpChunk* d_chunks, uint32_t nchunks, uint64_t total, uint32_t dims,
                            uint32_t scalar, const ZfpBrick* d_bricks, uint32_t fixed_maxbits, hipStream_t stream);
inline uint32_t zfp_fixed_maxbits(const std::vector<ZfpChunk>& ch) {
  uint32_t m = 0;
  for (const ZfpChunk& c : ch) {
    if (!c.fixed) return 0;
    m = c.maxbits > m ? c.maxbits : m;
  }
  return m;
}
// Variable-size modes, after the scan: writes each indexed chunk's block index.
void launch_zfp_index_write(const ZfpChunk* d_chunks, uint32_t nchunks, uint64_t total, const uint64_t* d_scan,
                            hipStream_t stream);

// Host plan of a brick batch (NdZfpBrickDecoder::decode_brick, chunk.rs:484,
// over BrickReader::fixed_rate's checks, lib.rs:509-599): per descriptor a
// status and message, and the launch entries of the good ones grouped by
// (dims, scalar) class.  heads holds kZfpHeadBytes stream bytes per brick.
struct BrickClass {
  uint32_t dims = 0, scalar = 0;
  std::vector<ZfpChunk> chunks;
  std::vector<ZfpBrick> bricks;
};
struct BrickPlan {
  std::vector<int> status;
  std::vector<std::string> msg;
  std::vector<BrickClass> classes;
};
void plan_bricks(const uint8_t* d_in, uint8_t* d_out, const ndic_zfp_brick_desc* descs, size_t n,
                 const uint8_t* heads, BrickPlan& bp);
// Writes each descriptor's status; returns the first failure (its message in err).
ndic_status report_bricks(ndic_zfp_brick_desc* descs, size_t n, const BrickPlan& bp, ndic_error* err);

}  // namespace ndic
