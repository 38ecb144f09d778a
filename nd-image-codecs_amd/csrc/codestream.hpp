// Device-side Part 1/15 codestream and ndht container handling (definitions
// in codestream.hip).
#pragma once

#include "common.hpp"
#include "ht.hpp"

namespace ndic {

// Geometry of one resolution's packet and its bands, shared by every plane of
// a chunk (crates/ndic-codestream/src/geometry.rs:77-127: Mallat positions,
// 64x64 code-blocks, maximal precincts).
struct GeoBand {
  uint32_t x0, y0, w, h;  // band rectangle in the Mallat plane
  uint32_t nx, ny;        // code-block grid
  uint32_t kidx;          // index into the plane's K_max table
  uint32_t block_first;   // first block of the band within the plane
};
struct GeoPacket {
  uint32_t band_first, nbands;
  uint32_t block_first, nblocks;
};

// Dequantisation of one irreversible (9/7) plane, filled by the parse: the
// step that turns the block decoder's aligned magnitude into a coefficient of
// the unscaled 9/7 lifting, per band index (0 LL, then HL, LH, HH of each
// resolution from the coarsest).
struct IrrevPlane {
  uint32_t on;  // the plane's codestream is 9/7 with an expounded QCD
  float scale[100];
};

// Decoded per-plane facts (written by the parse kernel / the encode planner).
struct DevPlane {
  int32_t depth;          // SIZ bit depth
  int32_t is_signed;
  uint8_t kmax[100];      // K_max per band index (1 + 3 * levels)
  uint32_t ok;
  uint32_t irrev;         // decode: the codestream is 9/7 (allowed by DevChunk::irrev)
  uint64_t cs_off;        // codestream offset within the chunk
  uint64_t cs_len;
  uint32_t packet_len[33];    // encode: bytes per packet (header + bodies)
  uint32_t packet_hdr[33];    // encode: packet header bytes
  uint32_t head_len;          // encode: main + tile-part header bytes
  uint64_t total_len;         // encode: codestream bytes
  int64_t lo, hi;             // encode: sample range (depth_for)
};

// One chunk of a codestream batch.
struct DevChunk {
  const uint8_t* in;      // decode: the compressed chunk
  uint64_t in_len;
  uint32_t ndim;
  uint32_t dims[32];
  uint32_t W, H, P, L;    // plane width/height, planes, decomposition levels
  int32_t max_res;        // decode: highest resolution decoded (L = all)
  uint32_t geo_packet;    // first GeoPacket of this chunk's geometry
  uint32_t geo_band;      // first GeoBand
  uint32_t blocks_per_plane;
  uint64_t task_base;     // first HT task
  uint64_t plane_base;    // first global plane
  int32_t* coef;          // Mallat planes (P x W x H)
  int32_t* err;           // {code, offset_lo, offset_hi, plane}
  uint8_t* tag;           // tag-tree scratch, tag_stride bytes per plane
  uint32_t tag_stride;
  uint32_t no_container;  // decode: bit 0 `in` is one bare codestream, bit 1 it may be a prefix
  const uint32_t* plane_sel;  // decode: container plane of each decoded plane (ascending), or null: plane p
  uint32_t P_all;             // decode: planes in the container (P when plane_sel is null)
  IrrevPlane* irrev;          // decode: 9/7 codestreams are decoded (one entry per plane), or refused (null)
  // encode
  int32_t* planes_in;     // the int32 spatial planes (P x W x H)
  int32_t sample_signed;  // the sample dtype is signed (SIZ signedness)
  int32_t progression;
  int32_t index;
  uint8_t* slots;         // HT slots base (per-task offsets in slot_off)
  uint8_t* head;          // per-plane header bytes, head_stride apart
  uint32_t head_stride;
  uint8_t* pkt_hdr;       // per-packet header bytes, pkt_stride apart
  uint32_t pkt_stride;
  uint32_t pkt_slots;     // packet slots per plane (the window's largest L + 1, shared by all chunks)
  uint8_t* chunk_head;    // container header + index bytes
  uint64_t chunk_head_len;
  uint64_t out_off;       // encode: byte offset of the chunk in the output
  uint64_t out_len;
};

// One chunk of a window copy: the top-left-anchored window (x0, y0, w, h)
// of each of `planes` int32 planes (W x H, row pitch W, plane_elems apart)
// written packed (plane, row, column) as `dtype` samples, with a range flag;
// range mode (dst null) folds the window's min / max into minmax[0..1].
struct WindowJob {
  const int32_t* src;
  uint8_t* dst;
  int32_t* flag;
  int64_t* minmax;
  uint64_t plane_elems;
  uint32_t W, x0, y0, w, h, planes;
  uint32_t dtype, dsize;
  uint64_t items, item_begin;  // rows: planes x h
};
void launch_window(const WindowJob* d_jobs, uint32_t njobs, uint64_t rows, hipStream_t stream);
// Sets `bytes` (a multiple of 4) at p to 32-bit `value`: a kernel rather
// than hipMemsetAsync, so that it can stand in a captured graph.
void launch_fill32(void* p, uint32_t value, uint64_t bytes, hipStream_t stream);

// Tag-tree scratch bytes a plane needs for a band grid of nx x ny blocks.
uint32_t tag_scratch_bytes(uint32_t nx, uint32_t ny);
// Sets every code-block task of the window to "skip" / "zero fill" (before
// launch_parse_planes, which marks the included ones).
void launch_init_tasks(const DevChunk* d_chunks, uint32_t nchunks, uint64_t ntasks, HtDecTask* tasks,
                       const GeoPacket* gp, const GeoBand* gb, hipStream_t stream);
// tag_stride: the window's tag-tree scratch per plane (DevChunk::tag_stride)
void launch_parse_planes(const DevChunk* d_chunks, uint32_t nchunks, uint64_t nplanes, DevPlane* planes,
                         HtDecTask* tasks, const GeoPacket* gp, const GeoBand* gb, uint32_t tag_stride,
                         hipStream_t stream);
void launch_walk_chunks(const DevChunk* d_chunks, uint32_t nchunks, DevPlane* planes, hipStream_t stream);
// Unsigned planes: DC un-shift with clamping (reader.rs:626-638), in place on
// each chunk's result planes (d_data[c], region d_dims[2c] x d_dims[2c+1]).
// Irreversible planes (IrrevPlane::on): the HT decoder's aligned magnitudes
// (sign and bits 30..0) to f32 coefficients of the unscaled 9/7 lifting, in
// place (T.800 E.1.1.2, the half-bin bit included).
void launch_dequantize(const DevChunk* d_chunks, uint32_t nchunks, uint64_t nplanes, hipStream_t stream);
// Irreversible planes: the synthesised f32 region of d_data[c] to the nearest
// int32 (ties to even) in the level-shifted sample range, in place.
void launch_irrev_round(const DevChunk* d_chunks, uint32_t nchunks, uint64_t nplanes, const DevPlane* planes,
                        int32_t* const* d_data, const uint64_t* d_dims, hipStream_t stream);
void launch_unshift(const DevChunk* d_chunks, uint32_t nchunks, uint64_t nplanes, const DevPlane* planes,
                    int32_t* const* d_data, const uint64_t* d_dims, hipStream_t stream);
void launch_plane_minmax(const DevChunk* d_chunks, uint32_t nchunks, uint64_t nplanes, DevPlane* planes,
                         hipStream_t stream);
void launch_plane_setup(const DevChunk* d_chunks, uint32_t nchunks, uint64_t nplanes, DevPlane* planes,
                        hipStream_t stream);
void launch_dc_shift(const DevChunk* d_chunks, uint32_t nchunks, uint64_t nplanes, const DevPlane* planes,
                     hipStream_t stream);
void launch_build_enc_tasks(const DevChunk* d_chunks, uint32_t nchunks, uint64_t nplanes, const DevPlane* planes,
                            const GeoPacket* gp, const GeoBand* gb, HtEncTask* tasks, uint32_t* lens,
                            uint64_t slot_stride, hipStream_t stream);
void launch_write_planes(const DevChunk* d_chunks, uint32_t nchunks, uint64_t nplanes, DevPlane* planes,
                         const HtEncTask* tasks, const uint32_t* lens, const GeoPacket* gp, const GeoBand* gb,
                         hipStream_t stream);
void launch_write_chunks(const DevChunk* d_chunks, uint32_t nchunks, DevPlane* planes, uint64_t* chunk_len,
                         hipStream_t stream);
void launch_assemble(const DevChunk* d_chunks, uint32_t nchunks, uint64_t nplanes, const DevPlane* planes,
                     const HtEncTask* tasks, const uint32_t* lens, const GeoPacket* gp, uint8_t* out,
                     hipStream_t stream);

}  // namespace ndic
