// C-ABI entry points of the htj2k / nd-lift-ht path.
//   htj2k_encode / htj2k_decode: crates/ndic-zarr/src/htj2k.rs:286-441
//   plane prefix decode:         crates/ndic-codestream/src/reader.rs:103-117, 478-645
//   batches:                     the fused nd_lift -> htj2k codec chain over device memory
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <sstream>

#include "codestream.hpp"
#include "engine.hpp"
#include "ht.hpp"
#include "../../include/ndic_hip_dev.h"
#include "items.hpp"
#include "tune.hpp"

namespace ndic {

uint64_t window_elems_knob() {
  const int64_t k = tune(kTuneWindowElems);
  if (k > 0) return (uint64_t)k;
  static uint64_t w = [] {
    const char* e = std::getenv("NDIC_WINDOW_ELEMS");
    uint64_t v = e ? std::strtoull(e, nullptr, 10) : 0;
    return v ? v : (uint64_t)1 << 31;
  }();
  return w;
}

namespace {

uint64_t window_elems() { return window_elems_knob(); }

bool ht_dtype_ok(int d) {
  return d == NDIC_U8 || d == NDIC_I8 || d == NDIC_U16 || d == NDIC_I16 || d == NDIC_U32 || d == NDIC_I32;
}

// plane_geometry (htj2k.rs:113-159)
void check_geometry(const size_t* shape, size_t ndim, int dtype, size_t len, uint64_t* elems) {
  if (ndim < 2) {
    std::ostringstream m;
    m << "htj2k needs a chunk with trailing 2D (y, x) planes; got " << ndim << " dimension(s)";
    fail(NDIC_INVALID_ARGUMENT, m.str());
  }
  if (ndim > 32) fail(NDIC_INVALID_ARGUMENT, "htj2k chunks carry at most 32 dimensions");
  for (size_t i = 0; i < ndim; ++i)
    if (shape[i] == 0) {
      std::ostringstream m;
      m << "htj2k needs non-zero chunk extents; dimension " << i << " is 0";
      fail(NDIC_INVALID_ARGUMENT, m.str());
    }
  uint64_t e = 1;
  for (size_t i = 0; i < ndim; ++i) {
    if (shape[i] > 0xFFFFFFFFull || e > UINT64_MAX / 8 / shape[i]) fail(NDIC_INVALID_ARGUMENT, "chunk element count overflows");
    e *= shape[i];
  }
  if (shape[ndim - 1] > 0xFFFFFFFFull || shape[ndim - 2] > 0xFFFFFFFFull)
    fail(NDIC_INVALID_ARGUMENT, "chunk extent exceeds u32");
  if (len != e * dtype_size(dtype)) {
    std::ostringstream m;
    m << "chunk needs " << e * dtype_size(dtype) << " bytes, got " << len;
    fail(NDIC_INVALID_ARGUMENT, m.str());
  }
  *elems = e;
}

void raise_item(int status, const std::string& msg, uint64_t offset) {
  if (status != NDIC_OK) fail(status, msg, offset);
}

uint64_t encode_capacity(uint64_t elems, uint64_t planes, uint32_t L) {
  // HT output never exceeds (K_max + 1) / 7 bytes per sample plus per-block
  // MEL/VLC, and the headers are a few hundred bytes per plane.
  uint64_t blocks = elems / 1024 + planes * (3 * (uint64_t)L + 1) * 4 + 16;
  return elems * 5 + blocks * 8192 + planes * (1024 + 16 * (uint64_t)L) + 65536;
}

}  // namespace

}  // namespace ndic

using namespace ndic;

extern "C" {

ndic_status ndic_htj2k_encode_chunk(const uint8_t* bytes, size_t len, const size_t* shape, size_t ndim,
                                    ndic_dtype dtype, const ndic_htj2k_config* config, ndic_buffer* out,
                                    ndic_error* err) {
  return guard_call(err, [&] {
    if (!out || !config) fail(NDIC_INVALID_ARGUMENT, "htj2k: null argument");
    out->data = nullptr;
    out->len = 0;
    if (config->xy_levels > 32) fail(NDIC_INVALID_ARGUMENT, "htj2k: xy_levels exceeds the 32-level J2K bound (T.800 SPcod)");
    if (config->progression < 0 || config->progression > 4) fail(NDIC_UNSUPPORTED, "htj2k: unknown progression order");
    if (!config->reversible) fail(NDIC_UNSUPPORTED, "htj2k: lossy (reversible: false) coding lands with a later phase");
    uint64_t elems;
    check_geometry(shape, ndim, dtype, len, &elems);
    if (!ht_dtype_ok(dtype)) fail(NDIC_UNSUPPORTED, std::string("htj2k does not support dtype ") + dtype_name(dtype));
    Context& ctx = Context::get();
    std::lock_guard<std::recursive_mutex> lock(ctx.mu);
    hipStream_t s = ctx.stream;
    uint8_t* d_in = (uint8_t*)ctx.in.ensure(len + 16);
    NDIC_HIP(hipMemcpyAsync(d_in, bytes, len, hipMemcpyHostToDevice, s));
    std::vector<EncodeItem> items(1);
    EncodeItem& it = items[0];
    it.d_in = d_in;
    it.shape.assign(shape, shape + ndim);
    it.in_dtype = dtype;
    it.lift = false;
    it.L = config->xy_levels;
    it.progression = config->progression;
    it.index = config->index != 0;
    uint64_t planes = elems / ((uint64_t)shape[ndim - 1] * shape[ndim - 2]);
    uint64_t cap = encode_capacity(elems, planes, it.L);
    uint8_t* d_out = (uint8_t*)ctx.out.ensure(cap);
    encode_items(ctx, items, d_out, cap, s, window_elems());
    raise_item(it.status, it.message, 0);
    emit_host(d_out + it.out_off, it.out_len, out, s);
  });
}

ndic_status ndic_htj2k_decode_chunk(const uint8_t* chunk, size_t len, const size_t* shape, size_t ndim,
                                    ndic_dtype dtype, ndic_buffer* out, ndic_error* err) {
  return guard_call(err, [&] {
    if (!out) fail(NDIC_INVALID_ARGUMENT, "htj2k: null output buffer");
    out->data = nullptr;
    out->len = 0;
    // header-level checks the reference makes before decoding (container.rs:191-230)
    if (len < 8) fail(NDIC_CODESTREAM, "chunk header prefix too short", 0);
    if (std::memcmp(chunk, "ndht", 4) != 0) fail(NDIC_CODESTREAM, "not an htj2k chunk (bad magic)", 0);
    if (chunk[6] > 32) fail(NDIC_CODESTREAM, "xy_levels exceeds the 32-level J2K bound (T.800 SPcod)", 6);
    if (ndim < 2 || ndim > 32) fail(NDIC_CODESTREAM, "chunk header declares a different shape than decode expects", 0);
    uint64_t elems = 1;
    for (size_t i = 0; i < ndim; ++i) elems *= shape[i];
    if (elems == 0) fail(NDIC_CODESTREAM, "chunk header declares a different shape than decode expects", 0);
    if (!ht_dtype_ok(dtype)) fail(NDIC_UNSUPPORTED, std::string("htj2k does not support dtype ") + dtype_name(dtype));
    Context& ctx = Context::get();
    std::lock_guard<std::recursive_mutex> lock(ctx.mu);
    hipStream_t s = ctx.stream;
    uint8_t* d_in = (uint8_t*)ctx.in.ensure(len + 16);
    NDIC_HIP(hipMemcpyAsync(d_in, chunk, len, hipMemcpyHostToDevice, s));
    uint64_t out_bytes = elems * dtype_size(dtype);
    uint8_t* d_out = (uint8_t*)ctx.out.ensure(out_bytes + 16);
    std::vector<DecodeItem> items(1);
    DecodeItem& it = items[0];
    it.d_in = d_in;
    it.len = len;
    it.shape.assign(shape, shape + ndim);
    it.L = chunk[6];
    it.ht_dtype = dtype;
    it.lift = false;
    it.d_out = d_out;
    decode_items(ctx, items, s, window_elems());
    raise_item(it.status, it.message, it.offset);
    emit_host(d_out, out_bytes, out, s);
  });
}

ndic_status ndic_htj2k_decode_plane(const uint8_t* cs, size_t len, int32_t max_res, int32_t prefix,
                                    ndic_buffer* out, size_t* width, size_t* height, ndic_error* err) {
  return guard_call(err, [&] {
    if (!out || !width || !height) fail(NDIC_INVALID_ARGUMENT, "htj2k: null output argument");
    out->data = nullptr;
    out->len = 0;
    // host peek at SIZ / COD for the plane geometry the device tables need
    auto rd16 = [&](size_t p) -> uint32_t {
      if (p + 2 > len) fail(NDIC_CODESTREAM, "unexpected end of codestream", p);
      return (uint32_t(cs[p]) << 8) | cs[p + 1];
    };
    if (rd16(0) != 0xFF4F) fail(NDIC_CODESTREAM, "missing SOC", 0);
    size_t pos = 2;
    uint32_t W = 0, H = 0, L = 0;
    bool siz = false, cod = false;
    for (;;) {
      uint32_t m = rd16(pos);
      if (m == 0xFF90) break;
      uint32_t l = rd16(pos + 2);
      if (m < 0xFF00 || l < 2 || pos + 2 + l > len) fail(NDIC_CODESTREAM, "marker segment length out of bounds", pos);
      const uint8_t* q = cs + pos + 4;
      if (m == 0xFF51 && l >= 40) {
        W = (uint32_t(q[2]) << 24) | (uint32_t(q[3]) << 16) | (uint32_t(q[4]) << 8) | q[5];
        H = (uint32_t(q[6]) << 24) | (uint32_t(q[7]) << 16) | (uint32_t(q[8]) << 8) | q[9];
        uint32_t xo = (uint32_t(q[10]) << 24) | (uint32_t(q[11]) << 16) | (uint32_t(q[12]) << 8) | q[13];
        uint32_t yo = (uint32_t(q[14]) << 24) | (uint32_t(q[15]) << 16) | (uint32_t(q[16]) << 8) | q[17];
        if (xo || yo) fail(NDIC_UNSUPPORTED, "multi-tile or offset-canvas codestreams");
        siz = true;
      } else if (m == 0xFF52 && l >= 12) {
        L = q[5];
        cod = true;
      }
      pos += 2 + l;
    }
    if (!siz) fail(NDIC_CODESTREAM, "missing SIZ", pos);
    if (!cod) fail(NDIC_CODESTREAM, "missing COD", pos);
    if (W == 0 || H == 0 || L > 32) fail(NDIC_CODESTREAM, "SIZ/COD geometry out of range", 0);
    Context& ctx = Context::get();
    std::lock_guard<std::recursive_mutex> lock(ctx.mu);
    hipStream_t s = ctx.stream;
    uint8_t* d_in = (uint8_t*)ctx.in.ensure(len + 16);
    NDIC_HIP(hipMemcpyAsync(d_in, cs, len, hipMemcpyHostToDevice, s));
    std::vector<DecodeItem> items(1);
    DecodeItem& it = items[0];
    it.d_in = d_in;
    it.len = len;
    it.shape = {H, W};
    it.L = L;
    it.bare = true;
    it.prefix = prefix != 0;
    it.max_res = max_res < 0 ? -1 : max_res;
    decode_items(ctx, items, s, window_elems());
    raise_item(it.status, it.message, it.offset);
    uint64_t ow = it.out_w, oh = it.out_h;
    out->data = (uint8_t*)std::malloc(ow * oh * 4 + 4);
    out->len = ow * oh * 4;
    if (ow * oh)
      NDIC_HIP(hipMemcpy2DAsync(out->data, ow * 4, it.d_result, (size_t)W * 4, ow * 4, oh, hipMemcpyDeviceToHost, s));
    NDIC_HIP(hipStreamSynchronize(s));
    *width = ow;
    *height = oh;
  });
}

ndic_status ndic_lift_ht_decode_batch(const uint8_t* d_in, uint8_t* d_out, ndic_chunk_desc* descs, size_t n,
                                      uint32_t flags, void* stream, ndic_error* err) {
  return guard_call(err, [&] {
    Context& ctx = Context::get();
    std::lock_guard<std::recursive_mutex> lock(ctx.mu);
    hipStream_t s = stream ? (hipStream_t)stream : ctx.stream;
    std::vector<DecodeItem> items(n);
    for (size_t i = 0; i < n; ++i) {
      ndic_chunk_desc& d = descs[i];
      DecodeItem& it = items[i];
      if (d.ndim < 2 || d.ndim > NDIC_MAX_NDIM || d.nsteps > NDIC_MAX_STEPS || !ht_dtype_ok(d.dtype) ||
          d.xy_levels > 32)
        fail(NDIC_INVALID_ARGUMENT, "batch descriptor out of range");
      lift_validate("0.1", d.steps, d.nsteps, d.ndim);
      if (d.out_offset % dtype_size(d.dtype)) fail(NDIC_INVALID_ARGUMENT, "batch output offsets must be aligned to the dtype");
      it.d_in = d_in + d.in_offset;
      it.len = d.in_len;
      it.shape.assign(d.shape, d.shape + d.ndim);
      it.L = d.xy_levels;
      it.lift = true;
      it.out_dtype = d.dtype;
      it.steps.assign(d.steps, d.steps + d.nsteps);
      it.d_out = d_out + d.out_offset;
    }
    if (flags & NDIC_BATCH_PLAIN) {
      for (auto& it : items) {
        std::vector<DecodeItem> one{it};
        decode_items(ctx, one, s, window_elems());
        it.status = one[0].status;
        it.message = one[0].message;
        it.offset = one[0].offset;
      }
    } else {
      decode_items(ctx, items, s, window_elems());
    }
    int first = -1;
    for (size_t i = 0; i < n; ++i) {
      descs[i].status = items[i].status;
      if (items[i].status != NDIC_OK && first < 0) first = (int)i;
    }
    if (first >= 0) {
      std::ostringstream m;
      m << "chunk " << first << ": " << items[first].message;
      fail(items[first].status, m.str(), items[first].offset);
    }
  });
}

ndic_status ndic_lift_ht_encode_batch(const uint8_t* d_in, uint8_t* d_out, size_t capacity, ndic_chunk_desc* descs,
                                      size_t n, uint32_t flags, void* stream, ndic_error* err) {
  return guard_call(err, [&] {
    Context& ctx = Context::get();
    std::lock_guard<std::recursive_mutex> lock(ctx.mu);
    hipStream_t s = stream ? (hipStream_t)stream : ctx.stream;
    std::vector<EncodeItem> items(n);
    for (size_t i = 0; i < n; ++i) {
      ndic_chunk_desc& d = descs[i];
      EncodeItem& it = items[i];
      if (d.ndim < 2 || d.ndim > NDIC_MAX_NDIM || d.nsteps > NDIC_MAX_STEPS || !ht_dtype_ok(d.dtype) ||
          d.xy_levels > 32)
        fail(NDIC_INVALID_ARGUMENT, "batch descriptor out of range");
      lift_validate("0.1", d.steps, d.nsteps, d.ndim);
      if (d.in_offset % dtype_size(d.dtype)) fail(NDIC_INVALID_ARGUMENT, "batch input offsets must be aligned to the dtype");
      it.d_in = d_in + d.in_offset;
      it.shape.assign(d.shape, d.shape + d.ndim);
      uint64_t e = 1;
      for (size_t k = 0; k < it.shape.size(); ++k) e *= it.shape[k];
      if (e == 0) fail(NDIC_INVALID_ARGUMENT, "htj2k needs non-zero chunk extents");
      it.in_dtype = d.dtype;
      it.lift = true;
      it.steps.assign(d.steps, d.steps + d.nsteps);
      it.L = d.xy_levels;
      it.progression = 2;
      it.index = true;
    }
    if (flags & NDIC_BATCH_PLAIN) {
      uint64_t cursor = 0;
      for (auto& it : items) {
        std::vector<EncodeItem> one{it};
        encode_items(ctx, one, d_out + cursor, capacity - cursor, s, window_elems());
        it.status = one[0].status;
        it.message = one[0].message;
        it.out_off = cursor + one[0].out_off;
        it.out_len = one[0].out_len;
        cursor += one[0].out_len;
      }
    } else {
      encode_items(ctx, items, d_out, capacity, s, window_elems());
    }
    int first = -1;
    for (size_t i = 0; i < n; ++i) {
      descs[i].status = items[i].status;
      descs[i].out_offset = items[i].out_off;
      descs[i].out_len = items[i].out_len;
      if (items[i].status != NDIC_OK && first < 0) first = (int)i;
    }
    if (first >= 0) {
      std::ostringstream m;
      m << "chunk " << first << ": " << items[first].message;
      fail(items[first].status, m.str());
    }
  });
}

// ---- device-kernel test hooks (include/ndic_hip_dev.h) ---------------------

ndic_status ndic_dev_ht_encode_block(const uint32_t* samples, size_t width, size_t height, uint32_t missing_msbs,
                                     ndic_buffer* out, ndic_error* err) {
  return guard_call(err, [&] {
    if (width == 0 || height == 0 || width > 64 || height > 64)
      fail(NDIC_INVALID_ARGUMENT, "device code-blocks are 1..=64 on each side");
    if (missing_msbs >= 30) fail(NDIC_INVALID_ARGUMENT, "missing_msbs must be < 30 for the 32-bit datapath");
    ht_init_tables();
    Context& ctx = Context::get();
    std::lock_guard<std::recursive_mutex> lock(ctx.mu);
    hipStream_t s = ctx.stream;
    uint32_t kmax = missing_msbs + 1;
    uint32_t ms_cap = ht_ms_cap((uint32_t)width, (uint32_t)height, 31);
    size_t slot = ms_cap + kMelCap + kVlcCap;
    uint8_t* base = (uint8_t*)ctx.work.ensure(width * height * 4 + slot + 64 + sizeof(HtEncTask) + 64);
    uint32_t* d_samples = (uint32_t*)base;
    uint8_t* d_slot = base + ((width * height * 4 + 63) & ~size_t(63));
    uint32_t* d_lens = (uint32_t*)(d_slot + ((slot + 63) & ~size_t(63)));
    HtEncTask* d_task = (HtEncTask*)(d_lens + 16);
    NDIC_HIP(hipMemcpyAsync(d_samples, samples, width * height * 4, hipMemcpyHostToDevice, s));
    HtEncTask t{};
    t.in = (const int32_t*)d_samples;
    t.slot = d_slot;
    t.lens = d_lens;
    t.err = (int32_t*)(d_lens + 8);
    t.stride = (uint32_t)width;
    t.ms_cap = ms_cap;
    t.w = (uint16_t)width;
    t.h = (uint16_t)height;
    t.kmax = (uint8_t)kmax;
    t.flags = kTaskSignMag;
    NDIC_HIP(hipMemsetAsync(d_lens, 0, 64, s));
    NDIC_HIP(hipMemcpyAsync(d_task, &t, sizeof(t), hipMemcpyHostToDevice, s));
    if (tune(kTuneHtEncodeWaveMaxBlocks) >= 1) launch_ht_encode_wave(d_task, 1, ctx.queue(1), s);
    else launch_ht_encode(d_task, 1, s);
    uint32_t lens[16];
    std::vector<uint8_t> host(slot);
    NDIC_HIP(hipMemcpyAsync(lens, d_lens, 64, hipMemcpyDeviceToHost, s));
    NDIC_HIP(hipMemcpyAsync(host.data(), d_slot, slot, hipMemcpyDeviceToHost, s));
    NDIC_HIP(hipStreamSynchronize(s));
    if (lens[8]) fail(NDIC_INVALID_ARGUMENT, "cleanup segment suffix out of range (Scup)");
    uint32_t total = lens[3] ? lens[0] + lens[1] + lens[2] : 0;
    out->data = (uint8_t*)std::malloc(total + 1);
    out->len = total;
    for (uint32_t k = 0; k < total; ++k) out->data[k] = ht_seg_byte(host.data(), ms_cap, lens, k);
  });
}

ndic_status ndic_dev_ht_decode_block(const uint8_t* coded, size_t len, uint32_t* samples, size_t width,
                                     size_t height, uint32_t missing_msbs, uint32_t num_passes, size_t len_cleanup,
                                     size_t len_refinement, int stripe_causal, ndic_error* err) {
  return guard_call(err, [&] {
    if (width == 0 || height == 0 || width > 64 || height > 64)
      fail(NDIC_INVALID_ARGUMENT, "device code-blocks are 1..=64 on each side");
    if (len_cleanup + len_refinement > len) fail(NDIC_CODESTREAM, "invalid code-block pass lengths");
    ht_init_tables();
    Context& ctx = Context::get();
    std::lock_guard<std::recursive_mutex> lock(ctx.mu);
    hipStream_t s = ctx.stream;
    uint8_t* base = (uint8_t*)ctx.work.ensure(len + width * height * 4 + 512);
    uint8_t* d_coded = base;
    uint32_t* d_out = (uint32_t*)(base + ((len + 63) & ~size_t(63)));
    int32_t* d_err = (int32_t*)(d_out + ((width * height + 15) & ~size_t(15)));
    HtDecTask* d_task = (HtDecTask*)(d_err + 16);
    NDIC_HIP(hipMemcpyAsync(d_coded, coded, len, hipMemcpyHostToDevice, s));
    NDIC_HIP(hipMemsetAsync(d_err, 0, 64, s));
    HtDecTask t{};
    t.seg = d_coded;
    t.out = d_out;
    t.err = d_err;
    t.stride = (uint32_t)width;
    t.len1 = (uint32_t)len_cleanup;
    t.len2 = (uint32_t)len_refinement;
    t.w = (uint16_t)width;
    t.h = (uint16_t)height;
    t.mmsb = (uint8_t)std::min<uint32_t>(missing_msbs, 255);
    t.passes = (uint8_t)std::min<uint32_t>(num_passes, 255);
    t.kmax = 31;
    t.flags = kTaskIncluded | kTaskSignMag | (stripe_causal ? kTaskCausal : 0);
    NDIC_HIP(hipMemcpyAsync(d_task, &t, sizeof(t), hipMemcpyHostToDevice, s));
    if (tune(kTuneHtDecodeWaveMaxBlocks) >= 1) launch_ht_decode_wave(d_task, 1, ctx.queue(0), s);
    else if (tune(kTuneHtDecodeSplit)) launch_ht_decode_split(d_task, 1, (uint32_t*)ctx.htrec.ensure(kHtRecBytes), s);
    else launch_ht_decode(d_task, nullptr, 1, s);
    launch_ht_refine(d_task, 1, s);
    int32_t e = 0;
    NDIC_HIP(hipMemcpyAsync(&e, d_err, 4, hipMemcpyDeviceToHost, s));
    NDIC_HIP(hipMemcpyAsync(samples, d_out, width * height * 4, hipMemcpyDeviceToHost, s));
    NDIC_HIP(hipStreamSynchronize(s));
    if (e) fail(NDIC_CODESTREAM, "malformed HT code-block segment");
  });
}

}  // extern "C"

// Free and total device memory of the GPU the calling thread's context uses
// (hipMemGetInfo): what a serving loop admits work against.
extern "C" ndic_status ndic_device_memory(size_t* free_bytes, size_t* total_bytes, ndic_error* err) {
  return ndic::guard_call(err, [&] {
    using namespace ndic;
    if (!free_bytes || !total_bytes) fail(NDIC_INVALID_ARGUMENT, "null output pointer");
    Context::get();
    NDIC_HIP(hipMemGetInfo(free_bytes, total_bytes));
  });
}

namespace {

ndic_status plane_batch(const uint8_t* d_in, uint8_t* d_out, ndic_plane_desc* descs, size_t n, void* stream,
                        ndic_error* err, bool irreversible) {
  return ndic::guard_call(err, [&] {
    using namespace ndic;
    if (n && !descs) fail(NDIC_INVALID_ARGUMENT, "null descriptors");
    Context& ctx = Context::get();
    std::lock_guard<std::recursive_mutex> lock(ctx.mu);
    hipStream_t s = stream ? (hipStream_t)stream : ctx.stream;
    std::vector<DecodeItem> items(n);
    for (size_t i = 0; i < n; ++i) {
      ndic_plane_desc& d = descs[i];
      DecodeItem& it = items[i];
      if (d.width == 0 || d.height == 0 || d.levels > 32) fail(NDIC_INVALID_ARGUMENT, "plane descriptor out of range");
      if (d.out_offset % 4) fail(NDIC_INVALID_ARGUMENT, "plane batch output offsets must be 4-byte aligned");
      it.d_in = d_in + d.in_offset;
      it.len = d.in_len;
      it.shape = {d.height, d.width};
      it.L = d.levels;
      it.bare = true;
      it.irreversible = irreversible;
      it.prefix = d.prefix != 0;
      it.max_res = d.max_res < 0 ? -1 : d.max_res;
      it.d_out = d_out + d.out_offset;
    }
    decode_items(ctx, items, s, window_elems());
    NDIC_HIP(hipStreamSynchronize(s));
    int first = -1;
    for (size_t i = 0; i < n; ++i) {
      descs[i].status = items[i].status;
      descs[i].out_width = (uint32_t)items[i].out_w;
      descs[i].out_height = (uint32_t)items[i].out_h;
      if (items[i].status != NDIC_OK && first < 0) first = (int)i;
    }
    if (first >= 0) {
      std::ostringstream m;
      m << "plane " << first << ": " << items[first].message;
      fail(items[first].status, m.str(), items[first].offset);
    }
  });
}

}  // namespace

extern "C" ndic_status ndic_htj2k_decode_plane_batch(const uint8_t* d_in, uint8_t* d_out, ndic_plane_desc* descs,
                                                     size_t n, void* stream, ndic_error* err) {
  return plane_batch(d_in, d_out, descs, n, stream, err, false);
}

extern "C" ndic_status ndic_htj2k_decode_plane_batch_irreversible(const uint8_t* d_in, uint8_t* d_out,
                                                                  ndic_plane_desc* descs, size_t n, void* stream,
                                                                  ndic_error* err) {
  return plane_batch(d_in, d_out, descs, n, stream, err, true);
}
