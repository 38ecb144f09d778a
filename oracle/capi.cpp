// Oracle (test infrastructure only): a flat C ABI over the restatement so the
// Python tests can call it through ctypes.  Output buffers are malloc'ed and
// released with orc_free.
#include <cstdlib>
#include <cstring>

#include "oracle.hpp"
#include "zfp_oracle.hpp"

extern "C" {

struct orc_err {
  int kind;
  size_t offset;
  char message[512];
};

// One nd_lift step as four ints: dimension, kind (0 delta, 1 haar, 2 lift53),
// levels, group.
}

namespace {

template <class F> int guard(orc_err* err, F&& f) {
  try {
    f();
    if (err) { err->kind = 0; err->offset = 0; err->message[0] = 0; }
    return 0;
  } catch (orc::Error& e) {
    if (err) {
      err->kind = e.kind;
      err->offset = e.offset;
      std::strncpy(err->message, e.message.c_str(), sizeof(err->message) - 1);
      err->message[sizeof(err->message) - 1] = 0;
    }
    return e.kind;
  }
}

void emit(const std::vector<uint8_t>& v, uint8_t** out, size_t* out_len) {
  *out = (uint8_t*)std::malloc(v.size() ? v.size() : 1);
  if (!v.empty()) std::memcpy(*out, v.data(), v.size());
  *out_len = v.size();
}

std::vector<orc::AxisStep> steps_of(const int* s, size_t n) {
  std::vector<orc::AxisStep> v;
  for (size_t i = 0; i < n; ++i)
    v.push_back({"", (size_t)s[4 * i], s[4 * i + 1], (unsigned)s[4 * i + 2], (uint32_t)s[4 * i + 3]});
  return v;
}

}  // namespace

extern "C" {

void orc_free(void* p) { std::free(p); }

int orc_lift_chunk(int forward, const uint8_t* in, size_t len, const size_t* shape, size_t ndim, int dtype,
                   const char* version, const int* steps, size_t nsteps, uint8_t** out, size_t* out_len, orc_err* err) {
  return guard(err, [&] {
    std::vector<size_t> sh(shape, shape + ndim);
    auto st = steps_of(steps, nsteps);
    emit(forward ? orc::lift_forward_chunk(in, len, sh, dtype, version, st)
                 : orc::lift_inverse_chunk(in, len, sh, dtype, version, st),
         out, out_len);
  });
}

int orc_lift_plane_i64(int forward, int64_t* data, const size_t* shape, size_t ndim, const int* steps, size_t nsteps,
                       orc_err* err) {
  return guard(err, [&] {
    size_t n = 1;
    for (size_t i = 0; i < ndim; ++i) n *= shape[i];
    std::vector<int64_t> v(data, data + n);
    std::vector<size_t> sh(shape, shape + ndim);
    auto st = steps_of(steps, nsteps);
    if (forward) orc::lift_forward_i64(v, sh, st);
    else orc::lift_inverse_i64(v, sh, st);
    std::memcpy(data, v.data(), n * 8);
  });
}

int orc_dwt53(int forward, int32_t* plane, size_t width, size_t height, unsigned levels, orc_err* err) {
  return guard(err, [&] {
    if (forward) orc::dwt53_forward(plane, width, height, levels);
    else orc::dwt53_inverse(plane, width, height, levels);
  });
}

void orc_ht_tables(uint16_t* dec0, uint16_t* dec1, uint16_t* enc0, uint16_t* enc1, uint16_t* uvlc0, uint16_t* uvlc1) {
  const orc::HtTables& t = orc::ht_tables();
  std::memcpy(dec0, t.dec0, sizeof(t.dec0));
  std::memcpy(dec1, t.dec1, sizeof(t.dec1));
  std::memcpy(enc0, t.enc0, sizeof(t.enc0));
  std::memcpy(enc1, t.enc1, sizeof(t.enc1));
  std::memcpy(uvlc0, t.uvlc0, sizeof(t.uvlc0));
  std::memcpy(uvlc1, t.uvlc1, sizeof(t.uvlc1));
}

int orc_ht_encode_block(const uint32_t* buf, size_t width, size_t height, uint32_t missing_msbs, uint8_t** out,
                        size_t* out_len, orc_err* err) {
  return guard(err, [&] { emit(orc::ht_encode_block(buf, width, height, width, missing_msbs), out, out_len); });
}

int orc_ht_decode_block(const uint8_t* coded, size_t coded_len, uint32_t* out, size_t width, size_t height,
                        uint32_t missing_msbs, uint32_t num_passes, size_t len1, size_t len2, int stripe_causal,
                        orc_err* err) {
  return guard(err, [&] {
    orc::ht_decode_block(coded, coded_len, out, width, height, width, missing_msbs, num_passes, len1, len2,
                         stripe_causal != 0);
  });
}

int orc_encode_image(const int32_t* samples, size_t width, size_t height, unsigned depth, int is_signed,
                     unsigned levels, uint8_t** out, size_t* out_len, orc_err* err) {
  return guard(err, [&] {
    emit(orc::encode_image_with_depth(samples, width, height, depth, is_signed != 0, levels, 2, true), out, out_len);
  });
}

int orc_dwt97(int forward, float* plane, size_t width, size_t height, unsigned levels, orc_err* err) {
  return guard(err, [&] {
    if (forward) orc::dwt97_forward(plane, width, height, levels);
    else orc::dwt97_inverse(plane, width, height, levels);
  });
}

int orc_encode_image_irreversible(const int32_t* samples, size_t width, size_t height, unsigned depth, int is_signed,
                                  unsigned levels, unsigned eps, unsigned mu, unsigned guard_bits, uint8_t** out,
                                  size_t* out_len, orc_err* err) {
  return guard(err, [&] {
    emit(orc::encode_image_irreversible(samples, width, height, depth, is_signed != 0, levels, eps, mu, guard_bits),
         out, out_len);
  });
}

int orc_decode_codestream(const uint8_t* data, size_t len, int max_res, int prefix, int irreversible, int32_t** out,
                          size_t* width, size_t* height, orc_err* err) {
  return guard(err, [&] {
    orc::Decoded d = orc::decode_codestream(data, len, max_res, prefix != 0, irreversible != 0);
    *out = (int32_t*)std::malloc(d.samples.size() * 4 + 4);
    std::memcpy(*out, d.samples.data(), d.samples.size() * 4);
    *width = d.width;
    *height = d.height;
  });
}

int orc_htj2k_encode_chunk(const uint8_t* in, size_t len, const size_t* shape, size_t ndim, int dtype,
                           unsigned xy_levels, int reversible, int progression, int index, uint8_t** out,
                           size_t* out_len, orc_err* err) {
  return guard(err, [&] {
    orc::Htj2kConfig c;
    c.xy_levels = xy_levels;
    c.reversible = reversible != 0;
    c.progression = progression;
    c.index = index != 0;
    emit(orc::htj2k_encode_chunk(in, len, std::vector<size_t>(shape, shape + ndim), dtype, c), out, out_len);
  });
}

int orc_htj2k_decode_chunk(const uint8_t* in, size_t len, const size_t* shape, size_t ndim, int dtype, uint8_t** out,
                           size_t* out_len, orc_err* err) {
  return guard(err, [&] {
    emit(orc::htj2k_decode_chunk(in, len, std::vector<size_t>(shape, shape + ndim), dtype), out, out_len);
  });
}

// ZFP: mode = {kind (0 rate, 1 precision, 2 accuracy, 3 reversible), value}.
// type: 1 int32, 2 int64, 3 float, 4 double (raw field) or, with chunk = 1,
// a dtype code (narrow integers promoted).
static orc::zfp::Mode zfp_mode_of(int kind, double value) {
  orc::zfp::Mode m;
  m.kind = (orc::zfp::Mode::Kind)kind;
  m.rate = value;
  m.tolerance = value;
  m.precision = (uint32_t)value;
  return m;
}

int orc_zfp_encode(int chunk, const uint8_t* in, size_t len, const size_t* shape, size_t ndim, int type, int kind,
                   double value, uint8_t** out, size_t* out_len, uint64_t* header_bits, orc_err* err) {
  return guard(err, [&] {
    std::vector<size_t> sh(shape, shape + ndim);
    orc::zfp::Mode m = zfp_mode_of(kind, value);
    if (chunk) {
      emit(orc::zfp::encode_chunk(in, len, sh, type, m), out, out_len);
      if (header_bits) *header_bits = 0;
    } else {
      emit(orc::zfp::compress(in, type, sh, m, header_bits), out, out_len);
    }
  });
}

int orc_zfp_decode(int chunk, const uint8_t* in, size_t len, const size_t* shape, size_t ndim, int type, int kind,
                   double value, uint8_t** out, size_t* out_len, orc_err* err) {
  return guard(err, [&] {
    std::vector<size_t> sh(shape, shape + ndim);
    orc::zfp::Mode m = zfp_mode_of(kind, value);
    if (chunk) {
      emit(orc::zfp::decode_chunk(in, len, sh, type, m), out, out_len);
    } else {
      size_t n = 1;
      for (size_t d : sh) n *= d;
      std::vector<uint8_t> v(n * (type == 1 || type == 3 ? 4 : 8));
      orc::zfp::decompress(in, len, type, sh, m, v.data());
      emit(v, out, out_len);
    }
  });
}

// BrickIndex (lib.rs:398-500): header bits, bits per brick, stream length
// and the grid; byte_range of brick k = [bit / 8, ceil((bit + len) / 8)).
int orc_zfp_brick_index(const size_t* shape, size_t ndim, int type, double rate, uint64_t* header_bits,
                        uint64_t* bits_per_brick, uint64_t* stream_len, uint64_t* grid, orc_err* err) {
  return guard(err, [&] {
    orc::zfp::BrickIndex bi = orc::zfp::brick_index(std::vector<size_t>(shape, shape + ndim), type, rate);
    *header_bits = bi.header_bits;
    *bits_per_brick = bi.bits_per_brick;
    *stream_len = bi.stream_len;
    for (size_t a = 0; a < bi.grid.size(); ++a) grid[a] = bi.grid[a];
  });
}

// chunk = 1: decode_chunk_brick (type is a dtype code); 0: decompress_brick (type is a zfp type)
int orc_zfp_decode_brick(int chunk, const uint8_t* in, size_t len, const size_t* shape, size_t ndim, int type,
                         double rate, const size_t* brick, uint8_t** out, size_t* out_len, size_t* brick_shape,
                         orc_err* err) {
  return guard(err, [&] {
    std::vector<size_t> sh(shape, shape + ndim), br(brick, brick + ndim), bs;
    std::vector<uint8_t> v = chunk ? orc::zfp::decode_chunk_brick(in, len, sh, type, rate, br, &bs)
                                   : orc::zfp::decompress_brick(in, len, type, sh, rate, br, &bs);
    for (size_t a = 0; a < bs.size(); ++a) brick_shape[a] = bs[a];
    emit(v, out, out_len);
  });
}

int orc_zfp_block_starts(const uint8_t* in, size_t len, const size_t* shape, size_t ndim, int type, int kind,
                         double value, uint8_t** out, size_t* out_len, orc_err* err) {
  return guard(err, [&] {
    std::vector<uint64_t> v =
        orc::zfp::block_starts(in, len, type, std::vector<size_t>(shape, shape + ndim), zfp_mode_of(kind, value));
    std::vector<uint8_t> b(v.size() * 8);
    std::memcpy(b.data(), v.data(), b.size());
    emit(b, out, out_len);
  });
}

void orc_zfp_perm(int dims, uint8_t* out) {
  const uint8_t* p = orc::zfp::perm_of(dims);
  std::memcpy(out, p, (size_t)1 << (2 * dims));
}

}  // extern "C"
