// CPU oracle for the nd-lift-ht / nd-zfp codec path: TEST INFRASTRUCTURE ONLY.
//
// A line-by-line C++ restatement of the reference's Rust CPU path, used by
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
// checker.  The product library (nd-image-codecs_amd/csrc) never includes,
// links or calls anything here.  Every function cites the reference file:line
// it restates.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace orc {

// ndic_core::Error variants (crates/ndic-core/src/error.rs:14-43).
enum ErrKind { kOk = 0, kInvalidArgument = 1, kCodestream = 2, kUnsupported = 3, kIo = 4 };

struct Error {
  int kind;
  size_t offset;
  std::string message;
};

[[noreturn]] void fail(int kind, const std::string& message, size_t offset = 0);

// Sample types shared by nd_lift (LiftDtype, crates/ndic-zarr/src/lift.rs:18-35)
// and htj2k (SampleType, crates/ndic-core/src/dtype.rs:12).
enum Dtype { kU8 = 0, kI8 = 1, kU16 = 2, kI16 = 3, kU32 = 4, kI32 = 5, kU64 = 6, kI64 = 7, kF32 = 8, kF64 = 9 };
size_t dtype_size(int dtype);
bool dtype_signed(int dtype);
const char* dtype_name(int dtype);

// ---- nd_lift (crates/ndic-lift) ---------------------------------------------
enum LiftKind { kDelta = 0, kHaar = 1, kLift53 = 2 };
struct AxisStep {
  std::string axis;
  size_t dimension;
  int kind;
  unsigned levels;  // u8 in the reference
  uint32_t group;
};

uint32_t lift_band_count(size_t n, unsigned levels);
void check_lift_version(const std::string& version);
std::vector<uint8_t> lift_forward_chunk(const uint8_t* bytes, size_t len, const std::vector<size_t>& shape,
                                        int dtype, const std::string& version,
                                        const std::vector<AxisStep>& steps);
std::vector<uint8_t> lift_inverse_chunk(const uint8_t* bytes, size_t len, const std::vector<size_t>& shape,
                                        int dtype, const std::string& version,
                                        const std::vector<AxisStep>& steps);
// In-place plane transforms (crates/ndic-lift/src/chunk.rs:30-80).
void lift_forward_i32(std::vector<int32_t>& chunk, const std::vector<size_t>& shape, const std::vector<AxisStep>& steps);
void lift_inverse_i32(std::vector<int32_t>& chunk, const std::vector<size_t>& shape, const std::vector<AxisStep>& steps);
void lift_forward_i64(std::vector<int64_t>& chunk, const std::vector<size_t>& shape, const std::vector<AxisStep>& steps);
void lift_inverse_i64(std::vector<int64_t>& chunk, const std::vector<size_t>& shape, const std::vector<AxisStep>& steps);

// ---- 5/3 DWT (crates/ndic-htj2k/src/dwt/mod.rs) -----------------------------
void level_dims(size_t w, size_t h, unsigned level, size_t* ow, size_t* oh);
void dwt53_forward(int32_t* plane, size_t width, size_t height, unsigned levels);
void dwt53_inverse(int32_t* plane, size_t width, size_t height, unsigned levels);
// 9/7 float lifting without K scaling (dwt/mod.rs:103-325)
void dwt97_forward(float* plane, size_t width, size_t height, unsigned levels);
void dwt97_inverse(float* plane, size_t width, size_t height, unsigned levels);

// ---- HT block coder (crates/ndic-htj2k/src/block) ---------------------------
struct HtTables {
  uint16_t dec0[1024], dec1[1024];
  uint16_t enc0[2048], enc1[2048];
  uint16_t uvlc0[320], uvlc1[256];
};
const HtTables& ht_tables();
std::vector<uint8_t> ht_encode_block(const uint32_t* buf, size_t width, size_t height, size_t stride,
                                     uint32_t missing_msbs);
void ht_decode_block(const uint8_t* coded, size_t coded_len, uint32_t* out, size_t width, size_t height,
                     size_t stride, uint32_t missing_msbs, uint32_t num_passes, size_t len_cleanup,
                     size_t len_refinement, bool stripe_causal);

// ---- codestream / container / chunk cores ----------------------------------
struct Htj2kConfig {
  unsigned xy_levels = 5;
  bool reversible = true;
  int progression = 2;  // 0 LRCP, 1 RLCP, 2 RPCL, 3 PCRL, 4 CPRL
  bool index = true;
};
std::vector<uint8_t> encode_image_with_depth(const int32_t* samples, size_t width, size_t height,
                                             unsigned bit_depth, bool is_signed, unsigned xy_levels,
                                             int progression, bool emit_tlm_plt);
struct Decoded {
  size_t width = 0, height = 0;
  std::vector<int32_t> samples;
};
// allow_irreversible: decode 9/7 codestreams (QCD style 2) instead of
// refusing them as the reference does (reader.rs:326-330)
Decoded decode_codestream(const uint8_t* data, size_t len, int max_res /* -1 = full */, bool prefix,
                          bool allow_irreversible = false);
std::vector<uint8_t> encode_image_irreversible(const int32_t* samples, size_t width, size_t height,
                                               unsigned bit_depth, bool is_signed, unsigned levels, unsigned eps,
                                               unsigned mu, unsigned guard);
std::vector<uint8_t> htj2k_encode_chunk(const uint8_t* bytes, size_t len, const std::vector<size_t>& shape,
                                        int dtype, const Htj2kConfig& cfg);
std::vector<uint8_t> htj2k_decode_chunk(const uint8_t* chunk, size_t len, const std::vector<size_t>& shape,
                                        int dtype);

}  // namespace orc
