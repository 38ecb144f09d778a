// Oracle (test infrastructure only): ZFP stream restatement.
#pragma once
#include <cstdint>
#include <limits>
#include <type_traits>
#include <vector>

#include "oracle.hpp"

namespace orc {
namespace zfp {

// zfp_type codes as stored in the 2-bit header field (value - 1)
enum ScalarType { kInt32 = 1, kInt64 = 2, kFloat = 3, kDouble = 4 };
constexpr uint32_t kMaxBits = 16657;  // ZFP_MAX_BITS of the C library 0.5.5
constexpr int kMinExp = -1074;         // ZFP_MIN_EXP

struct Mode {
  enum Kind { kRate = 0, kPrecision = 1, kAccuracy = 2, kReversible = 3 } kind = kReversible;
  double rate = 0;
  uint32_t precision = 0;
  double tolerance = 0;
};

// Row-major shape (slowest axis first); the stream's x axis is shape.back().
std::vector<uint8_t> compress(const void* data, int type, const std::vector<size_t>& shape, const Mode& mode,
                              uint64_t* header_bits);
void decompress(const uint8_t* bytes, size_t len, int type, const std::vector<size_t>& shape, const Mode& mode,
                void* out);
// Chunk level: narrow integers ride int32 (shifted into the high bits).
std::vector<uint8_t> encode_chunk(const uint8_t* bytes, size_t len, const std::vector<size_t>& shape, int dtype,
                                  const Mode& mode);
std::vector<uint8_t> decode_chunk(const uint8_t* bytes, size_t len, const std::vector<size_t>& shape, int dtype,
                                  const Mode& mode);
// Fixed-rate brick addressing and single-brick decodes (lib.rs:386-599,
// chunk.rs:446-540).  Coordinates and brick shapes are row-major.
struct BrickIndex {
  uint64_t header_bits = 0, bits_per_brick = 0;
  std::vector<size_t> grid;
  size_t stream_len = 0;
};
BrickIndex brick_index(const std::vector<size_t>& shape, int type, double rate);
uint64_t brick_linear(const BrickIndex& bi, const std::vector<size_t>& coords);
std::vector<uint8_t> decompress_brick(const uint8_t* bytes, size_t len, int type, const std::vector<size_t>& shape,
                                      double rate, const std::vector<size_t>& brick, std::vector<size_t>* brick_shape);
std::vector<uint8_t> decode_chunk_brick(const uint8_t* bytes, size_t len, const std::vector<size_t>& shape, int dtype,
                                        double rate, const std::vector<size_t>& brick,
                                        std::vector<size_t>* brick_shape);
// Absolute bit position of every block of a stream, then the end of the last.
std::vector<uint64_t> block_starts(const uint8_t* bytes, size_t len, int type, const std::vector<size_t>& shape,
                                   const Mode& mode);
// Order of the 4^d block coefficients (d = 1..4).
const uint8_t* perm_of(int dims);

}  // namespace zfp
}  // namespace orc
