// Shared host/device definitions of the MI355X nd-lift-ht / nd-zfp library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "../../include/ndic_hip.h"

namespace ndic {

// A host-side failure carrying the C-ABI status (ndic_core::Error variant).
struct Failure {
  int status;
  uint64_t offset;
  std::string message;
};

[[noreturn]] inline void fail(int status, const std::string& message, uint64_t offset = 0) {
  throw Failure{status, offset, message};
}

#define NDIC_HIP(call)                                                                       \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      ::ndic::fail(NDIC_IO, std::string("HIP error in ") + #call + ": " + hipGetErrorString(e_)); \
  } while (0)

inline size_t dtype_size(int d) {
  switch (d) {
    case NDIC_U8: case NDIC_I8: return 1;
    case NDIC_U16: case NDIC_I16: return 2;
    case NDIC_U32: case NDIC_I32: case NDIC_F32: return 4;
    case NDIC_U64: case NDIC_I64: case NDIC_F64: return 8;
  }
  fail(NDIC_INVALID_ARGUMENT, "unknown dtype code " + std::to_string(d));
}
inline bool dtype_signed(int d) {
  return d == NDIC_I8 || d == NDIC_I16 || d == NDIC_I32 || d == NDIC_I64 || d == NDIC_F32 || d == NDIC_F64;
}
inline const char* dtype_name(int d) {
  static const char* n[] = {"u8", "i8", "u16", "i16", "u32", "i32", "u64", "i64", "f32", "f64"};
  return (d >= 0 && d <= 9) ? n[d] : "?";
}

inline uint64_t ceil_div(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

// XCD-aware work-group order.  The dispatcher deals work-groups to the 8 XCDs
// of a gfx950 device round-robin by linear id, and each XCD has its own L2:
// with items in launch order, neighbouring tiles (which share the cache lines
// at their edges) land on different XCDs and each fetches the shared lines.
// xcd_item deals items to the XCDs in runs of kXcdRun consecutive items
// instead (run r on XCD r % 8), so neighbours meet in one L2 while every XCD
// still gets a fine-grained mix of the launch's work; xcd_grid sizes a 2-D
// grid whose work-group count is a multiple of 8 runs (x a multiple of 8).
// The map is a bijection on the grid; ids at or past the item count are idle.
constexpr uint32_t kXcds = 8, kXcdRun = 16;
inline dim3 xcd_grid(uint64_t items) {
  const uint64_t n = ceil_div(items, kXcds * kXcdRun) * kXcds * kXcdRun;
  const uint32_t gx = (uint32_t)(n < 65408 ? n : 65408);  // 65408 = 511 * 128
  return dim3(gx, (uint32_t)ceil_div(n, gx));
}
__device__ __forceinline__ uint64_t xcd_item() {
  const uint64_t id = blockIdx.x + (uint64_t)blockIdx.y * gridDim.x;
  const uint64_t k = id / kXcds;  // slot on the XCD
  return ((k / kXcdRun) * kXcds + id % kXcds) * kXcdRun + k % kXcdRun;
}

// Device-side per-item error record: a code from DevErr plus a byte offset.
// The host maps codes to the reference's messages.
enum DevErr : int32_t {
  kDevOk = 0,
  kErrNarrow = 1,          // nd_lift decode: coefficients do not narrow back to the dtype
  kErrHtNarrow,            // htj2k decode: decoded samples do not fit the dtype
  kErrChunkMagic,          // container: bad magic
  kErrChunkVersion,
  kErrChunkFlags,
  kErrChunkLevels,
  kErrChunkNdim,
  kErrChunkTruncated,
  kErrChunkZeroDim,
  kErrChunkShape,          // header dims differ from the expected shape
  kErrChunkIndex,          // offsets overlap / prefixes inconsistent
  kErrPlaneRange,          // plane byte range exceeds the chunk
  kErrCsMarker,            // codestream: malformed marker structure
  kErrCsUnsupported,       // codestream outside the device decoder's scope
  kErrCsGeometry,          // plane geometry differs from the chunk header
  kErrCsPlt,               // PLT does not cover the body / wrong packet count
  kErrPacket,              // malformed packet header
  kErrBlock,               // malformed code-block segment
  kErrCsLevels,            // COD levels differ from the chunk header
  kErrEncodeScup,          // encoder: Scup out of range
  kErrEncodeKmax,          // encoder: K_max > 30
  kErrCount
};

}  // namespace ndic
