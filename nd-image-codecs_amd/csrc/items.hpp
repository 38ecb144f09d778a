// Batch items of the htj2k / nd-lift-ht device pipeline.
#pragma once

#include <string>
#include <vector>

#include "common.hpp"

namespace ndic {

// One chunk of a device decode batch.
struct DecodeItem {
  const uint8_t* d_in = nullptr;
  uint64_t len = 0;
  std::vector<size_t> shape;
  uint32_t L = 0;
  int ht_dtype = NDIC_I32;  // the htj2k codec's dtype (int32 after nd_lift)
  bool lift = false;
  int out_dtype = NDIC_I32;
  std::vector<ndic_lift_step> steps;
  uint8_t* d_out = nullptr;
  int max_res = -1;
  bool bare = false;        // d_in is one plane codestream
  bool irreversible = false;  // ... which may be 9/7 (decoded, not refused)
  bool prefix = false;      // ... possibly cut short (a byte-range prefix)
  uint64_t out_w = 0, out_h = 0;
  int32_t* d_result = nullptr;  // bare: the decoded region (row pitch = plane width)
  // plane-selective decode: the container planes to decode (ascending;
  // empty = every plane), and, with `window`, the (x0, y0, w, h) window of
  // each decoded plane at the decoded resolution, written packed to d_out as
  // ht_dtype samples
  std::vector<uint32_t> plane_sel;
  bool window = false;
  uint32_t win[4] = {0, 0, 0, 0};
  // results
  int status = NDIC_OK;
  uint64_t offset = 0;
  std::string message;
};

struct EncodeItem {
  const uint8_t* d_in = nullptr;  // dtype samples
  std::vector<size_t> shape;
  int in_dtype = NDIC_U16;
  bool lift = false;
  std::vector<ndic_lift_step> steps;
  uint32_t L = 5;
  int progression = 2;
  bool index = true;
  // results
  uint64_t out_off = 0, out_len = 0;
  int status = NDIC_OK;
  std::string message;
};

}  // namespace ndic
