// Registry of the launch-parameter knobs (tune.hpp) and its C ABI.
#include "tune.hpp"

#include <atomic>
#include <cstring>

#include "../../include/ndic_hip.h"

namespace ndic {

namespace {

struct Knob {
  const char* name;
  int64_t def, lo, hi;
};

const Knob kKnobs[kTuneCount] = {
    {"ht_decode_wave_max_blocks", 8192, 0, INT64_MAX},
    {"ht_encode_wave_max_blocks", INT64_MAX, 0, INT64_MAX},
    {"parse_wave_max_planes", 4096, 0, INT64_MAX},
    {"ht_wave_grain", 1, 1, 64},
    {"fuse_dwt_lift", 4096, 0, INT64_MAX},
    {"pipeline_windows", 1, 0, 3},
    {"window_elems", 0, 0, (int64_t)1 << 40},
    {"ht_lane_waves", 0, 0, 4},
    {"graph_windows", 0, 0, 1},
    {"ht_decode_sort", 0, 0, 1},
    {"ht_decode_split", 1, 0, 1},
};

std::atomic<int64_t> g_vals[kTuneCount] = {};
std::atomic<bool> g_set[kTuneCount] = {};

int find(const char* key) {
  if (!key) return -1;
  for (int i = 0; i < kTuneCount; ++i)
    if (std::strcmp(kKnobs[i].name, key) == 0) return i;
  return -1;
}

}  // namespace

int64_t tune(TuneKey k) { return g_set[k].load() ? g_vals[k].load() : kKnobs[k].def; }

}  // namespace ndic

using namespace ndic;

extern "C" int32_t ndic_tune_set(const char* key, int64_t value) {
  const int i = find(key);
  if (i < 0) return -1;
  if (value < kKnobs[i].lo || value > kKnobs[i].hi) return -2;
  g_vals[i].store(value);
  g_set[i].store(true);
  return 0;
}

extern "C" int32_t ndic_tune_reset(const char* key) {
  const int i = find(key);
  if (i < 0) return -1;
  g_set[i].store(false);
  return 0;
}

extern "C" int64_t ndic_tune_get(const char* key) {
  const int i = find(key);
  return i < 0 ? -1 : tune((TuneKey)i);
}

extern "C" const char* ndic_tune_name(int32_t index) {
  return index >= 0 && index < kTuneCount ? kKnobs[index].name : nullptr;
}
