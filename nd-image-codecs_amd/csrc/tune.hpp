// Launch parameters chosen per shape class by the autotuner (and pinned by
// tests): named integer knobs with defaults, read by the orchestration code.
#pragma once

#include <cstdint>

namespace ndic {

enum TuneKey : int {
  // batch windows with at most this many code-blocks decode one block per
  // wavefront (ht_decode_wave_kernel); larger ones one block per lane
  kTuneHtDecodeWaveMaxBlocks = 0,
  // the same choice for the encoder (ht_encode_wave_kernel / ht_encode_kernel)
  kTuneHtEncodeWaveMaxBlocks,
  // windows with at most this many planes parse one plane per wavefront
  // (bytes and tag trees staged in LDS), larger ones one plane per thread
  kTuneParseWaveMaxPlanes,
  // code-blocks a wave of the persistent wave-per-block coders takes from
  // the work queue at a time
  kTuneHtWaveGrain,
  kTuneCount
};

int64_t tune(TuneKey k);

}  // namespace ndic
