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
  // the finest inverse DWT level, the inverse nd_lift step along a plane
  // axis and the narrowing store run as one pass (dwt_lift_kernel) in
  // windows with at least this many fused work items (64 x 32 tiles x lift
  // segments); 0: always as separate passes
  kTuneFuseDwtLift,
  // 1: the windows of a large decode batch alternate between two streams
  // (one window's parse / entropy decode overlaps the other's synthesis);
  // 0: one window after the other
  kTunePipelineWindows,
  // samples per batch window (bounds the intermediate planes' memory);
  // 0: NDIC_WINDOW_ELEMS from the environment, else 2^31
  kTuneWindowElems,
  // waves per workgroup of the lane-per-block HT decoder (they share one LDS
  // copy of the decode tables): 0 = 4 when the window has at least four
  // waves per CU, else 1; 1-3 = one wave, 4 = four
  kTuneHtLaneWaves,
  // 1: each decode window's launch sequence is captured into a HIP graph
  // and launched as one (instantiated graphs are cached per context and
  // updated in place); 0: launched one by one
  kTuneGraphWindows,
  // 1: the lane-per-block HT decoder takes its code-blocks heaviest first
  // (a device counting sort by cleanup-segment length) so that a wave's
  // blocks carry similar work; 0: in window order
  kTuneHtDecodeSort,
  // 1: windows above the wavefront crossover decode their code-blocks in two
  // passes (VLC / MEL one lane per block into per-quad records, then MagSgn
  // one wavefront per block); 0: the one-pass lane-per-block decoder
  kTuneHtDecodeSplit,
  kTuneCount
};

int64_t tune(TuneKey k);

}  // namespace ndic
