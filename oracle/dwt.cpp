// Oracle (test infrastructure only): reversible 5/3 DWT restated from the
// reference's scalar lane, crates/ndic-htj2k/src/dwt/mod.rs:33-199, and the
// irreversible 9/7 float lifting of dwt/mod.rs:103-325 (no K scaling: the
// reference leaves it to the per-subband quantisation step).  Built with
// -ffp-contract=off: every step rounds its sum, product and update to f32
// in the reference's order.
#include <algorithm>
#include <vector>

#include "oracle.hpp"

namespace orc {

// dwt/mod.rs:33-43
void level_dims(size_t w, size_t h, unsigned level, size_t* ow, size_t* oh) {
  for (unsigned i = 0; i < level; ++i) {
    w = (w + 1) / 2;
    h = (h + 1) / 2;
  }
  *ow = w;
  *oh = h;
}

namespace {

// dwt/mod.rs:51-74
void ana_53(const int32_t* x, size_t n, int32_t* low, int32_t* high) {
  if (n == 1) {
    low[0] = x[0];
    return;
  }
  size_t nl = (n + 1) / 2, nh = n / 2;
  for (size_t i = 0; i < nh; ++i) {
    int32_t right = (2 * i + 2 < n) ? x[2 * i + 2] : x[n - 2];
    high[i] = x[2 * i + 1] - ((x[2 * i] + right) >> 1);
  }
  for (size_t i = 0; i < nl; ++i) {
    int32_t dl = high[std::min(i == 0 ? 0 : i - 1, nh - 1)];
    int32_t dr = high[std::min(i, nh - 1)];
    low[i] = x[2 * i] + ((dl + dr + 2) >> 2);
  }
}

// dwt/mod.rs:77-101
void syn_53(const int32_t* low, const int32_t* high, size_t n, int32_t* x) {
  if (n == 1) {
    x[0] = low[0];
    return;
  }
  size_t nl = (n + 1) / 2, nh = n / 2;
  for (size_t i = 0; i < nl; ++i) {
    int32_t dl = high[std::min(i == 0 ? 0 : i - 1, nh - 1)];
    int32_t dr = high[std::min(i, nh - 1)];
    x[2 * i] = low[i] - ((dl + dr + 2) >> 2);
  }
  for (size_t i = 0; i < nh; ++i) {
    int32_t right = (2 * i + 2 < n) ? x[2 * i + 2] : x[n - 2];
    x[2 * i + 1] = high[i] + ((x[2 * i] + right) >> 1);
  }
}

}  // namespace

// dwt/mod.rs:133-160 (vertical pass first, then horizontal)
void dwt53_forward(int32_t* plane, size_t width, size_t height, unsigned levels) {
  std::vector<int32_t> sig(std::max(width, height)), lo(sig.size()), hi(sig.size());
  for (unsigned l = 0; l < levels; ++l) {
    size_t w, h;
    level_dims(width, height, l, &w, &h);
    if (w == 1 && h == 1) break;
    if (h > 1) {
      size_t nl = (h + 1) / 2;
      for (size_t x = 0; x < w; ++x) {
        for (size_t y = 0; y < h; ++y) sig[y] = plane[y * width + x];
        ana_53(sig.data(), h, lo.data(), hi.data());
        for (size_t y = 0; y < nl; ++y) plane[y * width + x] = lo[y];
        for (size_t y = nl; y < h; ++y) plane[y * width + x] = hi[y - nl];
      }
    }
    if (w > 1) {
      size_t nl = (w + 1) / 2;
      for (size_t y = 0; y < h; ++y) {
        int32_t* row = plane + y * width;
        std::copy(row, row + w, sig.begin());
        ana_53(sig.data(), w, lo.data(), hi.data());
        std::copy(lo.begin(), lo.begin() + nl, row);
        std::copy(hi.begin(), hi.begin() + (w - nl), row + nl);
      }
    }
  }
}

// dwt/mod.rs:167-198 (horizontal synthesis first, then vertical)
void dwt53_inverse(int32_t* plane, size_t width, size_t height, unsigned levels) {
  std::vector<int32_t> sig(std::max(width, height)), lo(sig.size()), hi(sig.size());
  for (unsigned l = levels; l-- > 0;) {
    size_t w, h;
    level_dims(width, height, l, &w, &h);
    if (w == 1 && h == 1) continue;
    if (w > 1) {
      size_t nl = (w + 1) / 2;
      for (size_t y = 0; y < h; ++y) {
        int32_t* row = plane + y * width;
        std::copy(row, row + nl, lo.begin());
        std::copy(row + nl, row + w, hi.begin());
        syn_53(lo.data(), hi.data(), w, sig.data());
        std::copy(sig.begin(), sig.begin() + w, row);
      }
    }
    if (h > 1) {
      size_t nl = (h + 1) / 2;
      for (size_t x = 0; x < w; ++x) {
        for (size_t y = 0; y < nl; ++y) lo[y] = plane[y * width + x];
        for (size_t y = nl; y < h; ++y) hi[y - nl] = plane[y * width + x];
        syn_53(lo.data(), hi.data(), h, sig.data());
        for (size_t y = 0; y < h; ++y) plane[y * width + x] = sig[y];
      }
    }
  }
}

// ---- irreversible 9/7 (dwt/mod.rs:103-325) ---------------------------------
namespace {

// T.800 Table F.4 lifting multipliers, analysis order, as the f32 values the
// reference's literals round to (dwt/mod.rs:104-107; its delta, 0.44350685,
// is one f32 step below T.800's 0.443506852)
constexpr float kAlpha = -1.5861343f;
constexpr float kBeta = -0.052980118f;
constexpr float kGamma = 0.8829111f;
constexpr float kDelta = 0.44350685f;

// dwt/mod.rs:110-131: one lifting step over the de-interleaved halves with
// the clamped (whole-sample symmetric) neighbours
void lift_97(float* low, float* high, size_t n, float c, bool high_from_low) {
  size_t nl = (n + 1) / 2, nh = n / 2;
  if (high_from_low) {
    for (size_t i = 0; i < nh; ++i) {
      float l = low[i], r = low[std::min(i + 1, nl - 1)];
      float t = l + r;
      float p = c * t;
      high[i] = high[i] + p;
    }
  } else {
    for (size_t i = 0; i < nl; ++i) {
      float l = high[std::min(i == 0 ? 0 : i - 1, nh - 1)], r = high[std::min(i, nh - 1)];
      float t = l + r;
      float p = c * t;
      low[i] = low[i] + p;
    }
  }
}

// dwt/mod.rs:286-304
void ana_97(const float* x, size_t n, float* low, float* high) {
  if (n == 1) {
    low[0] = x[0];
    return;
  }
  size_t nl = (n + 1) / 2, nh = n / 2;
  for (size_t i = 0; i < nl; ++i) low[i] = x[2 * i];
  for (size_t i = 0; i < nh; ++i) high[i] = x[2 * i + 1];
  lift_97(low, high, n, kAlpha, true);
  lift_97(low, high, n, kBeta, false);
  lift_97(low, high, n, kGamma, true);
  lift_97(low, high, n, kDelta, false);
}

// dwt/mod.rs:307-325
void syn_97(float* low, float* high, size_t n, float* x) {
  if (n == 1) {
    x[0] = low[0];
    return;
  }
  size_t nl = (n + 1) / 2, nh = n / 2;
  lift_97(low, high, n, -kDelta, false);
  lift_97(low, high, n, -kGamma, true);
  lift_97(low, high, n, -kBeta, false);
  lift_97(low, high, n, -kAlpha, true);
  for (size_t i = 0; i < nl; ++i) x[2 * i] = low[i];
  for (size_t i = 0; i < nh; ++i) x[2 * i + 1] = high[i];
}

}  // namespace

// dwt/mod.rs:205-240 (vertical pass first, then horizontal)
void dwt97_forward(float* plane, size_t width, size_t height, unsigned levels) {
  std::vector<float> sig(std::max(width, height)), lo(sig.size()), hi(sig.size());
  for (unsigned l = 0; l < levels; ++l) {
    size_t w, h;
    level_dims(width, height, l, &w, &h);
    if (w == 1 && h == 1) break;
    if (h > 1) {
      size_t nl = (h + 1) / 2;
      for (size_t x = 0; x < w; ++x) {
        for (size_t y = 0; y < h; ++y) sig[y] = plane[y * width + x];
        ana_97(sig.data(), h, lo.data(), hi.data());
        for (size_t y = 0; y < nl; ++y) plane[y * width + x] = lo[y];
        for (size_t y = nl; y < h; ++y) plane[y * width + x] = hi[y - nl];
      }
    }
    if (w > 1) {
      size_t nl = (w + 1) / 2;
      for (size_t y = 0; y < h; ++y) {
        float* row = plane + y * width;
        std::copy(row, row + w, sig.begin());
        ana_97(sig.data(), w, lo.data(), hi.data());
        std::copy(lo.begin(), lo.begin() + nl, row);
        std::copy(hi.begin(), hi.begin() + (w - nl), row + nl);
      }
    }
  }
}

// dwt/mod.rs:247-283 (horizontal synthesis first, then vertical)
void dwt97_inverse(float* plane, size_t width, size_t height, unsigned levels) {
  std::vector<float> sig(std::max(width, height)), lo(sig.size()), hi(sig.size());
  for (unsigned l = levels; l-- > 0;) {
    size_t w, h;
    level_dims(width, height, l, &w, &h);
    if (w == 1 && h == 1) continue;
    if (w > 1) {
      size_t nl = (w + 1) / 2;
      for (size_t y = 0; y < h; ++y) {
        float* row = plane + y * width;
        std::copy(row, row + nl, lo.begin());
        std::copy(row + nl, row + w, hi.begin());
        syn_97(lo.data(), hi.data(), w, sig.data());
        std::copy(sig.begin(), sig.begin() + w, row);
      }
    }
    if (h > 1) {
      size_t nl = (h + 1) / 2;
      for (size_t x = 0; x < w; ++x) {
        for (size_t y = 0; y < nl; ++y) lo[y] = plane[y * width + x];
        for (size_t y = nl; y < h; ++y) hi[y - nl] = plane[y * width + x];
        syn_97(lo.data(), hi.data(), h, sig.data());
        for (size_t y = 0; y < h; ++y) plane[y * width + x] = sig[y];
      }
    }
  }
}

}  // namespace orc
