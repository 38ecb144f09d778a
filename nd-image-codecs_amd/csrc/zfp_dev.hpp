// Device helpers shared by the two ZFP block-coder layouts (zfp.hip: one
// wavefront per block; zfp_lane.hip: one lane per block): scalar traits,
// the 4-point lifting steps, IEEE-754 helpers, element loads and stores with
// the chunk layer's dtype promotions, block geometry.  The algorithm is the
// published ZFP one (Lindstrom, IEEE TVCG 2014; LLNL ZFP stream format).
#pragma once

#include <type_traits>

#include "zfp.hpp"

namespace ndic {
namespace zfpdev {

template <class S> struct ZT;
template <> struct ZT<int32_t> {
  using Int = int32_t;
  using UInt = uint32_t;
  static constexpr int ebits = 0, ebias = 0;
  static constexpr bool fp = false;
};
template <> struct ZT<int64_t> {
  using Int = int64_t;
  using UInt = uint64_t;
  static constexpr int ebits = 0, ebias = 0;
  static constexpr bool fp = false;
};
template <> struct ZT<float> {
  using Int = int32_t;
  using UInt = uint32_t;
  static constexpr int ebits = 8, ebias = 127, mbits = 23;
  static constexpr bool fp = true;
};
template <> struct ZT<double> {
  using Int = int64_t;
  using UInt = uint64_t;
  static constexpr int ebits = 11, ebias = 1023, mbits = 52;
  static constexpr bool fp = true;
};

template <class Int> __device__ __forceinline__ Int wadd(Int a, Int b) {
  using U = typename std::make_unsigned<Int>::type;
  return (Int)((U)a + (U)b);
}
template <class Int> __device__ __forceinline__ Int wsub(Int a, Int b) {
  using U = typename std::make_unsigned<Int>::type;
  return (Int)((U)a - (U)b);
}
template <class Int> __device__ __forceinline__ Int wshl1(Int a) {
  using U = typename std::make_unsigned<Int>::type;
  return (Int)((U)a << 1);
}

template <class UInt> __device__ __forceinline__ UInt nbmask() { return (UInt)0xaaaaaaaaaaaaaaaaull; }

// ---------------------------------------------------------------------------
// 4-point lifting steps (forward / inverse, lossy / reversible)
// ---------------------------------------------------------------------------
template <class Int, bool Fwd, bool Rev> __device__ __forceinline__ void lift4(Int& x, Int& y, Int& z, Int& w) {
  if (Fwd && !Rev) {
    x = wadd(x, w); x >>= 1; w = wsub(w, x);
    z = wadd(z, y); z >>= 1; y = wsub(y, z);
    x = wadd(x, z); x >>= 1; z = wsub(z, x);
    w = wadd(w, y); w >>= 1; y = wsub(y, w);
    w = wadd(w, (Int)(y >> 1)); y = wsub(y, (Int)(w >> 1));
  } else if (!Fwd && !Rev) {
    y = wadd(y, (Int)(w >> 1)); w = wsub(w, (Int)(y >> 1));
    y = wadd(y, w); w = wshl1(w); w = wsub(w, y);
    z = wadd(z, x); x = wshl1(x); x = wsub(x, z);
    y = wadd(y, z); z = wshl1(z); z = wsub(z, y);
    w = wadd(w, x); x = wshl1(x); x = wsub(x, w);
  } else if (Fwd) {
    w = wsub(w, z); z = wsub(z, y); y = wsub(y, x);
    w = wsub(w, z); z = wsub(z, y);
    w = wsub(w, z);
  } else {
    w = wadd(w, z); z = wadd(z, y); y = wadd(y, x);
    w = wadd(w, z); z = wadd(z, y);
    w = wadd(w, z);
  }
}


// Stream words straight from global memory (the parallel block decoders).
// (global-space loads: a flat load would also hold up every later wait for
// scalar loads)
struct GlobalSrc {
  const uint8_t* base;
  uint64_t nbytes;
  __device__ uint64_t load(uint64_t w) {
    typedef const __attribute__((address_space(1))) uint64_t* g64;
    typedef const __attribute__((address_space(1))) uint8_t* g8;
    uint64_t b = w * 8;
    if (b + 8 <= nbytes) return *(g64)(uintptr_t)(base + b);
    uint64_t v = 0;
    for (uint64_t i = b; i < nbytes && i < b + 8; ++i) v |= (uint64_t)*(g8)(uintptr_t)(base + i) << (8 * (i - b));
    return v;
  }
};

// The stream seen through a window of 64 words, one per lane, read with
// readlane; the next 64 words are loaded as the window moves, so the serial
// block walk of the variable-size modes waits on memory once per 512 bytes

template <class S> __device__ __forceinline__ typename ZT<S>::UInt fbits(S x) {
  typename ZT<S>::UInt b;
  __builtin_memcpy(&b, &x, sizeof(b));
  return b;
}
template <class S> __device__ __forceinline__ S fromb(typename ZT<S>::UInt b) {
  S x;
  __builtin_memcpy(&x, &b, sizeof(b));
  return x;
}
// 2^k: +inf above the normal range, subnormal or 0 below it
template <class S> __device__ __forceinline__ S pow2(int k) {
  using U = typename ZT<S>::UInt;
  constexpr int bias = ZT<S>::ebias, mb = ZT<S>::mbits;
  if (k > bias) return fromb<S>((U)(2 * bias + 1) << mb);
  if (k >= 1 - bias) return fromb<S>((U)(k + bias) << mb);
  if (k >= 1 - bias - mb) return fromb<S>((U)1 << (k - (1 - bias - mb)));
  return (S)0;
}

template <class S> __device__ __forceinline__ typename ZT<S>::Int cast_int(S y) {
  using Int = typename ZT<S>::Int;
  constexpr int bits = sizeof(Int) * 8;
  const S lim = pow2<S>(bits - 1);
  if (!(y < lim && y >= -lim)) return (Int)((typename ZT<S>::UInt)1 << (bits - 1));
  return (Int)y;
}

// ---------------------------------------------------------------------------
// gather / scatter with the dtype promotions of the chunk layer
// (crates/ndic-zfp/src/chunk.rs:309-356)
// ---------------------------------------------------------------------------
template <class S> __device__ __forceinline__ S load_elem(const uint8_t* p, uint64_t i, int32_t dtype) {
  if constexpr (std::is_same<S, int32_t>::value) {
    switch (dtype) {
      case NDIC_U8: return ((int32_t)p[i] - 0x80) * (1 << 23);
      case NDIC_I8: return (int32_t)(int8_t)p[i] * (1 << 23);
      case NDIC_U16: return ((int32_t)((const uint16_t*)p)[i] - 0x8000) * (1 << 15);
      case NDIC_I16: return (int32_t)((const int16_t*)p)[i] * (1 << 15);
      default: return ((const int32_t*)p)[i];
    }
  } else {
    return ((const S*)p)[i];
  }
}
template <class S> __device__ __forceinline__ void store_elem(uint8_t* p, uint64_t i, int32_t dtype, S v) {
  if constexpr (std::is_same<S, int32_t>::value) {
    switch (dtype) {
      case NDIC_U8: {
        int32_t t = (v >> 23) + 0x80;
        p[i] = (uint8_t)(t < 0 ? 0 : t > 0xff ? 0xff : t);
        return;
      }
      case NDIC_I8: {
        int32_t t = v >> 23;
        p[i] = (uint8_t)(int8_t)(t < -0x80 ? -0x80 : t > 0x7f ? 0x7f : t);
        return;
      }
      case NDIC_U16: {
        int32_t t = (v >> 15) + 0x8000;
        ((uint16_t*)p)[i] = (uint16_t)(t < 0 ? 0 : t > 0xffff ? 0xffff : t);
        return;
      }
      case NDIC_I16: {
        int32_t t = v >> 15;
        ((int16_t*)p)[i] = (int16_t)(t < -0x8000 ? -0x8000 : t > 0x7fff ? 0x7fff : t);
        return;
      }
      default: ((int32_t*)p)[i] = v; return;
    }
  } else {
    ((S*)p)[i] = v;
  }
}

__device__ __forceinline__ uint32_t pad_src(uint32_t c, uint32_t ext) {
  // pad rule of a partial 4-vector: n=1 -> p0 p0 p0 p0, n=2 -> p0 p1 p1 p0, n=3 -> p0 p1 p2 p0
  if (c < ext) return c;
  if (ext == 2 && c == 2) return 1;
  return 0;
}

__device__ __forceinline__ void block_coords(const ZfpChunk& c, uint64_t k, uint32_t* b) {
  b[0] = (uint32_t)(k % c.nb[0]);
  uint64_t t = k / c.nb[0];
  b[1] = (uint32_t)(t % c.nb[1]);
  t /= c.nb[1];
  b[2] = (uint32_t)(t % c.nb[2]);
  b[3] = (uint32_t)(t / c.nb[2]);
}

__device__ __forceinline__ uint32_t find_chunk(const ZfpChunk* chunks, uint32_t n, uint64_t g) {
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) / 2;
    if (chunks[mid].block_begin <= g) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Coefficient orders: by total sequency i+j+k+l, then by sum of squares.
// The 1D-3D orders (rank -> index), written out per (i, j, k); the 4D one is
// in zfp.hip.
#define P2(i, j) ((i) + 4 * (j))
#define P3(i, j, k) ((i) + 4 * (j) + 16 * (k))
constexpr uint8_t kPerm2[16] = {P2(0, 0), P2(1, 0), P2(0, 1), P2(1, 1), P2(2, 0), P2(0, 2), P2(2, 1), P2(1, 2),
                            P2(3, 0), P2(0, 3), P2(2, 2), P2(3, 1), P2(1, 3), P2(3, 2), P2(2, 3), P2(3, 3)};
constexpr uint8_t kPerm3[64] = {
    P3(0, 0, 0), P3(1, 0, 0), P3(0, 1, 0), P3(0, 0, 1), P3(0, 1, 1), P3(1, 0, 1), P3(1, 1, 0), P3(2, 0, 0),
    P3(0, 2, 0), P3(0, 0, 2), P3(1, 1, 1), P3(2, 1, 0), P3(2, 0, 1), P3(0, 2, 1), P3(1, 2, 0), P3(1, 0, 2),
    P3(0, 1, 2), P3(3, 0, 0), P3(0, 3, 0), P3(0, 0, 3), P3(2, 1, 1), P3(1, 2, 1), P3(1, 1, 2), P3(0, 2, 2),
    P3(2, 0, 2), P3(2, 2, 0), P3(3, 1, 0), P3(3, 0, 1), P3(0, 3, 1), P3(1, 3, 0), P3(1, 0, 3), P3(0, 1, 3),
    P3(1, 2, 2), P3(2, 1, 2), P3(2, 2, 1), P3(3, 1, 1), P3(1, 3, 1), P3(1, 1, 3), P3(3, 2, 0), P3(3, 0, 2),
    P3(0, 3, 2), P3(2, 3, 0), P3(2, 0, 3), P3(0, 2, 3), P3(2, 2, 2), P3(3, 2, 1), P3(3, 1, 2), P3(1, 3, 2),
    P3(2, 3, 1), P3(2, 1, 3), P3(1, 2, 3), P3(0, 3, 3), P3(3, 0, 3), P3(3, 3, 0), P3(3, 2, 2), P3(2, 3, 2),
    P3(2, 2, 3), P3(1, 3, 3), P3(3, 1, 3), P3(3, 3, 1), P3(2, 3, 3), P3(3, 2, 3), P3(3, 3, 2), P3(3, 3, 3)};
#undef P2
#undef P3

}  // namespace zfpdev
}  // namespace ndic
