// Device helpers of the ZFP kernels (zfp.hip, zfp_lane_kernels.hpp): scalar
// traits, the 4-point lifting steps, IEEE-754 helpers, element loads and stores with
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
// The 1D-3D orders (rank -> index), written out per (i, j, k).
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
// The 4D coefficient order (rank -> index): the order the LLNL C library 0.5.5
// emits, ranks measured from its streams.
constexpr uint8_t kPerm4[256] = {
    0,   1,   4,   16,  64,  5,   80,  17,  68,  65,  20,  2,   8,   32,  128, 84,  81,  69,  21,  6,   18,  66,
    24,  72,  9,   96,  33,  36,  129, 132, 144, 3,   12,  48,  192, 85,  82,  70,  22,  73,  25,  88,  37,  100,
    97,  148, 145, 133, 10,  160, 34,  136, 130, 40,  7,   19,  67,  28,  76,  13,  112, 49,  52,  193, 196, 208,
    86,  89,  101, 149, 161, 137, 41,  134, 38,  164, 26,  152, 146, 104, 98,  74,  83,  71,  23,  77,  29,  92,
    53,  116, 113, 212, 209, 197, 11,  35,  131, 44,  140, 14,  176, 50,  56,  194, 200, 224, 90,  165, 102, 153,
    150, 105, 168, 162, 138, 42,  87,  93,  117, 213, 27,  75,  99,  39,  135, 147, 108, 45,  141, 156, 30,  78,
    177, 180, 54,  114, 120, 57,  198, 210, 216, 201, 225, 228, 15,  240, 51,  204, 195, 60,  169, 166, 154, 106,
    91,  103, 151, 109, 157, 94,  181, 118, 121, 214, 217, 229, 163, 139, 43,  142, 46,  172, 58,  184, 178, 232,
    226, 202, 241, 205, 61,  199, 55,  244, 31,  220, 211, 124, 115, 79,  170, 167, 155, 107, 158, 110, 173, 122,
    185, 182, 233, 230, 218, 95,  245, 119, 221, 215, 125, 242, 206, 62,  203, 59,  248, 47,  236, 227, 188, 179,
    143, 171, 174, 186, 234, 246, 222, 126, 219, 123, 249, 111, 237, 231, 189, 183, 159, 252, 243, 207, 63,  175,
    250, 187, 238, 235, 190, 253, 247, 223, 127, 254, 251, 239, 191, 255};

}  // namespace zfpdev
}  // namespace ndic
