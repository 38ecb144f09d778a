// ZFP block codec on the device: one lane per 4^d block (d = 1..3), four
// lanes per block for 4D fields.
//
// Each lane runs the whole block through the coder: gather (pad rule of
// partial blocks), block-floating-point cast, decorrelating lifts, negabinary
// ranks, embedded group-test bit-plane coder, and the inverse of all that.
// The 4^d coefficients of a lane's block live in its registers: every index
// into them is a compile-time constant (the lifts walk fixed lines, the rank
// order is a constexpr table resolved per rank), so nothing spills.  Bit
// planes are the coder's unit: the decoder fills plane words (bit r = the
// coefficient of rank r) one plane at a time, the encoder reads them, and
// one in-register bit-matrix transpose converts between plane words and
// coefficient magnitudes.
//
// A wave decodes 64 neighbouring blocks.  When their bits are one contiguous
// stretch of one stream (fixed-rate, or variable-size with a block index or
// walked offsets) the wave first copies that stretch into LDS with
// coalesced loads, and the lanes' bit readers take their words from there;
// otherwise (bricks of different streams, a stretch larger than the stage)
// they read global memory.
//
// 4D blocks (256 coefficients) are shared by the four lanes of a quad (see
// the 4D section below).  The algorithm and stream layout are the published ZFP ones (Lindstrom, IEEE
// TVCG 2014; LLNL ZFP stream format, codec 5), which the reference reaches
// through zfp-rs 0.1.1 (crates/ndic-zfp/src/lib.rs:249-376).
#pragma once

#include <utility>

#include "zfp.hpp"
#include "zfp_dev.hpp"

namespace ndic {

namespace {

using namespace zfpdev;

constexpr uint32_t kLaneThreads = 64;

template <int D> constexpr int perm_at(int r) {
  return D == 1 ? r : D == 2 ? (int)kPerm2[r] : (int)kPerm3[r];
}

__device__ __forceinline__ uint32_t ctz64(uint64_t x) { return x ? (uint32_t)__builtin_ctzll(x) : 64u; }
__device__ __forceinline__ uint64_t lowmask64(uint32_t n) { return n >= 64 ? ~0ull : ((1ull << n) - 1); }

// ---------------------------------------------------------------------------
// bit-matrix transpose of 32 words of 32 bits, in place: bit i of a[k]
// becomes bit k of a[i] (recursive exchange of off-diagonal blocks)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void transpose32(uint32_t* a) {
#pragma unroll
  for (int j = 16; j >= 1; j >>= 1) {
    const uint32_t m = 0xffffffffu / ((1u << j) + 1u);  // bits whose index has bit j clear
#pragma unroll
    for (int k = 0; k < 32; ++k) {
      if (k & j) continue;
      const uint32_t t = ((a[k] >> j) ^ a[k + j]) & m;
      a[k + j] ^= t;
      a[k] ^= t << j;
    }
  }
}

// Plane words <-> coefficient magnitudes for S = 4^D coefficients of IP bits.
// Planes: lo[k] holds bit k of coefficients 0..31 (bit r = rank r), hi[k] of
// coefficients 32..63 (S = 64 only).  Coefficients: u[r], r < S.
template <int S, int IP> struct PlaneSet {
  static constexpr int kHi = S > 32 ? IP : 1;
  uint32_t lo[IP];
  uint32_t hi[kHi];
};

template <int S, class UInt> __device__ __forceinline__ void planes_to_coeffs(PlaneSet<S, sizeof(UInt) * 8>& p,
                                                                               UInt* u) {
  constexpr int IP = sizeof(UInt) * 8;
  constexpr int S0 = S < 32 ? S : 32;  // coefficients in lo
  transpose32(p.lo);
  if constexpr (IP == 64) transpose32(p.lo + 32);
  if constexpr (S > 32) {
    transpose32(p.hi);
    if constexpr (IP == 64) transpose32(p.hi + 32);
  }
  // 64 planes: planes 32..63 (the upper coefficient bits) are their own
  // 32x32 blocks
#pragma unroll
  for (int r = 0; r < S0; ++r) {
    if constexpr (IP == 64) u[r] = ((uint64_t)p.lo[32 + r] << 32) | p.lo[r];
    else u[r] = p.lo[r];
  }
  if constexpr (S > 32) {
#pragma unroll
    for (int r = 0; r < 32; ++r) {
      if constexpr (IP == 64) u[32 + r] = ((uint64_t)p.hi[32 + r] << 32) | p.hi[r];
      else u[32 + r] = p.hi[r];
    }
  }
}

template <int S, class UInt> __device__ __forceinline__ void coeffs_to_planes(const UInt* u,
                                                                               PlaneSet<S, sizeof(UInt) * 8>& p) {
  constexpr int IP = sizeof(UInt) * 8;
  constexpr int S0 = S < 32 ? S : 32;
#pragma unroll
  for (int r = 0; r < IP; ++r) p.lo[r] = 0;
#pragma unroll
  for (int r = 0; r < S0; ++r) {
    p.lo[r] = (uint32_t)u[r];
    if constexpr (IP == 64) p.lo[32 + r] = (uint32_t)((uint64_t)u[r] >> 32);
  }
  if constexpr (S > 32) {
#pragma unroll
    for (int r = 0; r < 32; ++r) {
      p.hi[r] = (uint32_t)u[32 + r];
      if constexpr (IP == 64) p.hi[32 + r] = (uint32_t)((uint64_t)u[32 + r] >> 32);
    }
  }
  transpose32(p.lo);
  if constexpr (IP == 64) transpose32(p.lo + 32);
  if constexpr (S > 32) {
    transpose32(p.hi);
    if constexpr (IP == 64) transpose32(p.hi + 32);
  }
}

// ranks <-> coefficients in block order (negabinary), resolved per rank at
// compile time
template <int D, class Int, class UInt, int... R>
__device__ __forceinline__ void ranks_to_block(const UInt* u, Int* v, std::integer_sequence<int, R...>) {
  const UInt m = nbmask<UInt>();
  ((v[std::integral_constant<int, perm_at<D>(R)>::value] = (Int)((u[R] ^ m) - m)), ...);
}
template <int D, class Int, class UInt, int... R>
__device__ __forceinline__ void block_to_ranks(const Int* v, UInt* u, std::integer_sequence<int, R...>) {
  const UInt m = nbmask<UInt>();
  ((u[R] = ((UInt)v[std::integral_constant<int, perm_at<D>(R)>::value] + m) ^ m), ...);
}

// separable lifts along x, y, z of a block in registers (forward: x first)
template <int D, class Int, bool Fwd, bool Rev> __device__ __forceinline__ void lane_xform(Int* v) {
  constexpr int S = 1 << (2 * D);
#pragma unroll
  for (int t = 0; t < D; ++t) {
    const int a = Fwd ? t : D - 1 - t;
    const int s = 1 << (2 * a);
#pragma unroll
    for (int p = 0; p < S; ++p) {
      if ((p / s) % 4) continue;
      lift4<Int, Fwd, Rev>(v[p], v[p + s], v[p + 2 * s], v[p + 3 * s]);
    }
  }
}

// ---------------------------------------------------------------------------
// per-lane bit reader: words from the wave's LDS stage or global memory
// ---------------------------------------------------------------------------
extern __shared__ uint64_t zstage[];

__device__ __forceinline__ uint64_t global_word(const uint8_t* base, uint64_t nbytes, uint64_t w) {
  typedef const __attribute__((address_space(1))) uint64_t* g64;
  typedef const __attribute__((address_space(1))) uint8_t* g8;
  const uint64_t b = w * 8;
  if (b + 8 <= nbytes) return *(g64)(uintptr_t)(base + b);
  uint64_t v = 0;
  for (uint64_t i = b; i < nbytes && i < b + 8; ++i) v |= (uint64_t)*(g8)(uintptr_t)(base + i) << (8 * (i - b));
  return v;
}

// Keeps the 128 stream bits at its position in two words and the word after
// them in flight, so a word is asked for one word of parsing before it is
// needed.
struct LaneReader {
  const uint8_t* base;
  uint64_t nbytes;
  uint64_t s0, sn;   // the stage holds words s0 .. s0 + sn - 1
  uint64_t next;     // index of the word after w2
  uint64_t cur, w1, w2;
  uint32_t left;     // valid bits of cur (low end), 1..64

  __device__ __forceinline__ uint64_t load(uint64_t w) const {
    if (w - s0 < sn) return zstage[w - s0];
    return global_word(base, nbytes, w);
  }
  __device__ __forceinline__ void seek(uint64_t bit) {
    const uint64_t w = bit >> 6;
    const uint32_t o = (uint32_t)bit & 63;
    cur = load(w) >> o;
    left = 64 - o;
    w1 = load(w + 1);
    w2 = load(w + 2);
    next = w + 3;
  }
  // the next 64 bits, first bit lowest
  __device__ __forceinline__ uint64_t peek() const { return left == 64 ? cur : cur | (w1 << left); }
  // consumes n <= 64 bits
  __device__ __forceinline__ void skip(uint32_t n) {
    if (n < left) {
      cur >>= n;
      left -= n;
      return;
    }
    const uint32_t r = n - left;  // < 64: left >= 1
    cur = w1 >> r;
    left = 64 - r;
    w1 = w2;
    w2 = load(next++);
  }
  __device__ __forceinline__ uint64_t get(uint32_t n) {
    const uint64_t v = peek() & lowmask64(n);
    skip(n);
    return v;
  }
  __device__ __forceinline__ uint32_t bit() { return (uint32_t)get(1); }
};

// ---------------------------------------------------------------------------
// per-lane bit writer (LSB-first 64-bit words)
// ---------------------------------------------------------------------------
// Fixed-rate blocks share their edge words with the neighbouring blocks (OR
// them in); the stream is zeroed beforehand, so zero words need no store.
// A variable-size block writes every word of its private slot.
struct LaneWriter {
  uint64_t* out;
  uint64_t word;
  uint64_t acc;
  uint32_t fill;
  bool shared, first;
  uint64_t count;

  __device__ __forceinline__ void init(uint64_t* base, uint64_t bitpos, bool share) {
    out = base;
    word = bitpos >> 6;
    fill = (uint32_t)bitpos & 63;
    acc = 0;
    shared = share;
    first = true;
    count = 0;
  }
  __device__ __forceinline__ void flush() {
    if (shared && first) atomicOr((unsigned long long*)&out[word], (unsigned long long)acc);
    else if (!shared || acc) out[word] = acc;
    first = false;
    ++word;
    acc = 0;
  }
  // appends the low n <= 64 bits of v (v has no bits at or above n)
  __device__ __forceinline__ void put(uint64_t v, uint32_t n) {
    if (!n) return;
    count += n;
    acc |= v << fill;
    const uint32_t f = fill + n;
    if (f >= 64) {
      const uint32_t used = 64 - fill;
      flush();
      acc = used < 64 ? v >> used : 0;
      fill = f - 64;
    } else {
      fill = f;
    }
  }
  __device__ __forceinline__ void zeros(uint64_t n) {
    if (!n) return;
    count += n;
    if (fill + n < 64) {
      fill += (uint32_t)n;
      return;
    }
    n -= 64 - fill;
    flush();
    while (n >= 64) {
      if (!shared) out[word] = 0;
      ++word;
      n -= 64;
    }
    fill = (uint32_t)n;
  }
  __device__ __forceinline__ void finish() {
    if (!fill) return;
    if (shared) atomicOr((unsigned long long*)&out[word], (unsigned long long)acc);
    else out[word] = acc;
  }
};

// ---------------------------------------------------------------------------
// embedded bit-plane coder over one lane's ranked coefficients
// ---------------------------------------------------------------------------
// Planes KHI down to KLO (at most 32 of them, so that the loop unrolls and
// every plane word has a register of its own).
template <int S, int IP, int KHI, int KLO>
__device__ __forceinline__ void lane_decode_planes(LaneReader& r, uint32_t kmin, uint32_t& bits, uint32_t& n,
                                                   PlaneSet<S, IP>& p) {
#pragma unroll
  for (int k = KHI; k >= KLO; --k) {
    uint64_t x = 0;
    if ((uint32_t)k >= kmin && bits) {
      const uint32_t m = n < bits ? n : bits;
      bits -= m;
      if (m) x = r.get(m);
      while (n < (uint32_t)S && bits) {
        bits--;
        const uint64_t w = r.peek();
        if (!(w & 1)) {
          r.skip(1);
          break;
        }
        // a positive group test: zeros up to the next significant
        // coefficient (at most lim of them), then its one (implied at lim)
        uint32_t lim = S - 1 - n;
        if (lim > bits) lim = bits;
        uint32_t z = ctz64(w >> 1);
        if (z < lim) {
          r.skip(z + 2);
          bits -= z + 1;
        } else {
          z = lim;
          r.skip(lim + 1);
          bits -= lim;
        }
        n += z;
        x |= 1ull << n;
        n++;
      }
    }
    p.lo[k] = (uint32_t)x;
    if constexpr (S > 32) p.hi[k] = (uint32_t)(x >> 32);
  }
}

template <int S, int IP>
__device__ __forceinline__ uint32_t lane_decode_ints(LaneReader& r, uint32_t maxbits, uint32_t maxprec,
                                                     PlaneSet<S, IP>& p) {
  const uint32_t kmin = IP > (int)maxprec ? IP - maxprec : 0;
  uint32_t bits = maxbits, n = 0;
  if constexpr (IP == 64) lane_decode_planes<S, IP, 63, 32>(r, kmin, bits, n, p);
  lane_decode_planes<S, IP, 31, 0>(r, kmin, bits, n, p);
  return maxbits - bits;
}

template <int S, int IP, int KHI, int KLO>
__device__ __forceinline__ void lane_encode_planes(LaneWriter& w, uint32_t kmin, uint32_t& bits, uint32_t& n,
                                                   const PlaneSet<S, IP>& p) {
#pragma unroll
  for (int k = KHI; k >= KLO; --k) {
    if ((uint32_t)k >= kmin && bits) {
      uint64_t x = p.lo[k];
      if constexpr (S > 32) x |= (uint64_t)p.hi[k] << 32;
      const uint32_t m = n < bits ? n : bits;
      bits -= m;
      w.put(x & lowmask64(m), m);
      while (n < (uint32_t)S && bits) {
        bits--;
        const uint64_t rest = x >> n;  // n < S <= 64
        if (!rest) {
          w.put(0, 1);
          break;
        }
        const uint32_t z = (uint32_t)__builtin_ctzll(rest);
        w.put(1, 1);
        if (bits <= z) {
          w.zeros(bits);
          bits = 0;
          break;
        }
        bits -= z;
        n += z;
        if (n < (uint32_t)S - 1) {
          bits--;
          w.put(1ull << z, z + 1);
        } else {
          w.zeros(z);
        }
        n++;
      }
    }
  }
}

template <int S, int IP>
__device__ __forceinline__ uint32_t lane_encode_ints(LaneWriter& w, uint32_t maxbits, uint32_t maxprec,
                                                     const PlaneSet<S, IP>& p) {
  const uint32_t kmin = IP > (int)maxprec ? IP - maxprec : 0;
  uint32_t bits = maxbits, n = 0;
  if constexpr (IP == 64) lane_encode_planes<S, IP, 63, 32>(w, kmin, bits, n, p);
  lane_encode_planes<S, IP, 31, 0>(w, kmin, bits, n, p);
  return maxbits - bits;
}

// ---------------------------------------------------------------------------
// integer blocks
// ---------------------------------------------------------------------------
template <int D, class Int>
__device__ __forceinline__ uint32_t lane_decode_int_block(LaneReader& r, int minbits, int maxbits, int maxprec,
                                                          bool rev, Int* v) {
  using UInt = typename std::make_unsigned<Int>::type;
  constexpr int S = 1 << (2 * D);
  constexpr int IP = sizeof(Int) * 8;
  constexpr int pbits = sizeof(Int) == 4 ? 5 : 6;
  int bits = 0;
  uint32_t prec = (uint32_t)maxprec;
  if (rev) {
    prec = (uint32_t)r.get(pbits) + 1;
    bits += pbits;
  }
  PlaneSet<S, IP> p;
  bits += (int)lane_decode_ints<S, IP>(r, maxbits - bits > 0 ? (uint32_t)(maxbits - bits) : 0u, prec, p);
  if (bits < minbits) bits = minbits;
  UInt u[S];
  planes_to_coeffs<S, UInt>(p, u);
  ranks_to_block<D, Int, UInt>(u, v, std::make_integer_sequence<int, S>{});
  if (rev) lane_xform<D, Int, false, true>(v);
  else lane_xform<D, Int, false, false>(v);
  return (uint32_t)bits;
}

template <int D, class Int>
__device__ __forceinline__ uint32_t lane_encode_int_block(LaneWriter& w, int minbits, int maxbits, int maxprec,
                                                          bool rev, Int* v) {
  using UInt = typename std::make_unsigned<Int>::type;
  constexpr int S = 1 << (2 * D);
  constexpr int IP = sizeof(Int) * 8;
  constexpr int pbits = sizeof(Int) == 4 ? 5 : 6;
  if (rev) lane_xform<D, Int, true, true>(v);
  else lane_xform<D, Int, true, false>(v);
  UInt u[S];
  block_to_ranks<D, Int, UInt>(v, u, std::make_integer_sequence<int, S>{});
  int bits = 0;
  uint32_t prec = (uint32_t)maxprec;
  if (rev) {
    // precision: the planes down to the lowest nonzero bit of any coefficient
    UInt m = 0;
#pragma unroll
    for (int i = 0; i < S; ++i) m |= u[i];
    prec = m ? IP - (uint32_t)__builtin_ctzll((unsigned long long)m) : 0;
    if (prec > (uint32_t)maxprec) prec = (uint32_t)maxprec;
    if (prec < 1) prec = 1;
    w.put(prec - 1, pbits);
    bits = pbits;
  }
  PlaneSet<S, IP> p;
  coeffs_to_planes<S, UInt>(u, p);
  bits += (int)lane_encode_ints<S, IP>(w, maxbits - bits > 0 ? (uint32_t)(maxbits - bits) : 0u, prec, p);
  if (bits < minbits) {
    w.zeros((uint64_t)(minbits - bits));
    bits = minbits;
  }
  return (uint32_t)bits;
}

// ---------------------------------------------------------------------------
// floating-point blocks
// ---------------------------------------------------------------------------
// x = m 2^e, 0.5 <= m < 1, of the block's largest |value| (NaN ignored,
// infinity -> 0)
template <int S, class F> __device__ __forceinline__ int lane_exponent(const F* f) {
  using U = typename ZT<F>::UInt;
  constexpr int bias = ZT<F>::ebias, mb = ZT<F>::mbits;
  constexpr U absmask = ~(U)0 >> 1, inf = (U)(2 * bias + 1) << mb;
  U mx = 0;
#pragma unroll
  for (int j = 0; j < S; ++j) {
    U a = fbits(f[j]) & absmask;
    if (a > inf) a = 0;
    mx = a > mx ? a : mx;
  }
  if (mx == 0) return -bias;
  const int be = (int)(mx >> mb);
  if (be == 0) return 1 - bias;
  if (be == 2 * bias + 1) return 0;
  return be - (bias - 1);
}

template <int D, class F> __device__ __forceinline__ uint32_t lane_decode_float_block(LaneReader& r,
                                                                                   const ZfpChunk& c, F* f) {
  using Int = typename ZT<F>::Int;
  using U = typename ZT<F>::UInt;
  constexpr int S = 1 << (2 * D);
  constexpr int ebits = ZT<F>::ebits, ebias = ZT<F>::ebias;
  Int v[S];
  if (c.minexp < kZfpMinExp) {  // reversible
    uint32_t bits = 1;
    if (r.bit()) {
      bits++;
      if (r.bit()) {  // the samples' bits, coded losslessly
        bits += lane_decode_int_block<D, Int>(r, (int)c.minbits - (int)bits, (int)c.maxbits - (int)bits,
                                              (int)c.maxprec, true, v);
        const U mx = ~(U)0 >> 1;
#pragma unroll
        for (int j = 0; j < S; ++j) {
          const Int x = v[j] < 0 ? (Int)((U)v[j] ^ mx) : v[j];
          f[j] = fromb<F>((U)x);
        }
      } else {  // an exact block-floating-point cast
        bits += ebits;
        const int emax = (int)r.get(ebits) - ebias;
        bits += lane_decode_int_block<D, Int>(r, (int)c.minbits - (int)bits, (int)c.maxbits - (int)bits,
                                              (int)c.maxprec, true, v);
        // the lowest exponent marks an all-zero block: +0 whatever the
        // integers (the reversible inverse cast of the ZFP algorithm)
        const F s = pow2<F>(emax - ((int)sizeof(F) * 8 - 2));
        const bool zero = emax == -ebias;
#pragma unroll
        for (int j = 0; j < S; ++j) f[j] = zero ? (F)0 : s * (F)v[j];
      }
      return bits;
    }
#pragma unroll
    for (int j = 0; j < S; ++j) f[j] = 0;
    return bits;
  }
  if (r.bit()) {
    uint32_t bits = 1 + ebits;
    const int emax = (int)r.get(ebits) - ebias;
    const int dp = emax - c.minexp + 2 * (D + 1);
    const int maxprec = (int)c.maxprec < (dp > 0 ? dp : 0) ? (int)c.maxprec : (dp > 0 ? dp : 0);
    bits += lane_decode_int_block<D, Int>(r, (int)c.minbits - (int)bits, (int)c.maxbits - (int)bits, maxprec, false,
                                          v);
    const F s = pow2<F>(emax - ((int)sizeof(F) * 8 - 2));
#pragma unroll
    for (int j = 0; j < S; ++j) f[j] = s * (F)v[j];
    return bits;
  }
#pragma unroll
  for (int j = 0; j < S; ++j) f[j] = 0;
  return c.minbits > 1 ? c.minbits : 1;
}

template <int D, class F> __device__ __forceinline__ uint32_t lane_encode_float_block(LaneWriter& w,
                                                                                   const ZfpChunk& c,
                                                                                   const F* f) {
  using Int = typename ZT<F>::Int;
  using U = typename ZT<F>::UInt;
  constexpr int S = 1 << (2 * D);
  constexpr int ebits = ZT<F>::ebits, ebias = ZT<F>::ebias;
  Int v[S];
  const int emax = lane_exponent<S, F>(f);
  if (c.minexp < kZfpMinExp) {  // reversible
    uint32_t bits = 0;
    bool same = true;
    if (emax != -ebias) {
      const F s = pow2<F>((int)sizeof(F) * 8 - 2 - emax), si = pow2<F>(emax - ((int)sizeof(F) * 8 - 2));
#pragma unroll
      for (int j = 0; j < S; ++j) {
        v[j] = cast_int<F>(s * f[j]);
        same &= fbits(si * (F)v[j]) == fbits(f[j]);
      }
    } else {
#pragma unroll
      for (int j = 0; j < S; ++j) {
        v[j] = 0;
        same &= fbits(f[j]) == 0;
      }
    }
    if (same) {
      const uint32_t e = (uint32_t)(emax + ebias);
      if (!e) {
        w.put(0, 1);
        return 1;
      }
      w.put(1, 2);
      w.put(e, ebits);
      bits = 2 + ebits;
    } else {
      const U mx = ~(U)0 >> 1;
#pragma unroll
      for (int j = 0; j < S; ++j) {
        const U b = fbits(f[j]);
        v[j] = (Int)b < 0 ? (Int)(b ^ mx) : (Int)b;
      }
      w.put(3, 2);
      bits = 2;
    }
    return bits + lane_encode_int_block<D, Int>(w, (int)c.minbits - (int)bits, (int)c.maxbits - (int)bits,
                                                 (int)c.maxprec, true, v);
  }
  const int dp = emax - c.minexp + 2 * (D + 1);
  const int maxprec = (int)c.maxprec < (dp > 0 ? dp : 0) ? (int)c.maxprec : (dp > 0 ? dp : 0);
  const uint32_t e = maxprec ? (uint32_t)(emax + ebias) : 0;
  if (e) {
    const uint32_t bits = 1 + ebits;
    w.put(2 * (uint64_t)e + 1, bits);
    const F s = pow2<F>((int)sizeof(F) * 8 - 2 - emax);
#pragma unroll
    for (int j = 0; j < S; ++j) v[j] = cast_int<F>(s * f[j]);
    return bits + lane_encode_int_block<D, Int>(w, (int)c.minbits - (int)bits, (int)c.maxbits - (int)bits,
                                                 maxprec, false, v);
  }
  w.put(0, 1);
  if (c.minbits > 1) {
    w.zeros(c.minbits - 1);
    return c.minbits;
  }
  return 1;
}

template <int D, class S> __device__ __forceinline__ uint32_t lane_decode_block(LaneReader& r, const ZfpChunk& c,
                                                                               S* vals) {
  if constexpr (ZT<S>::fp) {
    return lane_decode_float_block<D, S>(r, c, vals);
  } else {
    return lane_decode_int_block<D, S>(r, (int)c.minbits, (int)c.maxbits, (int)c.maxprec, c.minexp < kZfpMinExp,
                                       vals);
  }
}

template <int D, class S> __device__ __forceinline__ uint32_t lane_encode_block(LaneWriter& w, const ZfpChunk& c,
                                                                               S* vals) {
  if constexpr (ZT<S>::fp) {
    return lane_encode_float_block<D, S>(w, c, vals);
  } else {
    return lane_encode_int_block<D, S>(w, (int)c.minbits, (int)c.maxbits, (int)c.maxprec, c.minexp < kZfpMinExp,
                                       vals);
  }
}

// ---------------------------------------------------------------------------
// gather / scatter of one lane's block (x fastest; rows of four samples)
// ---------------------------------------------------------------------------
template <class S> __device__ __forceinline__ bool vector_rows(const ZfpChunk& c, const uint8_t* p) {
  // whole 4-sample rows of a native 4- or 8-byte type start 16-byte aligned
  const bool native = !std::is_same<S, int32_t>::value || c.dtype == NDIC_I32;
  return native && (c.n[0] & 3) == 0 && ((uintptr_t)p & 15) == 0;
}

template <class S> __device__ __forceinline__ void load_row4(const S* p, S* o) {
  if constexpr (sizeof(S) == 4) {
    typedef __attribute__((ext_vector_type(4))) uint32_t u4;
    const u4 t = *(const u4*)p;
    uint32_t a[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
    for (int i = 0; i < 4; ++i) __builtin_memcpy(&o[i], &a[i], 4);
  } else {
    typedef __attribute__((ext_vector_type(2))) uint64_t u2;
    const u2 t0 = *(const u2*)p, t1 = *(const u2*)(p + 2);
    uint64_t a[4] = {t0.x, t0.y, t1.x, t1.y};
#pragma unroll
    for (int i = 0; i < 4; ++i) __builtin_memcpy(&o[i], &a[i], 8);
  }
}

template <class S> __device__ __forceinline__ void store_row4(S* p, const S* v) {
  if constexpr (sizeof(S) == 4) {
    typedef __attribute__((ext_vector_type(4))) uint32_t u4;
    uint32_t a[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) __builtin_memcpy(&a[i], &v[i], 4);
    *(u4*)p = u4{a[0], a[1], a[2], a[3]};
  } else {
    typedef __attribute__((ext_vector_type(2))) uint64_t u2;
    uint64_t a[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) __builtin_memcpy(&a[i], &v[i], 8);
    *(u2*)p = u2{a[0], a[1]};
    *(u2*)(p + 2) = u2{a[2], a[3]};
  }
}

template <int D> __device__ __forceinline__ void block_extents(const ZfpChunk& c, const uint32_t* b, uint32_t* ext) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const uint32_t e = a < D ? c.n[a] - 4 * b[a] : 1;
    ext[a] = e < 4 ? e : 4;
  }
}

template <int D, class S> __device__ __forceinline__ void lane_gather(const ZfpChunk& c, const uint32_t* b, S* vals) {
  uint32_t ext[3];
  block_extents<D>(c, b, ext);
  const bool vec = vector_rows<S>(c, c.src) && ext[0] == 4;
#pragma unroll
  for (int zz = 0; zz < (D > 2 ? 4 : 1); ++zz)
#pragma unroll
    for (int yy = 0; yy < (D > 1 ? 4 : 1); ++yy) {
      const uint64_t sz = D > 2 ? 4ull * b[2] + pad_src(zz, ext[2]) : 0;
      const uint64_t sy = D > 1 ? 4ull * b[1] + pad_src(yy, ext[1]) : 0;
      const uint64_t row = (sz * c.n[1] + sy) * c.n[0] + 4ull * b[0];
      S* o = vals + 4 * yy + 16 * zz;
      if (vec) {
        load_row4<S>((const S*)c.src + row, o);
      } else {
#pragma unroll
        for (int xx = 0; xx < 4; ++xx) o[xx] = load_elem<S>(c.src, row + pad_src(xx, ext[0]), c.dtype);
      }
    }
}

template <int D, class S>
__device__ __forceinline__ void lane_scatter(const ZfpChunk& c, const uint32_t* b, const S* vals, uint8_t* brick) {
  uint32_t ext[3];
  block_extents<D>(c, b, ext);
  uint8_t* dst = brick ? brick : c.dst;
  const bool vec = !brick && vector_rows<S>(c, dst) && ext[0] == 4;
#pragma unroll
  for (int zz = 0; zz < (D > 2 ? 4 : 1); ++zz)
#pragma unroll
    for (int yy = 0; yy < (D > 1 ? 4 : 1); ++yy) {
      if ((uint32_t)zz >= ext[2] || (uint32_t)yy >= ext[1]) continue;
      const S* v = vals + 4 * yy + 16 * zz;
      uint64_t row;
      if (brick) row = ((uint64_t)zz * ext[1] + yy) * ext[0];
      else row = ((4ull * b[2] + zz) * c.n[1] + 4ull * b[1] + yy) * c.n[0] + 4ull * b[0];
      if (vec) {
        store_row4<S>((S*)dst + row, v);
      } else {
#pragma unroll
        for (int xx = 0; xx < 4; ++xx)
          if ((uint32_t)xx < ext[0]) store_elem<S>(dst, row + xx, c.dtype, v[xx]);
      }
    }
}

__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    const uint64_t t = __shfl_xor(v, o);
    v = t < v ? t : v;
  }
  return v;
}
__device__ __forceinline__ uint64_t wave_max_u64(uint64_t v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    const uint64_t t = __shfl_xor(v, o);
    v = t > v ? t : v;
  }
  return v;
}

__device__ __forceinline__ uint64_t lane_block_id() {
  return (blockIdx.x + (uint64_t)blockIdx.y * gridDim.x) * kLaneThreads + threadIdx.x;
}

// ---------------------------------------------------------------------------
// kernels: one wave per workgroup, one block per lane
// ---------------------------------------------------------------------------
template <int D, class S>
__global__ __launch_bounds__(kLaneThreads) void zfp_lane_encode_kernel(const ZfpChunk* __restrict__ chunks,
                                                                      uint32_t nchunks, uint64_t total) {
  constexpr int SZ = 1 << (2 * D);
  const uint64_t g = lane_block_id();
  if (g >= total) return;
  const ZfpChunk& c = chunks[find_chunk(chunks, nchunks, g)];
  if (c.fixed ? !c.dst : !c.slots) return;  // refused chunk
  const uint64_t k = g - c.block_begin;
  uint32_t b[4];
  block_coords(c, k, b);
  S vals[SZ];
  lane_gather<D, S>(c, b, vals);
  LaneWriter w;
  if (c.fixed) w.init((uint64_t*)c.dst, c.header_bits + k * c.maxbits, true);
  else w.init(c.slots + g * c.slot_words, 0, false);
  lane_encode_block<D, S>(w, c, vals);
  w.finish();
  if (!c.fixed) c.offs[g] = w.count;
}

template <int D, class S>
__global__ __launch_bounds__(kLaneThreads) void zfp_lane_decode_kernel(const ZfpChunk* __restrict__ chunks,
                                                                      uint32_t nchunks, uint64_t total,
                                                                      const ZfpBrick* __restrict__ bricks,
                                                                      uint32_t stage_words) {
  constexpr int SZ = 1 << (2 * D);
  const uint64_t g = lane_block_id();
  const bool valid = g < total;
  uint32_t ci = 0;
  uint64_t k = 0;
  uint8_t* brick = nullptr;
  if (valid) {
    if (bricks) {
      ci = bricks[g].chunk;
      k = bricks[g].block;
      brick = bricks[g].out;
    } else {
      ci = find_chunk(chunks, nchunks, g);
      k = g - chunks[ci].block_begin;
    }
  }
  const ZfpChunk& c = chunks[ci];
  const bool indexed = !c.fixed && c.index;
  // this block's bits: [start, end)
  uint64_t start = 0, end = 0;
  if (valid) {
    if (c.fixed) {
      start = c.header_bits + k * c.maxbits;
      end = start + c.maxbits;
    } else if (indexed) {
      start = c.header_bits + c.index[k];
      end = c.header_bits + c.index[k + 1];
    } else {
      start = c.offs[c.block_begin + k];
      end = k + 1 < c.nblocks ? c.offs[c.block_begin + k + 1] : 8 * c.src_len;
    }
  }
  // Stage the wave's stretch of the stream in LDS when its blocks are one
  // stretch of one stream that fits (words lo .. hi + 2: the readers keep
  // two words ahead).
  const uint32_t ci0 = (uint32_t)__shfl((int)ci, 0);
  const bool one = __ballot(valid && ci != ci0) == 0;
  const uint64_t lo = wave_min_u64(valid ? start : ~0ull) >> 6;
  const uint64_t hi = ((wave_max_u64(valid ? end : 0) + 63) >> 6) + 2;
  const bool staged = one && hi > lo && hi - lo <= stage_words;
  LaneReader r;
  r.base = c.src;
  r.nbytes = c.src_len;
  r.s0 = 0;
  r.sn = 0;
  if (staged) {
    const ZfpChunk& c0 = chunks[ci0];
    for (uint64_t i = threadIdx.x; i < hi - lo; i += kLaneThreads) zstage[i] = global_word(c0.src, c0.src_len, lo + i);
    r.s0 = lo;
    r.sn = hi - lo;
  }
  __syncthreads();
  if (!valid) return;
  r.seek(start);
  S vals[SZ];
  const uint32_t used = lane_decode_block<D, S>(r, c, vals);
  if (indexed) {
    // the block must take exactly the bits the index gives it, the index
    // must start at the first block and end inside the stream
    const uint32_t ilo = c.index[k], ihi = c.index[k + 1];
    if (ihi < ilo || ihi - ilo != used || (k == 0 && ilo != 0) ||
        (k + 1 == c.nblocks && c.header_bits + (uint64_t)ihi > 8 * c.src_len))
      atomicOr(c.bad, 1u);
  }
  uint32_t b[4];
  block_coords(c, k, b);
  lane_scatter<D, S>(c, b, vals, brick);
}

// ---------------------------------------------------------------------------
// 4D fields: four lanes per block (sixteen blocks a wave)
// ---------------------------------------------------------------------------
// The 256 coefficients of a 4^4 block do not fit one lane's registers, so
// the four lanes of a quad share a block: lane s holds the w-slice w = s
// (64 samples) for the casts and the x/y/z lifts, and rank quarter s (ranks
// 64 s .. 64 s + 63) for the bit-plane coder.  The quad parses (decode) or
// codes (encode) the stream in lockstep -- the same bits, the same decisions
// -- each lane keeping its quarter of every plane word (encode: the quarters
// travel across the quad by DPP, plane by plane).  Ranks reach positions,
// and the w lift its 4-sample lines, through the block's 256 entries in LDS.
constexpr int kQuadDppBcast[4] = {0x00, 0x55, 0xAA, 0xFF};  // quad_perm [q,q,q,q]

__device__ __forceinline__ uint64_t dtype_bytes_dev(int32_t d) {
  return d == NDIC_U8 || d == NDIC_I8 ? 1 : d == NDIC_U16 || d == NDIC_I16 ? 2 : d == NDIC_I64 || d == NDIC_U64 || d == NDIC_F64 ? 8 : 4;
}

__constant__ __align__(16) uint8_t c_lane_perm4[256] = {
#define P(i) kPerm4[i]
    P(0), P(1), P(2), P(3), P(4), P(5), P(6), P(7), P(8), P(9), P(10), P(11), P(12), P(13), P(14), P(15),
    P(16), P(17), P(18), P(19), P(20), P(21), P(22), P(23), P(24), P(25), P(26), P(27), P(28), P(29), P(30), P(31),
    P(32), P(33), P(34), P(35), P(36), P(37), P(38), P(39), P(40), P(41), P(42), P(43), P(44), P(45), P(46), P(47),
    P(48), P(49), P(50), P(51), P(52), P(53), P(54), P(55), P(56), P(57), P(58), P(59), P(60), P(61), P(62), P(63),
    P(64), P(65), P(66), P(67), P(68), P(69), P(70), P(71), P(72), P(73), P(74), P(75), P(76), P(77), P(78), P(79),
    P(80), P(81), P(82), P(83), P(84), P(85), P(86), P(87), P(88), P(89), P(90), P(91), P(92), P(93), P(94), P(95),
    P(96), P(97), P(98), P(99), P(100), P(101), P(102), P(103), P(104), P(105), P(106), P(107), P(108), P(109),
    P(110), P(111), P(112), P(113), P(114), P(115), P(116), P(117), P(118), P(119), P(120), P(121), P(122), P(123),
    P(124), P(125), P(126), P(127), P(128), P(129), P(130), P(131), P(132), P(133), P(134), P(135), P(136), P(137),
    P(138), P(139), P(140), P(141), P(142), P(143), P(144), P(145), P(146), P(147), P(148), P(149), P(150), P(151),
    P(152), P(153), P(154), P(155), P(156), P(157), P(158), P(159), P(160), P(161), P(162), P(163), P(164), P(165),
    P(166), P(167), P(168), P(169), P(170), P(171), P(172), P(173), P(174), P(175), P(176), P(177), P(178), P(179),
    P(180), P(181), P(182), P(183), P(184), P(185), P(186), P(187), P(188), P(189), P(190), P(191), P(192), P(193),
    P(194), P(195), P(196), P(197), P(198), P(199), P(200), P(201), P(202), P(203), P(204), P(205), P(206), P(207),
    P(208), P(209), P(210), P(211), P(212), P(213), P(214), P(215), P(216), P(217), P(218), P(219), P(220), P(221),
    P(222), P(223), P(224), P(225), P(226), P(227), P(228), P(229), P(230), P(231), P(232), P(233), P(234), P(235),
    P(236), P(237), P(238), P(239), P(240), P(241), P(242), P(243), P(244), P(245), P(246), P(247), P(248), P(249),
    P(250), P(251), P(252), P(253), P(254), P(255)
#undef P
};

// ranks 64 s .. 64 s + 63 of the 4D order: byte r of word r / 4
struct QuadPerm {
  uint32_t w[16];
  __device__ __forceinline__ void load(uint32_t s) {
    const uint4* t = (const uint4*)(c_lane_perm4 + 64 * s);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const uint4 v = t[i];
      w[4 * i] = v.x;
      w[4 * i + 1] = v.y;
      w[4 * i + 2] = v.z;
      w[4 * i + 3] = v.w;
    }
  }
  __device__ __forceinline__ uint32_t operator[](int r) const { return (w[r >> 2] >> (8 * (r & 3))) & 0xFF; }
};

__device__ __forceinline__ void quad_sync() {
  // a wave-local ordering point between the quads' LDS writes and reads
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int Ctrl> __device__ __forceinline__ uint32_t quad_dpp(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, Ctrl, 0xF, 0xF, false);
}
template <int Ctrl> __device__ __forceinline__ uint64_t quad_dpp64(uint64_t v) {
  return ((uint64_t)quad_dpp<Ctrl>((uint32_t)(v >> 32)) << 32) | quad_dpp<Ctrl>((uint32_t)v);
}
// across the four lanes of the quad
__device__ __forceinline__ uint64_t quad_max_u64(uint64_t v) {
  uint64_t t = quad_dpp64<0xB1>(v);  // [1,0,3,2]
  v = t > v ? t : v;
  t = quad_dpp64<0x4E>(v);           // [2,3,0,1]
  return t > v ? t : v;
}
__device__ __forceinline__ uint64_t quad_or_u64(uint64_t v) {
  v |= quad_dpp64<0xB1>(v);
  return v | quad_dpp64<0x4E>(v);
}
__device__ __forceinline__ bool quad_all(bool b) {
  uint32_t v = b;
  v &= quad_dpp<0xB1>(v);
  v &= quad_dpp<0x4E>(v);
  return v != 0;
}

// Planes KHI down to KLO of a 256-coefficient block; this lane keeps
// quarter s of each plane word.
template <int IP, int KHI, int KLO>
__device__ __forceinline__ void quad_decode_planes(LaneReader& r, uint32_t kmin, uint32_t& bits, uint32_t& n,
                                                   uint32_t s, PlaneSet<64, IP>& p) {
#pragma unroll
  for (int k = KHI; k >= KLO; --k) {
    uint64_t x = 0;
    if ((uint32_t)k >= kmin && bits) {
      const uint32_t m = n < bits ? n : bits;
      bits -= m;
#pragma unroll
      for (uint32_t q = 0; q < 4; ++q) {
        const uint32_t cnt = m > 64 * q ? (m - 64 * q < 64 ? m - 64 * q : 64u) : 0u;
        if (cnt) {
          const uint64_t v = r.get(cnt);
          if (q == s) x = v;
        }
      }
      while (n < 256 && bits) {
        bits--;
        if (!(r.peek() & 1)) {
          r.skip(1);
          break;
        }
        r.skip(1);
        // zeros up to the next significant coefficient (at most lim of
        // them), then its one (implied at lim)
        uint32_t lim = 255 - n;
        if (lim > bits) lim = bits;
        uint32_t z = 0;
        for (;;) {
          const uint32_t room = lim - z;
          const uint32_t t = ctz64(r.peek());  // 64: the next 64 bits are zeros
          if (t < 64 && t < room) {
            r.skip(t + 1);
            bits -= t + 1;
            z += t;
            break;
          }
          if (room <= 64) {
            r.skip(room);
            bits -= room;
            z += room;
            break;
          }
          r.skip(64);
          bits -= 64;
          z += 64;
        }
        n += z;
        if ((n >> 6) == s) x |= 1ull << (n & 63);
        n++;
      }
    }
    p.lo[k] = (uint32_t)x;
    p.hi[k] = (uint32_t)(x >> 32);
  }
}

template <int IP, int KHI, int KLO>
__device__ __forceinline__ void quad_encode_planes(LaneWriter& w, uint32_t kmin, uint32_t& bits, uint32_t& n,
                                                   const PlaneSet<64, IP>& p) {
#pragma unroll
  for (int k = KHI; k >= KLO; --k) {
    if ((uint32_t)k >= kmin && bits) {
      const uint64_t mine = ((uint64_t)p.hi[k] << 32) | p.lo[k];
      const uint64_t x[4] = {quad_dpp64<kQuadDppBcast[0]>(mine), quad_dpp64<kQuadDppBcast[1]>(mine),
                             quad_dpp64<kQuadDppBcast[2]>(mine), quad_dpp64<kQuadDppBcast[3]>(mine)};
      const uint32_t m = n < bits ? n : bits;
      bits -= m;
#pragma unroll
      for (uint32_t q = 0; q < 4; ++q) {
        const uint32_t cnt = m > 64 * q ? (m - 64 * q < 64 ? m - 64 * q : 64u) : 0u;
        w.put(x[q] & lowmask64(cnt), cnt);
      }
      while (n < 256 && bits) {
        bits--;
        // the next significant coefficient at or above n
        uint32_t pnext = 256;
#pragma unroll
        for (int q = 3; q >= 0; --q) {
          uint64_t v = x[q];
          if (n >= 64u * (q + 1)) v = 0;
          else if (n > 64u * q) v &= ~0ull << (n - 64u * q);
          if (v) pnext = 64u * q + (uint32_t)__builtin_ctzll(v);
        }
        if (pnext >= 256) {
          w.put(0, 1);
          break;
        }
        const uint32_t z = pnext - n;
        w.put(1, 1);
        if (bits <= z) {
          w.zeros(bits);
          bits = 0;
          break;
        }
        bits -= z;
        n += z;
        w.zeros(z);
        if (n < 255) {
          bits--;
          w.put(1, 1);
        }
        n++;
      }
    }
  }
}

// The block's integer coefficients (decode: after the quad's parse; encode:
// before its coding) move through LDS: ex holds the block's 256 values.
template <class Int> __device__ __forceinline__ void quad_ranks_to_slice(const typename std::make_unsigned<Int>::type* u,
                                                                         Int* v, Int* ex, uint32_t s, bool rev) {
  using UInt = typename std::make_unsigned<Int>::type;
  const UInt m = nbmask<UInt>();
  QuadPerm perm;
  perm.load(s);
  quad_sync();  // the quads' stream reads (the stage under ex) are done
#pragma unroll
  for (int r = 0; r < 64; ++r) ex[perm[r]] = (Int)((u[r] ^ m) - m);
  quad_sync();
  // inverse lift along w: lane s takes the 16 lines of x,y,z positions
  // 16 s .. 16 s + 15
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const uint32_t p0 = 16 * s + i;
    Int a = ex[p0], b = ex[p0 + 64], cc = ex[p0 + 128], d = ex[p0 + 192];
    if (rev) lift4<Int, false, true>(a, b, cc, d);
    else lift4<Int, false, false>(a, b, cc, d);
    ex[p0] = a;
    ex[p0 + 64] = b;
    ex[p0 + 128] = cc;
    ex[p0 + 192] = d;
  }
  quad_sync();
#pragma unroll
  for (int i = 0; i < 64; ++i) v[i] = ex[64 * s + i];
  if (rev) lane_xform<3, Int, false, true>(v);
  else lane_xform<3, Int, false, false>(v);
}

template <class Int> __device__ __forceinline__ void quad_slice_to_ranks(Int* v, typename std::make_unsigned<Int>::type* u,
                                                                         Int* ex, uint32_t s, bool rev) {
  using UInt = typename std::make_unsigned<Int>::type;
  const UInt m = nbmask<UInt>();
  if (rev) lane_xform<3, Int, true, true>(v);
  else lane_xform<3, Int, true, false>(v);
#pragma unroll
  for (int i = 0; i < 64; ++i) ex[64 * s + i] = v[i];
  quad_sync();
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const uint32_t p0 = 16 * s + i;
    Int a = ex[p0], b = ex[p0 + 64], cc = ex[p0 + 128], d = ex[p0 + 192];
    if (rev) lift4<Int, true, true>(a, b, cc, d);
    else lift4<Int, true, false>(a, b, cc, d);
    ex[p0] = a;
    ex[p0 + 64] = b;
    ex[p0 + 128] = cc;
    ex[p0 + 192] = d;
  }
  quad_sync();
  QuadPerm perm;
  perm.load(s);
#pragma unroll
  for (int r = 0; r < 64; ++r) u[r] = ((UInt)ex[perm[r]] + m) ^ m;
  quad_sync();  // reads done before the LDS is used again
}

template <class Int>
__device__ __forceinline__ uint32_t quad_decode_int_block(LaneReader& r, int minbits, int maxbits, int maxprec,
                                                          bool rev, uint32_t s, Int* v, Int* ex) {
  using UInt = typename std::make_unsigned<Int>::type;
  constexpr int IP = sizeof(Int) * 8;
  constexpr int pbits = sizeof(Int) == 4 ? 5 : 6;
  int bits = 0;
  uint32_t prec = (uint32_t)maxprec;
  if (rev) {
    prec = (uint32_t)r.get(pbits) + 1;
    bits += pbits;
  }
  PlaneSet<64, IP> p;
  const uint32_t mb = maxbits - bits > 0 ? (uint32_t)(maxbits - bits) : 0u;
  const uint32_t kmin = IP > (int)prec ? IP - prec : 0;
  uint32_t left = mb, n = 0;
  // sixteen planes a loop, so that each loop unrolls (plane words in registers)
  if constexpr (IP == 64) {
    quad_decode_planes<IP, 63, 48>(r, kmin, left, n, s, p);
    quad_decode_planes<IP, 47, 32>(r, kmin, left, n, s, p);
  }
  quad_decode_planes<IP, 31, 16>(r, kmin, left, n, s, p);
  quad_decode_planes<IP, 15, 0>(r, kmin, left, n, s, p);
  bits += (int)(mb - left);
  if (bits < minbits) bits = minbits;
  UInt u[64];
  planes_to_coeffs<64, UInt>(p, u);
  quad_ranks_to_slice<Int>(u, v, ex, s, rev);
  return (uint32_t)bits;
}

template <class Int>
__device__ __forceinline__ uint32_t quad_encode_int_block(LaneWriter& w, int minbits, int maxbits, int maxprec,
                                                          bool rev, uint32_t s, Int* v, Int* ex) {
  using UInt = typename std::make_unsigned<Int>::type;
  constexpr int IP = sizeof(Int) * 8;
  constexpr int pbits = sizeof(Int) == 4 ? 5 : 6;
  UInt u[64];
  quad_slice_to_ranks<Int>(v, u, ex, s, rev);
  int bits = 0;
  uint32_t prec = (uint32_t)maxprec;
  if (rev) {
    UInt mo = 0;
#pragma unroll
    for (int i = 0; i < 64; ++i) mo |= u[i];
    const uint64_t all = quad_or_u64((uint64_t)mo);
    prec = all ? IP - (uint32_t)__builtin_ctzll((unsigned long long)all) : 0;
    if (prec > (uint32_t)maxprec) prec = (uint32_t)maxprec;
    if (prec < 1) prec = 1;
    w.put(prec - 1, pbits);
    bits = pbits;
  }
  PlaneSet<64, IP> p;
  coeffs_to_planes<64, UInt>(u, p);
  const uint32_t mb = maxbits - bits > 0 ? (uint32_t)(maxbits - bits) : 0u;
  const uint32_t kmin = IP > (int)prec ? IP - prec : 0;
  uint32_t left = mb, n = 0;
  if constexpr (IP == 64) {
    quad_encode_planes<IP, 63, 48>(w, kmin, left, n, p);
    quad_encode_planes<IP, 47, 32>(w, kmin, left, n, p);
  }
  quad_encode_planes<IP, 31, 16>(w, kmin, left, n, p);
  quad_encode_planes<IP, 15, 0>(w, kmin, left, n, p);
  bits += (int)(mb - left);
  if (bits < minbits) {
    w.zeros((uint64_t)(minbits - bits));
    bits = minbits;
  }
  return (uint32_t)bits;
}

// One call site of the integer block decode per block (the stage under the
// LDS exchange must be read by every quad before any quad writes it).
template <class S>
__device__ __forceinline__ uint32_t quad_decode_block(LaneReader& r, const ZfpChunk& c, uint32_t s, S* f, void* exv) {
  using Int = typename ZT<S>::Int;
  using U = typename ZT<S>::UInt;
  Int* ex = (Int*)exv;
  Int v[64];
  uint32_t bits = 0;
  bool coded = false, rawbits = false;
  int emax = 0, minb = (int)c.minbits, maxb = (int)c.maxbits, maxp = (int)c.maxprec;
  bool rev = c.minexp < kZfpMinExp;
  if constexpr (ZT<S>::fp) {
    constexpr int ebits = ZT<S>::ebits, ebias = ZT<S>::ebias;
    if (rev) {
      bits = 1;
      if (r.bit()) {
        bits++;
        coded = true;
        if (r.bit()) {
          rawbits = true;
        } else {
          bits += ebits;
          emax = (int)r.get(ebits) - ebias;
        }
      }
    } else if (r.bit()) {
      bits = 1 + ebits;
      emax = (int)r.get(ebits) - ebias;
      const int dp = emax - c.minexp + 2 * (4 + 1);
      maxp = (int)c.maxprec < (dp > 0 ? dp : 0) ? (int)c.maxprec : (dp > 0 ? dp : 0);
      coded = true;
    } else {
      bits = c.minbits > 1 ? c.minbits : 1;
    }
    minb -= (int)bits;
    maxb -= (int)bits;
  } else {
    coded = true;
  }
  if (coded) {
    bits += quad_decode_int_block<Int>(r, minb, maxb, maxp, rev, s, v, ex);
  } else {
#pragma unroll
    for (int j = 0; j < 64; ++j) v[j] = 0;
  }
  if constexpr (ZT<S>::fp) {
    if (rawbits) {
      const U mx = ~(U)0 >> 1;
#pragma unroll
      for (int j = 0; j < 64; ++j) {
        const Int x = v[j] < 0 ? (Int)((U)v[j] ^ mx) : v[j];
        f[j] = fromb<S>((U)x);
      }
    } else if (coded) {
      // reversible: the lowest exponent marks an all-zero block (+0)
      const S sc = pow2<S>(emax - ((int)sizeof(S) * 8 - 2));
      const bool zero = rev && emax == -ZT<S>::ebias;
#pragma unroll
      for (int j = 0; j < 64; ++j) f[j] = zero ? (S)0 : sc * (S)v[j];
    } else {
#pragma unroll
      for (int j = 0; j < 64; ++j) f[j] = 0;
    }
  } else {
#pragma unroll
    for (int j = 0; j < 64; ++j) f[j] = v[j];
  }
  return bits;
}

template <class S>
__device__ __forceinline__ uint32_t quad_encode_block(LaneWriter& w, const ZfpChunk& c, uint32_t s, const S* f,
                                                      void* exv) {
  using Int = typename ZT<S>::Int;
  using U = typename ZT<S>::UInt;
  Int* ex = (Int*)exv;
  Int v[64];
  const bool rev = c.minexp < kZfpMinExp;
  if constexpr (!ZT<S>::fp) {
#pragma unroll
    for (int j = 0; j < 64; ++j) v[j] = f[j];
    return quad_encode_int_block<Int>(w, (int)c.minbits, (int)c.maxbits, (int)c.maxprec, rev, s, v, ex);
  } else {
    constexpr int ebits = ZT<S>::ebits, ebias = ZT<S>::ebias, bias = ZT<S>::ebias, mb = ZT<S>::mbits;
    constexpr U absmask = ~(U)0 >> 1, inf = (U)(2 * bias + 1) << mb;
    // the block exponent: the largest |value| over the quad
    U mxl = 0;
#pragma unroll
    for (int j = 0; j < 64; ++j) {
      U a = fbits(f[j]) & absmask;
      if (a > inf) a = 0;
      mxl = a > mxl ? a : mxl;
    }
    const U mx = (U)quad_max_u64((uint64_t)mxl);
    int emax;
    if (mx == 0) emax = -bias;
    else if ((mx >> mb) == 0) emax = 1 - bias;
    else if ((int)(mx >> mb) == 2 * bias + 1) emax = 0;
    else emax = (int)(mx >> mb) - (bias - 1);
    if (rev) {
      bool same = true;
      if (emax != -ebias) {
        const S sc = pow2<S>((int)sizeof(S) * 8 - 2 - emax), si = pow2<S>(emax - ((int)sizeof(S) * 8 - 2));
#pragma unroll
        for (int j = 0; j < 64; ++j) {
          v[j] = cast_int<S>(sc * f[j]);
          same &= fbits(si * (S)v[j]) == fbits(f[j]);
        }
      } else {
#pragma unroll
        for (int j = 0; j < 64; ++j) {
          v[j] = 0;
          same &= fbits(f[j]) == 0;
        }
      }
      same = quad_all(same);
      uint32_t bits;
      if (same) {
        const uint32_t e = (uint32_t)(emax + ebias);
        if (!e) {
          w.put(0, 1);
          return 1;
        }
        w.put(1, 2);
        w.put(e, ebits);
        bits = 2 + ebits;
      } else {
        const U mxm = ~(U)0 >> 1;
#pragma unroll
        for (int j = 0; j < 64; ++j) {
          const U b = fbits(f[j]);
          v[j] = (Int)b < 0 ? (Int)(b ^ mxm) : (Int)b;
        }
        w.put(3, 2);
        bits = 2;
      }
      return bits + quad_encode_int_block<Int>(w, (int)c.minbits - (int)bits, (int)c.maxbits - (int)bits,
                                               (int)c.maxprec, true, s, v, ex);
    }
    const int dp = emax - c.minexp + 2 * (4 + 1);
    const int maxprec = (int)c.maxprec < (dp > 0 ? dp : 0) ? (int)c.maxprec : (dp > 0 ? dp : 0);
    const uint32_t e = maxprec ? (uint32_t)(emax + ebias) : 0;
    if (e) {
      const uint32_t bits = 1 + ebits;
      w.put(2 * (uint64_t)e + 1, bits);
      const S sc = pow2<S>((int)sizeof(S) * 8 - 2 - emax);
#pragma unroll
      for (int j = 0; j < 64; ++j) v[j] = cast_int<S>(sc * f[j]);
      return bits + quad_encode_int_block<Int>(w, (int)c.minbits - (int)bits, (int)c.maxbits - (int)bits, maxprec,
                                               false, s, v, ex);
    }
    w.put(0, 1);
    if (c.minbits > 1) {
      w.zeros(c.minbits - 1);
      return c.minbits;
    }
    return 1;
  }
}

// w-slice s of a 4D block: gather with the pad rule along every axis
template <class S> __device__ __forceinline__ void quad_gather(const ZfpChunk& c, const uint32_t* b, uint32_t s,
                                                               S* vals) {
  const uint32_t e3 = c.n[3] - 4 * b[3];
  const uint64_t sw = 4ull * b[3] + pad_src(s, e3 < 4 ? e3 : 4);
  ZfpChunk cs = c;  // the slice as a 3D field of the same row layout
  cs.src = c.src + sw * (uint64_t)c.n[2] * c.n[1] * c.n[0] * (uint64_t)dtype_bytes_dev(c.dtype);
  lane_gather<3, S>(cs, b, vals);
}
template <class S> __device__ __forceinline__ void quad_scatter(const ZfpChunk& c, const uint32_t* b, uint32_t s,
                                                                const S* vals, uint8_t* brick) {
  const uint32_t e3 = c.n[3] - 4 * b[3];
  if (s >= (e3 < 4 ? e3 : 4)) return;
  uint32_t ext[3];
  block_extents<3>(c, b, ext);
  ZfpChunk cs = c;
  if (brick) {
    brick += (uint64_t)s * ext[2] * ext[1] * ext[0] * dtype_bytes_dev(c.dtype);
  } else {
    cs.dst = c.dst + (4ull * b[3] + s) * (uint64_t)c.n[2] * c.n[1] * c.n[0] * (uint64_t)dtype_bytes_dev(c.dtype);
  }
  lane_scatter<3, S>(cs, b, vals, brick);
}

template <class S>
__global__ __launch_bounds__(kLaneThreads) void zfp_quad_encode_kernel(const ZfpChunk* __restrict__ chunks,
                                                                      uint32_t nchunks, uint64_t total) {
  const uint64_t g = lane_block_id() / 4;  // block of this quad
  const uint32_t s = threadIdx.x & 3;
  if (g >= total) return;
  const ZfpChunk& c = chunks[find_chunk(chunks, nchunks, g)];
  if (c.fixed ? !c.dst : !c.slots) return;  // refused chunk
  const uint64_t k = g - c.block_begin;
  uint32_t b[4];
  block_coords(c, k, b);
  S vals[64];
  quad_gather<S>(c, b, s, vals);
  LaneWriter w;  // the quad's lanes write the same words (OR / store of equal values)
  if (c.fixed) w.init((uint64_t*)c.dst, c.header_bits + k * c.maxbits, true);
  else w.init(c.slots + g * c.slot_words, 0, false);
  quad_encode_block<S>(w, c, s, vals, (uint8_t*)zstage + (threadIdx.x / 4) * 256 * sizeof(typename ZT<S>::Int));
  w.finish();
  if (!c.fixed && s == 0) c.offs[g] = w.count;
}

template <class S>
__global__ __launch_bounds__(kLaneThreads) void zfp_quad_decode_kernel(const ZfpChunk* __restrict__ chunks,
                                                                      uint32_t nchunks, uint64_t total,
                                                                      const ZfpBrick* __restrict__ bricks,
                                                                      uint32_t stage_words) {
  const uint64_t g = lane_block_id() / 4;
  const uint32_t s = threadIdx.x & 3;
  const bool valid = g < total;
  uint32_t ci = 0;
  uint64_t k = 0;
  uint8_t* brick = nullptr;
  if (valid) {
    if (bricks) {
      ci = bricks[g].chunk;
      k = bricks[g].block;
      brick = bricks[g].out;
    } else {
      ci = find_chunk(chunks, nchunks, g);
      k = g - chunks[ci].block_begin;
    }
  }
  const ZfpChunk& c = chunks[ci];
  const bool indexed = !c.fixed && c.index;
  uint64_t start = 0, end = 0;
  if (valid) {
    if (c.fixed) {
      start = c.header_bits + k * c.maxbits;
      end = start + c.maxbits;
    } else if (indexed) {
      start = c.header_bits + c.index[k];
      end = c.header_bits + c.index[k + 1];
    } else {
      start = c.offs[c.block_begin + k];
      end = k + 1 < c.nblocks ? c.offs[c.block_begin + k + 1] : 8 * c.src_len;
    }
  }
  const uint32_t ci0 = (uint32_t)__shfl((int)ci, 0);
  const bool one = __ballot(valid && ci != ci0) == 0;
  const uint64_t lo = wave_min_u64(valid ? start : ~0ull) >> 6;
  const uint64_t hi = ((wave_max_u64(valid ? end : 0) + 63) >> 6) + 2;
  const bool staged = one && hi > lo && hi - lo <= stage_words;
  LaneReader r;
  r.base = c.src;
  r.nbytes = c.src_len;
  r.s0 = 0;
  r.sn = 0;
  if (staged) {
    const ZfpChunk& c0 = chunks[ci0];
    for (uint64_t i = threadIdx.x; i < hi - lo; i += kLaneThreads) zstage[i] = global_word(c0.src, c0.src_len, lo + i);
    r.s0 = lo;
    r.sn = hi - lo;
  }
  __syncthreads();
  if (!valid) return;
  r.seek(start);
  S vals[64];
  const uint32_t used =
      quad_decode_block<S>(r, c, s, vals, (uint8_t*)zstage + (threadIdx.x / 4) * 256 * sizeof(typename ZT<S>::Int));
  if (indexed && s == 0) {
    const uint32_t ilo = c.index[k], ihi = c.index[k + 1];
    if (ihi < ilo || ihi - ilo != used || (k == 0 && ilo != 0) ||
        (k + 1 == c.nblocks && c.header_bits + (uint64_t)ihi > 8 * c.src_len))
      atomicOr(c.bad, 1u);
  }
  uint32_t b[4];
  block_coords(c, k, b);
  quad_scatter<S>(c, b, s, vals, brick);
}

inline dim3 lane_grid(uint64_t total) {
  const uint64_t wgs = ceil_div(total, kLaneThreads);
  const uint32_t gx = (uint32_t)(wgs < 65536 ? wgs : 65536);
  return dim3(gx, (uint32_t)ceil_div(wgs, gx));
}

}  // namespace

// Launchers of one scalar type's lane coders (instantiated per type in
// zfp_lane_<type>.hip, so that the four types compile in parallel).
template <class S> void zfp_lane_encode_launch(uint32_t dims, const ZfpChunk* d, uint32_t n, uint64_t total,
                                               hipStream_t s) {
  switch (dims) {
    case 1: hipLaunchKernelGGL((zfp_lane_encode_kernel<1, S>), lane_grid(total), dim3(kLaneThreads), 0, s, d, n, total); break;
    case 2: hipLaunchKernelGGL((zfp_lane_encode_kernel<2, S>), lane_grid(total), dim3(kLaneThreads), 0, s, d, n, total); break;
    case 3: hipLaunchKernelGGL((zfp_lane_encode_kernel<3, S>), lane_grid(total), dim3(kLaneThreads), 0, s, d, n, total); break;
    case 4: hipLaunchKernelGGL((zfp_quad_encode_kernel<S>), lane_grid(4 * total), dim3(kLaneThreads), 16 * 256 * sizeof(typename ZT<S>::Int), s, d, n, total); break;
    default: break;
  }
}
template <class S> void zfp_lane_decode_launch(uint32_t dims, const ZfpChunk* d, uint32_t n, uint64_t total,
                                               const ZfpBrick* br, uint32_t words, hipStream_t s) {
  const size_t lds = (size_t)words * 8;
  switch (dims) {
    case 1: hipLaunchKernelGGL((zfp_lane_decode_kernel<1, S>), lane_grid(total), dim3(kLaneThreads), lds, s, d, n, total, br, words); break;
    case 2: hipLaunchKernelGGL((zfp_lane_decode_kernel<2, S>), lane_grid(total), dim3(kLaneThreads), lds, s, d, n, total, br, words); break;
    case 3: hipLaunchKernelGGL((zfp_lane_decode_kernel<3, S>), lane_grid(total), dim3(kLaneThreads), lds, s, d, n, total, br, words); break;
    case 4: {
      const size_t ex = 16 * 256 * sizeof(typename ZT<S>::Int);  // the exchange overlays the stage
      hipLaunchKernelGGL((zfp_quad_decode_kernel<S>), lane_grid(4 * total), dim3(kLaneThreads), lds > ex ? lds : ex, s, d, n, total, br, words);
      break;
    }
    default: break;
  }
}

}  // namespace ndic
