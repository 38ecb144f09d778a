// ZFP block codec on the device for 4D fields, one wavefront per 4^4 block,
// and the serial walk of variable-size streams without a block index (every
// dimensionality).
//
// Layout of a block across the wave: lane l holds coefficient l (x = l & 3,
// y = (l >> 2) & 3, z = l >> 4) for d <= 3; for d = 4 each lane holds the four
// w-slices of its (x, y, z) position in registers v[0..3].  The separable
// decorrelating transform runs along x, y, z with lane shuffles (each lane
// rebuilds its 4-line and keeps its own output) and along w in registers.
// After the negabinary reorder, lane r holds the coefficient of rank r
// (ranks 64q + l in register q for d = 4); one ballot then yields a whole
// bit plane, so the embedded group-test coder runs wave-uniform on 64-bit
// plane words (ctz finds the next significant coefficient instead of a
// per-position loop) and every lane deposits its own bit when decoding.
//
// The algorithm and the stream layout are the published ZFP ones (Lindstrom,
// IEEE TVCG 2014; LLNL ZFP stream format, codec 5), which the reference
// reaches through zfp-rs 0.1.1 (crates/ndic-zfp/src/lib.rs:249-376): block
// order x fastest, partial blocks padded by the pad rule, block-floating-
// point cast at 2^(p-2-emax), integer lifting, negabinary, embedded coder,
// reversible variant (Lorenzo lifting, precision prefix, float reinterpret).
#include "zfp.hpp"
#include "zfp_dev.hpp"

#include <mutex>

#include <hipcub/hipcub.hpp>

namespace ndic {

namespace {

using namespace zfpdev;

// coefficient order per dimensionality (rank -> index) and its inverse
__constant__ uint8_t c_perm[4][256];
__constant__ uint8_t c_iperm[4][256];

__device__ __forceinline__ uint32_t lane() { return __lane_id(); }

template <class Int, bool Fwd, bool Rev> __device__ __forceinline__ Int lift_lanes(Int v, uint32_t axis) {
  uint32_t l = lane(), sh = 2 * axis, s = 1u << sh;
  uint32_t pos = (l >> sh) & 3, base = l - pos * s;
  Int x = __shfl(v, (int)base), y = __shfl(v, (int)(base + s)), z = __shfl(v, (int)(base + 2 * s)),
      w = __shfl(v, (int)(base + 3 * s));
  lift4<Int, Fwd, Rev>(x, y, z, w);
  return pos == 0 ? x : pos == 1 ? y : pos == 2 ? z : w;
}

template <int D, class Int, bool Fwd, bool Rev> __device__ __forceinline__ void xform(Int* v) {
  constexpr int R = D == 4 ? 4 : 1;
  if (Fwd) {
#pragma unroll
    for (int a = 0; a < (D < 3 ? D : 3); ++a)
#pragma unroll
      for (int j = 0; j < R; ++j) v[j] = lift_lanes<Int, Fwd, Rev>(v[j], a);
    if (D == 4) lift4<Int, Fwd, Rev>(v[0], v[1], v[2], v[R - 1]);
  } else {
    if (D == 4) lift4<Int, Fwd, Rev>(v[0], v[1], v[2], v[R - 1]);
#pragma unroll
    for (int a = (D < 3 ? D : 3) - 1; a >= 0; --a)
#pragma unroll
      for (int j = 0; j < R; ++j) v[j] = lift_lanes<Int, Fwd, Rev>(v[j], a);
  }
}

// coefficients -> ranked negabinary (lane r, register q holds rank 64q + r)
template <int D, class Int, class UInt> __device__ __forceinline__ void to_ranks(const Int* v, UInt* u) {
  constexpr uint32_t S = 1u << (2 * D);
  const UInt m = nbmask<UInt>();
  uint32_t l = lane();
  if constexpr (D < 4) {
    uint32_t c = l < S ? c_perm[D - 1][l] : 0;
    Int t = __shfl(v[0], (int)c);
    u[0] = l < S ? (((UInt)t + m) ^ m) : 0;
  } else {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      uint32_t c = c_perm[3][l + 64 * q], sl = c & 63, sr = c >> 6;
      Int t0 = __shfl(v[0], (int)sl), t1 = __shfl(v[1], (int)sl), t2 = __shfl(v[2], (int)sl),
          t3 = __shfl(v[3], (int)sl);
      Int t = sr == 0 ? t0 : sr == 1 ? t1 : sr == 2 ? t2 : t3;
      u[q] = ((UInt)t + m) ^ m;
    }
  }
}

template <int D, class Int, class UInt> __device__ __forceinline__ void from_ranks(const UInt* u, Int* v) {
  constexpr uint32_t S = 1u << (2 * D);
  const UInt m = nbmask<UInt>();
  uint32_t l = lane();
  if constexpr (D < 4) {
    uint32_t r = l < S ? c_iperm[D - 1][l] : 0;
    UInt t = __shfl(u[0], (int)r);
    v[0] = (Int)((t ^ m) - m);
  } else {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      uint32_t r = c_iperm[3][l + 64 * j], sl = r & 63, sq = r >> 6;
      UInt t0 = __shfl(u[0], (int)sl), t1 = __shfl(u[1], (int)sl), t2 = __shfl(u[2], (int)sl),
           t3 = __shfl(u[3], (int)sl);
      UInt t = sq == 0 ? t0 : sq == 1 ? t1 : sq == 2 ? t2 : t3;
      v[j] = (Int)((t ^ m) - m);
    }
  }
}

// ---------------------------------------------------------------------------
// wave-uniform bit I/O (LSB-first 64-bit words)
// ---------------------------------------------------------------------------
struct BitWriter {
  uint64_t* out;
  uint64_t word;
  uint64_t acc;
  uint32_t fill;
  bool shared;  // edge words may be shared with neighbouring blocks (OR them in)
  bool first;
  uint64_t count;

  __device__ void init(uint64_t* base, uint64_t bitpos, bool share) {
    out = base;
    word = bitpos >> 6;
    fill = (uint32_t)(bitpos & 63);
    acc = 0;
    shared = share;
    first = true;
    count = 0;
  }
  __device__ void store_word() {
    if (lane() == 0) {
      if (shared && first) atomicOr((unsigned long long*)&out[word], (unsigned long long)acc);
      else out[word] = acc;
    }
    first = false;
    ++word;
    acc = 0;
  }
  // appends the low n bits of v (v has no bits at or above n)
  __device__ void put(uint64_t v, uint32_t n) {
    if (n == 0) return;
    count += n;
    acc |= v << fill;
    uint32_t f = fill + n;
    if (f >= 64) {
      uint32_t used = 64 - fill;
      store_word();
      acc = used < 64 ? v >> used : 0;
      fill = f - 64;
    } else {
      fill = f;
    }
  }
  __device__ void zeros(uint64_t n) {
    while (n) {
      uint32_t k = n < 64 ? (uint32_t)n : 64;
      put(0, k);
      n -= k;
    }
  }
  __device__ void finish() {
    if (fill && lane() == 0) {
      if (shared) atomicOr((unsigned long long*)&out[word], (unsigned long long)acc);
      else out[word] = acc;
    }
  }
};

// instead of once per word.
struct WindowSrc {
  GlobalSrc g;
  uint64_t first;      // first word of the window (wave-uniform)
  uint64_t cur, nxt;   // lane l: words first + l and first + 64 + l
  __device__ void init(const uint8_t* base, uint64_t nbytes, uint64_t w) {
    g.base = base;
    g.nbytes = nbytes;
    first = w;
    cur = g.load(w + lane());
    nxt = g.load(w + 64 + lane());
  }
  __device__ uint64_t load(uint64_t w) {
    if (w - first >= 64) {
      if (w - first < 128) {
        cur = nxt;
        first += 64;
      } else {
        first = w;
        cur = g.load(w + lane());
      }
      nxt = g.load(first + 64 + lane());
    }
    const int i = (int)(w - first);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)cur, i);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(cur >> 32), i);
    return ((uint64_t)hi << 32) | lo;
  }
};

template <class Src> struct BitReaderT {
  Src src;
  uint64_t next;
  uint64_t cur;
  uint32_t left;

  __device__ uint64_t load(uint64_t w) { return src.load(w); }
  __device__ void seek(uint64_t bit) {
    next = bit >> 6;
    uint32_t o = (uint32_t)(bit & 63);
    cur = load(next++) >> o;
    left = 64 - o;
  }
  __device__ uint64_t pos() const { return next * 64 - left; }
  __device__ uint64_t get(uint32_t n) {
    if (n == 0) return 0;
    if (n <= left) {
      uint64_t v = n == 64 ? cur : cur & ((1ull << n) - 1);
      cur = n == 64 ? 0 : cur >> n;
      left -= n;
      return v;
    }
    uint64_t v = cur;
    uint32_t got = left;
    uint64_t w = load(next++);
    uint32_t rem = n - got;
    v |= w << got;
    cur = rem == 64 ? 0 : w >> rem;
    left = 64 - rem;
    return n == 64 ? v : v & ((1ull << n) - 1);
  }
  __device__ uint32_t bit() { return (uint32_t)get(1); }
  // consumes up to lim consecutive zero bits; stops before a one bit
  __device__ uint32_t zeros(uint32_t lim) {
    uint32_t z = 0;
    while (z < lim) {
      if (left == 0) {
        cur = load(next++);
        left = 64;
      }
      uint32_t t = cur ? (uint32_t)__builtin_ctzll(cur) : 64;
      if (t > left) t = left;
      uint32_t take = t < lim - z ? t : lim - z;
      cur = take == 64 ? 0 : cur >> take;
      left -= take;
      z += take;
      if (take < t || left > 0) break;
    }
    return z;
  }
  __device__ void skip(uint64_t n) { seek(pos() + n); }
};
using BitReader = BitReaderT<GlobalSrc>;

// ---------------------------------------------------------------------------
// embedded bit-plane coder over ranked coefficients
// ---------------------------------------------------------------------------
template <int R> __device__ __forceinline__ uint32_t next_one(const uint64_t* x, uint32_t n, uint32_t S) {
#pragma unroll
  for (int q = 0; q < R; ++q) {
    if (64u * q + 63 < n) continue;
    uint64_t v = x[q];
    if (64u * q < n) v &= ~0ull << (n - 64u * q);
    if (v) return 64u * q + (uint32_t)__builtin_ctzll(v);
  }
  return S;
}

template <int D, class UInt>
__device__ uint32_t encode_ints(BitWriter& w, uint32_t maxbits, uint32_t maxprec, const UInt* u) {
  constexpr int R = D == 4 ? 4 : 1;
  constexpr uint32_t S = 1u << (2 * D);
  constexpr uint32_t intprec = sizeof(UInt) * 8;
  uint32_t kmin = intprec > maxprec ? intprec - maxprec : 0;
  uint32_t bits = maxbits, n = 0;
  for (uint32_t k = intprec; bits && k-- > kmin;) {
    uint64_t x[R];
#pragma unroll
    for (int q = 0; q < R; ++q) x[q] = __ballot((int)((u[q] >> k) & 1));
    uint32_t m = n < bits ? n : bits;
    bits -= m;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      if (m > 64u * q) {
        uint32_t c = m - 64u * q;
        if (c >= 64) w.put(x[q], 64);
        else w.put(x[q] & ((1ull << c) - 1), c);
      }
    }
    while (n < S && bits) {
      bits--;
      uint32_t p = next_one<R>(x, n, S);
      if (p >= S) {
        w.put(0, 1);
        break;
      }
      w.put(1, 1);
      uint32_t z = p - n;
      if (bits <= z) {
        w.zeros(bits);
        bits = 0;
        break;
      }
      w.zeros(z);
      bits -= z;
      n = p;
      if (n < S - 1) {
        bits--;
        w.put(1, 1);
      }
      n++;
    }
  }
  return maxbits - bits;
}

template <int D, class UInt, bool Parse, class Rd>
__device__ uint32_t decode_ints(Rd& r, uint32_t maxbits, uint32_t maxprec, UInt* u) {
  constexpr int R = D == 4 ? 4 : 1;
  constexpr uint32_t S = 1u << (2 * D);
  constexpr uint32_t intprec = sizeof(UInt) * 8;
  uint32_t kmin = intprec > maxprec ? intprec - maxprec : 0;
  uint32_t bits = maxbits, n = 0;
#pragma unroll
  for (int q = 0; q < R; ++q) u[q] = 0;
  uint32_t l = lane();
  for (uint32_t k = intprec; bits && k-- > kmin;) {
    if (Parse && n == S) {
      // every coefficient is significant: each plane left (k .. kmin) is S
      // verbatim bits and nothing else, up to the bit budget
      const uint64_t need = (uint64_t)(k + 1 - kmin) * S;
      const uint32_t take = need < bits ? (uint32_t)need : bits;
      r.skip(take);
      bits -= take;
      break;
    }
    uint64_t x[R];
#pragma unroll
    for (int q = 0; q < R; ++q) x[q] = 0;
    uint32_t m = n < bits ? n : bits;
    bits -= m;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      if (m > 64u * q) {
        uint32_t c = m - 64u * q;
        x[q] = r.get(c >= 64 ? 64 : c);
      }
    }
    while (n < S && bits) {
      bits--;
      if (!r.bit()) break;
      uint32_t lim = S - 1 - n;
      if (lim > bits) lim = bits;
      uint32_t z = r.zeros(lim);
      if (z < lim) {
        (void)r.get(1);
        bits -= z + 1;
      } else {
        bits -= z;
      }
      n += z;
      if (!Parse) {
#pragma unroll
        for (int q = 0; q < R; ++q)
          if ((n >> 6) == (uint32_t)q) x[q] |= 1ull << (n & 63);
      }
      n++;
    }
    if (!Parse) {
#pragma unroll
      for (int q = 0; q < R; ++q) u[q] |= (UInt)((x[q] >> l) & 1) << k;
    }
  }
  return maxbits - bits;
}

// ---------------------------------------------------------------------------
// integer blocks
// ---------------------------------------------------------------------------
template <int D, class Int>
__device__ uint32_t encode_int_block(BitWriter& w, int minbits, int maxbits, int maxprec, Int* v) {
  using UInt = typename std::make_unsigned<Int>::type;
  constexpr int R = D == 4 ? 4 : 1;
  xform<D, Int, true, false>(v);
  UInt u[R];
  to_ranks<D, Int, UInt>(v, u);
  int bits = (int)encode_ints<D, UInt>(w, maxbits > 0 ? (uint32_t)maxbits : 0u, (uint32_t)maxprec, u);
  if (bits < minbits) {
    w.zeros((uint64_t)(minbits - bits));
    bits = minbits;
  }
  return (uint32_t)bits;
}

template <class UInt> __device__ __forceinline__ UInt wave_or(UInt v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v |= __shfl_xor(v, o);
  return v;
}

template <int D, class Int>
__device__ uint32_t rev_encode_int_block(BitWriter& w, int minbits, int maxbits, int maxprec, Int* v) {
  using UInt = typename std::make_unsigned<Int>::type;
  constexpr int R = D == 4 ? 4 : 1;
  constexpr int pbits = sizeof(Int) == 4 ? 5 : 6;
  constexpr uint32_t intprec = sizeof(Int) * 8;
  xform<D, Int, true, true>(v);
  UInt u[R];
  to_ranks<D, Int, UInt>(v, u);
  UInt m = 0;
#pragma unroll
  for (int q = 0; q < R; ++q) m |= u[q];
  m = wave_or(m);
  uint32_t prec = m ? intprec - (uint32_t)__builtin_ctzll((unsigned long long)m) : 0;
  if (prec > (uint32_t)maxprec) prec = (uint32_t)maxprec;
  if (prec < 1) prec = 1;
  w.put(prec - 1, pbits);
  int bits = pbits;
  bits += (int)encode_ints<D, UInt>(w, maxbits - bits > 0 ? (uint32_t)(maxbits - bits) : 0u, prec, u);
  if (bits < minbits) {
    w.zeros((uint64_t)(minbits - bits));
    bits = minbits;
  }
  return (uint32_t)bits;
}

template <int D, class Int, bool Parse, class Rd>
__device__ uint32_t decode_int_block(Rd& r, int minbits, int maxbits, int maxprec, bool rev, Int* v) {
  using UInt = typename std::make_unsigned<Int>::type;
  constexpr int R = D == 4 ? 4 : 1;
  constexpr int pbits = sizeof(Int) == 4 ? 5 : 6;
  int bits = 0;
  uint32_t prec = (uint32_t)maxprec;
  if (rev) {
    prec = (uint32_t)r.get(pbits) + 1;
    bits += pbits;
  }
  UInt u[R];
  bits += (int)decode_ints<D, UInt, Parse>(r, maxbits - bits > 0 ? (uint32_t)(maxbits - bits) : 0u, prec, u);
  if (bits < minbits) {
    r.skip((uint64_t)(minbits - bits));
    bits = minbits;
  }
  if (!Parse) {
    from_ranks<D, Int, UInt>(u, v);
    if (rev) xform<D, Int, false, true>(v);
    else xform<D, Int, false, false>(v);
  }
  return (uint32_t)bits;
}

// ---------------------------------------------------------------------------
// floating-point blocks
// ---------------------------------------------------------------------------

template <class UInt> __device__ __forceinline__ UInt wave_max(UInt v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    UInt t = __shfl_xor(v, o);
    v = t > v ? t : v;
  }
  return v;
}

// x = m 2^e, 0.5 <= m < 1, of the largest |value| (NaN ignored, infinity -> 0)
template <int D, class S> __device__ __forceinline__ int block_exponent(const S* f) {
  using U = typename ZT<S>::UInt;
  constexpr int R = D == 4 ? 4 : 1;
  constexpr int bias = ZT<S>::ebias, mb = ZT<S>::mbits;
  constexpr U absmask = ~(U)0 >> 1, inf = (U)(2 * bias + 1) << mb;
  U mx = 0;
#pragma unroll
  for (int j = 0; j < R; ++j) {
    U a = fbits(f[j]) & absmask;
    if (a > inf) a = 0;
    mx = a > mx ? a : mx;
  }
  mx = wave_max(mx);
  if (mx == 0) return -bias;
  int be = (int)(mx >> mb);
  if (be == 0) return 1 - bias;
  if (be == 2 * bias + 1) return 0;
  return be - (bias - 1);
}


template <int D, class S> __device__ __forceinline__ void fwd_cast(const S* f, typename ZT<S>::Int* v, int emax) {
  constexpr int R = D == 4 ? 4 : 1;
  S s = pow2<S>((int)sizeof(S) * 8 - 2 - emax);
#pragma unroll
  for (int j = 0; j < R; ++j) v[j] = cast_int<S>(s * f[j]);
}
template <int D, class S> __device__ __forceinline__ void inv_cast(const typename ZT<S>::Int* v, S* f, int emax) {
  constexpr int R = D == 4 ? 4 : 1;
  S s = pow2<S>(emax - ((int)sizeof(S) * 8 - 2));
#pragma unroll
  for (int j = 0; j < R; ++j) f[j] = s * (S)v[j];
}

template <int D, class S> __device__ uint32_t encode_float_block(BitWriter& w, const ZfpChunk& c, const S* f) {
  using Int = typename ZT<S>::Int;
  using U = typename ZT<S>::UInt;
  constexpr int R = D == 4 ? 4 : 1;
  constexpr uint32_t SZ = 1u << (2 * D);
  constexpr int ebits = ZT<S>::ebits, ebias = ZT<S>::ebias;
  Int v[R];
  int emax = block_exponent<D, S>(f);
  if (c.minexp < kZfpMinExp) {  // reversible
    uint32_t bits = 0;
    if (emax != -ebias) fwd_cast<D, S>(f, v, emax);
    else
#pragma unroll
      for (int j = 0; j < R; ++j) v[j] = 0;
    S g[R];
    if (emax != -ebias) inv_cast<D, S>(v, g, emax);
    else
#pragma unroll
      for (int j = 0; j < R; ++j) g[j] = 0;
    bool bad = false;
#pragma unroll
    for (int j = 0; j < R; ++j) bad |= (D == 4 || lane() < SZ) && fbits(g[j]) != fbits(f[j]);
    bool same = __ballot(bad) == 0;
    if (same) {
      uint32_t e = (uint32_t)(emax + ebias);
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
      for (int j = 0; j < R; ++j) {
        U b = fbits(f[j]);
        v[j] = (Int)b < 0 ? (Int)(b ^ mx) : (Int)b;
      }
      w.put(3, 2);
      bits = 2;
    }
    return bits + rev_encode_int_block<D, Int>(w, (int)c.minbits - (int)bits, (int)c.maxbits - (int)bits,
                                                (int)c.maxprec, v);
  }
  int dp = emax - c.minexp + 2 * (D + 1);
  int maxprec = (int)c.maxprec < (dp > 0 ? dp : 0) ? (int)c.maxprec : (dp > 0 ? dp : 0);
  uint32_t e = maxprec ? (uint32_t)(emax + ebias) : 0;
  if (e) {
    uint32_t bits = 1 + ebits;
    w.put(2 * (uint64_t)e + 1, bits);
    fwd_cast<D, S>(f, v, emax);
    return bits + encode_int_block<D, Int>(w, (int)c.minbits - (int)bits, (int)c.maxbits - (int)bits, maxprec, v);
  }
  w.put(0, 1);
  if (c.minbits > 1) {
    w.zeros(c.minbits - 1);
    return c.minbits;
  }
  return 1;
}

template <int D, class S, bool Parse, class Rd> __device__ uint32_t decode_float_block(Rd& r, const ZfpChunk& c, S* f) {
  using Int = typename ZT<S>::Int;
  using U = typename ZT<S>::UInt;
  constexpr int R = D == 4 ? 4 : 1;
  constexpr int ebits = ZT<S>::ebits, ebias = ZT<S>::ebias;
  Int v[R];
  if (c.minexp < kZfpMinExp) {  // reversible
    uint32_t bits = 1;
    if (r.bit()) {
      bits++;
      if (r.bit()) {
        bits += decode_int_block<D, Int, Parse>(r, (int)c.minbits - (int)bits, (int)c.maxbits - (int)bits,
                                                (int)c.maxprec, true, v);
        if (!Parse) {
          const U mx = ~(U)0 >> 1;
#pragma unroll
          for (int j = 0; j < R; ++j) {
            Int x = v[j] < 0 ? (Int)((U)v[j] ^ mx) : v[j];
            f[j] = fromb<S>((U)x);
          }
        }
      } else {
        bits += ebits;
        int emax = (int)r.get(ebits) - ebias;
        bits += decode_int_block<D, Int, Parse>(r, (int)c.minbits - (int)bits, (int)c.maxbits - (int)bits,
                                                (int)c.maxprec, true, v);
        if (!Parse) inv_cast<D, S>(v, f, emax);
      }
      return bits;
    }
    if (!Parse)
#pragma unroll
      for (int j = 0; j < R; ++j) f[j] = 0;
    return bits;
  }
  if (r.bit()) {
    uint32_t bits = 1 + ebits;
    int emax = (int)r.get(ebits) - ebias;
    int dp = emax - c.minexp + 2 * (D + 1);
    int maxprec = (int)c.maxprec < (dp > 0 ? dp : 0) ? (int)c.maxprec : (dp > 0 ? dp : 0);
    bits += decode_int_block<D, Int, Parse>(r, (int)c.minbits - (int)bits, (int)c.maxbits - (int)bits, maxprec,
                                            false, v);
    if (!Parse) inv_cast<D, S>(v, f, emax);
    return bits;
  }
  if (!Parse)
#pragma unroll
    for (int j = 0; j < R; ++j) f[j] = 0;
  if (c.minbits > 1) {
    r.skip(c.minbits - 1);
    return c.minbits;
  }
  return 1;
}

template <int D, class S> __device__ uint32_t encode_block(BitWriter& w, const ZfpChunk& c, S* vals) {
  if constexpr (ZT<S>::fp) {
    return encode_float_block<D, S>(w, c, vals);
  } else {
    if (c.minexp < kZfpMinExp)
      return rev_encode_int_block<D, S>(w, (int)c.minbits, (int)c.maxbits, (int)c.maxprec, vals);
    return encode_int_block<D, S>(w, (int)c.minbits, (int)c.maxbits, (int)c.maxprec, vals);
  }
}

template <int D, class S, bool Parse, class Rd> __device__ uint32_t decode_block(Rd& r, const ZfpChunk& c, S* vals) {
  if constexpr (ZT<S>::fp) {
    return decode_float_block<D, S, Parse>(r, c, vals);
  } else {
    return decode_int_block<D, S, Parse>(r, (int)c.minbits, (int)c.maxbits, (int)c.maxprec,
                                         c.minexp < kZfpMinExp, vals);
  }
}


template <int D, class S> __device__ void gather(const ZfpChunk& c, const uint32_t* b, S* vals) {
  constexpr int R = D == 4 ? 4 : 1;
  constexpr uint32_t SZ = 1u << (2 * D);
  uint32_t l = lane();
  uint32_t cc[3] = {l & 3, (l >> 2) & 3, (l >> 4) & 3};
  uint32_t ext[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    uint32_t e = a < D ? c.n[a] - 4 * b[a] : 1;
    ext[a] = e < 4 ? e : 4;
  }
  uint64_t sx = 4ull * b[0] + pad_src(cc[0], ext[0]);
  uint64_t sy = D > 1 ? 4ull * b[1] + pad_src(cc[1], ext[1]) : 0;
  uint64_t sz = D > 2 ? 4ull * b[2] + pad_src(cc[2], ext[2]) : 0;
#pragma unroll
  for (int j = 0; j < R; ++j) {
    if (D < 4 && l >= SZ) {
      vals[j] = 0;
      continue;
    }
    uint64_t sw = D > 3 ? 4ull * b[3] + pad_src((uint32_t)j, ext[3]) : 0;
    uint64_t idx = ((sw * c.n[2] + sz) * c.n[1] + sy) * c.n[0] + sx;
    vals[j] = load_elem<S>(c.src, idx, c.dtype);
  }
}

// Writes the block's samples that fall inside the field (or, for a brick,
// into its own clipped C-ordered array).
template <int D, class S> __device__ void scatter(const ZfpChunk& c, const uint32_t* b, const S* vals, uint8_t* brick) {
  constexpr int R = D == 4 ? 4 : 1;
  constexpr uint32_t SZ = 1u << (2 * D);
  uint32_t l = lane();
  if (D < 4 && l >= SZ) return;
  uint32_t cc[3] = {l & 3, (l >> 2) & 3, (l >> 4) & 3};
  uint32_t ext[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    uint32_t e = a < D ? c.n[a] - 4 * b[a] : 1;
    ext[a] = e < 4 ? e : 4;
  }
  for (int a = 0; a < (D < 3 ? D : 3); ++a)
    if (cc[a] >= ext[a]) return;
#pragma unroll
  for (int j = 0; j < R; ++j) {
    if (D == 4 && (uint32_t)j >= ext[3]) continue;
    uint64_t idx;
    if (brick) {
      uint64_t ey = D > 1 ? ext[1] : 1, ez = D > 2 ? ext[2] : 1;
      idx = ((uint64_t)(D > 3 ? j : 0) * ez + (D > 2 ? cc[2] : 0)) * ey * ext[0] + (D > 1 ? cc[1] : 0) * ext[0] +
            cc[0];
      store_elem<S>(brick, idx, c.dtype, vals[j]);
    } else {
      uint64_t x = 4ull * b[0] + cc[0], y = D > 1 ? 4ull * b[1] + cc[1] : 0, z = D > 2 ? 4ull * b[2] + cc[2] : 0,
               w = D > 3 ? 4ull * b[3] + j : 0;
      idx = ((w * c.n[2] + z) * c.n[1] + y) * c.n[0] + x;
      store_elem<S>(c.dst, idx, c.dtype, vals[j]);
    }
  }
}


// ---------------------------------------------------------------------------
// kernels: four wavefronts (blocks) per workgroup
// ---------------------------------------------------------------------------
template <int D, class S>
__global__ __launch_bounds__(256) void zfp_encode_kernel(const ZfpChunk* __restrict__ chunks, uint32_t nchunks,
                                                         uint64_t total) {
  constexpr int R = D == 4 ? 4 : 1;
  uint64_t g = xcd_item() * 4 + (threadIdx.x >> 6);  // neighbouring blocks share input lines: one L2
  if (g >= total) return;
  const ZfpChunk& c = chunks[find_chunk(chunks, nchunks, g)];
  if (c.fixed ? !c.dst : !c.slots) return;  // refused chunk
  uint64_t k = g - c.block_begin;
  uint32_t b[4];
  block_coords(c, k, b);
  S vals[R];
  gather<D, S>(c, b, vals);
  BitWriter w;
  if (c.fixed) w.init((uint64_t*)c.dst, c.header_bits + k * c.maxbits, true);
  else w.init(c.slots + g * c.slot_words, 0, false);
  encode_block<D, S>(w, c, vals);
  w.finish();
  if (!c.fixed && lane() == 0) c.offs[g] = w.count;
}

// Variable-size modes: moves block g's bits from its scratch slot to its
// place in the stream (header + exclusive prefix sum of the lengths).
__global__ __launch_bounds__(256) void zfp_place_kernel(const ZfpChunk* __restrict__ chunks, uint32_t nchunks,
                                                        uint64_t total, const uint64_t* __restrict__ scan) {
  uint64_t g = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (g >= total) return;
  const ZfpChunk& c = chunks[find_chunk(chunks, nchunks, g)];
  if (c.fixed || !c.offs || !c.dst) return;  // fixed-rate (written in place) or refused
  uint64_t len = c.offs[g];
  if (len == 0) return;
  uint64_t d0 = c.header_bits + scan[g] - scan[c.block_begin];
  const uint64_t* src = c.slots + g * c.slot_words;
  uint64_t* dst = (uint64_t*)c.dst;
  uint64_t first = d0 >> 6, last = (d0 + len - 1) >> 6;
  for (uint64_t dw = first + lane(); dw <= last; dw += 64) {
    uint64_t wb = dw * 64;
    uint64_t lo = wb > d0 ? wb : d0, hi = wb + 64 < d0 + len ? wb + 64 : d0 + len;
    uint64_t sb = lo - d0;
    uint32_t cnt = (uint32_t)(hi - lo);
    uint64_t sw = sb >> 6;
    uint32_t so = (uint32_t)(sb & 63);
    uint64_t v = src[sw] >> so;
    if (so && so + cnt > 64) v |= src[sw + 1] << (64 - so);
    if (cnt < 64) v &= (1ull << cnt) - 1;
    v <<= (uint32_t)(lo - wb);
    if (cnt == 64) dst[dw] = v;
    else atomicOr((unsigned long long*)&dst[dw], (unsigned long long)v);
  }
}

// ---------------------------------------------------------------------------
// The serial walk of a variable-size stream (which has no block index): where
// each block begins, found by parsing the blocks' bits without decoding them.
// One wavefront; every value but the window is wave-uniform.
//
// The stream is read through a window of 127 words: lane l holds words
// first + l (cur) and first + 64 + l (nxt, loaded while cur is parsed).  When
// a read reaches cur's last word the window moves on by 63 words, cur taking
// its own last word and nxt's first 63.  The walk is one loop with one read of
// 64 stream bits at its head, each turn consuming at most those 64 bits, so
// the window registers change in one place only.
//
// A bit plane's group tests are parsed a word at a time rather than a test at
// a time: a run of significance tokens "1 0^z 1" ends at the first token whose
// closing one is followed by a zero, which a prefix parity of the ones finds.
// Where that cannot decide (a token longer than the word, the last
// coefficient, a bit budget below 64) a turn parses one test.  The walk
// consumes exactly the bits the block decoder does (decode_float_block /
// decode_int_block / decode_ints with Parse), budgets and minimum sizes
// included.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t readlane64(uint64_t v, uint32_t i) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, (int)i);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), (int)i);
  return ((uint64_t)hi << 32) | lo;
}

// inclusive prefix parity: bit p of the result is the parity of x's bits 0..p
__device__ __forceinline__ uint64_t prefix_parity(uint64_t x) {
  x ^= x << 1;
  x ^= x << 2;
  x ^= x << 4;
  x ^= x << 8;
  x ^= x << 16;
  x ^= x << 32;
  return x;
}

struct WalkWindow {
  GlobalSrc g;
  uint64_t first;     // first word of cur (wave-uniform)
  uint64_t cur, nxt;  // lane l: words first + l and first + 64 + l

  // cur's load is waited for at once (the empty asm reads it), so that only
  // nxt's is ever in flight: a wait for cur, where the compiler copies it
  // between registers, then never waits for nxt as well
  __device__ void load_cur(uint64_t w) {
    cur = g.load(w + lane());
    __asm__ volatile("" : : "v"(cur));
  }
  __device__ void init(const uint8_t* base, uint64_t nbytes, uint64_t bit) {
    g.base = base;
    g.nbytes = nbytes;
    first = bit >> 6;
    load_cur(first);
    nxt = g.load(first + 64 + lane());
  }
  // the 64 stream bits from bit `pos` on, the first in bit 0 (pos never
  // moves back, and never by 2^38 bits at once)
  __device__ uint64_t peek(uint64_t pos) {
    const uint64_t w = pos >> 6;
    uint32_t i = (uint32_t)(w - first);
    if (i >= 63) {     // word w + 1 is past cur
      if (i < 126) {   // within the window: move on by 63 words
        const uint64_t last = readlane64(cur, 63);
        const uint64_t up = __shfl_up(nxt, 1);
        cur = lane() == 0 ? last : up;
        first += 63;
      } else {
        first = w;
        load_cur(w);
      }
      nxt = g.load(first + 64 + lane());
      i = (uint32_t)(w - first);
    }
    return peek_at(pos, i);
  }
  // the same where word pos / 64 is word i < 63 of cur (any i < 64 reads
  // within cur)
  __device__ uint64_t peek_at(uint64_t pos, uint32_t i) const {
    const uint64_t lo = readlane64(cur, i);
    const uint64_t hi = readlane64(cur, (i + 1) & 63);
    const uint32_t s = (uint32_t)pos & 63;
    return (lo >> s) | ((hi << 1) << (63 - s));
  }
};

// One word of a bit plane's group tests, from coefficient count n.  Ones at
// odd ranks open tokens "1 0^z 1", ones at even ranks close them; a closing
// one followed by a zero ends the plane, the zero being its last test.  The
// prefix parity of the ones is taken across the wave: lane j holds the count
// of ones in bits 0..j (mlo / mhi are lane j's masks of bits 0..j).
struct TestWord {
  uint64_t closing;      // the closing ones (bits 0..62)
  uint32_t n2, used, p;  // count after the word's tests, bits they take, last closing one taken
  bool z0, end;          // first test negative; the plane ends in the word
  __device__ __forceinline__ TestWord(uint64_t x, uint32_t n, uint32_t mlo, uint32_t mhi) {
    z0 = !(x & 1);
    const uint32_t cnt = __builtin_popcount((uint32_t)x & mlo) + __builtin_popcount((uint32_t)(x >> 32) & mhi);
    closing = x & ~__ballot(cnt & 1) & ~(1ull << 63);
    const uint64_t stop = closing & ~(x >> 1);
    const uint32_t pa = (uint32_t)__builtin_ctzll(stop | (1ull << 63));
    const uint32_t pb = 63 - (uint32_t)__builtin_clzll(closing | 1);
    p = stop ? pa : pb;
    const uint32_t tokens = (uint32_t)__builtin_amdgcn_readlane((int)cnt, (int)p) >> 1;
    n2 = z0 ? n : n + p + 1 - tokens;
    used = z0 ? 1 : (stop ? p + 2 : p + 1);
    end = z0 | (stop != 0);
  }
};

// The test word at every stream position, computed ahead of the walk.  For
// the group tests of a plane starting at bit t, what the word at t decides
// does not depend on the coefficient count n (only its last-coefficient
// check does), so other waves compute it for every t in parallel while the
// walking wave only chains lookups.  Entry (16 bits): kind in bits 14..15 --
// 0: the first test is negative (bits 7..13: the run of zeros, at most 63),
// 1: the plane's tests end in the word, 2: tokens run past the word, 3: a
// token longer than the word; bits 7..13 the bits taken, bits 0..6 the count
// of coefficients they make significant.
enum : uint32_t { kWordZero = 0, kWordEnd = 1, kWordOn = 2, kWordLong = 3 };

__device__ __forceinline__ uint32_t test_word_entry(GlobalSrc& g, uint64_t t) {
  const uint64_t w = t >> 6;
  const uint32_t sh = (uint32_t)t & 63;
  const uint64_t x = (g.load(w) >> sh) | ((g.load(w + 1) << 1) << (63 - sh));
  if (!(x & 1)) return (kWordZero << 14) | ((uint32_t)__builtin_ctzll(x | (1ull << 63)) << 7);
  const uint64_t closing = x & ~prefix_parity(x) & ~(1ull << 63);
  if (!closing) return kWordLong << 14;
  const uint64_t stop = closing & ~(x >> 1);
  const uint32_t p = stop ? (uint32_t)__builtin_ctzll(stop) : 63 - (uint32_t)__builtin_clzll(closing);
  const uint32_t tokens = (uint32_t)__builtin_popcountll(x & (~0ull >> (63 - p))) >> 1;
  return ((stop ? kWordEnd : kWordOn) << 14) | ((stop ? p + 2 : p + 1) << 7) | (p + 1 - tokens);
}

// The ring of test-word entries in LDS: kRingSlots slots of kSlotBits
// positions; chunk c (positions base + c kSlotBits ..) goes to slot
// c % kRingSlots.  Producers take 64-position batches from a shared counter
// (so all of them work on the chunks nearest the walk) and write a batch
// only while its chunk is within kRingSlots of the chunk the walk reads
// (`reading`); when the walk moves on, it raises the counter past the
// batches it no longer needs.  Each entry carries its chunk number in its
// upper 32 bits, and the walk takes an entry only when the number is its
// own, so that an entry not yet written, or one overwritten by a producer
// that fell behind, can cost time (a short wait, then that turn parses the
// stream words) but never the result.
constexpr uint32_t kWalkProducers = 15;
constexpr uint32_t kWalkThreads = 64 * (1 + kWalkProducers);
constexpr uint32_t kRingSlots = 16, kSlotBits = 512;
constexpr uint32_t kWalkSpins = 1u << 22;

__device__ __forceinline__ uint32_t lds_acquire(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_release(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Variable-size modes: one workgroup per field walks its blocks in stream
// order and records where each begins (the format has no block index).  Wave
// 0 walks; waves 1.. fill the ring of test words ahead of it.  Lane j of the
// walking wave keeps the offset of block 64 q + j; it stores 64 at a time.
//
// The walk is a loop of turns.  The common turn, branch-free, takes one word
// of a plane's group tests from the ring, and when the plane ends, the next
// plane's verbatim bits (a run of negative first tests with no significant
// coefficients ends as many planes at once).  The general turn reads the
// stream words and takes the rest: block headers, the last coefficient
// becoming significant (the plane's tests end there and the planes left are
// verbatim), ends of blocks, and single tests where the word cannot decide (a
// token longer than the word, a bit budget below 64).
template <int D, class S>
__global__ __launch_bounds__(kWalkThreads) void zfp_index_kernel(const ZfpChunk* __restrict__ chunks,
                                                                 uint32_t nchunks) {
  typedef __attribute__((address_space(1))) uint64_t* g64;
  constexpr bool fp = ZT<S>::fp;
  using BInt = typename ZT<S>::Int;  // the coded integers
  constexpr uint32_t SZ = 1u << (2 * D);
  constexpr uint32_t intprec = sizeof(BInt) * 8;
  constexpr uint32_t pbits = sizeof(BInt) == 4 ? 5 : 6;
  constexpr uint32_t ebits = ZT<S>::ebits;
  constexpr int ebias = ZT<S>::ebias;
  enum { kHeader, kTests, kZeroRun };

  uint32_t ci = blockIdx.x;
  if (ci >= nchunks) return;
  const ZfpChunk& c = chunks[ci];
  if (c.fixed || c.index) return;  // computed addressing, or a block index
  const int minbits = (int)c.minbits, maxbits = (int)c.maxbits, cmaxprec = (int)c.maxprec, minexp = c.minexp;
  const bool rev = minexp < kZfpMinExp;
  const uint64_t nblocks = c.nblocks;
  const g64 offs = (g64)(uintptr_t)(c.offs + c.block_begin);
  const uint32_t l = lane();
  const uint64_t base = c.header_bits;
  const uint64_t stream_bits = c.src_len * 8;
  const uint32_t nring = stream_bits > base ? (uint32_t)((stream_bits - base + kSlotBits - 1) / kSlotBits) : 0u;

  __shared__ uint64_t ring[kRingSlots * kSlotBits];
  __shared__ uint32_t reading, next_batch, done;
  for (uint32_t i = threadIdx.x; i < kRingSlots * kSlotBits; i += kWalkThreads)
    ring[i] = ~0ull << 32 | kWordLong << 14;  // no chunk's; undecided
  if (threadIdx.x == 0) reading = next_batch = done = 0;
  __syncthreads();

  constexpr uint32_t kBatches = kSlotBits / 64;  // batches per chunk
  if (threadIdx.x >= 64) {  // a producer
    GlobalSrc g{c.src, c.src_len};
    const uint64_t nbatches = (uint64_t)nring * kBatches;
    for (;;) {
      uint32_t b = 0;
      if (l == 0) b = __hip_atomic_fetch_add(&next_batch, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      b = (uint32_t)__builtin_amdgcn_readfirstlane((int)b);
      if (b >= nbatches || lds_acquire(&done)) break;
      const uint32_t ch = b / kBatches;
      bool go = false;
      for (uint32_t spin = 0; spin < kWalkSpins; ++spin) {
        const uint32_t at = lds_acquire(&reading);
        if (at > ch || lds_acquire(&done)) break;  // passed by the walk: not needed
        if (at + kRingSlots > ch) {
          go = true;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
      if (!go) continue;
      const uint32_t r = (b % kBatches) * 64 + l;
      __hip_atomic_store(&ring[(ch % kRingSlots) * kSlotBits + r],
                         (uint64_t)ch << 32 | test_word_entry(g, base + (uint64_t)ch * kSlotBits + r),
                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    return;
  }

  const uint32_t mlo = l < 32 ? (2u << l) - 1 : ~0u;
  const uint32_t mhi = l < 32 ? 0u : (2u << (l - 32)) - 1;
  WalkWindow win;
  win.init(c.src, c.src_len, c.header_bits);

  uint64_t pos = base, blk_start = pos, kb = 0, mine = 0;
  uint32_t bits = 0, n = 0, k = 0, kmin = 0, zacc = 0, zlim = 0;
  uint32_t cur_ch = ~0u;
  int st = kHeader;
  for (;;) {
    if (st == kTests) {  // common turns (n < SZ here)
      for (;;) {
        const uint64_t rel = pos - base;
        const uint64_t te = __hip_atomic_load(&ring[(uint32_t)rel % (kRingSlots * kSlotBits)], __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_WORKGROUP);
        const uint32_t e = (uint32_t)te, kind = e >> 14, uf = (e >> 7) & 127;
        const bool z0 = kind == kWordZero;
        const uint32_t used = (z0 & (n != 0)) ? 1u : uf;
        const uint32_t n2 = n + (e & 127);
        const uint32_t ends = z0 ? used : 2 - kind;  // planes ended: kWordEnd 1, kWordOn 0
        const uint32_t bits1 = bits - used;
        const uint32_t madd = kind <= kWordEnd ? (n2 < bits1 ? n2 : bits1) : 0u;  // the next plane's verbatim bits
        // the entry is this chunk's and decides; the budget is above the
        // word; the planes ended exist
        if (!(((uint32_t)(te >> 32) == (uint32_t)(rel / kSlotBits)) & (kind != kWordLong) & (n2 < SZ) &
              (bits > 64) & (k - kmin >= ends)))
          break;
        n = n2;
        pos += used + madd;
        bits = bits1 - madd;
        k -= ends;
      }
    }
    // the general turn
    {
      const uint32_t ch = (uint32_t)((pos - base) / kSlotBits);
      if (ch != cur_ch) {  // let the producers move on
        lds_release(&reading, ch);
        __hip_atomic_fetch_max(&next_batch, ch * kBatches, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        cur_ch = ch;
      }
    }
    if (st == kHeader) {
      if (kb == nblocks) break;
      const uint32_t j = (uint32_t)kb & 63;
      if (l == j) mine = pos;
      if (j == 63) offs[kb - 63 + l] = mine;
      blk_start = pos;
    }
    const uint64_t x = win.peek(pos);
    bool plane_end = false;
    if (st == kHeader) {
      int hdr = 0, prec = cmaxprec;
      if constexpr (fp) {
        if (!(x & 1)) {  // every value zero: one bit (reversible) or minbits
          pos += rev || minbits <= 1 ? 1u : (uint32_t)minbits;
          ++kb;
          continue;
        }
        if (rev) {  // 1 1: raw bits; 1 0: a block exponent follows
          hdr = (x & 2) ? 2 : 2 + (int)ebits;
          prec = (int)((x >> hdr) & ((1u << pbits) - 1)) + 1;
          hdr += (int)pbits;
        } else {
          const int emax = (int)((x >> 1) & ((1u << ebits) - 1)) - ebias;
          const int dp = emax - minexp + 2 * (D + 1);
          prec = cmaxprec < (dp > 0 ? dp : 0) ? cmaxprec : (dp > 0 ? dp : 0);
          hdr = 1 + (int)ebits;
        }
      } else if (rev) {
        prec = (int)(x & ((1u << pbits) - 1)) + 1;
        hdr = (int)pbits;
      }
      pos += (uint32_t)hdr;
      bits = maxbits - hdr > 0 ? (uint32_t)(maxbits - hdr) : 0u;
      n = 0;
      k = intprec;
      kmin = intprec > (uint32_t)prec ? intprec - (uint32_t)prec : 0;
      plane_end = true;
    } else if (st == kZeroRun) {  // a test's zeros that ran past the last word
      const uint32_t t = x ? (uint32_t)__builtin_ctzll(x) : 64u;
      if (t < zlim && t < 64) {  // the closing one
        pos += t + 1;
        bits -= t + 1;
        n += zacc + t + 1;
        st = kTests;
      } else if (zlim <= 64) {  // the zeros reach the limit: no closing one
        pos += zlim;
        bits -= zlim;
        n += zacc + zlim + 1;
        st = kTests;
      } else {
        pos += 64;
        bits -= 64;
        zacc += 64;
        zlim -= 64;
      }
    } else if (n >= SZ || !bits) {
      plane_end = true;
    } else {
      const TestWord tw(x, n, mlo, mhi);
      if (bits >= 64 && tw.z0) {
        pos += 1;
        bits -= 1;
        plane_end = true;
      } else if (bits >= 64 && tw.closing && tw.n2 >= SZ) {
        // the last coefficient becomes significant inside the word: its
        // token is the one at which the count of bits other than closing
        // ones reaches SZ - n (lim zeros after the opening one, no closing)
        const uint64_t other = ~tw.closing;
        const uint32_t cnt =
            __builtin_popcount((uint32_t)other & mlo) + __builtin_popcount((uint32_t)(other >> 32) & mhi);
        const uint32_t q = (uint32_t)__builtin_ctzll(__ballot(cnt == SZ - n) & other);
        pos += q + 1;
        bits -= q + 1;
        n = SZ;
        plane_end = true;
      } else if (bits >= 64 && tw.closing) {
        n = tw.n2;
        pos += tw.used;
        bits -= tw.used;
        plane_end = tw.end;
      } else {  // one test: a budget below 64, or a token longer than the word
        bits--;
        if (!(x & 1)) {
          pos += 1;
          plane_end = true;
        } else {
          uint32_t lim = SZ - 1 - n;
          if (lim > bits) lim = bits;
          const uint64_t y = x >> 1;
          const uint32_t t = y ? (uint32_t)__builtin_ctzll(y) : 64u;
          if (t < lim && t < 64) {  // the closing one, in view
            pos += t + 2;
            bits -= t + 1;
            n += t + 1;
          } else if (lim <= 63) {  // the zeros reach the limit in view
            pos += lim + 1;
            bits -= lim;
            n += lim + 1;
          } else {  // 63 zeros in view, more to come
            pos += 64;
            bits -= 63;
            zacc = 63;
            zlim = lim - 63;
            st = kZeroRun;
          }
        }
      }
    }
    plane_end |= (st == kTests) & (n >= SZ);  // the tests end with the last coefficient
    if (plane_end) {
      bool block_end = true;
      if (bits && k > kmin) {
        --k;
        if (n == SZ) {  // every coefficient significant: the planes left are SZ verbatim bits each
          const uint64_t need = (uint64_t)(k + 1 - kmin) * SZ;
          const uint32_t take = need < bits ? (uint32_t)need : bits;
          pos += take;
          bits -= take;
        } else {
          const uint32_t m = n < bits ? n : bits;
          bits -= m;
          pos += m;
          st = kTests;
          block_end = false;
        }
      }
      if (block_end) {
        if (pos - blk_start < (uint64_t)(minbits > 0 ? minbits : 0)) pos = blk_start + (uint32_t)minbits;
        ++kb;
        st = kHeader;
      }
    }
  }
  lds_release(&done, 1);
  const uint64_t tail = nblocks & 63;
  if (l < tail) offs[nblocks - tail + l] = mine;
}

template <int D, class S>
__global__ __launch_bounds__(256) void zfp_decode_kernel(const ZfpChunk* __restrict__ chunks, uint32_t nchunks,
                                                         uint64_t total, const ZfpBrick* __restrict__ bricks) {
  constexpr int R = D == 4 ? 4 : 1;
  uint64_t g = xcd_item() * 4 + (threadIdx.x >> 6);  // neighbouring blocks share output lines: one L2
  if (g >= total) return;
  uint32_t ci;
  uint64_t k;
  uint8_t* brick = nullptr;
  if (bricks) {
    ci = bricks[g].chunk;
    k = bricks[g].block;
    brick = bricks[g].out;
  } else {
    ci = find_chunk(chunks, nchunks, g);
    k = g - chunks[ci].block_begin;
  }
  const ZfpChunk& c = chunks[ci];
  BitReader r;
  r.src.base = c.src;
  r.src.nbytes = c.src_len;
  const bool indexed = !c.fixed && c.index;
  r.seek(c.fixed ? c.header_bits + k * c.maxbits : indexed ? c.header_bits + c.index[k] : c.offs[c.block_begin + k]);
  S vals[R];
  const uint32_t used = decode_block<D, S, false>(r, c, vals);
  if (indexed && lane() == 0) {
    // the block must take exactly the bits the index gives it, the index
    // must start at the first block and end inside the stream
    const uint32_t lo = c.index[k], hi = c.index[k + 1];
    if (hi < lo || hi - lo != used || (k == 0 && lo != 0) ||
        (k + 1 == c.nblocks && c.header_bits + (uint64_t)hi > 8 * c.src_len))
      atomicOr(c.bad, 1u);
  }
  uint32_t b[4];
  block_coords(c, k, b);
  scatter<D, S>(c, b, vals, brick);
}

// Block index of each variable-size chunk that asks for one: entry k is
// block k's bit offset from the first block, entry nblocks the end.
__global__ __launch_bounds__(256) void zfp_index_write_kernel(const ZfpChunk* __restrict__ chunks, uint32_t nchunks,
                                                              uint64_t total, const uint64_t* __restrict__ scan) {
  const uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (g >= total) return;
  const ZfpChunk& c = chunks[find_chunk(chunks, nchunks, g)];
  if (c.fixed || !c.index || !c.dst || !c.offs) return;
  const uint64_t k = g - c.block_begin;
  const uint64_t at = scan[g] - scan[c.block_begin];
  c.index[k] = (uint32_t)at;
  if (k + 1 == c.nblocks) c.index[k + 1] = (uint32_t)(at + c.offs[g]);
}

template <int D, class S> void enc(const ZfpChunk* d, uint32_t n, uint64_t total, hipStream_t s) {
  hipLaunchKernelGGL((zfp_encode_kernel<D, S>), xcd_grid(ceil_div(total, 4)), dim3(256), 0, s, d, n, total);
}
template <int D, class S> void dec(const ZfpChunk* d, uint32_t n, uint64_t total, const ZfpBrick* br, hipStream_t s) {
  hipLaunchKernelGGL((zfp_decode_kernel<D, S>), xcd_grid(ceil_div(total, 4)), dim3(256), 0, s, d, n, total, br);
}
template <int D, class S> void idx(const ZfpChunk* d, uint32_t n, hipStream_t s) {
  hipLaunchKernelGGL((zfp_index_kernel<D, S>), dim3(n), dim3(kWalkThreads), 0, s, d, n);
}

#define NDIC_ZFP_DISPATCH(FN, ...)                                                     \
  switch (dims * 8 + scalar) {                                                         \
    case 1 * 8 + kZfpInt32: FN<1, int32_t>(__VA_ARGS__); break;                        \
    case 1 * 8 + kZfpInt64: FN<1, int64_t>(__VA_ARGS__); break;                        \
    case 1 * 8 + kZfpFloat: FN<1, float>(__VA_ARGS__); break;                          \
    case 1 * 8 + kZfpDouble: FN<1, double>(__VA_ARGS__); break;                        \
    case 2 * 8 + kZfpInt32: FN<2, int32_t>(__VA_ARGS__); break;                        \
    case 2 * 8 + kZfpInt64: FN<2, int64_t>(__VA_ARGS__); break;                        \
    case 2 * 8 + kZfpFloat: FN<2, float>(__VA_ARGS__); break;                          \
    case 2 * 8 + kZfpDouble: FN<2, double>(__VA_ARGS__); break;                        \
    case 3 * 8 + kZfpInt32: FN<3, int32_t>(__VA_ARGS__); break;                        \
    case 3 * 8 + kZfpInt64: FN<3, int64_t>(__VA_ARGS__); break;                        \
    case 3 * 8 + kZfpFloat: FN<3, float>(__VA_ARGS__); break;                          \
    case 3 * 8 + kZfpDouble: FN<3, double>(__VA_ARGS__); break;                        \
    case 4 * 8 + kZfpInt32: FN<4, int32_t>(__VA_ARGS__); break;                        \
    case 4 * 8 + kZfpInt64: FN<4, int64_t>(__VA_ARGS__); break;                        \
    case 4 * 8 + kZfpFloat: FN<4, float>(__VA_ARGS__); break;                          \
    case 4 * 8 + kZfpDouble: FN<4, double>(__VA_ARGS__); break;                        \
    default: fail(NDIC_INVALID_ARGUMENT, "zfp: unsupported field class");              \
  }

// The block coders of 4D fields (256 coefficients a block: one wavefront per
// block); 1D-3D fields take one lane per block (zfp_lane.hip).
#define NDIC_ZFP_DISPATCH_4D(FN, ...)                                                  \
  switch (dims * 8 + scalar) {                                                         \
    case 4 * 8 + kZfpInt32: FN<4, int32_t>(__VA_ARGS__); break;                        \
    case 4 * 8 + kZfpInt64: FN<4, int64_t>(__VA_ARGS__); break;                        \
    case 4 * 8 + kZfpFloat: FN<4, float>(__VA_ARGS__); break;                          \
    case 4 * 8 + kZfpDouble: FN<4, double>(__VA_ARGS__); break;                        \
    default: fail(NDIC_INVALID_ARGUMENT, "zfp: unsupported field class");              \
  }



__global__ void zfp_headers_kernel(const ZfpHeader* __restrict__ h, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (uint32_t k = 0; k < h[i].nwords; ++k) h[i].dst[k] = h[i].w[k];
}

__global__ void zfp_totals_kernel(const ZfpChunk* __restrict__ chunks, uint32_t n, const uint64_t* __restrict__ scan,
                                  uint64_t* totals) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const ZfpChunk& c = chunks[i];
  if (c.fixed || c.nblocks == 0) {
    totals[i] = 0;
    return;
  }
  uint64_t last = c.block_begin + c.nblocks - 1;
  totals[i] = scan[last] + c.offs[last] - scan[c.block_begin];
}

__global__ void zfp_heads_kernel(const ZfpHead* __restrict__ h, uint32_t n, uint8_t* out) {
  uint32_t i = blockIdx.x, b = threadIdx.x;
  if (i >= n || b >= kZfpHeadBytes) return;
  out[(uint64_t)i * kZfpHeadBytes + b] = b < h[i].len ? h[i].src[b] : 0;
}

}  // namespace

size_t zfp_scan_temp_bytes(uint64_t n) {
  size_t bytes = 0;
  NDIC_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, (const uint64_t*)nullptr, (uint64_t*)nullptr, (int)n));
  return bytes;
}

void zfp_exclusive_scan(void* temp, size_t temp_bytes, const uint64_t* d_in, uint64_t* d_out, uint64_t n,
                        hipStream_t stream) {
  if (n == 0) return;
  if (n > 0x7fffffffull) fail(NDIC_INVALID_ARGUMENT, "zfp: too many blocks in one batch");
  NDIC_HIP(hipcub::DeviceScan::ExclusiveSum(temp, temp_bytes, d_in, d_out, (int)n, stream));
}

void launch_zfp_headers(const ZfpHeader* d_headers, uint32_t n, hipStream_t stream) {
  if (!n) return;
  hipLaunchKernelGGL(zfp_headers_kernel, dim3((n + 63) / 64), dim3(64), 0, stream, d_headers, n);
  NDIC_HIP(hipGetLastError());
}

void launch_zfp_totals(const ZfpChunk* d_chunks, uint32_t nchunks, const uint64_t* d_scan, uint64_t* d_totals,
                       hipStream_t stream) {
  if (!nchunks) return;
  hipLaunchKernelGGL(zfp_totals_kernel, dim3((nchunks + 63) / 64), dim3(64), 0, stream, d_chunks, nchunks, d_scan,
                     d_totals);
  NDIC_HIP(hipGetLastError());
}

void launch_zfp_heads(const ZfpHead* d_heads, uint32_t n, uint8_t* d_out, hipStream_t stream) {
  if (!n) return;
  hipLaunchKernelGGL(zfp_heads_kernel, dim3(n), dim3(64), 0, stream, d_heads, n, d_out);
  NDIC_HIP(hipGetLastError());
}

void zfp_init_tables() {
  static std::mutex mu;
  static bool done = false;
  std::lock_guard<std::mutex> lock(mu);
  if (done) return;
  uint8_t perm[4][256] = {}, iperm[4][256] = {};
  for (int i = 0; i < 4; ++i) perm[0][i] = (uint8_t)i;
  for (int i = 0; i < 16; ++i) perm[1][i] = kPerm2[i];
  for (int i = 0; i < 64; ++i) perm[2][i] = kPerm3[i];
  for (int i = 0; i < 256; ++i) perm[3][i] = kPerm4[i];
  for (int d = 0; d < 4; ++d)
    for (int r = 0; r < (1 << (2 * (d + 1))); ++r) iperm[d][perm[d][r]] = (uint8_t)r;
  NDIC_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_perm), perm, sizeof(perm)));
  NDIC_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_iperm), iperm, sizeof(iperm)));
  done = true;
}

void launch_zfp_encode(const ZfpChunk* d_chunks, uint32_t nchunks, uint64_t total, uint32_t dims, uint32_t scalar,
                       hipStream_t stream) {
  if (total == 0) return;
  if (launch_zfp_lane_encode(d_chunks, nchunks, total, dims, scalar, stream)) return;
  fail(NDIC_INVALID_ARGUMENT, "zfp: unsupported field class");
  NDIC_HIP(hipGetLastError());
}

void launch_zfp_place(const ZfpChunk* d_chunks, uint32_t nchunks, uint64_t total, const uint64_t* d_scan,
                      hipStream_t stream) {
  if (total == 0) return;
  hipLaunchKernelGGL(zfp_place_kernel, dim3((uint32_t)ceil_div(total, 4)), dim3(256), 0, stream, d_chunks, nchunks,
                     total, d_scan);
  NDIC_HIP(hipGetLastError());
}

void launch_zfp_index(const ZfpChunk* d_chunks, uint32_t nchunks, uint32_t dims, uint32_t scalar,
                      hipStream_t stream) {
  if (nchunks == 0) return;
  NDIC_ZFP_DISPATCH(idx, d_chunks, nchunks, stream);
  NDIC_HIP(hipGetLastError());
}

void launch_zfp_index_write(const ZfpChunk* d_chunks, uint32_t nchunks, uint64_t total, const uint64_t* d_scan,
                            hipStream_t stream) {
  if (total == 0) return;
  hipLaunchKernelGGL(zfp_index_write_kernel, dim3((uint32_t)ceil_div(total, 256)), dim3(256), 0, stream, d_chunks,
                     nchunks, total, d_scan);
  NDIC_HIP(hipGetLastError());
}

void launch_zfp_decode(const ZfpChunk* d_chunks, uint32_t nchunks, uint64_t total, uint32_t dims, uint32_t scalar,
                       const ZfpBrick* d_bricks, uint32_t fixed_maxbits, hipStream_t stream) {
  if (total == 0) return;
  if (launch_zfp_lane_decode(d_chunks, nchunks, total, dims, scalar, d_bricks, fixed_maxbits, stream)) return;
  fail(NDIC_INVALID_ARGUMENT, "zfp: unsupported field class");
  NDIC_HIP(hipGetLastError());
}

}  // namespace ndic
