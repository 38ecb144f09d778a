// Oracle (test infrastructure only): ZFP (fixed-rate / fixed-precision /
// fixed-accuracy / reversible) restated from the published LLNL ZFP
// algorithm (P. Lindstrom, "Fixed-Rate Compressed Floating-Point Arrays",
// IEEE TVCG 2014, and the ZFP stream format, codec version 5), which the
// reference reaches through the un-vendored crate zfp-rs 0.1.1
// (Cargo.lock:2562-2565; call sites crates/ndic-zfp/src/lib.rs:249-376,
// 525-577), plus the ndic-zfp wrapper semantics (crates/ndic-zfp/src/lib.rs
// :75-599, chunk.rs:246-417).  Pinned by fixtures/zfp/checksums.json (104
// stream hashes) and tiny-chunk-4x8x8-rate8.zfp.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <sstream>

#include "zfp_oracle.hpp"

namespace orc {
namespace zfp {
namespace {

// ---------------------------------------------------------------------------
// bit stream: LSB-first bits, flushed to whole 64-bit little-endian words
// ---------------------------------------------------------------------------
struct BitWriter {
  std::vector<uint8_t> bytes;
  uint64_t nbits = 0;
  void bit(uint32_t b) {
    if ((nbits & 7) == 0) bytes.push_back(0);
    bytes.back() |= uint8_t((b & 1) << (nbits & 7));
    ++nbits;
  }
  uint64_t bits(uint64_t v, uint32_t n) {
    for (uint32_t i = 0; i < n; ++i) bit((uint32_t)(v >> i) & 1);
    return n >= 64 ? 0 : v >> n;
  }
  void pad(uint64_t n) {
    for (uint64_t i = 0; i < n; ++i) bit(0);
  }
  void flush() {
    while (nbits % 64) bit(0);
  }
};

struct BitReader {
  const uint8_t* data;
  uint64_t nbytes, pos = 0;
  BitReader(const uint8_t* d, uint64_t n) : data(d), nbytes(n) {}
  uint32_t bit() {
    uint64_t byte = pos >> 3;
    uint32_t b = byte < nbytes ? (data[byte] >> (pos & 7)) & 1 : 0;
    ++pos;
    return b;
  }
  uint64_t bits(uint32_t n) {
    uint64_t v = 0;
    for (uint32_t i = 0; i < n; ++i) v |= uint64_t(bit()) << i;
    return v;
  }
  void skip(uint64_t n) { pos += n; }
};

// ---------------------------------------------------------------------------
// coefficient orders: total sequency, then squared distance from the origin
// ---------------------------------------------------------------------------
const uint8_t PERM1[4] = {0, 1, 2, 3};
#define I2(i, j) ((i) + 4 * (j))
const uint8_t PERM2[16] = {I2(0, 0), I2(1, 0), I2(0, 1), I2(1, 1), I2(2, 0), I2(0, 2), I2(2, 1), I2(1, 2),
                           I2(3, 0), I2(0, 3), I2(2, 2), I2(3, 1), I2(1, 3), I2(3, 2), I2(2, 3), I2(3, 3)};
#define I3(i, j, k) ((i) + 4 * (j) + 16 * (k))
const uint8_t PERM3[64] = {
    I3(0, 0, 0), I3(1, 0, 0), I3(0, 1, 0), I3(0, 0, 1), I3(0, 1, 1), I3(1, 0, 1), I3(1, 1, 0), I3(2, 0, 0),
    I3(0, 2, 0), I3(0, 0, 2), I3(1, 1, 1), I3(2, 1, 0), I3(2, 0, 1), I3(0, 2, 1), I3(1, 2, 0), I3(1, 0, 2),
    I3(0, 1, 2), I3(3, 0, 0), I3(0, 3, 0), I3(0, 0, 3), I3(2, 1, 1), I3(1, 2, 1), I3(1, 1, 2), I3(0, 2, 2),
    I3(2, 0, 2), I3(2, 2, 0), I3(3, 1, 0), I3(3, 0, 1), I3(0, 3, 1), I3(1, 3, 0), I3(1, 0, 3), I3(0, 1, 3),
    I3(1, 2, 2), I3(2, 1, 2), I3(2, 2, 1), I3(3, 1, 1), I3(1, 3, 1), I3(1, 1, 3), I3(3, 2, 0), I3(3, 0, 2),
    I3(0, 3, 2), I3(2, 3, 0), I3(2, 0, 3), I3(0, 2, 3), I3(2, 2, 2), I3(3, 2, 1), I3(3, 1, 2), I3(1, 3, 2),
    I3(2, 3, 1), I3(2, 1, 3), I3(1, 2, 3), I3(0, 3, 3), I3(3, 0, 3), I3(3, 3, 0), I3(3, 2, 2), I3(2, 3, 2),
    I3(2, 2, 3), I3(1, 3, 3), I3(3, 1, 3), I3(3, 3, 1), I3(2, 3, 3), I3(3, 2, 3), I3(3, 3, 2), I3(3, 3, 3)};
// 4D order (i + 4j + 16k + 64l), as the C library 0.5.5 emits it: measured by
// encoding single-coefficient reversible blocks and reading the run length
// of the lone one bit (tests/golden/make_zfp_vectors.py); grouped by total
// sequency i+j+k+l like the lower-dimensional orders.
const uint8_t PERM4[256] = {
      0,   1,   4,  16,  64,   5,  80,  17,  68,  65,  20,   2,   8,  32, 128,  84,
     81,  69,  21,   6,  18,  66,  24,  72,   9,  96,  33,  36, 129, 132, 144,   3,
     12,  48, 192,  85,  82,  70,  22,  73,  25,  88,  37, 100,  97, 148, 145, 133,
     10, 160,  34, 136, 130,  40,   7,  19,  67,  28,  76,  13, 112,  49,  52, 193,
    196, 208,  86,  89, 101, 149, 161, 137,  41, 134,  38, 164,  26, 152, 146, 104,
     98,  74,  83,  71,  23,  77,  29,  92,  53, 116, 113, 212, 209, 197,  11,  35,
    131,  44, 140,  14, 176,  50,  56, 194, 200, 224,  90, 165, 102, 153, 150, 105,
    168, 162, 138,  42,  87,  93, 117, 213,  27,  75,  99,  39, 135, 147, 108,  45,
    141, 156,  30,  78, 177, 180,  54, 114, 120,  57, 198, 210, 216, 201, 225, 228,
     15, 240,  51, 204, 195,  60, 169, 166, 154, 106,  91, 103, 151, 109, 157,  94,
    181, 118, 121, 214, 217, 229, 163, 139,  43, 142,  46, 172,  58, 184, 178, 232,
    226, 202, 241, 205,  61, 199,  55, 244,  31, 220, 211, 124, 115,  79, 170, 167,
    155, 107, 158, 110, 173, 122, 185, 182, 233, 230, 218,  95, 245, 119, 221, 215,
    125, 242, 206,  62, 203,  59, 248,  47, 236, 227, 188, 179, 143, 171, 174, 186,
    234, 246, 222, 126, 219, 123, 249, 111, 237, 231, 189, 183, 159, 252, 243, 207,
     63, 175, 250, 187, 238, 235, 190, 253, 247, 223, 127, 254, 251, 239, 191, 255,
};

const uint8_t* perm_table(int dims) {
  if (dims == 4) return PERM4;
  return dims == 1 ? PERM1 : dims == 2 ? PERM2 : PERM3;
}

// ---------------------------------------------------------------------------
// scalar traits
// ---------------------------------------------------------------------------
template <class S> struct Tr;
template <> struct Tr<int32_t> {
  typedef int32_t Int;
  typedef uint32_t UInt;
  static constexpr int prec = 32, pbits = 5;
  static constexpr UInt nbmask = 0xaaaaaaaau;
};
template <> struct Tr<int64_t> {
  typedef int64_t Int;
  typedef uint64_t UInt;
  static constexpr int prec = 64, pbits = 6;
  static constexpr UInt nbmask = 0xaaaaaaaaaaaaaaaaull;
};
template <> struct Tr<float> {
  typedef int32_t Int;
  typedef uint32_t UInt;
  static constexpr int prec = 32, pbits = 5, ebits = 8, ebias = 127;
  static constexpr UInt nbmask = 0xaaaaaaaau;
};
template <> struct Tr<double> {
  typedef int64_t Int;
  typedef uint64_t UInt;
  static constexpr int prec = 64, pbits = 6, ebits = 11, ebias = 1023;
  static constexpr UInt nbmask = 0xaaaaaaaaaaaaaaaaull;
};

template <class Int> inline Int wadd(Int a, Int b) { return (Int)((typename std::make_unsigned<Int>::type)a + (typename std::make_unsigned<Int>::type)b); }
template <class Int> inline Int wsub(Int a, Int b) { return (Int)((typename std::make_unsigned<Int>::type)a - (typename std::make_unsigned<Int>::type)b); }
template <class Int> inline Int wshl1(Int a) { return (Int)((typename std::make_unsigned<Int>::type)a << 1); }

// decorrelating lifts of a 4-vector with stride s
template <class Int> void fwd_lift(Int* p, int s) {
  Int x = p[0], y = p[s], z = p[2 * s], w = p[3 * s];
  x = wadd(x, w); x >>= 1; w = wsub(w, x);
  z = wadd(z, y); z >>= 1; y = wsub(y, z);
  x = wadd(x, z); x >>= 1; z = wsub(z, x);
  w = wadd(w, y); w >>= 1; y = wsub(y, w);
  w = wadd(w, (Int)(y >> 1)); y = wsub(y, (Int)(w >> 1));
  p[0] = x; p[s] = y; p[2 * s] = z; p[3 * s] = w;
}
template <class Int> void inv_lift(Int* p, int s) {
  Int x = p[0], y = p[s], z = p[2 * s], w = p[3 * s];
  y = wadd(y, (Int)(w >> 1)); w = wsub(w, (Int)(y >> 1));
  y = wadd(y, w); w = wshl1(w); w = wsub(w, y);
  z = wadd(z, x); x = wshl1(x); x = wsub(x, z);
  y = wadd(y, z); z = wshl1(z); z = wsub(z, y);
  w = wadd(w, x); x = wshl1(x); x = wsub(x, w);
  p[0] = x; p[s] = y; p[2 * s] = z; p[3 * s] = w;
}
template <class Int> void rev_fwd_lift(Int* p, int s) {
  Int x = p[0], y = p[s], z = p[2 * s], w = p[3 * s];
  w = wsub(w, z); z = wsub(z, y); y = wsub(y, x);
  w = wsub(w, z); z = wsub(z, y);
  w = wsub(w, z);
  p[0] = x; p[s] = y; p[2 * s] = z; p[3 * s] = w;
}
template <class Int> void rev_inv_lift(Int* p, int s) {
  Int x = p[0], y = p[s], z = p[2 * s], w = p[3 * s];
  w = wadd(w, z); z = wadd(z, y); y = wadd(y, x);
  w = wadd(w, z); z = wadd(z, y);
  w = wadd(w, z);
  p[0] = x; p[s] = y; p[2 * s] = z; p[3 * s] = w;
}

// the lift along each axis in turn: x, y, z, w forward; w, z, y, x inverse
template <class Int> void xform(Int* p, int dims, void (*lift)(Int*, int), bool forward) {
  int n = 1 << (2 * dims);
  for (int a = 0; a < dims; ++a) {
    int axis = forward ? a : dims - 1 - a;
    int stride = 1 << (2 * axis);
    for (int i = 0; i < n; ++i)
      if (((i >> (2 * axis)) & 3) == 0) lift(p + i, stride);
  }
}

template <class Int, class UInt> UInt int2uint(Int x, UInt mask) { return ((UInt)x + mask) ^ mask; }
template <class Int, class UInt> Int uint2int(UInt x, UInt mask) { return (Int)((x ^ mask) - mask); }

// embedded bit-plane coder (group tests + unary run lengths), MSB plane first
template <class UInt> uint32_t encode_ints(BitWriter& s, uint32_t maxbits, uint32_t maxprec, const UInt* data, uint32_t size) {
  const uint32_t intprec = sizeof(UInt) * 8;
  uint32_t kmin = intprec > maxprec ? intprec - maxprec : 0;
  uint32_t bits = maxbits, n = 0;
  for (uint32_t k = intprec; bits && k-- > kmin;) {
    std::vector<uint8_t> plane(size);
    for (uint32_t i = 0; i < size; ++i) plane[i] = (data[i] >> k) & 1;
    uint32_t m = std::min(n, bits);
    bits -= m;
    for (uint32_t i = 0; i < m; ++i) s.bit(plane[i]);
    // remaining ones in positions [n, size)
    uint32_t i = n;
    auto any_from = [&](uint32_t from) {
      for (uint32_t t = from; t < size; ++t)
        if (plane[t]) return 1u;
      return 0u;
    };
    // step 3 of the reference loop structure
    for (; n < size && bits;) {
      --bits;
      uint32_t b = any_from(i);
      s.bit(b);
      if (!b) break;
      for (; n < size - 1 && bits;) {
        --bits;
        uint32_t v = plane[n];
        s.bit(v);
        if (v) break;
        ++n;
        ++i;
      }
      // the one at position n has been located (written, or implied at n = size - 1)
      ++n;
      i = n;
    }
  }
  return maxbits - bits;
}

template <class UInt> uint32_t decode_ints(BitReader& s, uint32_t maxbits, uint32_t maxprec, UInt* data, uint32_t size) {
  const uint32_t intprec = sizeof(UInt) * 8;
  uint32_t kmin = intprec > maxprec ? intprec - maxprec : 0;
  uint32_t bits = maxbits, n = 0;
  for (uint32_t i = 0; i < size; ++i) data[i] = 0;
  for (uint32_t k = intprec; bits && k-- > kmin;) {
    std::vector<uint8_t> plane(size, 0);
    uint32_t m = std::min(n, bits);
    bits -= m;
    for (uint32_t i = 0; i < m; ++i) plane[i] = (uint8_t)s.bit();
    for (; n < size && bits;) {
      --bits;
      if (!s.bit()) break;
      for (; n < size - 1 && bits;) {
        --bits;
        if (s.bit()) break;
        ++n;
      }
      plane[n] = 1;
      ++n;
    }
    for (uint32_t i = 0; i < size; ++i)
      if (plane[i]) data[i] += (UInt)1 << k;
  }
  return maxbits - bits;
}

// ---------------------------------------------------------------------------
// IEEE-754 helpers (bit manipulation, no math library internals)
// ---------------------------------------------------------------------------
template <class S> struct Fp;
template <> struct Fp<float> {
  typedef uint32_t Bits;
  static constexpr int mbits = 23, emask = 0xFF, bias = 127;
};
template <> struct Fp<double> {
  typedef uint64_t Bits;
  static constexpr int mbits = 52, emask = 0x7FF, bias = 1023;
};
template <class S> typename Fp<S>::Bits to_bits(S x) {
  typename Fp<S>::Bits b;
  std::memcpy(&b, &x, sizeof(b));
  return b;
}
template <class S> S from_bits(typename Fp<S>::Bits b) {
  S x;
  std::memcpy(&x, &b, sizeof(b));
  return x;
}
// 2^k as S (0 below the subnormal range, +inf above the normal range)
template <class S> S pow2(int k) {
  typedef typename Fp<S>::Bits B;
  const int bias = Fp<S>::bias, mb = Fp<S>::mbits;
  if (k > bias) return from_bits<S>((B)Fp<S>::emask << mb);
  if (k >= 1 - bias) return from_bits<S>((B)(k + bias) << mb);
  if (k >= 1 - bias - mb) return from_bits<S>((B)1 << (k - (1 - bias - mb)));
  return (S)0;
}
// exponent e of max|x| with x = m 2^e, 0.5 <= m < 1; subnormals clamp to 1 - bias
template <class S> int block_exponent(const S* p, int n) {
  S mx = 0;
  for (int i = 0; i < n; ++i) {
    S f = p[i] < 0 ? -p[i] : p[i];
    if (mx < f) mx = f;
  }
  if (!(mx > 0)) return -Fp<S>::bias;
  typename Fp<S>::Bits b = to_bits(mx);
  int be = (int)((b >> Fp<S>::mbits) & Fp<S>::emask);
  if (be == 0) return 1 - Fp<S>::bias;
  if (be == Fp<S>::emask) return 0;  // infinity: frexp reports exponent 0
  return be - (Fp<S>::bias - 1);
}
// truncating conversion with the x86 out-of-range result
template <class Int, class S> Int trunc_int(S y) {
  const S lim = pow2<S>((int)sizeof(Int) * 8 - 1);
  if (!(y < lim && y >= -lim)) return std::numeric_limits<Int>::min();
  return (Int)y;
}

template <class S> void fwd_cast(typename Tr<S>::Int* ib, const S* fb, int n, int emax) {
  S s = pow2<S>((int)sizeof(S) * 8 - 2 - emax);
  for (int i = 0; i < n; ++i) ib[i] = trunc_int<typename Tr<S>::Int, S>(s * fb[i]);
}
template <class S> void inv_cast(const typename Tr<S>::Int* ib, S* fb, int n, int emax) {
  S s = pow2<S>(emax - ((int)sizeof(S) * 8 - 2));
  for (int i = 0; i < n; ++i) fb[i] = (S)(s * (S)ib[i]);
}

// ---------------------------------------------------------------------------
// block coders
// ---------------------------------------------------------------------------
template <class Int> uint32_t encode_int_block(BitWriter& s, int minbits, int maxbits, int maxprec, Int* ib, int dims) {
  typedef typename std::make_unsigned<Int>::type UInt;
  const UInt mask = sizeof(Int) == 4 ? (UInt)0xaaaaaaaau : (UInt)0xaaaaaaaaaaaaaaaaull;
  int n = 1 << (2 * dims);
  xform<Int>(ib, dims, fwd_lift<Int>, true);
  const uint8_t* perm = perm_table(dims);
  std::vector<UInt> ub(n);
  for (int i = 0; i < n; ++i) ub[i] = int2uint<Int, UInt>(ib[perm[i]], mask);
  int bits = (int)encode_ints<UInt>(s, (uint32_t)std::max(maxbits, 0), (uint32_t)maxprec, ub.data(), n);
  if (bits < minbits) {
    s.pad(minbits - bits);
    bits = minbits;
  }
  return (uint32_t)bits;
}

template <class Int> uint32_t rev_encode_int_block(BitWriter& s, int minbits, int maxbits, int maxprec, Int* ib, int dims) {
  typedef typename std::make_unsigned<Int>::type UInt;
  const UInt mask = sizeof(Int) == 4 ? (UInt)0xaaaaaaaau : (UInt)0xaaaaaaaaaaaaaaaaull;
  const int pbits = sizeof(Int) == 4 ? 5 : 6;
  int n = 1 << (2 * dims);
  int bits = pbits;
  xform<Int>(ib, dims, rev_fwd_lift<Int>, true);
  const uint8_t* perm = perm_table(dims);
  std::vector<UInt> ub(n);
  UInt m = 0;
  for (int i = 0; i < n; ++i) {
    ub[i] = int2uint<Int, UInt>(ib[perm[i]], mask);
    m |= ub[i];
  }
  // planes from the top down to the lowest set bit of any coefficient
  uint32_t prec = m ? (uint32_t)(sizeof(UInt) * 8 - __builtin_ctzll((unsigned long long)m)) : 0;
  prec = std::min(prec, (uint32_t)maxprec);
  prec = std::max(prec, 1u);
  s.bits(prec - 1, pbits);
  bits += (int)encode_ints<UInt>(s, (uint32_t)std::max(maxbits - bits, 0), prec, ub.data(), n);
  if (bits < minbits) {
    s.pad(minbits - bits);
    bits = minbits;
  }
  return (uint32_t)bits;
}

template <class Int> uint32_t decode_int_block(BitReader& s, int minbits, int maxbits, int maxprec, Int* ib, int dims, bool rev) {
  typedef typename std::make_unsigned<Int>::type UInt;
  const UInt mask = sizeof(Int) == 4 ? (UInt)0xaaaaaaaau : (UInt)0xaaaaaaaaaaaaaaaaull;
  const int pbits = sizeof(Int) == 4 ? 5 : 6;
  int n = 1 << (2 * dims);
  int bits = 0;
  uint32_t prec = (uint32_t)maxprec;
  if (rev) {
    prec = (uint32_t)s.bits(pbits) + 1;
    bits += pbits;
  }
  std::vector<UInt> ub(n);
  bits += (int)decode_ints<UInt>(s, (uint32_t)std::max(maxbits - bits, 0), prec, ub.data(), n);
  if (bits < minbits) {
    s.skip(minbits - bits);
    bits = minbits;
  }
  const uint8_t* perm = perm_table(dims);
  for (int i = 0; i < n; ++i) ib[perm[i]] = uint2int<Int, UInt>(ub[i], mask);
  xform<Int>(ib, dims, rev ? rev_inv_lift<Int> : inv_lift<Int>, false);
  return (uint32_t)bits;
}

struct Cfg {
  uint32_t minbits, maxbits, maxprec;
  int minexp;
  bool reversible() const { return minexp < kMinExp; }
};

template <class S> void encode_float_block(BitWriter& s, const Cfg& c, const S* fb, int dims) {
  typedef typename Tr<S>::Int Int;
  const int ebits = Tr<S>::ebits, ebias = Tr<S>::ebias;
  int n = 1 << (2 * dims);
  std::vector<Int> ib(n);
  int emax = block_exponent<S>(fb, n);
  if (c.reversible()) {
    uint32_t bits = 0;
    if (emax != -ebias) fwd_cast<S>(ib.data(), fb, n, emax);
    else std::fill(ib.begin(), ib.end(), (Int)0);
    std::vector<S> g(n);
    if (emax != -ebias) inv_cast<S>(ib.data(), g.data(), n, emax);
    else std::fill(g.begin(), g.end(), (S)0);
    if (std::memcmp(g.data(), fb, n * sizeof(S)) == 0) {
      uint32_t e = (uint32_t)(emax + ebias);
      if (e) {
        bits += 2;
        s.bits(1, 2);
        bits += ebits;
        s.bits(e, ebits);
      } else {
        s.bit(0);
        return;
      }
    } else {
      // sign-magnitude bit patterns to two's complement (negatives: x ^ INT_MAX,
      // so that -0.0 maps to -1 and stays distinct from +0.0)
      const Int mx = std::numeric_limits<Int>::max();
      for (int i = 0; i < n; ++i) {
        Int x;
        std::memcpy(&x, &fb[i], sizeof(x));
        ib[i] = x < 0 ? (Int)(x ^ mx) : x;
      }
      bits += 2;
      s.bits(3, 2);
    }
    rev_encode_int_block<Int>(s, (int)c.minbits - (int)bits, (int)c.maxbits - (int)bits, (int)c.maxprec, ib.data(), dims);
    return;
  }
  int maxprec = std::min<int>((int)c.maxprec, std::max(0, emax - c.minexp + 2 * (dims + 1)));
  uint32_t e = maxprec ? (uint32_t)(emax + ebias) : 0;
  uint32_t bits = 1;
  if (e) {
    bits += ebits;
    s.bits(2 * (uint64_t)e + 1, bits);
    fwd_cast<S>(ib.data(), fb, n, emax);
    encode_int_block<Int>(s, (int)c.minbits - (int)bits, (int)c.maxbits - (int)bits, maxprec, ib.data(), dims);
  } else {
    s.bit(0);
    if (c.minbits > bits) s.pad(c.minbits - bits);
  }
}

template <class S> void decode_float_block(BitReader& s, const Cfg& c, S* fb, int dims) {
  typedef typename Tr<S>::Int Int;
  const int ebits = Tr<S>::ebits, ebias = Tr<S>::ebias;
  int n = 1 << (2 * dims);
  std::vector<Int> ib(n);
  if (c.reversible()) {
    uint32_t bits = 1;
    if (s.bit()) {
      bits++;
      if (s.bit()) {
        decode_int_block<Int>(s, (int)c.minbits - (int)bits, (int)c.maxbits - (int)bits, (int)c.maxprec, ib.data(), dims, true);
        const Int mx = std::numeric_limits<Int>::max();
        for (int i = 0; i < n; ++i) {
          Int x = ib[i] < 0 ? (Int)(ib[i] ^ mx) : ib[i];
          std::memcpy(&fb[i], &x, sizeof(x));
        }
      } else {
        bits += ebits;
        int emax = (int)s.bits(ebits) - ebias;
        decode_int_block<Int>(s, (int)c.minbits - (int)bits, (int)c.maxbits - (int)bits, (int)c.maxprec, ib.data(), dims, true);
        if (emax != -ebias) inv_cast<S>(ib.data(), fb, n, emax);
        else for (int i = 0; i < n; ++i) fb[i] = 0;
      }
    } else {
      for (int i = 0; i < n; ++i) fb[i] = 0;
    }
    return;
  }
  uint32_t bits = 1;
  if (s.bit()) {
    bits += ebits;
    int emax = (int)s.bits(ebits) - ebias;
    int maxprec = std::min<int>((int)c.maxprec, std::max(0, emax - c.minexp + 2 * (dims + 1)));
    decode_int_block<Int>(s, (int)c.minbits - (int)bits, (int)c.maxbits - (int)bits, maxprec, ib.data(), dims, false);
    inv_cast<S>(ib.data(), fb, n, emax);
  } else {
    for (int i = 0; i < n; ++i) fb[i] = 0;
    if (c.minbits > bits) s.skip(c.minbits - bits);
  }
}

// gather with the reference padding rule for partial blocks
template <class T> void pad_line(T* p, uint32_t n, int s) {
  switch (n) {
    case 0: p[0 * s] = 0; [[fallthrough]];
    case 1: p[1 * s] = p[0 * s]; [[fallthrough]];
    case 2: p[2 * s] = p[1 * s]; [[fallthrough]];
    case 3: p[3 * s] = p[0 * s]; [[fallthrough]];
    default: break;
  }
}

template <class T> void gather(T* q, const T* data, const uint32_t* nd, const uint32_t* org, const uint32_t* ext, int dims) {
  int n = 1 << (2 * dims);
  std::fill(q, q + n, (T)0);
  uint64_t st[4] = {1, nd[0], (uint64_t)nd[0] * nd[1], (uint64_t)nd[0] * nd[1] * nd[2]};
  for (int i = 0; i < n; ++i) {
    uint32_t c[4] = {(uint32_t)(i & 3), (uint32_t)((i >> 2) & 3), (uint32_t)((i >> 4) & 3), (uint32_t)((i >> 6) & 3)};
    bool in = true;
    for (int a = 0; a < dims; ++a) in &= c[a] < ext[a];
    if (!in) continue;
    uint64_t o = 0;
    for (int a = 0; a < dims; ++a) o += (uint64_t)(org[a] + c[a]) * st[a];
    q[i] = data[o];
  }
  // pad along x (each defined line), then y, z, w (every line)
  for (int a = 0; a < dims; ++a) {
    int s = 1 << (2 * a);
    for (int i = 0; i < n; ++i) {
      if (((i >> (2 * a)) & 3) != 0) continue;
      bool line_defined = true;
      for (int b = a + 1; b < dims; ++b) line_defined &= (uint32_t)((i >> (2 * b)) & 3) < ext[b];
      if (!line_defined) continue;
      pad_line(q + i, ext[a], s);
    }
  }
}

// ---------------------------------------------------------------------------
// configurations, header and whole streams
// ---------------------------------------------------------------------------
Cfg cfg_of(const Mode& m, int type, int dims) {
  Cfg c{1, kMaxBits, 64, kMinExp};
  switch (m.kind) {
    case Mode::kRate: {
      uint32_t n = 1u << (2 * dims);
      uint32_t bits = (uint32_t)std::floor(n * m.rate + 0.5);
      if (type == kFloat) bits = std::max(bits, 1u + 8u);
      if (type == kDouble) bits = std::max(bits, 1u + 11u);
      // the header's long mode form holds 1..32768 bits per block
      if (bits < 1 || bits > 0x8000) fail(kInvalidArgument, "nd_zfp: fixed-rate mode needs 1..=32768 bits per block");
      c.minbits = c.maxbits = bits;
      c.maxprec = 64;
      c.minexp = kMinExp;
      break;
    }
    case Mode::kPrecision:
      c.maxprec = m.precision;
      break;
    case Mode::kAccuracy: {
      int emin = kMinExp;
      if (m.tolerance > 0) {
        // tolerance = x 2^e with 0.5 <= x < 1 (exact for normal doubles)
        uint64_t b = to_bits<double>(m.tolerance);
        int be = (int)((b >> 52) & 0x7FF);
        int e = be ? be - 1022 : 0;
        if (!be) {
          double t = m.tolerance;
          e = 0;
          while (t < 0.5) {
            t *= 2;
            --e;
          }
        }
        emin = e - 1;
      }
      c.minexp = emin;
      break;
    }
    case Mode::kReversible:
      c.minexp = kMinExp - 1;
      break;
  }
  return c;
}

uint64_t mode_word(const Mode& m, const Cfg& c, uint32_t* nbits) {
  *nbits = 12;
  // the all-default parameter set (precision 64, or tolerance 0) is the
  // library's "expert" mode and always takes the long form
  bool defaults = c.minbits == 1 && c.maxbits == kMaxBits && c.maxprec == 64 && c.minexp == kMinExp;
  if (!defaults) switch (m.kind) {
    case Mode::kRate:
      if (c.maxbits <= 2048) return c.maxbits - 1;
      break;
    case Mode::kPrecision:
      if (c.maxprec <= 128) return (c.maxprec - 1) + 2048;
      break;
    case Mode::kAccuracy:
      if (c.minexp <= 843) return (uint64_t)(c.minexp - kMinExp) + (2048 + 128 + 1);
      break;
    case Mode::kReversible:
      return 2048 + 128;
  }
  *nbits = 64;
  uint64_t mode = 0;
  uint64_t minbits = std::max(1u, std::min(c.minbits, 0x8000u)) - 1;
  uint64_t maxbits = std::max(1u, std::min(c.maxbits, 0x8000u)) - 1;
  uint64_t maxprec = std::max(1u, std::min(c.maxprec, 0x80u)) - 1;
  uint64_t minexp = (uint64_t)std::max(0, std::min(c.minexp + 16495, 0x7fff));
  mode <<= 15; mode += minexp;
  mode <<= 7; mode += maxprec;
  mode <<= 15; mode += maxbits;
  mode <<= 15; mode += minbits;
  mode <<= 12; mode += 0xfffu;
  return mode;
}

uint64_t field_meta(int type, int dims, const uint32_t* n) {
  uint64_t meta = 0;
  int fb = dims == 1 ? 48 : dims == 2 ? 24 : dims == 3 ? 16 : 12;
  for (int a = dims - 1; a >= 0; --a) {
    if (fb < 48 && ((uint64_t)(n[a] - 1) >> fb)) fail(kInvalidArgument, "zfp: dimensions exceed the ZFP header metadata bounds");
    meta <<= fb;
    meta += n[a] - 1;
  }
  meta <<= 2;
  meta += dims - 1;
  meta <<= 2;
  meta += type - 1;
  return meta;
}

template <class S> void compress_t(BitWriter& w, const Cfg& c, const S* data, int dims, const uint32_t* nd, bool is_float) {
  int n = 1 << (2 * dims);
  std::vector<S> q(n);
  uint32_t nb[4] = {1, 1, 1, 1};
  for (int a = 0; a < dims; ++a) nb[a] = (nd[a] + 3) / 4;
  for (uint32_t bw = 0; bw < nb[3]; ++bw)
    for (uint32_t bz = 0; bz < nb[2]; ++bz)
      for (uint32_t by = 0; by < nb[1]; ++by)
        for (uint32_t bx = 0; bx < nb[0]; ++bx) {
          uint32_t org[4] = {4 * bx, 4 * by, 4 * bz, 4 * bw}, ext[4];
          for (int a = 0; a < 4; ++a) ext[a] = a < dims ? std::min(4u, nd[a] - org[a]) : 4;
          gather(q.data(), data, nd, org, ext, dims);
          if constexpr (std::is_floating_point<S>::value) {
            encode_float_block<S>(w, c, q.data(), dims);
          } else {
            (void)is_float;
            if (c.reversible()) rev_encode_int_block<S>(w, (int)c.minbits, (int)c.maxbits, (int)c.maxprec, q.data(), dims);
            else encode_int_block<S>(w, (int)c.minbits, (int)c.maxbits, (int)c.maxprec, q.data(), dims);
          }
        }
}

template <class S> void decompress_block_t(BitReader& r, const Cfg& c, S* q, int dims) {
  if constexpr (std::is_floating_point<S>::value) {
    decode_float_block<S>(r, c, q, dims);
  } else {
    decode_int_block<S>(r, (int)c.minbits, (int)c.maxbits, (int)c.maxprec, q, dims, c.reversible());
  }
}

template <class S> void decompress_t(BitReader& r, const Cfg& c, S* out, int dims, const uint32_t* nd) {
  int n = 1 << (2 * dims);
  std::vector<S> q(n);
  uint32_t nb[4] = {1, 1, 1, 1};
  for (int a = 0; a < dims; ++a) nb[a] = (nd[a] + 3) / 4;
  uint64_t st[4] = {1, nd[0], (uint64_t)nd[0] * nd[1], (uint64_t)nd[0] * nd[1] * nd[2]};
  for (uint32_t bw = 0; bw < nb[3]; ++bw)
    for (uint32_t bz = 0; bz < nb[2]; ++bz)
      for (uint32_t by = 0; by < nb[1]; ++by)
        for (uint32_t bx = 0; bx < nb[0]; ++bx) {
          decompress_block_t<S>(r, c, q.data(), dims);
          uint32_t org[4] = {4 * bx, 4 * by, 4 * bz, 4 * bw};
          for (int i = 0; i < n; ++i) {
            uint32_t cc[4] = {(uint32_t)(i & 3), (uint32_t)((i >> 2) & 3), (uint32_t)((i >> 4) & 3), (uint32_t)((i >> 6) & 3)};
            bool in = true;
            uint64_t o = 0;
            for (int a = 0; a < dims; ++a) {
              in &= org[a] + cc[a] < nd[a];
              o += (uint64_t)(org[a] + cc[a]) * st[a];
            }
            if (in) out[o] = q[i];
          }
        }
}

void rev_dims(const std::vector<size_t>& shape, uint32_t* nd) {
  for (int a = 0; a < 4; ++a) nd[a] = 0;
  for (size_t a = 0; a < shape.size(); ++a) nd[a] = (uint32_t)shape[shape.size() - 1 - a];
}

}  // namespace

const uint8_t* perm_of(int dims) { return perm_table(dims); }

std::vector<uint8_t> compress(const void* data, int type, const std::vector<size_t>& shape, const Mode& mode,
                              uint64_t* header_bits) {
  if (shape.empty() || shape.size() > 4) fail(kInvalidArgument, "nd_zfp: shape must have 1..=4 dimensions");
  for (size_t d : shape)
    if (d == 0) fail(kInvalidArgument, "nd_zfp: shape has a zero extent");
  int dims = (int)shape.size();
  uint32_t nd[4];
  rev_dims(shape, nd);
  Cfg c = cfg_of(mode, type, dims);
  BitWriter w;
  uint64_t meta = field_meta(type, dims, nd);
  w.bits('z', 8);
  w.bits('f', 8);
  w.bits('p', 8);
  w.bits(5, 8);
  w.bits(meta, 52);
  uint32_t mb;
  uint64_t mw = mode_word(mode, c, &mb);
  w.bits(mw, mb);
  if (header_bits) *header_bits = w.nbits;
  switch (type) {
    case kInt32: compress_t<int32_t>(w, c, (const int32_t*)data, dims, nd, false); break;
    case kInt64: compress_t<int64_t>(w, c, (const int64_t*)data, dims, nd, false); break;
    case kFloat: compress_t<float>(w, c, (const float*)data, dims, nd, true); break;
    default: compress_t<double>(w, c, (const double*)data, dims, nd, true); break;
  }
  w.flush();
  return w.bytes;
}

void decompress(const uint8_t* bytes, size_t len, int type, const std::vector<size_t>& shape, const Mode& mode,
                void* out) {
  if (shape.empty() || shape.size() > 4) fail(kInvalidArgument, "nd_zfp: shape must have 1..=4 dimensions");
  int dims = (int)shape.size();
  uint32_t nd[4];
  rev_dims(shape, nd);
  Cfg c = cfg_of(mode, type, dims);
  BitReader r(bytes, len);
  if (r.bits(8) != 'z' || r.bits(8) != 'f' || r.bits(8) != 'p' || r.bits(8) != 5)
    fail(kCodestream, "nd_zfp: bad ZFP magic");
  uint64_t meta = r.bits(52);
  if (meta != field_meta(type, dims, nd)) fail(kCodestream, "nd_zfp: stream header does not declare the expected field");
  uint32_t mb;
  uint64_t mw = mode_word(mode, c, &mb);
  uint64_t got = r.bits(12);
  if (got == 0xfff) got |= r.bits(52) << 12;
  if (got != mw) fail(kCodestream, "nd_zfp: stream compression parameters do not match the configuration");
  if (mode.kind == Mode::kRate) {
    uint64_t blocks = 1;
    for (int a = 0; a < dims; ++a) blocks *= (nd[a] + 3) / 4;
    uint64_t total = r.pos + blocks * c.maxbits;
    if (len != ((total + 63) / 64) * 8) fail(kCodestream, "nd_zfp: fixed-rate stream length differs from the computed size");
  }
  switch (type) {
    case kInt32: decompress_t<int32_t>(r, c, (int32_t*)out, dims, nd); break;
    case kInt64: decompress_t<int64_t>(r, c, (int64_t*)out, dims, nd); break;
    case kFloat: decompress_t<float>(r, c, (float*)out, dims, nd); break;
    default: decompress_t<double>(r, c, (double*)out, dims, nd); break;
  }
}

// Where each block of a stream begins (absolute bit positions, found by
// walking the blocks the way decompress does), with the end of the last
// block as a final entry: what a block index of the stream records.
std::vector<uint64_t> block_starts(const uint8_t* bytes, size_t len, int type, const std::vector<size_t>& shape,
                                   const Mode& mode) {
  if (shape.empty() || shape.size() > 4) fail(kInvalidArgument, "nd_zfp: shape must have 1..=4 dimensions");
  int dims = (int)shape.size();
  uint32_t nd[4];
  rev_dims(shape, nd);
  Cfg c = cfg_of(mode, type, dims);
  BitReader r(bytes, len);
  if (r.bits(8) != 'z' || r.bits(8) != 'f' || r.bits(8) != 'p' || r.bits(8) != 5)
    fail(kCodestream, "nd_zfp: bad ZFP magic");
  if (r.bits(52) != field_meta(type, dims, nd)) fail(kCodestream, "nd_zfp: stream header does not declare the expected field");
  uint64_t got = r.bits(12);
  if (got == 0xfff) got |= r.bits(52) << 12;
  uint64_t blocks = 1;
  for (int a = 0; a < dims; ++a) blocks *= (nd[a] + 3) / 4;
  std::vector<uint64_t> out;
  auto walk = [&](auto* q) {
    using S = std::remove_pointer_t<decltype(q)>;
    for (uint64_t k = 0; k < blocks; ++k) {
      out.push_back(r.pos);
      decompress_block_t<S>(r, c, q, dims);
    }
    out.push_back(r.pos);
  };
  switch (type) {
    case kInt32: { int32_t q[256]; walk(q); break; }
    case kInt64: { int64_t q[256]; walk(q); break; }
    case kFloat: { float q[256]; walk(q); break; }
    default: { double q[256]; walk(q); break; }
  }
  return out;
}

// lib.rs:386-500 BrickIndex::fixed_rate: brick k spans bits
// [header_bits + k * bits_per_brick, + bits_per_brick); the stream is padded
// to a 64-bit word.
BrickIndex brick_index(const std::vector<size_t>& shape, int type, double rate) {
  if (shape.empty() || shape.size() > 4) fail(kInvalidArgument, "nd_zfp: shape must have 1..=4 dimensions");
  for (size_t d : shape)
    if (d == 0) fail(kInvalidArgument, "nd_zfp: shape has a zero extent");
  Mode m;
  m.kind = Mode::kRate;
  m.rate = rate;
  int dims = (int)shape.size();
  Cfg c = cfg_of(m, type, dims);
  uint32_t mb;
  (void)mode_word(m, c, &mb);
  BrickIndex bi;
  bi.header_bits = 32 + 52 + mb;
  bi.bits_per_brick = c.maxbits;
  uint64_t nbricks = 1;
  for (size_t d : shape) {
    bi.grid.push_back((d + 3) / 4);
    nbricks *= (d + 3) / 4;
  }
  bi.stream_len = (size_t)(((bi.header_bits + nbricks * bi.bits_per_brick + 63) / 64) * 8);
  return bi;
}

// BrickIndex::linear (lib.rs:453-474): row-major over the grid, last axis fastest
uint64_t brick_linear(const BrickIndex& bi, const std::vector<size_t>& coords) {
  if (coords.size() != bi.grid.size()) fail(kInvalidArgument, "nd_zfp: brick coordinates do not match the grid");
  uint64_t k = 0;
  for (size_t a = 0; a < coords.size(); ++a) {
    if (coords[a] >= bi.grid[a]) fail(kInvalidArgument, "nd_zfp: brick coordinates exceed the grid");
    k = k * bi.grid[a] + coords[a];
  }
  return k;
}

// BrickReader::fixed_rate + brick (lib.rs:509-578): the stream length and
// full header are checked as decompress checks them, then a seek to brick k
// and one block decode, stored clipped to the array bounds in C order.
std::vector<uint8_t> decompress_brick(const uint8_t* bytes, size_t len, int type, const std::vector<size_t>& shape,
                                      double rate, const std::vector<size_t>& brick, std::vector<size_t>* brick_shape) {
  BrickIndex bi = brick_index(shape, type, rate);
  if (len != bi.stream_len) fail(kCodestream, "nd_zfp: fixed-rate stream length differs from the computed size");
  int dims = (int)shape.size();
  uint32_t nd[4];
  rev_dims(shape, nd);
  Mode m;
  m.kind = Mode::kRate;
  m.rate = rate;
  Cfg c = cfg_of(m, type, dims);
  BitReader r(bytes, len);
  if (r.bits(8) != 'z' || r.bits(8) != 'f' || r.bits(8) != 'p' || r.bits(8) != 5)
    fail(kCodestream, "nd_zfp: bad ZFP magic");
  if (r.bits(52) != field_meta(type, dims, nd)) fail(kCodestream, "nd_zfp: stream header does not declare the expected field");
  uint32_t mb;
  uint64_t mw = mode_word(m, c, &mb);
  uint64_t got = r.bits(12);
  if (got == 0xfff) got |= r.bits(52) << 12;
  if (got != mw) fail(kCodestream, "nd_zfp: stream compression parameters do not match the configuration");
  if (r.pos != bi.header_bits) fail(kCodestream, "nd_zfp: stream header has an unexpected length");
  uint64_t k = brick_linear(bi, brick);
  r.pos = bi.header_bits + k * bi.bits_per_brick;
  brick_shape->clear();
  size_t n = 1;
  for (size_t a = 0; a < shape.size(); ++a) {
    brick_shape->push_back(std::min<size_t>(4, shape[a] - 4 * brick[a]));
    n *= brick_shape->back();
  }
  size_t es = (type == kInt32 || type == kFloat) ? 4 : 8;
  std::vector<uint8_t> out(n * es);
  auto store = [&](auto* q) {
    using S = std::remove_pointer_t<decltype(q)>;
    decompress_block_t<S>(r, c, q, dims);
    // block value i sits at (x, y, z, w) = (i & 3, i >> 2 & 3, ...), x = the last axis
    size_t o = 0;
    std::vector<size_t> idx(shape.size(), 0);
    for (size_t j = 0; j < n; ++j) {
      uint32_t i = 0;
      for (size_t a = 0; a < shape.size(); ++a) i |= (uint32_t)idx[a] << (2 * (shape.size() - 1 - a));
      std::memcpy(&out[o], &q[i], es);
      o += es;
      for (size_t a = shape.size(); a-- > 0;) {
        if (++idx[a] < (*brick_shape)[a]) break;
        idx[a] = 0;
      }
    }
  };
  switch (type) {
    case kInt32: { int32_t q[256]; store(q); break; }
    case kInt64: { int64_t q[256]; store(q); break; }
    case kFloat: { float q[256]; store(q); break; }
    default: { double q[256]; store(q); break; }
  }
  return out;
}

// chunk.rs:446-540 NdZfpBrickDecoder / decode_chunk_brick: the brick over the
// effective shape, narrow integers demoted from the int32 reader.
std::vector<uint8_t> decode_chunk_brick(const uint8_t* bytes, size_t len, const std::vector<size_t>& shape, int dtype,
                                        double rate, const std::vector<size_t>& brick,
                                        std::vector<size_t>* brick_shape) {
  if (dtype == kF32) return decompress_brick(bytes, len, kFloat, shape, rate, brick, brick_shape);
  if (dtype == kF64) return decompress_brick(bytes, len, kDouble, shape, rate, brick, brick_shape);
  if (dtype == kI32) return decompress_brick(bytes, len, kInt32, shape, rate, brick, brick_shape);
  if (dtype == kI64) return decompress_brick(bytes, len, kInt64, shape, rate, brick, brick_shape);
  std::vector<uint8_t> v = decompress_brick(bytes, len, kInt32, shape, rate, brick, brick_shape);
  size_t n = v.size() / 4, es = dtype_size(dtype);
  std::vector<uint8_t> out(n * es);
  for (size_t i = 0; i < n; ++i) {
    int32_t x;
    std::memcpy(&x, &v[4 * i], 4);
    switch (dtype) {
      case kU8: out[i] = (uint8_t)std::min(std::max((x >> 23) + 0x80, 0), 0xff); break;
      case kI8: out[i] = (uint8_t)(int8_t)std::min(std::max(x >> 23, -0x80), 0x7f); break;
      case kU16: {
        uint16_t u = (uint16_t)std::min(std::max((x >> 15) + 0x8000, 0), 0xffff);
        out[2 * i] = u & 0xff;
        out[2 * i + 1] = u >> 8;
        break;
      }
      case kI16: {
        uint16_t u = (uint16_t)(int16_t)std::min(std::max(x >> 15, -0x8000), 0x7fff);
        out[2 * i] = u & 0xff;
        out[2 * i + 1] = u >> 8;
        break;
      }
      default: fail(kUnsupported, "zfp has no path for this dtype");
    }
  }
  return out;
}

// chunk.rs:305-345: narrow integers promoted to int32 (and back)
std::vector<uint8_t> encode_chunk(const uint8_t* bytes, size_t len, const std::vector<size_t>& shape, int dtype,
                                  const Mode& mode) {
  size_t n = 1;
  for (size_t d : shape) n *= d;
  size_t es = dtype_size(dtype);
  if (len != n * es) fail(kInvalidArgument, "zfp: chunk byte length does not match its shape and dtype");
  if (dtype == kF32) return compress(bytes, kFloat, shape, mode, nullptr);
  if (dtype == kF64) return compress(bytes, kDouble, shape, mode, nullptr);
  if (dtype == kI32) return compress(bytes, kInt32, shape, mode, nullptr);
  if (dtype == kI64) return compress(bytes, kInt64, shape, mode, nullptr);
  std::vector<int32_t> v(n);
  for (size_t i = 0; i < n; ++i) {
    const uint8_t* p = bytes + i * es;
    switch (dtype) {
      case kU8: v[i] = ((int32_t)p[0] - 0x80) * (1 << 23); break;
      case kI8: v[i] = (int32_t)(int8_t)p[0] * (1 << 23); break;
      case kU16: v[i] = ((int32_t)(p[0] | p[1] << 8) - 0x8000) * (1 << 15); break;
      case kI16: v[i] = (int32_t)(int16_t)(p[0] | p[1] << 8) * (1 << 15); break;
      default: fail(kUnsupported, "zfp has no path for this dtype");
    }
  }
  return compress(v.data(), kInt32, shape, mode, nullptr);
}

std::vector<uint8_t> decode_chunk(const uint8_t* bytes, size_t len, const std::vector<size_t>& shape, int dtype,
                                  const Mode& mode) {
  size_t n = 1;
  for (size_t d : shape) n *= d;
  size_t es = dtype_size(dtype);
  std::vector<uint8_t> out(n * es);
  if (dtype == kF32) decompress(bytes, len, kFloat, shape, mode, out.data());
  else if (dtype == kF64) decompress(bytes, len, kDouble, shape, mode, out.data());
  else if (dtype == kI32) decompress(bytes, len, kInt32, shape, mode, out.data());
  else if (dtype == kI64) decompress(bytes, len, kInt64, shape, mode, out.data());
  else {
    std::vector<int32_t> v(n);
    decompress(bytes, len, kInt32, shape, mode, v.data());
    for (size_t i = 0; i < n; ++i) {
      int32_t x = v[i];
      switch (dtype) {
        case kU8: out[i] = (uint8_t)std::min(std::max((x >> 23) + 0x80, 0), 0xff); break;
        case kI8: out[i] = (uint8_t)(int8_t)std::min(std::max(x >> 23, -0x80), 0x7f); break;
        case kU16: {
          uint16_t u = (uint16_t)std::min(std::max((x >> 15) + 0x8000, 0), 0xffff);
          out[2 * i] = u & 0xff;
          out[2 * i + 1] = u >> 8;
          break;
        }
        default: {
          uint16_t u = (uint16_t)(int16_t)std::min(std::max(x >> 15, -0x8000), 0x7fff);
          out[2 * i] = u & 0xff;
          out[2 * i + 1] = u >> 8;
        }
      }
    }
  }
  return out;
}

}  // namespace zfp
}  // namespace orc
