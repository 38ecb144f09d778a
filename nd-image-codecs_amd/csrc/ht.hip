// HT (FBCOT) cleanup-pass block coder on the device, one lane per code-block.
//
// Decode restates crates/ndic-htj2k/src/block/decode.rs:54-396 with its two
// steps fused per quad pair: the VLC/MEL stream (read backward from the end
// of the cleanup segment) and the MagSgn stream (read forward from its start)
// are independent, so decoding a pair's VLC codewords and then its MagSgn
// bits gives the reference's output while keeping only one quad row of state
// per lane: the row above's significance patterns (context c_q) and the
// bit-lengths of its bottom-sample magnitudes (exponent bound E_max).  Both
// rows live in LDS laid out [entry][lane], so a wave touching entry k hits
// consecutive bytes.  The pair loop is uniform across the wave; lanes whose
// block is narrower or shorter are masked.
//
// Encode restates crates/ndic-htj2k/src/block/encode.rs:113-331 with the MEL,
// VLC and MagSgn packers of mel.rs:175-262 and streams.rs:267-441, writing
// the three byte streams into a per-block slot; the codestream assembler
// stitches them together (ht_seg_byte).
//
// The CxtVLC/UVLC tables are derived on the host from the T.814 rows and
// live in constant memory.
#include <algorithm>
#include <mutex>
#include <vector>

#include "ht.hpp"
#include "codestream.hpp"
#include "tune.hpp"

namespace ndic {

namespace {
#include "tables/ht_cxtvlc_rows.inc"
}

__constant__ uint16_t c_dec0[1024], c_dec1[1024], c_enc0[2048], c_enc1[2048], c_uvlc0[320], c_uvlc1[256];
// 32-bit copies of the decode tables for the wave-uniform decoder: a scalar
// load reads whole dwords, so its lookups stay on the scalar unit
__constant__ uint32_t c_dec0w[1024], c_dec1w[1024], c_uvlc0w[320], c_uvlc1w[256];

namespace {

__constant__ uint8_t c_mel_exp[13] = {0, 0, 0, 1, 1, 1, 2, 2, 2, 3, 3, 4, 5};
__constant__ uint32_t c_mel_expw[13] = {0, 0, 0, 1, 1, 1, 2, 2, 2, 3, 3, 4, 5};

// ---------------------------------------------------------------------------
// host: table derivation (T.814 Annex C rows -> lookup tables)
// ---------------------------------------------------------------------------
template <size_t N> void derive_dec(const unsigned char (&src)[N][7], uint16_t* tbl) {
  for (int i = 0; i < 1024; ++i) {
    int cwd = i & 0x7F, cq = i >> 7;
    tbl[i] = 0;
    for (size_t r = 0; r < N; ++r)
      if (src[r][0] == cq && src[r][5] == (cwd & ((1 << src[r][6]) - 1)))
        tbl[i] = uint16_t((src[r][1] << 4) | (src[r][2] << 3) | (src[r][3] << 12) | (src[r][4] << 8) | src[r][6]);
  }
}

template <size_t N> void derive_enc(const unsigned char (&src)[N][7], uint16_t* tbl) {
  for (int i = 0; i < 2048; ++i) {
    int cq = i >> 8, rho = (i >> 4) & 0xF, emb = i & 0xF;
    tbl[i] = 0;
    if ((emb & rho) != emb || (rho == 0 && cq == 0)) continue;
    const unsigned char* best = nullptr;
    int best_ones = -1;
    for (size_t r = 0; r < N; ++r) {
      const unsigned char* e = src[r];
      if (e[0] != cq || e[1] != rho) continue;
      if (emb) {
        if (e[2] == 1 && (emb & e[3]) == e[4] && __builtin_popcount(e[3]) >= best_ones) {
          best = e;
          best_ones = __builtin_popcount(e[3]);
        }
      } else if (e[2] == 0) {
        best = e;
        break;
      }
    }
    if (best) tbl[i] = uint16_t((best[5] << 8) | (best[6] << 4) | best[3]);
  }
}

// UVLC prefixes of T.814: "1" -> 1, "01" -> 2, "001"+1 bit -> 3..4, "000"+5 bits -> 5..36.
void uvlc_pfx(int b, int* len, int* suf, int* u) {
  if (b & 1) { *len = 1; *suf = 0; *u = 1; }
  else if (b & 2) { *len = 2; *suf = 0; *u = 2; }
  else if (b & 4) { *len = 3; *suf = 1; *u = 3; }
  else { *len = 3; *suf = 5; *u = 5; }
}

void derive_uvlc(uint16_t* tbl, int modes, bool initial) {
  auto ent = [](int tp, int ts, int s0, int u0, int u1) {
    return uint16_t(tp | (ts << 3) | (s0 << 7) | (u0 << 10) | (u1 << 13));
  };
  for (int i = 0; i < modes * 64; ++i) {
    int mode = i >> 6, v = i & 0x3F;
    tbl[i] = 0;
    if (mode == 0) continue;
    int l0, s0, u0, l1, s1, u1;
    uvlc_pfx(v & 7, &l0, &s0, &u0);
    if (mode <= 2) {
      tbl[i] = mode == 1 ? ent(l0, s0, s0, u0, 0) : ent(l0, s0, 0, 0, u0);
      continue;
    }
    uvlc_pfx((v >> l0) & 7, &l1, &s1, &u1);
    if (mode == 3 && initial && l0 == 3) tbl[i] = ent(l0 + 1, s0, s0, u0, ((v >> l0) & 1) + 1);
    else if (mode == 4) tbl[i] = ent(l0 + l1, s0 + s1, s0, u0 + 2, u1 + 2);
    else tbl[i] = ent(l0 + l1, s0 + s1, s0, u0, u1);
  }
}

// ---------------------------------------------------------------------------
// device: stream readers (mel.rs:14-173, streams.rs:20-265)
// ---------------------------------------------------------------------------
// Segment reads (global memory; a 32-bit read need not be aligned).
typedef const __attribute__((address_space(1))) uint32_t* gwords;
__device__ __forceinline__ uint32_t ld32(const uint8_t* p) { return *(gwords)(uintptr_t)p; }
__device__ __forceinline__ uint32_t ld8(const uint8_t* p) {
  return *(const __attribute__((address_space(1))) uint8_t*)(uintptr_t)p;
}
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ u32x4 ld128(const uint8_t* p) {
  return *(const __attribute__((address_space(1))) u32x4*)(uintptr_t)p;
}
__device__ __forceinline__ void st8(uint8_t* p, uint8_t v) { *(__attribute__((address_space(1))) uint8_t*)(uintptr_t)p = v; }
__device__ __forceinline__ void st32(void* p, uint32_t v) { *(__attribute__((address_space(1))) uint32_t*)(uintptr_t)p = v; }
__device__ __forceinline__ int32_t ldi32(const int32_t* p) { return *(const __attribute__((address_space(1))) int32_t*)(uintptr_t)p; }
__device__ __forceinline__ uint32_t ldi32u(const uint32_t* p) { return *(const __attribute__((address_space(1))) uint32_t*)(uintptr_t)p; }

struct MelDec {
  const uint8_t* data;
  int size, bits, k;
  bool unstuff;
  uint64_t tmp, runs;
  uint32_t num_runs;
  __device__ void init(const uint8_t* coded, uint32_t lcup, uint32_t scup) {
    data = coded + lcup - scup;
    size = (int)scup - 1;
    bits = 0;
    k = 0;
    unstuff = false;
    tmp = 0;
    runs = 0;
    num_runs = 0;
  }
  __device__ void read() {
    if (bits > 32) return;
    uint32_t val = 0xFFFFFFFFu;
    if (size > 4) {
      val = ld32(data);
      data += 4;
      size -= 4;
    } else if (size > 0) {
      uint32_t i = 0;
      while (size > 1) {
        uint32_t v = ld8(data++);
        uint32_t m = ~(0xFFu << i);
        val = (val & m) | (v << i);
        size -= 1;
        i += 8;
      }
      uint32_t v = ld8(data++) | 0xF;
      uint32_t m = ~(0xFFu << i);
      val = (val & m) | (v << i);
      size -= 1;
    }
    int nb = 32 - int(unstuff);
    uint32_t t = val & 0xFF;
    bool us = (val & 0xFF) == 0xFF;
    nb -= int(us);
    t <<= 8 - int(us);
    t |= (val >> 8) & 0xFF;
    us = ((val >> 8) & 0xFF) == 0xFF;
    nb -= int(us);
    t <<= 8 - int(us);
    t |= (val >> 16) & 0xFF;
    us = ((val >> 16) & 0xFF) == 0xFF;
    nb -= int(us);
    t <<= 8 - int(us);
    t |= (val >> 24) & 0xFF;
    unstuff = ((val >> 24) & 0xFF) == 0xFF;
    tmp |= uint64_t(t) << (64 - nb - bits);
    bits += nb;
  }
  __device__ void decode() {
    if (bits < 6) read();
    while (bits >= 6 && num_runs < 8) {
      uint32_t e = c_mel_exp[k];
      uint32_t run;
      if (tmp & (1ull << 63)) {
        run = ((1u << e) - 1) << 1;
        k = k + 1 < 12 ? k + 1 : 12;
        tmp <<= 1;
        bits -= 1;
      } else {
        uint32_t count = uint32_t(tmp >> (63 - e)) & ((1u << e) - 1);
        run = (count << 1) | 1;
        k = k > 0 ? k - 1 : 0;
        tmp <<= e + 1;
        bits -= int(e) + 1;
      }
      uint32_t sh = num_runs * 7;
      runs &= ~(0x3Full << sh);
      runs |= uint64_t(run) << sh;
      num_runs += 1;
    }
  }
  __device__ int get_run() {
    if (num_runs == 0) decode();
    int t = int(runs & 0x7F);
    runs >>= 7;
    num_runs -= 1;
    return t;
  }
};

// The MagSgn and VLC readers take their bytes 16 at a time: a queue of four
// 32-bit words (q0 first) is refilled with one 16-byte load as soon as it
// empties, so the load is in flight while the words before it are decoded,
// and a lane touches each cache line of its stream a quarter as often.  The
// last < 16 bytes of a stream go through the 4-byte / byte path as before.
struct RevDec {
  const uint8_t* data;
  long pos;
  int size;
  uint64_t tmp;
  uint32_t bits;
  bool unstuff;
  uint32_t q0, q1, q2, q3, nq;  // bytes pos+1.. backward: q0 = the 4 highest
  __device__ void refill() {
    if (size >= 16 && pos >= 15) {
      u32x4 v = ld128(data + pos - 15);
      q0 = v.w;
      q1 = v.z;
      q2 = v.y;
      q3 = v.x;
      nq = 4;
      pos -= 16;
      size -= 16;
    }
  }
  __device__ void init_vlc(const uint8_t* coded, uint32_t lcup, uint32_t scup) {
    data = coded;
    uint8_t d = (uint8_t)ld8(coded + lcup - 2);
    tmp = d >> 4;
    pos = (long)lcup - 3;
    size = (int)scup - 2;
    bits = 4 - uint32_t((tmp & 7) == 7);
    unstuff = (d | 0xF) > 0x8F;
    nq = 0;
    refill();
    read();
  }
  __device__ void init_mrp(const uint8_t* coded, uint32_t lcup, uint32_t len2) {
    data = coded;
    pos = (long)(lcup + len2) - 1;
    size = (int)len2;
    tmp = 0;
    bits = 0;
    unstuff = true;
    nq = 0;
    refill();
    read();
  }
  __device__ void read() {
    if (bits > 32) return;
    uint32_t val = 0;
    if (nq) {
      val = q0;
      q0 = q1;
      q1 = q2;
      q2 = q3;
      if (--nq == 0) refill();
    } else if (size > 3 && pos >= 3) {
      val = ld32(data + pos - 3);
      pos -= 4;
      size -= 4;
    } else if (size > 0) {
      int i = 24;
      while (size > 0 && i >= 0) {
        uint32_t v = pos >= 0 ? ld8(data + pos) : 0;
        pos -= 1;
        val |= v << i;
        size -= 1;
        i -= 8;
      }
    }
    uint32_t t = val >> 24;
    uint32_t nb = 8 - uint32_t(unstuff && ((val >> 24) & 0x7F) == 0x7F);
    bool us = (val >> 24) > 0x8F;
    t |= ((val >> 16) & 0xFF) << nb;
    nb += 8 - uint32_t(us && ((val >> 16) & 0x7F) == 0x7F);
    us = ((val >> 16) & 0xFF) > 0x8F;
    t |= ((val >> 8) & 0xFF) << nb;
    nb += 8 - uint32_t(us && ((val >> 8) & 0x7F) == 0x7F);
    us = ((val >> 8) & 0xFF) > 0x8F;
    t |= (val & 0xFF) << nb;
    nb += 8 - uint32_t(us && (val & 0x7F) == 0x7F);
    us = (val & 0xFF) > 0x8F;
    tmp |= uint64_t(t) << bits;
    bits += nb;
    unstuff = us;
  }
  __device__ uint32_t fetch() {
    if (bits < 32) {
      read();
      if (bits < 32) read();
    }
    return uint32_t(tmp);
  }
  __device__ uint32_t advance(uint32_t n) {
    tmp = n >= 64 ? 0 : tmp >> n;
    bits = bits > n ? bits - n : 0;
    return uint32_t(tmp);
  }
};

template <uint32_t FILL> struct FwdDec {
  const uint8_t* data;
  int size;
  uint64_t tmp;
  uint32_t bits;
  bool unstuff;
  uint32_t q0, q1, q2, q3, nq;  // the 16 bytes before data, q0 first
  __device__ void refill() {
    if (size >= 16) {
      u32x4 v = ld128(data);
      q0 = v.x;
      q1 = v.y;
      q2 = v.z;
      q3 = v.w;
      nq = 4;
      data += 16;
      size -= 16;
    }
  }
  __device__ void init(const uint8_t* d, uint32_t len) {
    data = d;
    size = (int)len;
    tmp = 0;
    bits = 0;
    unstuff = false;
    nq = 0;
    refill();
    read();
  }
  __device__ void read() {
    if (bits > 32) return;
    uint32_t val;
    if (nq) {
      val = q0;
      q0 = q1;
      q1 = q2;
      q2 = q3;
      if (--nq == 0) refill();
    } else if (size > 3) {
      val = ld32(data);
      data += 4;
      size -= 4;
    } else if (size > 0) {
      uint32_t v = FILL, i = 0;
      while (size > 0) {
        uint32_t b = ld8(data++);
        uint32_t m = ~(0xFFu << i);
        v = (v & m) | (b << i);
        size -= 1;
        i += 8;
      }
      val = v;
    } else {
      val = FILL;
    }
    uint32_t nb = 8 - uint32_t(unstuff);
    uint32_t t = val & 0xFF;
    bool us = (val & 0xFF) == 0xFF;
    t |= ((val >> 8) & 0xFF) << nb;
    nb += 8 - uint32_t(us);
    us = ((val >> 8) & 0xFF) == 0xFF;
    t |= ((val >> 16) & 0xFF) << nb;
    nb += 8 - uint32_t(us);
    us = ((val >> 16) & 0xFF) == 0xFF;
    t |= ((val >> 24) & 0xFF) << nb;
    nb += 8 - uint32_t(us);
    unstuff = ((val >> 24) & 0xFF) == 0xFF;
    tmp |= uint64_t(t) << bits;
    bits += nb;
  }
  __device__ uint32_t fetch() {
    if (bits < 32) {
      read();
      if (bits < 32) read();
    }
    return uint32_t(tmp);
  }
  __device__ void advance(uint32_t n) {
    tmp = n >= 64 ? 0 : tmp >> n;
    bits = bits > n ? bits - n : 0;
  }
};

__device__ __forceinline__ uint32_t bitlen(uint32_t v) { return 32 - __clz(v); }

__device__ void report(const HtDecTask& t, int32_t code) {
  if (t.err) atomicCAS(t.err, 0, code);
}

constexpr int kRowCap = 34;  // quads per row (64 columns) + terminators
#ifndef NDIC_HT_STAGE_COLS
#define NDIC_HT_STAGE_COLS 8
#endif
// output columns staged per lane before a wave-wide flush (4, 8, 16 or 32)
constexpr uint32_t kStageCols = NDIC_HT_STAGE_COLS;
constexpr int kStagePitch = 65;  // LDS pitch of a staged sample row (bank-conflict free)

// Where a decoder looks up the CxtVLC and UVLC prefix tables: the constant
// arrays, or a workgroup's copy in LDS (a few hundred cycles closer than a
// vector-L1 line the lanes' byte streams keep evicting).
struct ConstTables {
  __device__ __forceinline__ uint32_t dec(bool initial, uint32_t i) const { return initial ? c_dec0[i] : c_dec1[i]; }
  __device__ __forceinline__ uint32_t uvlc(bool initial, uint32_t i) const { return initial ? c_uvlc0[i] : c_uvlc1[i]; }
};
typedef const __attribute__((address_space(3))) uint16_t* lds_u16p;
constexpr uint32_t kTabDec1 = 1024, kTabUv0 = 2048, kTabUv1 = 2048 + 320, kTabAll = 2048 + 320 + 256;
struct LdsTables {
  lds_u16p t;  // dec0 | dec1 | uvlc0 | uvlc1
  __device__ __forceinline__ uint32_t dec(bool initial, uint32_t i) const { return t[(initial ? 0 : kTabDec1) + i]; }
  __device__ __forceinline__ uint32_t uvlc(bool initial, uint32_t i) const {
    return t[(initial ? kTabUv0 : kTabUv1) + i];
  }
};

// Per-lane decoder of one code-block's cleanup pass (decode.rs:54-396).  The
// caller drives it through uniform row-pair / column-pair loops so that the
// whole wave can flush decoded rows together.
template <class Tab>
struct BlockDec {
  Tab tab;
  HtDecTask t;
  MelDec mel;
  RevDec vlc;
  FwdDec<0xFFFFFFFFu> ms;
  uint32_t p, shift, mmsbp2;
  int run;
  bool active;  // decoding (else the block's samples are zero)
  bool bad;
  bool refine;  // SigProp/MagRef follow: store raw sign-magnitude samples
  // row-pair state
  uint32_t c_q, prev_v_n, vp, q;

  // decode.rs:54-110: validation and stream set-up
  __device__ void start(const HtDecTask& task) {
    t = task;
    active = false;
    bad = false;
    refine = false;
    if (!(t.flags & kTaskIncluded) || !t.seg || t.w == 0 || t.h == 0) return;
    uint32_t num_passes = t.passes, missing_msbs = t.mmsb;
    if (num_passes > 1 && t.len2 == 0) num_passes = 1;
    if (num_passes > 3 || missing_msbs >= 30 || t.len1 < 2 || t.w > 64 || t.h > 64) {
      report(t, kErrBlock);
      return;
    }
    if (missing_msbs == 29) num_passes = 1;
    // blocks with SigProp/MagRef passes keep raw sign-magnitude samples
    // here; ht_refine_kernel finishes them
    refine = num_passes > 1;
    const uint8_t* coded = t.seg;
    const uint32_t lcup = t.len1;
    const uint32_t scup = (ld8(coded + lcup - 1) << 4) + (ld8(coded + lcup - 2) & 0xF);
    if (scup < 2 || scup > lcup || scup > 4079) {
      report(t, kErrBlock);
      return;
    }
    p = 30 - missing_msbs;
    shift = 31 - t.kmax;
    mmsbp2 = missing_msbs + 2;
    mel.init(coded, lcup, scup);
    vlc.init_vlc(coded, lcup, scup);
    ms.init(coded, lcup - scup);
    run = mel.get_run();
    active = true;
  }

  __device__ uint32_t value(uint32_t sm) const {
    if ((t.flags & kTaskSignMag) || refine) return sm;
    int32_t mag = (int32_t)((sm & 0x7FFFFFFFu) >> shift);
    return (uint32_t)((sm & 0x80000000u) ? -mag : mag);
  }

  __device__ void row_begin() {
    c_q = 0;
    prev_v_n = 0;
    vp = 0;
    q = 0;
  }

  // One pair of quads (columns x..x+3, rows y, y+1): decode.rs:115-396.
  // Samples go to stage[(r * kStageCols + col) * kStagePitch] (r = row - y,
  // col = column - window base).
  __device__ void pair(uint32_t x, uint32_t y, uint8_t* lrho, uint8_t* lbl, uint32_t* stage, uint32_t xbase) {
    auto RHO = [&](int k) -> uint8_t& { return lrho[k * 64]; };
    auto BL = [&](int k) -> uint8_t& { return lbl[k * 64]; };
    const uint32_t width = t.w, height = t.h;
    const bool initial = y == 0;
    uint32_t a0 = 0, a1 = 0, a2 = 0;
    if (!initial) {
      a0 = uint32_t(RHO(q)) << 4;
      a1 = uint32_t(RHO(q + 1)) << 4;
      a2 = uint32_t(RHO(q + 2)) << 4;
      c_q |= (a0 & 0xA0) << 2;
      c_q |= (a1 & 0x20) << 4;
    }
    uint32_t vlc_val = vlc.fetch();
    uint32_t t0 = tab.dec(initial, c_q | (vlc_val & 0x7F));
    if (c_q == 0) {
      run -= 2;
      t0 = run == -1 ? t0 : 0;
      if (run < 0) run = mel.get_run();
    }
    if (initial) {
      c_q = ((t0 & 0x10) << 3) | ((t0 & 0xE0) << 2);
    } else {
      c_q = ((t0 & 0x40) << 2) | ((t0 & 0x80) << 1);
      c_q |= a0 & 0x80;
      c_q |= (a1 & 0xA0) << 2;
      c_q |= (a2 & 0x20) << 4;
    }
    vlc_val = vlc.advance(t0 & 0x7);
    uint32_t t1 = tab.dec(initial, c_q | (vlc_val & 0x7F));
    if (c_q == 0 && x + 2 < width) {
      run -= 2;
      t1 = run == -1 ? t1 : 0;
      if (run < 0) run = mel.get_run();
    }
    if (x + 2 >= width) t1 = 0;
    if (initial) {
      c_q = ((t1 & 0x10) << 3) | ((t1 & 0xE0) << 2);
    } else {
      c_q = ((t1 & 0x40) << 2) | ((t1 & 0x80) << 1);
      c_q |= a1 & 0x80;
    }
    vlc_val = vlc.advance(t1 & 0x7);
    uint32_t uvlc_mode = ((t0 & 0x8) << 3) | ((t1 & 0x8) << 4);
    if (initial && uvlc_mode == 0xC0) {
      run -= 2;
      uvlc_mode += run == -1 ? 0x40 : 0;
      if (run < 0) run = mel.get_run();
    }
    uint32_t e = tab.uvlc(initial, uvlc_mode + (vlc_val & 0x3F));
    vlc_val = vlc.advance(e & 0x7);
    e >>= 3;
    uint32_t len = e & 0xF;
    uint32_t tmp = vlc_val & ((1u << len) - 1);
    vlc.advance(len);
    e >>= 4;
    uint32_t len0 = e & 0x7;
    e >>= 3;
    uint32_t base = initial ? 1 : 0;
    uint32_t uq[2] = {base + (e & 7) + (tmp & ~(0xFFu << len0)), base + (e >> 3) + (tmp >> len0)};
    uint32_t tq[2] = {t0, t1};
    // this row's patterns become the next row's context (reads above done)
    RHO(q) = uint8_t(t0 >> 4);
    RHO(q + 1) = uint8_t(t1 >> 4);

    for (int qi = 0; qi < 2; ++qi) {
      uint32_t cx = x + 2 * qi;
      if (cx >= width) break;
      uint32_t inf = tq[qi], u_q = uq[qi];
      if (!initial) {
        uint32_t gamma = inf & 0xF0;
        gamma &= gamma - 0x10;
        uint32_t emax = max(max(uint32_t(BL(vp)), uint32_t(BL(vp + 1))), 2u) - 1;
        u_q += gamma != 0 ? emax : 1;
      }
      if (u_q > mmsbp2) bad = true;
      uint32_t v_n[4];
      v_n[1] = v_n[3] = 0;
      for (uint32_t bit = 0; bit < 4; ++bit) {
        uint32_t sx = cx + (bit >> 1), sy = y + (bit & 1);
        if (bit >= 2 && sx >= width) break;
        v_n[bit] = 0;
        uint32_t val = 0;
        if (inf & (1u << (4 + bit))) {
          uint32_t ms_val = ms.fetch();
          uint32_t m_n = u_q - ((inf >> (12 + bit)) & 1);
          ms.advance(m_n);
          uint32_t vv = ms_val & ((1u << (m_n & 31)) - 1);
          vv |= ((inf >> (8 + bit)) & 1) << (m_n & 31);
          vv |= 1;
          v_n[bit] = vv;
          val = (ms_val << 31) | ((vv + 2) << ((p - 1) & 31));
        }
        if (sy < height) stage[((bit & 1) * kStageCols + (sx - xbase)) * kStagePitch] = value(val);
      }
      BL(vp) = uint8_t(bitlen(prev_v_n | v_n[1]));
      prev_v_n = 0;
      vp += 1;
      if (cx + 1 >= width) break;
      prev_v_n = v_n[3];
    }
    q += 2;
  }

  // terminators: the quad after the row reads as insignificant
  __device__ void row_end(uint8_t* lrho, uint8_t* lbl) {
    lrho[q * 64] = 0;
    lrho[(q + 1) * 64] = 0;
    lbl[vp * 64] = uint8_t(bitlen(prev_v_n));
  }
};

// Wave-wide store of the staged window (columns xbase..xbase+kStageCols-1
// of rows y, y+1) of every lane's block: each store instruction writes
// kStageCols-column row segments of 32 / kStageCols blocks, so lines are
// written in runs instead of 4 bytes per lane per instruction.
__device__ __forceinline__ void flush_stage(const uint32_t* stage_all, uint64_t my_out, uint32_t my_stride,
                                            uint32_t my_w, uint32_t my_h, uint32_t y, uint32_t xbase) {
  constexpr uint32_t kSeg = 2 * kStageCols, kPer = 64 / kSeg;  // lanes per block, blocks per store
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t r = (lane / kStageCols) & 1, col = lane % kStageCols;
#pragma unroll 4
  for (uint32_t L0 = 0; L0 < 64; L0 += kPer) {
    uint32_t L = L0 + lane / kSeg;
    uint64_t out = __shfl(my_out, (int)L);
    uint32_t stride = __shfl(my_stride, (int)L), w = __shfl(my_w, (int)L), h = __shfl(my_h, (int)L);
    uint32_t gx = xbase + col, gy = y + r;
    if (out && gx < w && gy < h)
      st32((uint32_t*)out + (uint64_t)gy * stride + gx, stage_all[(r * kStageCols + col) * kStagePitch + L]);
  }
}

__device__ void fill_zero(const HtDecTask& t) {
  for (uint32_t y = 0; y < t.h; ++y)
    for (uint32_t x = 0; x < t.w; ++x) st32((uint32_t*)t.out + (uint64_t)y * t.stride + x, 0);
}

// A wave-local ordering point: the lanes' LDS writes before it are seen by
// reads after it (a wave's LDS instructions execute in order; this keeps the
// compiler from moving memory operations across).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// kW independent waves per workgroup; lane l of wave w decodes block
// 64 (kW b + w) + l (in `order`).  The workgroup shares one LDS copy of the
// decode tables; each wave has its own context rows and stage.  Four waves
// share a copy on large windows; a window with fewer waves than four per CU
// runs one wave per workgroup, spread over every CU.
template <uint32_t kLaneWaves>
__global__ __launch_bounds__(64 * kLaneWaves) void ht_decode_kernel(const HtDecTask* __restrict__ tasks,
                                                                   const uint32_t* __restrict__ order, uint64_t n) {
  __shared__ __align__(16) uint16_t tabs[kTabAll];
  __shared__ uint8_t lrho_all[kLaneWaves][kRowCap * 64];
  __shared__ uint8_t lbl_all[kLaneWaves][kRowCap * 64];
  __shared__ uint32_t stage_all[kLaneWaves][2 * kStageCols * kStagePitch];
  {
    uint32_t* tw = (uint32_t*)tabs;
    const uint32_t* s0 = (const uint32_t*)c_dec0;
    const uint32_t* s1 = (const uint32_t*)c_dec1;
    const uint32_t* s2 = (const uint32_t*)c_uvlc0;
    const uint32_t* s3 = (const uint32_t*)c_uvlc1;
    for (uint32_t k = threadIdx.x; k < kTabAll / 2; k += blockDim.x) {
      uint32_t v;
      if (k < kTabDec1 / 2) v = s0[k];
      else if (k < kTabUv0 / 2) v = s1[k - kTabDec1 / 2];
      else if (k < kTabUv1 / 2) v = s2[k - kTabUv0 / 2];
      else v = s3[k - kTabUv1 / 2];
      tw[k] = v;
    }
  }
  __syncthreads();
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  uint8_t* lrho = lrho_all[wv];
  uint8_t* lbl = lbl_all[wv];
  uint32_t* stage = stage_all[wv];
  uint64_t i = blockIdx.x * (uint64_t)(64 * kLaneWaves) + threadIdx.x;
  HtDecTask task{};
  if (i < n) {
    if (order) i = order[i];
    task = tasks[i];
  }
  for (int k = 0; k < kRowCap; ++k) {
    lrho[k * 64 + lane] = 0;
    lbl[k * 64 + lane] = 0;
  }
  BlockDec<LdsTables> d;
  d.tab.t = (lds_u16p)tabs;
  d.start(task);
  const uint32_t w = task.w, h = task.h;
  const uint64_t out = i < n ? (uint64_t)task.out : 0;
  uint32_t* my_stage = stage + lane;
  for (uint32_t y = 0; y < 64; y += 2) {
    if (__ballot(y < h) == 0) break;
    if (d.active && y < h) d.row_begin();
    for (uint32_t xb = 0; xb < 64; xb += kStageCols) {
      if (__ballot(y < h && xb < w) == 0) break;
      for (uint32_t x = xb; x < xb + kStageCols; x += 4) {
        if (y < h && x < w) {
          if (d.active) {
            d.pair(x, y, lrho + lane, lbl + lane, my_stage, xb);
          } else {
            for (uint32_t s = 0; s < 8; ++s) {
              uint32_t sx = x + (s >> 1), sy = y + (s & 1);
              if (sx < w && sy < h) my_stage[((s & 1) * kStageCols + (sx - xb)) * kStagePitch] = 0;
            }
          }
        }
      }
      wave_sync();  // the staged window is complete
      flush_stage(stage, out, task.stride, w, h, y, xb);
      wave_sync();  // reads done before the next window overwrites the stage
    }
    if (d.active && y < h) d.row_end(lrho + lane, lbl + lane);
  }
  // a magnitude bound violated mid-block: the whole block reads as zero
  if (d.active && d.bad) {
    report(task, kErrBlock);
    __threadfence();  // the flushed rows land before they are overwritten
    fill_zero(task);
  }
  // blocks the refinement kernel must finish: clear the mark of failed ones
  if (i < n && task.passes > 1 && (!d.active || d.bad)) ((HtDecTask*)tasks)[i].passes = 1;
}

// ---------------------------------------------------------------------------
// Wave-parallel cleanup decode: one wavefront per code-block.
//
// 1. The three byte streams are unstuffed at once into LDS bit arrays (word w
//    holds stream bits 32w..32w+31, first bit lowest).  A byte's position is
//    a wave prefix sum of the widths its stuffing rule gives it, a function
//    of the byte and its predecessor only: MagSgn 8 - [b(k-1) == 0xFF]; VLC
//    (read backward) 8 - [b(k-1) > 0x8F and b(k) & 0x7F == 0x7F]; MEL (MSB
//    first, so stored bit-reversed) 8 - [b(k) == 0xFF].  Bytes are OR-ed into
//    place, which is what the serial readers do with a stuffed byte's spare
//    top bit.  Past its end MagSgn and MEL read ones, VLC zeros, as there.
// 2. Per quad row, the VLC/MEL decode, the one serial dependency (a
//    codeword's context is the significance of the codewords before it),
//    runs wave-uniform and leaves each quad's CxtVLC entry and residual
//    exponent offset u in the two lanes of its columns.
// 3. The row's MagSgn samples are decoded by all lanes at once, lane c owning
//    column c: the exponent bound from the row above comes from neighbouring
//    lanes, the bit offsets from a wave prefix sum of the bit counts, each
//    sample from one 64-bit LDS read; rows are stored 64 columns wide.
// The result equals the lane-per-block decoder's sample for sample, errors
// included.  Segments whose MagSgn part does not fit the LDS array are
// decoded by lane 0 with that serial decoder.
// ---------------------------------------------------------------------------
constexpr uint32_t kWaveMsWords = 3072;  // MagSgn bits: segments up to 12,280 MagSgn bytes
constexpr uint32_t kWaveSWords = 1024;   // MEL / VLC bits: Scup <= 4079 bytes

__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v, uint32_t* total) {
  const uint32_t lane = threadIdx.x & 63;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t t = __shfl_up(x, o);
    if (lane >= (uint32_t)o) x += t;
  }
  *total = __shfl(x, 63);
  return x - v;
}

__device__ __forceinline__ uint32_t uread(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }
__device__ __forceinline__ uint32_t lane_of(uint32_t v, uint32_t l) {
  return l < 64 ? (uint32_t)__builtin_amdgcn_readlane((int)v, (int)l) : 0u;
}

enum { kUnMs = 0, kUnVlc = 1, kUnMel = 2 };

// Unstuffs n bytes (byte k = at(k)) into buf starting at bit p0; prev0 stands
// for the byte before the first.  Returns the end position.
template <int Kind, class At>
__device__ uint32_t unstuff(uint32_t* buf, uint32_t n, uint32_t p0, uint32_t prev0, At at) {
  constexpr int kPer = 16;  // bytes per lane per step: 16 loads in flight
  const uint32_t lane = threadIdx.x & 63;
  uint32_t base = p0;
  for (uint32_t c0 = 0; c0 < n; c0 += 64 * kPer) {
    const uint32_t k0 = c0 + kPer * lane;
    uint32_t b[kPer];
#pragma unroll
    for (int j = 0; j < kPer; ++j) b[j] = k0 + j < n ? at(k0 + j) : 0u;
    uint32_t prev = k0 == 0 ? prev0 : (k0 - 1 < n ? at(k0 - 1) : 0u);
    uint32_t sum = 0;
    uint32_t w[kPer];
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      uint32_t wd;
      if (Kind == kUnMs) wd = 8 - uint32_t(prev == 0xFF);
      else if (Kind == kUnVlc) wd = 8 - uint32_t(prev > 0x8F && (b[j] & 0x7F) == 0x7F);
      else wd = 8 - uint32_t(b[j] == 0xFF);
      w[j] = k0 + j < n ? wd : 0u;
      sum += w[j];
      prev = b[j];
    }
    uint32_t total;
    uint32_t pos = base + wave_excl_scan(sum, &total);
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      if (k0 + j < n) {
        const uint32_t v = Kind == kUnMel ? (__builtin_bitreverse32(b[j]) >> 24) : b[j];
        const uint32_t sh = pos & 31, wi = pos >> 5;
        atomicOr(&buf[wi], v << sh);
        if (sh > 24) atomicOr(&buf[wi + 1], v >> (32 - sh));
      }
      pos += w[j];
    }
    base += total;
  }
  return base;
}

// Wave-uniform reader of an unstuffed LDS bit array (words past nvalid read
// as `fill`).
struct WaveBits {
  const uint32_t* buf;
  uint32_t nvalid, fill, next, have;
  uint64_t win;  // bits [pos, pos + have)
  __device__ void init(const uint32_t* b, uint32_t nv, uint32_t f) {
    buf = b;
    nvalid = nv;
    fill = f;
    next = 0;
    have = 0;
    win = 0;
  }
  __device__ uint32_t peek() {  // at least 33 bits valid afterwards
    while (have <= 32) {
      const uint32_t v = next < nvalid ? uread(buf[next]) : fill;
      win |= (uint64_t)v << have;
      have += 32;
      next += 1;
    }
    return (uint32_t)win;
  }
  __device__ uint32_t skip(uint32_t n) {  // n <= have
    win >>= n;
    have -= n;
    return (uint32_t)win;
  }
};

// The MEL symbol decoder of mel.rs:14-173 over the unstuffed MSB-first bits.
struct WaveMel {
  WaveBits b;
  int k;
  __device__ int next_run() {
    const uint32_t v = b.peek();
    const uint32_t e = uread(c_mel_expw[k]);
    if (v & 1) {
      b.skip(1);
      k = k + 1 < 12 ? k + 1 : 12;
      return (int)(((1u << e) - 1) << 1);
    }
    const uint32_t cnt = e ? (__builtin_bitreverse32(v >> 1) >> (32 - e)) : 0u;
    b.skip(e + 1);
    k = k > 0 ? k - 1 : 0;
    return (int)((cnt << 1) | 1);
  }
};

__device__ __forceinline__ uint32_t lds_bits32(const uint32_t* buf, uint32_t pos, uint32_t nvalid) {
  const uint32_t wi = pos >> 5;
  const uint32_t lo = wi < nvalid ? buf[wi] : ~0u;
  const uint32_t hi = wi + 1 < nvalid ? buf[wi + 1] : ~0u;
  return __builtin_amdgcn_alignbit(hi, lo, pos & 31);
}

// Lane 0 alone: the serial decoder over the segment in global memory, for
// segments too large for the LDS bit arrays.  lds holds rho / bit-length
// rows and a one-lane stage (laid out as in ht_decode_kernel).
__device__ void decode_block_serial(const HtDecTask& task, uint8_t* lds, HtDecTask* slot) {
  uint8_t* lrho = lds;
  uint8_t* lbl = lds + kRowCap * 64;
  uint32_t* stage = (uint32_t*)(lds + 2 * kRowCap * 64);
  for (int k = 0; k < kRowCap; ++k) lrho[k * 64] = lbl[k * 64] = 0;
  BlockDec<ConstTables> d;
  d.start(task);
  const uint32_t w = task.w, h = task.h;
  uint32_t* out = (uint32_t*)task.out;
  for (uint32_t y = 0; y < 64 && y < h; y += 2) {
    if (d.active) d.row_begin();
    for (uint32_t xb = 0; xb < 64 && xb < w; xb += kStageCols) {
      for (uint32_t x = xb; x < xb + kStageCols; x += 4) {
        if (x >= w) break;
        if (d.active) {
          d.pair(x, y, lrho, lbl, stage, xb);
        } else {
          for (uint32_t s = 0; s < 8; ++s) stage[((s & 1) * kStageCols + (x - xb + (s >> 1))) * kStagePitch] = 0;
        }
      }
      for (uint32_t r = 0; r < 2; ++r)
        for (uint32_t c = 0; c < kStageCols; ++c)
          if (xb + c < w && y + r < h)
            st32(out + (uint64_t)(y + r) * task.stride + xb + c, stage[(r * kStageCols + c) * kStagePitch]);
    }
    if (d.active) d.row_end(lrho, lbl);
  }
  if (d.active && d.bad) {
    report(task, kErrBlock);
    fill_zero(task);
  }
  if (task.passes > 1 && (!d.active || d.bad)) slot->passes = 1;
}

// Decodes task i (the whole wave; returns together).
__device__ void ht_decode_wave_block(const HtDecTask* __restrict__ tasks, uint64_t i, uint32_t* msbuf,
                                     uint32_t* melbuf, uint32_t* vlcbuf, uint32_t* blrow) {
  const uint32_t lane = threadIdx.x;
  const HtDecTask t = tasks[i];
  HtDecTask* slot = (HtDecTask*)tasks + i;
  const uint32_t width = t.w, height = t.h;
  uint32_t* out = (uint32_t*)t.out;
  const uint64_t stride = t.stride;
  // decode.rs:54-110, as BlockDec::start
  bool active = false, refine = false;
  uint32_t p = 0, shift = 0, mmsbp2 = 0, lcup = 0, scup = 0;
  if ((t.flags & kTaskIncluded) && t.seg && width && height) {
    uint32_t num_passes = t.passes;
    const uint32_t missing_msbs = t.mmsb;
    if (num_passes > 1 && t.len2 == 0) num_passes = 1;
    if (num_passes > 3 || missing_msbs >= 30 || t.len1 < 2 || width > 64 || height > 64) {
      if (lane == 0) report(t, kErrBlock);
    } else {
      if (missing_msbs == 29) num_passes = 1;
      refine = num_passes > 1;
      lcup = t.len1;
      scup = (ld8(t.seg + lcup - 1) << 4) + (ld8(t.seg + lcup - 2) & 0xF);
      if (scup < 2 || scup > lcup || scup > 4079) {
        if (lane == 0) report(t, kErrBlock);
      } else {
        active = true;
        p = 30 - missing_msbs;
        shift = 31 - t.kmax;
        mmsbp2 = missing_msbs + 2;
      }
    }
  }
  const uint32_t wcap = width < 64 ? width : 64, hcap = height < 64 ? height : 64;
  if (!active) {
    if (out)
      for (uint32_t y = 0; y < hcap; ++y)
        if (lane < wcap) st32(out + (uint64_t)y * stride + lane, 0);
    if (lane == 0 && t.passes > 1) slot->passes = 1;
    return;
  }
  const uint32_t n_ms = lcup - scup;
  if (n_ms > 4 * (kWaveMsWords - 2)) {
    if (lane == 0) decode_block_serial(t, (uint8_t*)msbuf, slot);
    return;
  }
  const uint8_t* seg = t.seg;
  // ---- 1. unstuff
  const uint32_t ms_words = ((8 * n_ms) >> 5) + 2, s_words = ((8 * scup + 4) >> 5) + 2;
  for (uint32_t w = lane; w < ms_words; w += 64) msbuf[w] = 0;
  for (uint32_t w = lane; w < s_words; w += 64) {
    melbuf[w] = 0;
    vlcbuf[w] = 0;
  }
  __syncthreads();
  const uint32_t t_ms = unstuff<kUnMs>(msbuf, n_ms, 0, 0, [&](uint32_t k) { return ld8(seg + k); });
  const uint32_t mel0 = lcup - scup, nmel = scup - 1;
  const uint32_t t_mel = unstuff<kUnMel>(melbuf, nmel, 0, 0, [&](uint32_t k) {
    const uint32_t v = ld8(seg + mel0 + k);
    return k + 1 == nmel ? (v | 0xF) : v;
  });
  const uint32_t d0 = ld8(seg + lcup - 2);
  if (lane == 0) atomicOr(&vlcbuf[0], d0 >> 4);
  const uint32_t t_vlc = unstuff<kUnVlc>(vlcbuf, scup - 2, 4 - uint32_t(((d0 >> 4) & 7) == 7), d0 | 0xF,
                                         [&](uint32_t k) { return ld8(seg + lcup - 3 - k); });
  if (lane == 0) {
    msbuf[t_ms >> 5] |= ~0u << (t_ms & 31);
    melbuf[t_mel >> 5] |= ~0u << (t_mel & 31);
  }
  if (lane < 68) blrow[lane] = 0;
  if (lane < 4) blrow[64 + lane] = 0;
  __syncthreads();
  const uint32_t ms_nvalid = (t_ms >> 5) + 1;
  WaveBits vlc;
  vlc.init(vlcbuf, (t_vlc >> 5) + 1, 0u);
  WaveMel mel;
  mel.b.init(melbuf, (t_mel >> 5) + 1, ~0u);
  mel.k = 0;
  int run = mel.next_run();
  const bool raw = (t.flags & kTaskSignMag) || refine;
  uint32_t prev_q = 0;  // lane c: the row above's entry of quad c / 2
  uint32_t msbase = 0;
  bool bad = false;
  for (uint32_t y = 0; y < height; y += 2) {
    const bool initial = y == 0;
    // ---- 2. VLC / MEL, wave-uniform (decode.rs:115-300)
    uint32_t qinfo = 0, c_q = 0;
    for (uint32_t x = 0, l4 = 0; x < width; x += 4, l4 += 4) {
      uint32_t a0 = 0, a1 = 0, a2 = 0;
      if (!initial) {
        a0 = lane_of(prev_q, l4) & 0xF0;
        a1 = lane_of(prev_q, l4 + 2) & 0xF0;
        a2 = lane_of(prev_q, l4 + 4) & 0xF0;
        c_q |= (a0 & 0xA0) << 2;
        c_q |= (a1 & 0x20) << 4;
      }
      uint32_t vlc_val = vlc.peek();
      uint32_t t0 = uread(initial ? c_dec0w[c_q | (vlc_val & 0x7F)] : c_dec1w[c_q | (vlc_val & 0x7F)]);
      if (c_q == 0) {
        run -= 2;
        t0 = run == -1 ? t0 : 0;
        if (run < 0) run = mel.next_run();
      }
      if (initial) {
        c_q = ((t0 & 0x10) << 3) | ((t0 & 0xE0) << 2);
      } else {
        c_q = ((t0 & 0x40) << 2) | ((t0 & 0x80) << 1);
        c_q |= a0 & 0x80;
        c_q |= (a1 & 0xA0) << 2;
        c_q |= (a2 & 0x20) << 4;
      }
      vlc_val = vlc.skip(t0 & 0x7);
      uint32_t t1 = uread(initial ? c_dec0w[c_q | (vlc_val & 0x7F)] : c_dec1w[c_q | (vlc_val & 0x7F)]);
      if (c_q == 0 && x + 2 < width) {
        run -= 2;
        t1 = run == -1 ? t1 : 0;
        if (run < 0) run = mel.next_run();
      }
      if (x + 2 >= width) t1 = 0;
      if (initial) {
        c_q = ((t1 & 0x10) << 3) | ((t1 & 0xE0) << 2);
      } else {
        c_q = ((t1 & 0x40) << 2) | ((t1 & 0x80) << 1);
        c_q |= a1 & 0x80;
      }
      vlc_val = vlc.skip(t1 & 0x7);
      uint32_t uvlc_mode = ((t0 & 0x8) << 3) | ((t1 & 0x8) << 4);
      if (initial && uvlc_mode == 0xC0) {
        run -= 2;
        uvlc_mode += run == -1 ? 0x40 : 0;
        if (run < 0) run = mel.next_run();
      }
      uint32_t e = uread(initial ? c_uvlc0w[uvlc_mode + (vlc_val & 0x3F)] : c_uvlc1w[uvlc_mode + (vlc_val & 0x3F)]);
      vlc_val = vlc.skip(e & 0x7);
      e >>= 3;
      const uint32_t len = e & 0xF;
      const uint32_t tmp = vlc_val & ((1u << len) - 1);
      vlc.skip(len);
      e >>= 4;
      const uint32_t len0 = e & 0x7;
      e >>= 3;
      const uint32_t base = initial ? 1 : 0;
      const uint32_t q0 = t0 | ((base + (e & 7) + (tmp & ~(0xFFu << len0))) << 16);
      const uint32_t q1 = t1 | ((base + (e >> 3) + (tmp >> len0)) << 16);
      if ((lane & ~3u) == l4) qinfo = (lane & 2) ? q1 : q0;
    }
    // ---- 3. MagSgn of rows y, y+1, lane = column (decode.rs:302-396)
    const bool col = lane < width;
    const uint32_t tq = qinfo & 0xFFFF;
    uint32_t U = qinfo >> 16;
    if (!initial) {
      uint32_t gamma = tq & 0xF0;
      gamma &= gamma - 0x10;
      const uint32_t q2 = lane & ~1u;  // columns 2q-1 .. 2q+2 of the row above: blrow[2q .. 2q+3]
      const uint32_t e4 = max(max(blrow[q2], blrow[q2 + 1]), max(blrow[q2 + 2], blrow[q2 + 3]));
      U += gamma != 0 ? max(e4, 2u) - 1 : 1;
    }
    if (col && !(lane & 1) && U > mmsbp2) bad = true;
    const uint32_t nb0 = 2 * (lane & 1);
    uint32_t m[2], sig[2];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const uint32_t nbit = nb0 + r;
      sig[r] = col ? (tq >> (4 + nbit)) & 1 : 0u;
      m[r] = U - ((tq >> (12 + nbit)) & 1);
    }
    const uint32_t c0 = sig[0] ? m[0] : 0u, c1 = sig[1] ? m[1] : 0u;
    uint32_t total;
    const uint32_t off = msbase + wave_excl_scan(c0 + c1, &total);
    uint32_t blnew = 0;
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const uint32_t nbit = nb0 + r;
      uint32_t val = 0, vv = 0;
      if (sig[r]) {
        const uint32_t ms_val = lds_bits32(msbuf, off + (r ? c0 : 0u), ms_nvalid);
        vv = ms_val & ((1u << (m[r] & 31)) - 1);
        vv |= ((tq >> (8 + nbit)) & 1) << (m[r] & 31);
        vv |= 1;
        val = (ms_val << 31) | ((vv + 2) << ((p - 1) & 31));
      }
      if (r == 1) blnew = bitlen(vv);
      if (col && y + r < height) {
        uint32_t o = val;
        if (!raw) {
          const int32_t mag = (int32_t)((val & 0x7FFFFFFFu) >> shift);
          o = (uint32_t)((val & 0x80000000u) ? -mag : mag);
        }
        st32(out + (uint64_t)(y + r) * stride + lane, o);
      }
    }
    msbase += total;
    blrow[lane + 1] = blnew;
    prev_q = qinfo;
  }
  // a magnitude bound violated mid-block: the whole block reads as zero
  if (__ballot(bad)) {
    if (lane == 0) report(t, kErrBlock);
    for (uint32_t y = 0; y < hcap; ++y)
      if (lane < wcap) st32(out + (uint64_t)y * stride + lane, 0);
    if (lane == 0 && t.passes > 1) slot->passes = 1;
  }
}

// Work queue of a persistent launch: q[0] hands out task indices `grain` at
// a time, q[1] counts the waves that found it drained; the last of them
// resets both, so the next launch on the stream starts from zero.
__device__ __forceinline__ uint64_t queue_pop(uint64_t* q, uint32_t grain, uint64_t* s_base) {
  __syncthreads();
  if (threadIdx.x == 0) *s_base = atomicAdd((unsigned long long*)q, (unsigned long long)grain);
  __syncthreads();
  return *s_base;
}
__device__ __forceinline__ void queue_done(uint64_t* q) {
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd((unsigned long long*)(q + 1), 1ull) == gridDim.x - 1) {
      atomicExch((unsigned long long*)q, 0ull);
      atomicExch((unsigned long long*)(q + 1), 0ull);
    }
  }
}

// Persistent: each wave takes `grain` code-blocks at a time from the queue
// until it is drained.
__global__ __launch_bounds__(64) void ht_decode_wave_kernel(const HtDecTask* __restrict__ tasks, uint64_t n,
                                                           uint64_t* queue, uint32_t grain) {
  __shared__ uint32_t msbuf[kWaveMsWords];
  __shared__ uint32_t melbuf[kWaveSWords];
  __shared__ uint32_t vlcbuf[kWaveSWords];
  __shared__ uint32_t blrow[68];
  __shared__ uint64_t s_base;
  for (;;) {
    const uint64_t base = queue_pop(queue, grain, &s_base);
    if (base >= n) break;
    const uint64_t end = base + grain < n ? base + grain : n;
    for (uint64_t i = base; i < end; ++i) {
      ht_decode_wave_block(tasks, i, msbuf, melbuf, vlcbuf, blrow);
      __syncthreads();
    }
  }
  queue_done(queue);
}

// ---------------------------------------------------------------------------
// Two-pass cleanup decode (large windows).
//
// Pass 1, ht_vlc_kernel: one lane per code-block runs only the VLC / MEL
// decode (decode.rs:115-300), the serial chain whose context is the quad
// significance of the row above, and writes per quad the CxtVLC entry and
// the UVLC exponent offset before the row-above correction, u << 16 | entry
// (what the wavefront decoder's step 2 leaves in its lanes).  A quad row of
// the wave's 64 blocks is staged in LDS and stored 32 quads (128 bytes) per
// block at a time.  The VLC and MEL streams are the short end of a segment.
// Pass 2, ht_magsgn_kernel: one wavefront per code-block unstuffs the MagSgn
// bytes into LDS with coalesced reads, loads the block's quad records, and
// decodes each row pair lane-per-column (the wavefront decoder's step 3):
// 64-column rows stored as full lines.  The result equals the other
// decoders' sample for sample, errors included.
// ---------------------------------------------------------------------------
constexpr uint32_t kRecWords = 32 * 32;  // quad records per block: 32 quad rows x 32 quads

#ifndef NDIC_HT_REC_SEG
#define NDIC_HT_REC_SEG 16
#endif
constexpr uint32_t kRecSeg = NDIC_HT_REC_SEG;  // quads of a row segment: the unit staged in LDS and stored
constexpr uint32_t kRecPitch = kRecSeg + 1;    // LDS pitch of a lane's staged segment
static_assert(kRecSeg == 8 || kRecSeg == 16 || kRecSeg == 32, "a row of 32 quads in whole segments");

template <class Tab>
struct VlcMelDec {
  Tab tab;
  MelDec mel;
  RevDec vlc;
  int run;
  bool active;
  uint32_t c_q, q;

  // decode.rs:54-110, the checks that decide whether the streams are read
  __device__ void start(const HtDecTask& t) {
    active = false;
    if (!(t.flags & kTaskIncluded) || !t.seg || t.w == 0 || t.h == 0) return;
    uint32_t num_passes = t.passes;
    if (num_passes > 1 && t.len2 == 0) num_passes = 1;
    if (num_passes > 3 || t.mmsb >= 30 || t.len1 < 2 || t.w > 64 || t.h > 64) return;
    const uint32_t lcup = t.len1;
    const uint32_t scup = (ld8(t.seg + lcup - 1) << 4) + (ld8(t.seg + lcup - 2) & 0xF);
    if (scup < 2 || scup > lcup || scup > 4079) return;
    mel.init(t.seg, lcup, scup);
    vlc.init_vlc(t.seg, lcup, scup);
    run = mel.get_run();
    active = true;
  }

  __device__ void row_begin() {
    c_q = 0;
    q = 0;
  }

  // One pair of quads (columns x..x+3, rows y, y+1): their records.
  __device__ void pair(uint32_t x, uint32_t y, uint32_t width, uint8_t* lrho, uint32_t* r0, uint32_t* r1) {
    auto RHO = [&](int k) -> uint8_t& { return lrho[k * 64]; };
    const bool initial = y == 0;
    uint32_t a0 = 0, a1 = 0, a2 = 0;
    if (!initial) {
      a0 = uint32_t(RHO(q)) << 4;
      a1 = uint32_t(RHO(q + 1)) << 4;
      a2 = uint32_t(RHO(q + 2)) << 4;
      c_q |= (a0 & 0xA0) << 2;
      c_q |= (a1 & 0x20) << 4;
    }
    uint32_t vlc_val = vlc.fetch();
    uint32_t t0 = tab.dec(initial, c_q | (vlc_val & 0x7F));
    if (c_q == 0) {
      run -= 2;
      t0 = run == -1 ? t0 : 0;
      if (run < 0) run = mel.get_run();
    }
    if (initial) {
      c_q = ((t0 & 0x10) << 3) | ((t0 & 0xE0) << 2);
    } else {
      c_q = ((t0 & 0x40) << 2) | ((t0 & 0x80) << 1);
      c_q |= a0 & 0x80;
      c_q |= (a1 & 0xA0) << 2;
      c_q |= (a2 & 0x20) << 4;
    }
    vlc_val = vlc.advance(t0 & 0x7);
    uint32_t t1 = tab.dec(initial, c_q | (vlc_val & 0x7F));
    if (c_q == 0 && x + 2 < width) {
      run -= 2;
      t1 = run == -1 ? t1 : 0;
      if (run < 0) run = mel.get_run();
    }
    if (x + 2 >= width) t1 = 0;
    if (initial) {
      c_q = ((t1 & 0x10) << 3) | ((t1 & 0xE0) << 2);
    } else {
      c_q = ((t1 & 0x40) << 2) | ((t1 & 0x80) << 1);
      c_q |= a1 & 0x80;
    }
    vlc_val = vlc.advance(t1 & 0x7);
    uint32_t uvlc_mode = ((t0 & 0x8) << 3) | ((t1 & 0x8) << 4);
    if (initial && uvlc_mode == 0xC0) {
      run -= 2;
      uvlc_mode += run == -1 ? 0x40 : 0;
      if (run < 0) run = mel.get_run();
    }
    uint32_t e = tab.uvlc(initial, uvlc_mode + (vlc_val & 0x3F));
    vlc_val = vlc.advance(e & 0x7);
    e >>= 3;
    const uint32_t len = e & 0xF;
    const uint32_t tmp = vlc_val & ((1u << len) - 1);
    vlc.advance(len);
    e >>= 4;
    const uint32_t len0 = e & 0x7;
    e >>= 3;
    const uint32_t base = initial ? 1 : 0;
    *r0 = t0 | ((base + (e & 7) + (tmp & ~(0xFFu << len0))) << 16);
    *r1 = t1 | ((base + (e >> 3) + (tmp >> len0)) << 16);
    RHO(q) = uint8_t(t0 >> 4);
    RHO(q + 1) = uint8_t(t1 >> 4);
    q += 2;
  }

  // the quads after the row read as insignificant
  __device__ void row_end(uint8_t* lrho) {
    lrho[q * 64] = 0;
    lrho[(q + 1) * 64] = 0;
  }
};

// kW waves per workgroup share one LDS copy of the decode tables; lane l of
// wave w decodes task 64 (kW b + w) + l; its records go to rec + task * kRecWords.
template <uint32_t kLaneWaves>
__global__ __launch_bounds__(64 * kLaneWaves) void ht_vlc_kernel(const HtDecTask* __restrict__ tasks, uint64_t n,
                                                                uint32_t* __restrict__ rec) {
  __shared__ __align__(16) uint16_t tabs[kTabAll];
  __shared__ uint8_t lrho_all[kLaneWaves][kRowCap * 64];
  __shared__ uint32_t recst_all[kLaneWaves][64 * kRecPitch];
  {
    uint32_t* tw = (uint32_t*)tabs;
    const uint32_t* s0 = (const uint32_t*)c_dec0;
    const uint32_t* s1 = (const uint32_t*)c_dec1;
    const uint32_t* s2 = (const uint32_t*)c_uvlc0;
    const uint32_t* s3 = (const uint32_t*)c_uvlc1;
    for (uint32_t k = threadIdx.x; k < kTabAll / 2; k += blockDim.x) {
      uint32_t v;
      if (k < kTabDec1 / 2) v = s0[k];
      else if (k < kTabUv0 / 2) v = s1[k - kTabDec1 / 2];
      else if (k < kTabUv1 / 2) v = s2[k - kTabUv0 / 2];
      else v = s3[k - kTabUv1 / 2];
      tw[k] = v;
    }
  }
  __syncthreads();
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  uint8_t* lrho = lrho_all[wv];
  uint32_t* recst = recst_all[wv];
  const uint64_t i = blockIdx.x * (uint64_t)(64 * kLaneWaves) + threadIdx.x;
  HtDecTask task{};
  if (i < n) task = tasks[i];
  for (int k = 0; k < kRowCap; ++k) lrho[k * 64 + lane] = 0;
  VlcMelDec<LdsTables> d;
  d.tab.t = (lds_u16p)tabs;
  d.start(task);
  const uint32_t w = d.active ? task.w : 0, h = d.active ? task.h : 0;
  const uint32_t nq = (w + 1) / 2;  // quads per row
  for (uint32_t y = 0; y < 64; y += 2) {
    if (__ballot(y < h) == 0) break;
    if (y < h) d.row_begin();
    // each row segment (kRecSeg quads) staged in LDS, then stored 64 /
    // kRecSeg blocks per store instruction: lanes kRecSeg k .. kRecSeg k +
    // kRecSeg - 1 the quads of block j + k
    constexpr uint32_t kPer = 64 / kRecSeg;
#pragma unroll 1
    for (uint32_t hx = 0; hx < 32 / kRecSeg; ++hx) {
      const uint32_t x0 = hx * 2 * kRecSeg;  // first column of the segment
      if (__ballot(y < h && x0 < w) == 0) break;
      if (y < h)
        for (uint32_t x = x0; x < w && x < x0 + 2 * kRecSeg; x += 4) {
          uint32_t r0, r1;
          d.pair(x, y, w, lrho + lane, &r0, &r1);
          recst[lane * kRecPitch + (x - x0) / 2] = r0;
          recst[lane * kRecPitch + (x - x0) / 2 + 1] = r1;
        }
      wave_sync();  // the segment of every lane's block is staged
      const uint32_t part = lane / kRecSeg, qi = lane % kRecSeg;
#pragma unroll 4
      for (uint32_t j = 0; j < 64; j += kPer) {
        const uint32_t b = j + part;
        const uint32_t bh = __shfl(h, (int)b), bq = __shfl(nq, (int)b);
        const uint64_t bi = (uint64_t)__shfl((long long)i, (int)b);  // the task index of lane b
        if (y < bh && x0 / 2 + qi < bq)
          rec[bi * kRecWords + (y / 2) * 32 + x0 / 2 + qi] = recst[b * kRecPitch + qi];
      }
      wave_sync();  // reads done before the next segment overwrites the stage
    }
    if (y < h) d.row_end(lrho + lane);
  }
}

// The MagSgn stream of one block unstuffed into an LDS ring of kMsRing words
// as the row decode advances (bit b of the stream at word (b >> 5) mod
// kMsRing), 1024 bytes per refill: lane l unstuffs bytes 16l..16l+15 of the
// chunk, its bit offset a wave prefix sum of the widths the stuffing rule
// gives (8 - [previous byte == 0xFF]).  The next chunk's bytes are loaded one
// refill ahead.  Past the stream's end the bits read as ones.
constexpr uint32_t kMsRing = 512;  // words: 16 Kbit, more than the 12 Kbit live window
static_assert((kMsRing & (kMsRing - 1)) == 0, "the ring index is a mask");

// Inclusive prefix sum over the wavefront from DPP moves (GFX9 forms): a
// Hillis-Steele scan inside each 16-lane row (row_shr 1, 2, 4, 8; lanes whose
// source is outside the row add 0), then the row totals carried across rows
// (row_bcast 15 into rows 1 and 3, row_bcast 31 into rows 2 and 3; rows the
// mask leaves out add 0).
// DPP lane moves (GFX9 encodings): source lanes outside the wave read 0.
constexpr int kDppWaveShl1 = 0x130, kDppWaveShr1 = 0x138;  // lane l reads lane l + 1 / l - 1
constexpr int kDppSwapPair = 0xB1, kDppEvenOfPair = 0xA0, kDppOddOfPair = 0xF5;  // quad_perm [1,0,3,2] [0,0,2,2] [1,1,3,3]
template <int Ctrl> __device__ __forceinline__ uint32_t dpp_mov(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, Ctrl, 0xF, 0xF, true);
}
#define dpp(v, ctrl) dpp_mov<ctrl>(v)

__device__ __forceinline__ uint32_t wave_scan_incl(uint32_t v) {
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, true);
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, true);
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, true);
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, true);
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false);
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false);
  return v;
}

struct MsRing {
  uint32_t* ring;
  const uint8_t* seg;
  uint32_t n;      // MagSgn bytes
  uint32_t next;   // first byte of the next chunk
  uint32_t end;    // bits unstuffed so far
  uint32_t prev;   // last byte of the previous chunk (0 before the first)
  uint32_t nvalid; // words that hold stream bits, once all are in (else ~0)
  uint32_t ones;   // once all are in: words nvalid .. ones - 1 are set to all ones
  uint32_t q[4];   // this lane's 16 bytes of the next chunk (byte j: bits 8 (j % 4) of q[j / 4])

  __device__ void load_next() {
    const uint32_t k0 = next + 16 * (threadIdx.x & 63);
    if (k0 + 16 <= n) {
      const u32x4 v = ld128(seg + k0);  // a 16-byte read need not be aligned
      q[0] = v.x;
      q[1] = v.y;
      q[2] = v.z;
      q[3] = v.w;
    } else {  // the stream's last bytes; past it zeros
      q[0] = q[1] = q[2] = q[3] = 0;
#pragma unroll
      for (uint32_t j = 0; j < 16; ++j)
        if (k0 + j < n) q[j >> 2] |= ld8(seg + k0 + j) << (8 * (j & 3));
    }
  }
  __device__ void init(uint32_t* r, const uint8_t* s, uint32_t nbytes) {
    ring = r;
    seg = s;
    n = nbytes;
    next = 0;
    end = 0;
    prev = 0;
    nvalid = ~0u;
    ones = 0;
    load_next();
  }
  // Unstuffs the next 1024 bytes into the ring.  A byte after 0xFF brings 7
  // bits, any other 8, and each byte is OR-ed in whole at its bit position
  // (a stuffed byte's spare top bit lands on the next byte's first bit, as
  // the serial reader has it).  A lane packs its 16 bytes in four registers
  // -- their plain concatenation, with the bytes after each 7-bit byte moved
  // down one bit -- and ORs them into the ring at its offset (a wave prefix
  // sum of the lanes' bit counts).
  __device__ void refill() {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t k0 = next + 16 * lane;
    uint32_t prev_l = __shfl(q[3] >> 24, (int)lane - 1);
    if (lane == 0) prev_l = prev;
    const uint32_t nin = k0 >= n ? 0u : (n - k0 < 16 ? n - k0 : 16u);
    // bit j: byte j is 0xFF (no carries cross bytes: low seven bits + 1 <= 0x80)
    uint32_t ff = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint32_t t = ((q[k] & 0x7F7F7F7Fu) + 0x01010101u) & q[k] & 0x80808080u;
      ff |= ((t >> 7) & 1) << (4 * k) | ((t >> 15) & 1) << (4 * k + 1) | ((t >> 23) & 1) << (4 * k + 2) |
            ((t >> 31) & 1) << (4 * k + 3);
    }
    // bit j: byte j brings 7 bits (the byte before it is 0xFF)
    const uint32_t inmask = nin >= 16 ? 0xFFFFu : (1u << nin) - 1;
    uint32_t seven = ((ff << 1) | uint32_t(prev_l == 0xFF)) & inmask;
    const uint32_t sum = 8 * nin - (uint32_t)__builtin_popcount(seven);
    uint32_t v0 = q[0], v1 = q[1], v2 = q[2], v3 = q[3];
    while (seven) {  // highest first: bytes above byte i move down one bit
      const uint32_t i = 31 - (uint32_t)__builtin_clz(seven);
      seven &= ~(1u << i);
      const uint32_t bb = 8 * (i + 1);  // bits kept in place
      const uint32_t m0 = bb >= 32 ? ~0u : (1u << bb) - 1;
      const uint32_t m1 = bb >= 64 ? ~0u : bb <= 32 ? 0u : (1u << (bb - 32)) - 1;
      const uint32_t m2 = bb >= 96 ? ~0u : bb <= 64 ? 0u : (1u << (bb - 64)) - 1;
      const uint32_t m3 = bb >= 128 ? ~0u : bb <= 96 ? 0u : (1u << (bb - 96)) - 1;
      const uint32_t h0 = v0 & ~m0, h1 = v1 & ~m1, h2 = v2 & ~m2, h3 = v3 & ~m3;
      v0 = (v0 & m0) | __builtin_amdgcn_alignbit(h1, h0, 1);
      v1 = (v1 & m1) | __builtin_amdgcn_alignbit(h2, h1, 1);
      v2 = (v2 & m2) | __builtin_amdgcn_alignbit(h3, h2, 1);
      v3 = (v3 & m3) | (h3 >> 1);
    }
    // exclusive prefix of the lane sums (<= 128)
    const uint32_t incl = wave_scan_incl(sum);
    const uint32_t excl = incl - sum, total = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
    // clear the words this chunk reaches past the partly filled one (their
    // previous contents are long consumed), then OR the lanes' bits in
    const uint32_t w0 = (end >> 5) + 1, w1 = ((end + total) >> 5) + 1;
    for (uint32_t k = w0 + lane; k <= w1; k += 64) ring[k % kMsRing] = 0;
    wave_sync();
    const uint32_t pos = end + excl, sh = pos & 31, wi = pos >> 5;
    const uint32_t r = 32 - sh;  // sh == 0: each word lands whole
    atomicOr(&ring[wi % kMsRing], v0 << sh);
    atomicOr(&ring[(wi + 1) % kMsRing], sh ? __builtin_amdgcn_alignbit(v1, v0, r) : v1);
    atomicOr(&ring[(wi + 2) % kMsRing], sh ? __builtin_amdgcn_alignbit(v2, v1, r) : v2);
    atomicOr(&ring[(wi + 3) % kMsRing], sh ? __builtin_amdgcn_alignbit(v3, v2, r) : v3);
    atomicOr(&ring[(wi + 4) % kMsRing], sh ? v3 >> r : 0u);
    prev = __shfl(q[3] >> 24, 63);
    if (next + 1024 > n) prev = 0;  // unused: the stream ends in this chunk
    end += total;
    next += 1024;
    if (next < n) {
      load_next();
    } else {
      // the bits past the end read as ones
      if (lane == 0) atomicOr(&ring[(end >> 5) % kMsRing], ~0u << (end & 31));
      nvalid = (end >> 5) + 1;
      ones = nvalid;
    }
    wave_sync();
  }
  // bits [from, from + need) readable: the stream holds them, or past its
  // end the ring words hold ones (filled ahead of the reads, never over the
  // live window: reads stay within 2 Kbit of `from`, far less than the ring)
  __device__ void ensure(uint32_t from, uint32_t need) {
    while (end < from + need && nvalid == ~0u) refill();
    if (nvalid != ~0u) {
      const uint32_t last = ((from + need) >> 5) + 1;
      if (ones <= last) {
        for (uint32_t k = ones + (threadIdx.x & 63); k <= last; k += 64) ring[k & (kMsRing - 1)] = ~0u;
        ones = last + 1;
        wave_sync();
      }
    }
  }
  __device__ uint32_t bits32(uint32_t pos) const {
    const uint32_t wi = pos >> 5;
    return __builtin_amdgcn_alignbit(ring[(wi + 1) & (kMsRing - 1)], ring[wi & (kMsRing - 1)], pos & 31);
  }
};

// Task i with its quad records (the whole wave; returns together).
__device__ void ht_magsgn_block(const HtDecTask* __restrict__ tasks, uint64_t i, const uint32_t* __restrict__ rec,
                                uint32_t* ringbuf, uint32_t* recs) {
  const uint32_t lane = threadIdx.x;
  const HtDecTask t = tasks[i];
  HtDecTask* slot = (HtDecTask*)tasks + i;
  const uint32_t width = t.w, height = t.h;
  uint32_t* out = (uint32_t*)t.out;
  const uint64_t stride = t.stride;
  // decode.rs:54-110, as BlockDec::start
  bool active = false, refine = false;
  uint32_t p = 0, shift = 0, mmsbp2 = 0, lcup = 0, scup = 0;
  if ((t.flags & kTaskIncluded) && t.seg && width && height) {
    uint32_t num_passes = t.passes;
    const uint32_t missing_msbs = t.mmsb;
    if (num_passes > 1 && t.len2 == 0) num_passes = 1;
    if (num_passes > 3 || missing_msbs >= 30 || t.len1 < 2 || width > 64 || height > 64) {
      if (lane == 0) report(t, kErrBlock);
    } else {
      if (missing_msbs == 29) num_passes = 1;
      refine = num_passes > 1;
      lcup = t.len1;
      scup = (ld8(t.seg + lcup - 1) << 4) + (ld8(t.seg + lcup - 2) & 0xF);
      if (scup < 2 || scup > lcup || scup > 4079) {
        if (lane == 0) report(t, kErrBlock);
      } else {
        active = true;
        p = 30 - missing_msbs;
        shift = 31 - t.kmax;
        mmsbp2 = missing_msbs + 2;
      }
    }
  }
  const uint32_t wcap = width < 64 ? width : 64, hcap = height < 64 ? height : 64;
  if (!active) {
    if (out)
      for (uint32_t y = 0; y < hcap; ++y)
        if (lane < wcap) st32(out + (uint64_t)y * stride + lane, 0);
    if (lane == 0 && t.passes > 1) slot->passes = 1;
    return;
  }
  const bool raw = (t.flags & kTaskSignMag) || refine;
  const bool col = lane < width;
  MsRing ms;
  ms.init(ringbuf, t.seg, lcup - scup);
  if (lane == 0) ringbuf[0] = 0;  // the first refill clears from word 1 on
  wave_sync();
  uint32_t msbase = 0;
  bool bad = false;
  uint32_t blprev = 0;  // bit length of this column's sample in the row above
  {
    // the block's quad records into LDS at once (16 bytes per load), so the
    // row loop reads them there instead of waiting on global loads that
    // queue behind its own stores
    const uint32_t words = ((height + 1) / 2) * 32;
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
      const uint32_t w = lane * 16 + 4 * k;
      if (w < words) *(uint4*)(recs + w) = *(const uint4*)(rec + w);
    }
    wave_sync();
  }
  uint32_t* orow = out + lane;  // row y of this lane's column
  for (uint32_t y = 0; y < height; y += 2) {
    const bool initial = y == 0;
    // ---- MagSgn of rows y, y+1, lane = column (decode.rs:302-396); the
    // next row's records are in flight meanwhile
    const uint32_t qinfo = col ? recs[(y / 2) * 32 + lane / 2] : 0u;
    const uint32_t tq = qinfo & 0xFFFF;
    uint32_t U = qinfo >> 16;
    if (!initial) {
      uint32_t gamma = tq & 0xF0;
      gamma &= gamma - 0x10;
      // columns 2q-1 .. 2q+2 of the row above (q = lane / 2); past either
      // edge 0: the pair's own two columns, the column left of the pair
      // (wave shift right by one, from the pair's even lane) and the one
      // right of it (wave shift left by one, from the pair's odd lane)
      const uint32_t pair = max(blprev, dpp(blprev, kDppSwapPair));
      const uint32_t left = dpp(dpp(blprev, kDppWaveShr1), kDppEvenOfPair);
      const uint32_t right = dpp(dpp(blprev, kDppWaveShl1), kDppOddOfPair);
      const uint32_t e4 = max(pair, max(left, right));
      U += gamma != 0 ? max(e4, 2u) - 1 : 1;
    }
    if (col && !(lane & 1) && U > mmsbp2) bad = true;
    const uint32_t nb0 = 2 * (lane & 1);
    uint32_t m[2], sig[2];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const uint32_t nbit = nb0 + r;
      sig[r] = col ? (tq >> (4 + nbit)) & 1 : 0u;
      m[r] = (U - ((tq >> (12 + nbit)) & 1)) & 31;
    }
    const uint32_t c0 = sig[0] ? m[0] : 0u, c1 = sig[1] ? m[1] : 0u;
    // exclusive prefix sum of the bit counts over the lanes
    const uint32_t cnt = c0 + c1;  // < 64
    const uint32_t incl = wave_scan_incl(cnt);
    const uint32_t excl = incl - cnt, total = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
    ms.ensure(msbase, total + 64);
    const uint32_t off = msbase + excl;
    uint32_t blnew = 0;
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const uint32_t nbit = nb0 + r;
      // read for every lane (the ring index is always in bounds), kept where significant
      const uint32_t ms_val = ms.bits32(off + (r ? c0 : 0u));
      uint32_t vv = ms_val & ((1u << m[r]) - 1);
      vv |= ((tq >> (8 + nbit)) & 1) << m[r];
      vv |= 1;
      const uint32_t val = (ms_val << 31) | ((vv + 2) << ((p - 1) & 31));  // the half-bin bit included
      uint32_t o = val;
      if (!raw) {
        const uint32_t mag = (val & 0x7FFFFFFFu) >> shift;
        o = (val >> 31) ? 0u - mag : mag;
      }
      if (!sig[r]) vv = o = 0;
      if (r == 1) blnew = bitlen(vv);
      if (col && y + r < height) st32(r ? orow + stride : orow, o);
    }
    msbase += total;
    blprev = blnew;
    orow += 2 * stride;
  }
  // a magnitude bound violated mid-block: the whole block reads as zero
  if (__ballot(bad)) {
    if (lane == 0) report(t, kErrBlock);
    for (uint32_t y = 0; y < hcap; ++y)
      if (lane < wcap) st32(out + (uint64_t)y * stride + lane, 0);
    if (lane == 0 && t.passes > 1) slot->passes = 1;
  }
}

__global__ __launch_bounds__(64) void ht_magsgn_kernel(const HtDecTask* __restrict__ tasks, uint64_t n,
                                                       const uint32_t* __restrict__ rec) {
  __shared__ uint32_t ring[kMsRing];
  __shared__ __align__(16) uint32_t recs[kRecWords];
  const uint64_t i = blockIdx.x + (uint64_t)blockIdx.y * gridDim.x;
  if (i >= n) return;
  ht_magsgn_block(tasks, i, rec + i * kRecWords, ring, recs);
}

// SigProp and MagRef passes (decode.rs:398-600) of the blocks that carry
// them, one lane per block, in place on the raw sign-magnitude samples the
// cleanup decode left; then the samples become coefficients.  The cleanup
// significance (what both passes read) is captured first, as one 64-bit row
// mask per block row in LDS.
__constant__ uint32_t c_spp_prop[4] = {0x33, 0x76, 0xEC, 0xC8};

__device__ __forceinline__ void refine_block(const HtDecTask& t, uint64_t* sigrow, uint16_t* prevrow, uint32_t lane) {
  if (!(t.flags & kTaskIncluded) || !t.seg || t.passes < 2 || t.len2 == 0 || t.mmsb >= 29) return;
  const uint32_t num_passes = t.passes;
  const uint32_t width = t.w, height = t.h;
  if (width == 0 || height == 0 || width > 64 || height > 64) return;
  uint32_t* out = (uint32_t*)t.out;
  const uint64_t stride = t.stride;
  const uint32_t p = 30 - t.mmsb;
  const bool causal = (t.flags & kTaskCausal) != 0;
  auto SIG = [&](uint32_t y) -> uint64_t& { return sigrow[y * 64 + lane]; };
  for (uint32_t y = 0; y < height; ++y) {
    uint64_t m = 0;
    for (uint32_t x = 0; x < width; ++x) m |= uint64_t(out[y * stride + x] != 0) << x;
    SIG(y) = m;
  }
  // 16-bit group masks: bit 4c + r is sample (row 4s + r, column 4g + c)
  auto group = [&](uint32_t s, uint32_t g) -> uint32_t {
    uint32_t b = 0;
    if (4 * g >= width) return 0;
    for (uint32_t r = 0; r < 4; ++r) {
      uint32_t y = 4 * s + r;
      if (y >= height) break;
      uint32_t bits = uint32_t(SIG(y) >> (4 * g)) & 0xF;
      for (uint32_t c = 0; c < 4; ++c)
        if (bits & (1u << c)) b |= 1u << (4 * c + r);
    }
    return b;
  };
  const uint32_t groups = (width + 3) / 4;
  for (uint32_t g = 0; g < groups + 2; ++g) prevrow[g * 64 + lane] = 0;
  // ---- SigProp (decode.rs:440-551)
  {
    FwdDec<0u> sp;
    sp.init(t.seg + t.len1, t.len2);
    for (uint32_t y = 0; y < height; y += 4) {
      uint32_t pattern = 0xFFFF;
      switch (height - y) {
        case 1: pattern = 0x1111; break;
        case 2: pattern = 0x3333; break;
        case 3: pattern = 0x7777; break;
        default: break;
      }
      uint32_t prev = 0;
      const uint32_t s = y / 4;
      for (uint32_t gi = 0, x = 0; x < width; x += 4, ++gi) {
        uint32_t overhang = x + 4 > width ? x + 4 - width : 0;
        pattern >>= overhang * 4;
        uint32_t ps = uint32_t(prevrow[gi * 64 + lane]) | (uint32_t(prevrow[(gi + 1) * 64 + lane]) << 16);
        uint32_t ns = group(s + 1, gi) | (group(s + 1, gi + 1) << 16);
        uint32_t cs = group(s, gi) | (group(s, gi + 1) << 16);
        uint32_t u = (ps & 0x88888888u) >> 3;
        if (!causal) u |= (ns & 0x11111111u) << 3;
        uint32_t mbr = cs;
        mbr |= (cs & 0x77777777u) << 1;
        mbr |= (cs & 0xEEEEEEEEu) >> 1;
        mbr |= u;
        uint32_t tt = mbr;
        mbr |= tt << 4;
        mbr |= tt >> 4;
        mbr |= prev >> 12;
        mbr &= pattern;
        mbr &= ~cs;
        uint32_t new_sig = mbr;
        if (new_sig != 0) {
          uint32_t cwd = sp.fetch();
          uint32_t cnt = 0;
          uint32_t inv_sig = ~cs & pattern;
          uint32_t col_mask = 0xF;
          for (uint32_t b = 0; b < 16; b += 4) {
            if (col_mask & new_sig) {
              uint32_t sample_mask = 0x1111 & col_mask;
              for (uint32_t row = 0; row < 4; ++row) {
                if (new_sig & sample_mask) {
                  new_sig &= ~sample_mask;
                  if (cwd & 1) new_sig |= (c_spp_prop[row] << b) & inv_sig;
                  cwd >>= 1;
                  cnt += 1;
                }
                sample_mask <<= 1;
              }
            }
            col_mask <<= 4;
          }
          if (new_sig != 0) {
            uint32_t val = 3u << (p - 2);
            uint32_t cm = 0xF;
            for (uint32_t col = 0; col < 4; ++col) {
              if (cm & new_sig) {
                uint32_t sample_mask = 0x1111 & cm;
                for (uint32_t row = 0; row < 4; ++row) {
                  if (new_sig & sample_mask) {
                    out[(uint64_t)(y + row) * stride + x + col] = (cwd << 31) | val;
                    cwd >>= 1;
                    cnt += 1;
                  }
                  sample_mask <<= 1;
                }
              }
              cm <<= 4;
            }
          }
          sp.advance(cnt);
        }
        new_sig |= cs;
        prevrow[gi * 64 + lane] = uint16_t(new_sig & 0xFFFF);
        uint32_t t2 = new_sig;
        new_sig |= (t2 & 0x7777) << 1;
        new_sig |= (t2 & 0xEEEE) >> 1;
        prev = (new_sig | u) & 0xF000;
      }
    }
  }
  // ---- MagRef (decode.rs:558-600)
  if (num_passes > 2) {
    RevDec mr;
    mr.init_mrp(t.seg, t.len1, t.len2);
    const uint32_t half = 1u << (p - 2);
    for (uint32_t y = 0; y < height; y += 4) {
      const uint32_t s = y / 4;
      for (uint32_t i8 = 0, gi = 0; i8 < width; i8 += 8, gi += 2) {
        uint32_t cwd = mr.fetch();
        uint32_t sig = group(s, gi) | (group(s, gi + 1) << 16);
        if (sig != 0) {
          uint32_t col_mask = 0xF;
          for (uint32_t j = 0; j < 8; ++j) {
            if (sig & col_mask) {
              uint32_t sample_mask = 0x11111111u & col_mask;
              for (uint32_t k = 0; k < 4; ++k) {
                if (sig & sample_mask) {
                  if (i8 + j < width && y + k < height) {
                    uint32_t sym = ((1 - (cwd & 1)) << (p - 1)) | half;
                    out[(uint64_t)(y + k) * stride + i8 + j] ^= sym;
                  }
                  cwd >>= 1;
                }
                sample_mask <<= 1;
              }
            }
            col_mask <<= 4;
          }
        }
        mr.advance(__builtin_popcount(sig));
      }
    }
  }
  // sign-magnitude -> coefficients
  if (!(t.flags & kTaskSignMag)) {
    const uint32_t shift = 31 - t.kmax;
    for (uint32_t y = 0; y < height; ++y)
      for (uint32_t x = 0; x < width; ++x) {
        uint32_t sm = out[y * stride + x];
        int32_t mag = (int32_t)((sm & 0x7FFFFFFFu) >> shift);
        out[y * stride + x] = (uint32_t)((sm & 0x80000000u) ? -mag : mag);
      }
  }
}

// Grid-stride over the blocks: the launch holds at most what the GPU keeps
// resident at once, so that it is cheap to place next to another stream's
// kernels when (as in cleanup-only streams) no block has the passes.
__global__ __launch_bounds__(64) void ht_refine_kernel(const HtDecTask* __restrict__ tasks, uint64_t n) {
  __shared__ uint64_t sigrow[64 * 64];
  __shared__ uint16_t prevrow[18 * 64];
  const uint32_t lane = threadIdx.x;
  for (uint64_t i = blockIdx.x * (uint64_t)64 + lane; i < n; i += (uint64_t)gridDim.x * 64) {
    const HtDecTask t = tasks[i];
    refine_block(t, sigrow, prevrow, lane);
  }
}

// ---------------------------------------------------------------------------
// device: encoders (mel.rs:175-262, streams.rs:267-441, encode.rs:24-331)
// ---------------------------------------------------------------------------
struct MelEnc {
  uint8_t* buf;
  uint32_t len, tmp, remaining, run, threshold;
  int k;
  __device__ void init(uint8_t* b) {
    buf = b;
    len = 0;
    tmp = 0;
    remaining = 8;
    run = 0;
    k = 0;
    threshold = 1;
  }
  __device__ void emit_bit(uint32_t v) {
    tmp = (tmp << 1) + v;
    remaining -= 1;
    if (remaining == 0) {
      if (len < kMelCap) st8(buf + len, (uint8_t)tmp);
      len += 1;
      remaining = tmp == 0xFF ? 7 : 8;
      tmp = 0;
    }
  }
  __device__ void encode(bool bit) {
    if (bit) {
      emit_bit(0);
      int t = c_mel_exp[k];
      while (t > 0) {
        t -= 1;
        emit_bit((run >> t) & 1);
      }
      run = 0;
      k = k > 0 ? k - 1 : 0;
      threshold = 1u << c_mel_exp[k];
    } else {
      run += 1;
      if (run >= threshold) {
        emit_bit(1);
        run = 0;
        k = k + 1 < 12 ? k + 1 : 12;
        threshold = 1u << c_mel_exp[k];
      }
    }
  }
};

struct VlcEnc {
  uint8_t* buf;  // production order; buf[0] is the 0xFF placeholder
  uint32_t len, tmp, used;
  bool gt8f;
  __device__ void init(uint8_t* b) {
    buf = b;
    st8(buf, 0xFF);
    len = 1;
    tmp = 0xF;
    used = 4;
    gt8f = true;
  }
  __device__ void encode(uint32_t cwd, uint32_t cwd_len) {
    while (cwd_len > 0) {
      uint32_t avail = 8 - (gt8f ? 1 : 0) - used;
      uint32_t t = avail < cwd_len ? avail : cwd_len;
      tmp |= (cwd & ((1u << t) - 1)) << used;
      used += t;
      avail -= t;
      cwd_len -= t;
      cwd >>= t;
      if (avail == 0) {
        if (gt8f && tmp != 0x7F) {
          gt8f = false;
          continue;
        }
        if (len < kVlcCap) st8(buf + len, (uint8_t)tmp);
        len += 1;
        gt8f = tmp > 0x8F;
        tmp = 0;
        used = 0;
      }
    }
  }
};

struct MsEnc {
  uint8_t* buf;
  uint32_t len, cap, tmp, used, max_bits;
  __device__ void init(uint8_t* b, uint32_t c) {
    buf = b;
    cap = c;
    len = 0;
    tmp = 0;
    used = 0;
    max_bits = 8;
  }
  __device__ void encode(uint32_t cwd, uint32_t cwd_len) {
    while (cwd_len > 0) {
      uint32_t room = max_bits - used;
      uint32_t t = room < cwd_len ? room : cwd_len;
      tmp |= (cwd & ((1u << t) - 1)) << used;
      used += t;
      cwd >>= t;
      cwd_len -= t;
      if (used >= max_bits) {
        if (len < cap) st8(buf + len, (uint8_t)tmp);
        len += 1;
        max_bits = tmp == 0xFF ? 7 : 8;
        tmp = 0;
        used = 0;
      }
    }
  }
  __device__ void terminate() {
    if (used > 0) {
      uint32_t t = max_bits - used;
      tmp |= (0xFFu & ((1u << t) - 1)) << used;
      used += t;
      if (tmp != 0xFF) {
        if (len < cap) st8(buf + len, (uint8_t)tmp);
        len += 1;
      }
    } else if (max_bits == 7) {
      len -= 1;
    }
  }
};

struct EQuad {
  uint32_t rho, e_q[4], e_qmax, s[4];
};

__device__ __forceinline__ EQuad load_quad(const HtEncTask& t, uint32_t x, uint32_t y, uint32_t p, uint32_t shift) {
  EQuad q;
  q.rho = 0;
  q.e_qmax = 0;
  for (int n = 0; n < 4; ++n) {
    q.e_q[n] = 0;
    q.s[n] = 0;
    uint32_t sx = x + n / 2, sy = y + n % 2;
    if (sx >= t.w || sy >= t.h) continue;
    uint32_t raw;
    if (t.flags & kTaskSignMag) {
      raw = ((const uint32_t*)t.in)[(uint64_t)sy * t.stride + sx];
    } else {
      int32_t v = ldi32(t.in + (uint64_t)sy * t.stride + sx);
      raw = (v < 0 ? 0x80000000u : 0u) | ((v < 0 ? 0u - (uint32_t)v : (uint32_t)v) << shift);
    }
    uint32_t val = ((raw + raw) >> p) & ~1u;
    if (val != 0) {
      q.rho |= 1u << n;
      val -= 1;
      q.e_q[n] = 32 - __clz(val);
      q.e_qmax = q.e_qmax > q.e_q[n] ? q.e_qmax : q.e_q[n];
      q.s[n] = (val - 1) + (raw >> 31);
    }
  }
  return q;
}

__device__ __forceinline__ uint32_t quad_eps(const EQuad& q, uint32_t u_q) {
  if (u_q == 0) return 0;
  uint32_t e = 0;
  for (int n = 0; n < 4; ++n) e |= uint32_t((q.rho & (1u << n)) != 0 && q.e_q[n] == q.e_qmax) << n;
  return e;
}

__device__ __forceinline__ void emit_ms(MsEnc& ms, const EQuad& q, uint32_t tuple, uint32_t uq) {
  for (uint32_t n = 0; n < 4; ++n)
    if (q.rho & (1u << n)) {
      uint32_t m = uq - ((tuple >> n) & 1);
      uint32_t mask = m >= 32 ? 0xFFFFFFFFu : (1u << m) - 1;
      ms.encode(q.s[n] & mask, m);
    }
}

// tables.rs:39-54 (u -> prefix, prefix_len, suffix, suffix_len)
__device__ __forceinline__ void uvlc_code(uint32_t u, uint32_t* pf, uint32_t* pl, uint32_t* sf, uint32_t* sl) {
  if (u == 0) { *pf = 0; *pl = 0; *sf = 0; *sl = 0; }
  else if (u == 1) { *pf = 1; *pl = 1; *sf = 0; *sl = 0; }
  else if (u == 2) { *pf = 2; *pl = 2; *sf = 0; *sl = 0; }
  else if (u <= 4) { *pf = 4; *pl = 3; *sf = u - 3; *sl = 1; }
  else { *pf = 0; *pl = 3; *sf = u - 5; *sl = 5; }
}

__device__ __forceinline__ void vlc_pair(VlcEnc& v, uint32_t u0, uint32_t u1) {
  uint32_t p0, l0, s0, sl0, p1, l1, s1, sl1;
  uvlc_code(u0, &p0, &l0, &s0, &sl0);
  uvlc_code(u1, &p1, &l1, &s1, &sl1);
  v.encode(p0, l0);
  v.encode(p1, l1);
  v.encode(s0, sl0);
  v.encode(s1, sl1);
}

__device__ void encode_one(const HtEncTask& t, uint8_t* le, uint8_t* lc) {
  auto EV = [&](uint32_t k) -> uint8_t& { return le[k * 64]; };
  auto CX = [&](uint32_t k) -> uint8_t& { return lc[k * 64]; };
  const uint32_t width = t.w, height = t.h;
  const uint32_t missing_msbs = (uint32_t)t.kmax - 1;
  const uint32_t p = 30 - missing_msbs, shift = 31 - t.kmax;
  MelEnc mel;
  mel.init(t.slot + t.ms_cap);
  VlcEnc vlc;
  vlc.init(t.slot + t.ms_cap + kMelCap);
  MsEnc ms;
  ms.init(t.slot, t.ms_cap);
  uint32_t cols = (width + 1) / 2;
  for (uint32_t k = 0; k < cols + 4; ++k) {
    EV(k) = 0;
    CX(k) = 0;
  }
  uint32_t any = 0;
  // ---- initial row (kappa = 1), encode.rs:150-238
  uint32_t c_q0 = 0, lep = 0, lcxp = 0;
  for (uint32_t x = 0; x < width; x += 4) {
    EQuad q0 = load_quad(t, x, 0, p, shift);
    any |= q0.rho;
    uint32_t uq0 = q0.e_qmax > 1 ? q0.e_qmax : 1;
    uint32_t u_q0 = uq0 - 1, u_q1 = 0;
    uint32_t eps0 = quad_eps(q0, u_q0);
    EV(lep) = max(uint32_t(EV(lep)), q0.e_q[1]);
    lep += 1;
    EV(lep) = (uint8_t)q0.e_q[3];
    CX(lcxp) |= (uint8_t)((q0.rho & 2) >> 1);
    lcxp += 1;
    CX(lcxp) = (uint8_t)((q0.rho & 8) >> 3);
    uint32_t tuple0 = c_enc0[(c_q0 << 8) + (q0.rho << 4) + eps0];
    vlc.encode(tuple0 >> 8, (tuple0 >> 4) & 7);
    if (c_q0 == 0) mel.encode(q0.rho != 0);
    emit_ms(ms, q0, tuple0, uq0);
    if (x + 2 < width) {
      EQuad q1 = load_quad(t, x + 2, 0, p, shift);
      any |= q1.rho;
      uint32_t c_q1 = (q0.rho >> 1) | (q0.rho & 1);
      uint32_t uq1 = q1.e_qmax > 1 ? q1.e_qmax : 1;
      u_q1 = uq1 - 1;
      uint32_t eps1 = quad_eps(q1, u_q1);
      EV(lep) = max(uint32_t(EV(lep)), q1.e_q[1]);
      lep += 1;
      EV(lep) = (uint8_t)q1.e_q[3];
      CX(lcxp) |= (uint8_t)((q1.rho & 2) >> 1);
      lcxp += 1;
      CX(lcxp) = (uint8_t)((q1.rho & 8) >> 3);
      uint32_t tuple1 = c_enc0[(c_q1 << 8) + (q1.rho << 4) + eps1];
      vlc.encode(tuple1 >> 8, (tuple1 >> 4) & 7);
      if (c_q1 == 0) mel.encode(q1.rho != 0);
      emit_ms(ms, q1, tuple1, uq1);
      c_q0 = (q1.rho >> 1) | (q1.rho & 1);
    } else {
      c_q0 = 0;
    }
    if (u_q0 > 0 && u_q1 > 0) mel.encode((u_q0 < u_q1 ? u_q0 : u_q1) > 2);
    if (u_q0 > 2 && u_q1 > 2) {
      vlc_pair(vlc, u_q0 - 2, u_q1 - 2);
    } else if (u_q0 > 2 && u_q1 > 0) {
      uint32_t pf, pl, sf, sl;
      uvlc_code(u_q0, &pf, &pl, &sf, &sl);
      vlc.encode(pf, pl);
      vlc.encode(u_q1 - 1, 1);
      vlc.encode(sf, sl);
    } else {
      vlc_pair(vlc, u_q0, u_q1);
    }
  }
  EV(lep + 1) = 0;
  // ---- non-initial rows, encode.rs:241-309
  for (uint32_t y = 2; y < height; y += 2) {
    lep = 0;
    int max_e = int(max(uint32_t(EV(0)), uint32_t(EV(1)))) - 1;
    EV(0) = 0;
    lcxp = 0;
    c_q0 = uint32_t(CX(0)) + (uint32_t(CX(1)) << 2);
    CX(0) = 0;
    for (uint32_t x = 0; x < width; x += 4) {
      EQuad q0 = load_quad(t, x, y, p, shift);
      any |= q0.rho;
      uint32_t kappa = (q0.rho & (q0.rho - 1)) != 0 ? uint32_t(max_e > 1 ? max_e : 1) : 1;
      uint32_t uq0 = q0.e_qmax > kappa ? q0.e_qmax : kappa;
      uint32_t u_q0 = uq0 - kappa, u_q1 = 0;
      uint32_t eps0 = quad_eps(q0, u_q0);
      EV(lep) = max(uint32_t(EV(lep)), q0.e_q[1]);
      lep += 1;
      max_e = int(max(uint32_t(EV(lep)), uint32_t(EV(lep + 1)))) - 1;
      EV(lep) = (uint8_t)q0.e_q[3];
      CX(lcxp) |= (uint8_t)((q0.rho & 2) >> 1);
      lcxp += 1;
      uint32_t c_q1 = uint32_t(CX(lcxp)) + (uint32_t(CX(lcxp + 1)) << 2);
      CX(lcxp) = (uint8_t)((q0.rho & 8) >> 3);
      uint32_t tuple0 = c_enc1[(c_q0 << 8) + (q0.rho << 4) + eps0];
      vlc.encode(tuple0 >> 8, (tuple0 >> 4) & 7);
      if (c_q0 == 0) mel.encode(q0.rho != 0);
      emit_ms(ms, q0, tuple0, uq0);
      if (x + 2 < width) {
        EQuad q1 = load_quad(t, x + 2, y, p, shift);
        any |= q1.rho;
        uint32_t kappa1 = (q1.rho & (q1.rho - 1)) != 0 ? uint32_t(max_e > 1 ? max_e : 1) : 1;
        c_q1 |= ((q0.rho & 4) >> 1) | ((q0.rho & 8) >> 2);
        uint32_t uq1 = q1.e_qmax > kappa1 ? q1.e_qmax : kappa1;
        u_q1 = uq1 - kappa1;
        uint32_t eps1 = quad_eps(q1, u_q1);
        EV(lep) = max(uint32_t(EV(lep)), q1.e_q[1]);
        lep += 1;
        max_e = int(max(uint32_t(EV(lep)), uint32_t(EV(lep + 1)))) - 1;
        EV(lep) = (uint8_t)q1.e_q[3];
        CX(lcxp) |= (uint8_t)((q1.rho & 2) >> 1);
        lcxp += 1;
        c_q0 = uint32_t(CX(lcxp)) + (uint32_t(CX(lcxp + 1)) << 2);
        CX(lcxp) = (uint8_t)((q1.rho & 8) >> 3);
        uint32_t tuple1 = c_enc1[(c_q1 << 8) + (q1.rho << 4) + eps1];
        vlc.encode(tuple1 >> 8, (tuple1 >> 4) & 7);
        if (c_q1 == 0) mel.encode(q1.rho != 0);
        emit_ms(ms, q1, tuple1, uq1);
        c_q0 |= ((q1.rho & 4) >> 1) | ((q1.rho & 8) >> 2);
      } else {
        c_q0 = c_q1;
      }
      vlc_pair(vlc, u_q0, u_q1);
    }
  }
  if (!any) {
    t.lens[0] = t.lens[1] = t.lens[2] = t.lens[3] = 0;
    return;
  }
  // ---- terminate (mel.rs:247-261, streams.rs:348-372, 421-440)
  if (mel.run > 0) mel.emit_bit(1);
  uint32_t mel_tmp = mel.tmp << mel.remaining;
  uint32_t mel_mask = (0xFFu << mel.remaining) & 0xFF;
  uint32_t vlc_mask = vlc.used == 0 ? 0 : (0xFFu >> (8 - vlc.used));
  if ((mel_mask | vlc_mask) != 0) {
    uint32_t fuse = mel_tmp | vlc.tmp;
    if ((((fuse ^ mel_tmp) & mel_mask) | ((fuse ^ vlc.tmp) & vlc_mask)) == 0 && fuse != 0xFF && vlc.len > 1) {
      if (mel.len < kMelCap) mel.buf[mel.len] = (uint8_t)fuse;
      mel.len += 1;
    } else {
      if (mel.len < kMelCap) mel.buf[mel.len] = (uint8_t)mel_tmp;
      mel.len += 1;
      if (vlc.len < kVlcCap) vlc.buf[vlc.len] = (uint8_t)vlc.tmp;
      vlc.len += 1;
    }
  }
  ms.terminate();
  uint32_t scup = mel.len + vlc.len;
  if (scup < 2 || scup > 4079 || mel.len > kMelCap || vlc.len > kVlcCap || ms.len > ms.cap) {
    if (t.err) atomicCAS(t.err, 0, kErrEncodeScup);
    t.lens[0] = t.lens[1] = t.lens[2] = t.lens[3] = 0;
    return;
  }
  // Scup patch on the last two bytes of [MagSgn][MEL][VLC reversed]: the last
  // byte is the VLC placeholder (production index 0).
  uint32_t lens[3] = {ms.len, mel.len, vlc.len};
  uint32_t total = ms.len + mel.len + vlc.len;
  vlc.buf[0] = uint8_t(scup >> 4);
  uint32_t k2 = total - 2;
  uint8_t* b2;
  if (k2 < lens[0]) b2 = t.slot + k2;
  else if (k2 < lens[0] + lens[1]) b2 = t.slot + t.ms_cap + (k2 - lens[0]);
  else b2 = vlc.buf + (lens[2] - 1 - (k2 - lens[0] - lens[1]));
  *b2 = uint8_t((*b2 & 0xF0) | (scup & 0xF));
  t.lens[0] = ms.len;
  t.lens[1] = mel.len;
  t.lens[2] = vlc.len;
  t.lens[3] = 1;
}

__global__ __launch_bounds__(64) void ht_encode_kernel(const HtEncTask* __restrict__ tasks, uint64_t n) {
  __shared__ uint8_t le[36 * 64];
  __shared__ uint8_t lc[36 * 64];
  uint64_t i = blockIdx.x * (uint64_t)64 + threadIdx.x;
  if (i >= n) return;
  HtEncTask t = tasks[i];
  encode_one(t, le + threadIdx.x, lc + threadIdx.x);
}


// ---------------------------------------------------------------------------
// Wave-parallel cleanup encode: one wavefront per code-block, lane c owning
// column c of each quad row (rows y, y+1).
//
// Every quad's significance pattern rho, exponents, exponent bound from the
// row above (neighbouring lanes), context, CxtVLC codeword, u and embedded-MSB
// mask are computed by its two lanes at once.  The MagSgn bits of a row are
// placed by a wave prefix sum of the lanes' bit counts and OR-ed into an LDS
// bit array; so are each quad pair's VLC bits (codewords, then the UVLC
// prefixes and suffixes, assembled by the pair's last lane).  MEL events are
// kept as two ballots per row and coded after the last row by a wave-uniform
// loop that takes runs of zero events in bulk.  The raw MagSgn and VLC bits
// then become stuffed bytes up to 256 at a time: each step assumes 8-bit
// bytes, finds by a ballot the first byte that the stuffing rule shortens
// (after 0xFF for MagSgn; seven ones after a byte above 0x8F for VLC),
// commits the bytes before it and restarts there.  Slot, lengths and errors
// equal ht_encode_kernel's.
// ---------------------------------------------------------------------------
constexpr uint32_t kEncMsWords = 3976;  // 4096 samples x 31 bits + slack
constexpr uint32_t kEncVlcWords = 488;  // 4 + 512 pairs x 30 bits + slack

// ORs the low n (<= 64) bits of v into the LDS bit array at bit pos.
__device__ __forceinline__ void lds_or_bits(uint32_t* buf, uint32_t pos, uint64_t v, uint32_t n) {
  if (n == 0) return;
  const uint32_t s = pos & 31, w = pos >> 5;
  atomicOr(&buf[w], (uint32_t)(v << s));
  if (s + n > 32) atomicOr(&buf[w + 1], (uint32_t)(s ? v >> (32 - s) : v >> 32));
  if (s + n > 64) atomicOr(&buf[w + 2], (uint32_t)(v >> (64 - s)));
}

// Bits [pos, pos + 32) of an LDS bit array of `words` words (reads past the
// end are clamped: callers never use such bits).
__device__ __forceinline__ uint32_t lds_window(const uint32_t* buf, uint32_t pos, uint32_t words) {
  uint32_t w = pos >> 5;
  w = w < words - 1 ? w : words - 2;
  return __builtin_amdgcn_alignbit(buf[w + 1], buf[w], pos & 31);
}

// Emits the full bytes of raw bits [0, T) of R under the MagSgn (Vlc =
// false: a byte after 0xFF carries 7 bits) or VLC (a byte after one above
// 0x8F carries 7 bits when they are all ones) stuffing rule: byte k goes to
// dst[k] for k < cap, except byte 0 when hold_first (its value is returned
// in *first).  prev0 stands for the byte before the first.  Returns the byte
// count; *endpos is the first raw bit not in a full byte, *last the last
// byte (prev0 if none).
template <bool Vlc>
__device__ __forceinline__ uint32_t stuff_bytes(const uint32_t* R, uint32_t words, uint32_t T, uint32_t prev0, uint8_t* dst,
                                uint32_t cap, bool hold_first, uint32_t* endpos, uint32_t* last, uint32_t* first) {
  const uint32_t lane = threadIdx.x & 63;
  uint32_t pos = 0, ob = 0, prev = prev0;
  *first = 0;
  while (true) {
    uint32_t w0 = 8;
    if (Vlc) {
      if (prev > 0x8F && pos + 7 <= T && (lds_window(R, pos, words) & 0x7F) == 0x7F) w0 = 7;
    } else if (prev == 0xFF) {
      w0 = 7;
    }
    if (pos + w0 > T) break;  // the next byte is partial
    const uint32_t nfull = min(256u, 1u + (T - pos - w0) / 8);
    // byte 0 at pos (w0 bits), byte j >= 1 at pos + w0 + 8 (j - 1)
    const uint32_t j0 = 4 * lane;
    const uint32_t s0 = j0 == 0 ? pos : pos + w0 + 8 * (j0 - 1);
    const uint32_t win = lds_window(R, s0, words);
    const uint32_t sh = j0 == 0 ? w0 : 8;
    uint32_t v[4];
    v[0] = win & (sh == 8 ? 0xFFu : 0x7Fu);
    v[1] = (win >> sh) & 0xFF;
    v[2] = (win >> (sh + 8)) & 0xFF;
    v[3] = (win >> (sh + 16)) & 0xFF;
    const uint32_t before = __shfl_up(v[3], 1);
    // the first byte j in 1..min(nfull, 255) of this lane that its
    // predecessor shortens
    uint32_t ev = 4;
#pragma unroll
    for (int k = 3; k >= 0; --k) {
      const uint32_t j = j0 + k;
      if (j == 0 || j > nfull) continue;
      const uint32_t pv = k == 0 ? before : v[k - 1];
      const bool e = Vlc ? (pv > 0x8F && (v[k] & 0x7F) == 0x7F) : (pv == 0xFF);
      if (e) ev = (uint32_t)k;
    }
    const uint64_t evm = __ballot(ev < 4);
    uint32_t commit = nfull;
    if (evm) {
      const uint32_t fl = (uint32_t)__builtin_ctzll(evm);
      commit = 4 * fl + (uint32_t)__builtin_amdgcn_readlane((int)ev, (int)fl);
    }
    if (ob == 0 && hold_first) *first = (uint32_t)__builtin_amdgcn_readlane((int)v[0], 0);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint32_t j = j0 + k, o = ob + j;
      if (j < commit && o < cap && !(hold_first && o == 0)) st8(dst + o, (uint8_t)v[k]);
    }
    const uint32_t lj = commit - 1;
    uint32_t lv = v[0];
    if ((lj & 3) == 1) lv = v[1];
    else if ((lj & 3) == 2) lv = v[2];
    else if ((lj & 3) == 3) lv = v[3];
    prev = (uint32_t)__builtin_amdgcn_readlane((int)lv, (int)(lj >> 2));
    pos = pos + w0 + 8 * lj;
    ob += commit;
  }
  *endpos = pos;
  *last = prev;
  return ob;
}

// MEL coder of mel.rs:175-262, wave-uniform; lane 0 stores the bytes.
struct WaveMelEnc {
  uint8_t* buf;
  uint32_t len, tmp, remaining, run, threshold, k, lastb;
  __device__ void init(uint8_t* b) {
    buf = b;
    len = 0;
    lastb = 0;
    tmp = 0;
    remaining = 8;
    run = 0;
    k = 0;
    threshold = 1;
  }
  __device__ void bit(uint32_t v) {
    tmp = (tmp << 1) + v;
    remaining -= 1;
    if (remaining == 0) {
      if (len < kMelCap && (threadIdx.x & 63) == 0) st8(buf + len, (uint8_t)tmp);
      lastb = tmp;
      len += 1;
      remaining = tmp == 0xFF ? 7 : 8;
      tmp = 0;
    }
  }
  // z zero events in a row
  __device__ void zeros(uint32_t z) {
    while (z) {
      const uint32_t need = threshold - run;
      if (z < need) {
        run += z;
        return;
      }
      z -= need;
      bit(1);
      run = 0;
      k = k + 1 < 12 ? k + 1 : 12;
      threshold = 1u << c_mel_exp[k];
    }
  }
  __device__ void one() {
    bit(0);
    for (int t = (int)c_mel_exp[k]; t > 0;) {
      t -= 1;
      bit((run >> t) & 1);
    }
    run = 0;
    k = k > 0 ? k - 1 : 0;
    threshold = 1u << c_mel_exp[k];
  }
};

// Encodes task i (the whole wave; returns together).  erow / srow: 1 +
// column, the row above's bottom-sample exponent / significance.
__device__ void ht_encode_wave_block(const HtEncTask* __restrict__ tasks, uint64_t i, uint32_t* msraw,
                                     uint32_t* vlcraw, uint32_t* erow, uint32_t* srow, uint64_t* melhas,
                                     uint64_t* melbit) {
  const uint32_t lane = threadIdx.x;
  const HtEncTask t = tasks[i];
  const uint32_t width = t.w, height = t.h;
  const uint32_t p = 31 - (uint32_t)t.kmax;  // 30 - missing_msbs
  const uint32_t shift = 31 - (uint32_t)t.kmax;
  const uint32_t ms_words = min(kEncMsWords, (width * height * ((uint32_t)t.kmax + 1) + 31) / 32 + 3);
  for (uint32_t w = lane; w < ms_words; w += 64) msraw[w] = 0;
  for (uint32_t w = lane; w < kEncVlcWords; w += 64) vlcraw[w] = 0;
  if (lane < 32) melhas[lane] = melbit[lane] = 0;
  erow[lane] = 0;
  srow[lane] = 0;
  if (lane < 4) erow[64 + lane] = srow[64 + lane] = 0;
  __syncthreads();
  if (lane == 0) vlcraw[0] = 0xF;  // VlcEnc starts with four ones (streams.rs:267-300)
  const uint32_t c = lane, j = lane & 1, q = lane >> 1;
  const uint32_t lq = (lane & ~1u) - 2;  // first lane of the left quad
  uint32_t mspos = 0, vlcpos = 4;
  uint64_t anysig = 0;
  for (uint32_t y = 0, r = 0; y < height; y += 2, ++r) {
    const bool initial = y == 0;
    // ---- the two samples of this column (encode.rs:24-60)
    uint32_t sig[2], e[2], s[2];
#pragma unroll
    for (int rr = 0; rr < 2; ++rr) {
      sig[rr] = e[rr] = s[rr] = 0;
      if (c < width && y + rr < height) {
        uint32_t raw;
        if (t.flags & kTaskSignMag) {
          raw = ((const uint32_t*)t.in)[(uint64_t)(y + rr) * t.stride + c];
        } else {
          const int32_t v = ldi32(t.in + (uint64_t)(y + rr) * t.stride + c);
          raw = (v < 0 ? 0x80000000u : 0u) | ((v < 0 ? 0u - (uint32_t)v : (uint32_t)v) << shift);
        }
        uint32_t val = ((raw + raw) >> p) & ~1u;
        if (val != 0) {
          sig[rr] = 1;
          val -= 1;
          e[rr] = 32 - __clz(val);
          s[rr] = (val - 1) + (raw >> 31);
        }
      }
    }
    // ---- the quad (lanes 2q, 2q+1): sample n = 2 j + row
    const uint32_t mine = (sig[0] | (sig[1] << 1)) << (2 * j);
    const uint32_t rho = mine | __shfl_xor(mine, 1);
    const uint32_t emine = max(e[0], e[1]);
    const uint32_t eqmax = max(emine, (uint32_t)__shfl_xor((int)emine, 1));
    anysig |= __ballot(rho != 0);
    const bool exists = 2 * q < width;
    const uint32_t rl = (uint32_t)__shfl((int)rho, (int)(lq & 63));
    const uint32_t rho_left = q > 0 ? rl : 0u;
    uint32_t cq, uq, kappa = 1;
    if (initial) {
      cq = q > 0 ? ((rho_left >> 1) | (rho_left & 1)) : 0u;
    } else {
      // context: the row above's bottom significance in columns 2q-1 .. 2q+2
      // and the left quad's right column (encode.rs:241-309)
      const uint32_t b = 2 * q;  // srow[col + 1]
      cq = (srow[b] | srow[b + 1]) | ((srow[b + 2] | srow[b + 3]) << 2) | ((((rho_left >> 2) | (rho_left >> 3)) & 1) << 1);
      const int max_e = (int)max(max(erow[b], erow[b + 1]), max(erow[b + 2], erow[b + 3])) - 1;
      if (rho & (rho - 1)) kappa = (uint32_t)(max_e > 1 ? max_e : 1);
    }
    uq = max(eqmax, kappa);
    const uint32_t u = uq - kappa;
    const uint32_t em = (uint32_t(sig[0] && e[0] == eqmax) | (uint32_t(sig[1] && e[1] == eqmax) << 1)) << (2 * j);
    const uint32_t emq = em | __shfl_xor(em, 1);
    const uint32_t eps = u > 0 ? emq : 0u;
    const uint32_t tuple = initial ? c_enc0[(cq << 8) + (rho << 4) + eps] : c_enc1[(cq << 8) + (rho << 4) + eps];
    // ---- MagSgn bits of this column's samples
    uint32_t m[2], nb[2];
#pragma unroll
    for (int rr = 0; rr < 2; ++rr) {
      m[rr] = uq - ((tuple >> (2 * j + rr)) & 1);
      nb[rr] = (exists && sig[rr]) ? m[rr] : 0u;
    }
    const uint64_t p0 = nb[0] ? (uint64_t)(s[0] & (m[0] >= 32 ? ~0u : (1u << m[0]) - 1)) : 0ull;
    const uint64_t p1 = nb[1] ? (uint64_t)(s[1] & (m[1] >= 32 ? ~0u : (1u << m[1]) - 1)) : 0ull;
    uint32_t mstot;
    const uint32_t moff = mspos + wave_excl_scan(nb[0] + nb[1], &mstot);
    lds_or_bits(msraw, moff, p0 | (p1 << nb[0]), nb[0] + nb[1]);
    mspos += mstot;
    // ---- MEL events: quad 2p at lane 4p, quad 2p+1 at lane 4p+2, the
    // initial row's UVLC event of the pair at lane 4p+3
    const uint32_t l4 = lane & 3;
    const uint32_t u_other = (uint32_t)__shfl((int)u, (int)(lane & ~3u));  // u of the pair's first quad
    const bool q1ex = (lane & ~3u) + 2 < width;
    const uint32_t u_second = (uint32_t)__shfl((int)u, (int)((lane & ~3u) + 2));
    const uint32_t u0 = u_other, u1 = q1ex ? u_second : 0u;
    bool has = false, bitv = false;
    if (exists && (l4 == 0 || l4 == 2) && cq == 0) {
      has = true;
      bitv = rho != 0;
    }
    if (initial && l4 == 3 && (lane & ~3u) < width && u0 > 0 && u1 > 0) {
      has = true;
      bitv = (u0 < u1 ? u0 : u1) > 2;
    }
    const uint64_t hm = __ballot(has), bm = __ballot(bitv);
    if (lane == 0) {
      melhas[r] = hm;
      melbit[r] = bm;
    }
    // ---- VLC bits of the pair, assembled at its last lane
    const uint32_t cw = tuple >> 8, cl = (tuple >> 4) & 7;
    const uint32_t cw0 = (uint32_t)__shfl((int)cw, (int)(lane & ~3u)), cl0 = (uint32_t)__shfl((int)cl, (int)(lane & ~3u));
    const uint32_t cw1 = (uint32_t)__shfl((int)cw, (int)((lane & ~3u) + 2)),
                   cl1 = (uint32_t)__shfl((int)cl, (int)((lane & ~3u) + 2));
    uint64_t vb = 0;
    uint32_t vn = 0;
    if (l4 == 3 && (lane & ~3u) < width) {
      auto put = [&](uint32_t bits, uint32_t len) {
        vb |= (uint64_t)(bits & ((1u << len) - 1)) << vn;
        vn += len;
      };
      put(cw0, cl0);
      if (q1ex) put(cw1, cl1);
      auto code = [](uint32_t uu, uint32_t* pf, uint32_t* pl, uint32_t* sf, uint32_t* sl) {
        if (uu == 0) { *pf = 0; *pl = 0; *sf = 0; *sl = 0; }
        else if (uu == 1) { *pf = 1; *pl = 1; *sf = 0; *sl = 0; }
        else if (uu == 2) { *pf = 2; *pl = 2; *sf = 0; *sl = 0; }
        else if (uu <= 4) { *pf = 4; *pl = 3; *sf = uu - 3; *sl = 1; }
        else { *pf = 0; *pl = 3; *sf = uu - 5; *sl = 5; }
      };
      uint32_t pf0, pl0, sf0, sl0, pf1, pl1, sf1, sl1;
      if (initial && u0 > 2 && u1 > 2) {
        code(u0 - 2, &pf0, &pl0, &sf0, &sl0);
        code(u1 - 2, &pf1, &pl1, &sf1, &sl1);
        put(pf0, pl0);
        put(pf1, pl1);
        put(sf0, sl0);
        put(sf1, sl1);
      } else if (initial && u0 > 2 && u1 > 0) {
        code(u0, &pf0, &pl0, &sf0, &sl0);
        put(pf0, pl0);
        put(u1 - 1, 1);
        put(sf0, sl0);
      } else {
        code(u0, &pf0, &pl0, &sf0, &sl0);
        code(u1, &pf1, &pl1, &sf1, &sl1);
        put(pf0, pl0);
        put(pf1, pl1);
        put(sf0, sl0);
        put(sf1, sl1);
      }
    }
    uint32_t vtot;
    const uint32_t voff = vlcpos + wave_excl_scan(vn, &vtot);
    lds_or_bits(vlcraw, voff, vb, vn);
    vlcpos += vtot;
    // ---- this row's bottom samples are the next row's context
    erow[c + 1] = (c < width) ? e[1] : 0u;
    srow[c + 1] = (c < width) ? sig[1] : 0u;
    __syncthreads();
  }
  if (!anysig) {
    if (lane == 0) t.lens[0] = t.lens[1] = t.lens[2] = t.lens[3] = 0;
    return;
  }
  // ---- MEL (wave-uniform)
  uint8_t* melbuf = t.slot + t.ms_cap;
  uint8_t* vlcbuf = t.slot + t.ms_cap + kMelCap;
  WaveMelEnc mel;
  mel.init(melbuf);
  const uint32_t rows = (height + 1) / 2;
  for (uint32_t r = 0; r < rows; ++r) {
    uint64_t has = melhas[r], bits = melbit[r];
    has = ((uint64_t)__builtin_amdgcn_readfirstlane((int)(has >> 32)) << 32) |
          (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)has);
    bits = ((uint64_t)__builtin_amdgcn_readfirstlane((int)(bits >> 32)) << 32) |
           (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)bits);
    bits &= has;
    while (has) {
      const uint64_t ones = bits;
      const uint32_t first1 = ones ? (uint32_t)__builtin_ctzll(ones) : 64u;
      const uint64_t below = first1 >= 64 ? ~0ull : ((1ull << first1) - 1);
      const uint32_t z = (uint32_t)__builtin_popcountll(has & below);
      mel.zeros(z);
      has &= ~below;
      if (first1 < 64) {
        mel.one();
        has &= ~(1ull << first1);
        bits &= ~(1ull << first1);
      }
    }
  }
  __syncthreads();
  // ---- MagSgn bytes (streams.rs:375-441)
  uint32_t ms_end, ms_last, unused;
  uint32_t ms_len = stuff_bytes<false>(msraw, kEncMsWords, mspos, 0u, t.slot, t.ms_cap, false, &ms_end, &ms_last,
                                       &unused);
  {
    const uint32_t rem = mspos - ms_end;
    const uint32_t maxb = ms_last == 0xFF ? 7u : 8u;
    if (rem > 0) {
      uint32_t tmp = lds_window(msraw, ms_end, kEncMsWords) & ((1u << rem) - 1);
      tmp |= (0xFFu & ((1u << (maxb - rem)) - 1)) << rem;
      if (tmp != 0xFF) {
        if (lane == 0 && ms_len < t.ms_cap) st8(t.slot + ms_len, (uint8_t)tmp);
        ms_len += 1;
      }
    } else if (maxb == 7) {
      ms_len -= 1;
    }
  }
  // ---- VLC bytes (streams.rs:267-372): production order after the 0xFF
  // placeholder; the first byte is held for the Scup patch
  uint32_t v_end, v_last, v_first;
  const uint32_t nv = stuff_bytes<true>(vlcraw, kEncVlcWords, vlcpos, 0xFFu, vlcbuf + 1, kVlcCap - 1, true, &v_end,
                                        &v_last, &v_first);
  uint32_t vlc_len = 1 + nv;
  const uint32_t vlc_used = vlcpos - v_end;
  const uint32_t vlc_tmp = vlc_used ? (lds_window(vlcraw, v_end, kEncVlcWords) & ((1u << vlc_used) - 1)) : 0u;
  // ---- terminate MEL and fuse its last byte with VLC's (mel.rs:247-261)
  if (mel.run > 0) mel.bit(1);
  const uint32_t mel_tmp = (mel.tmp << mel.remaining) & 0xFF;
  const uint32_t mel_mask = (0xFFu << mel.remaining) & 0xFF;
  const uint32_t vlc_mask = vlc_used == 0 ? 0u : (0xFFu >> (8 - vlc_used));
  if ((mel_mask | vlc_mask) != 0) {
    const uint32_t fuse = mel_tmp | vlc_tmp;
    if ((((fuse ^ mel_tmp) & mel_mask) | ((fuse ^ vlc_tmp) & vlc_mask)) == 0 && fuse != 0xFF && vlc_len > 1) {
      if (lane == 0 && mel.len < kMelCap) st8(melbuf + mel.len, (uint8_t)fuse);
      mel.lastb = fuse;
      mel.len += 1;
    } else {
      if (lane == 0 && mel.len < kMelCap) st8(melbuf + mel.len, (uint8_t)mel_tmp);
      mel.lastb = mel_tmp;
      mel.len += 1;
      if (vlc_len == 1) v_first = vlc_tmp;
      else if (lane == 0 && vlc_len < kVlcCap) st8(vlcbuf + vlc_len, (uint8_t)vlc_tmp);
      vlc_len += 1;
    }
  }
  const uint32_t scup = mel.len + vlc_len;
  if (scup < 2 || scup > 4079 || mel.len > kMelCap || vlc_len > kVlcCap || ms_len > t.ms_cap) {
    if (lane == 0) {
      if (t.err) atomicCAS(t.err, 0, kErrEncodeScup);
      t.lens[0] = t.lens[1] = t.lens[2] = t.lens[3] = 0;
    }
    return;
  }
  if (lane == 0) {
    // Scup in the last two bytes of [MagSgn][MEL][VLC reversed]: the last is
    // the placeholder (production index 0), the one before it production
    // index 1, or the last MEL byte when VLC has no other byte
    st8(vlcbuf, (uint8_t)(scup >> 4));
    if (vlc_len >= 2) {
      st8(vlcbuf + 1, (uint8_t)((v_first & 0xF0) | (scup & 0xF)));
    } else {
      st8(melbuf + mel.len - 1, (uint8_t)((mel.lastb & 0xF0) | (scup & 0xF)));
    }
    t.lens[0] = ms_len;
    t.lens[1] = mel.len;
    t.lens[2] = vlc_len;
    t.lens[3] = 1;
  }
}

// Persistent, fed by the same kind of queue as ht_decode_wave_kernel.
__global__ __launch_bounds__(64) void ht_encode_wave_kernel(const HtEncTask* __restrict__ tasks, uint64_t n,
                                                           uint64_t* queue, uint32_t grain) {
  __shared__ uint32_t msraw[kEncMsWords];
  __shared__ uint32_t vlcraw[kEncVlcWords];
  __shared__ uint32_t erow[68];
  __shared__ uint32_t srow[68];
  __shared__ uint64_t melhas[32], melbit[32];
  __shared__ uint64_t s_base;
  for (;;) {
    const uint64_t base = queue_pop(queue, grain, &s_base);
    if (base >= n) break;
    const uint64_t end = base + grain < n ? base + grain : n;
    for (uint64_t i = base; i < end; ++i) {
      ht_encode_wave_block(tasks, i, msraw, vlcraw, erow, srow, melhas, melbit);
      __syncthreads();
    }
  }
  queue_done(queue);
}

}  // namespace

void ht_init_tables() {
  static std::mutex mu;
  static bool done = false;
  std::lock_guard<std::mutex> lock(mu);
  if (done) return;
  std::vector<uint16_t> d0(1024), d1(1024), e0(2048), e1(2048), u0(320), u1(256);
  derive_dec(kCxtVlcRows0, d0.data());
  derive_dec(kCxtVlcRows1, d1.data());
  derive_enc(kCxtVlcRows0, e0.data());
  derive_enc(kCxtVlcRows1, e1.data());
  derive_uvlc(u0.data(), 5, true);
  derive_uvlc(u1.data(), 4, false);
  NDIC_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_dec0), d0.data(), 2048));
  NDIC_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_dec1), d1.data(), 2048));
  NDIC_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_enc0), e0.data(), 4096));
  NDIC_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_enc1), e1.data(), 4096));
  NDIC_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_uvlc0), u0.data(), 640));
  NDIC_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_uvlc1), u1.data(), 512));
  std::vector<uint32_t> w0(d0.begin(), d0.end()), w1(d1.begin(), d1.end()), wu0(u0.begin(), u0.end()),
      wu1(u1.begin(), u1.end());
  NDIC_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_dec0w), w0.data(), 4096));
  NDIC_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_dec1w), w1.data(), 4096));
  NDIC_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_uvlc0w), wu0.data(), 1280));
  NDIC_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_uvlc1w), wu1.data(), 1024));
  done = true;
}

namespace {
// Work class of a task: 0 heaviest ... kHtClasses - 1 (left out).
__device__ __forceinline__ uint32_t ht_class(const HtDecTask& t) {
  if (!(t.flags & kTaskIncluded) || !t.seg || t.len1 == 0) return kHtClasses - 1;
  const uint32_t bits = 32 - __clz(t.len1);                        // 1..16 (len1 < 65535)
  const uint32_t half = (t.len1 >> (bits > 1 ? bits - 2 : 0)) & 1;  // the bit below the top one
  return 33 - (2 * bits + half);                                     // 0 (longest) .. 31
}

__global__ __launch_bounds__(256) void ht_class_count(const HtDecTask* __restrict__ tasks, uint64_t n,
                                                      uint32_t* counts) {
  __shared__ uint32_t h[kHtClasses];
  for (uint32_t k = threadIdx.x; k < kHtClasses; k += blockDim.x) h[k] = 0;
  __syncthreads();
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    atomicAdd(&h[ht_class(tasks[i])], 1u);
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < kHtClasses; k += blockDim.x)
    if (h[k]) atomicAdd(&counts[k], h[k]);
}

__global__ void ht_class_scan(uint32_t* counts) {
  uint32_t acc = 0;
  for (uint32_t k = 0; k < kHtClasses; ++k) {
    const uint32_t c = counts[k];
    counts[k] = acc;
    acc += c;
  }
}

// One wave per 64 consecutive tasks: each class's lanes take consecutive
// slots (one atomic per class per wave), so neighbouring blocks of a class
// stay neighbours in the order.
__global__ __launch_bounds__(64) void ht_class_scatter(const HtDecTask* __restrict__ tasks, uint64_t n,
                                                       uint32_t* offsets, uint32_t* order) {
  const uint64_t i = blockIdx.x * (uint64_t)64 + threadIdx.x;
  const uint32_t lane = threadIdx.x;
  const uint32_t cls = i < n ? ht_class(tasks[i]) : 0xFFFFFFFFu;
  uint64_t pending = __ballot(i < n);
  while (pending) {
    const uint32_t leader = __ffsll((unsigned long long)pending) - 1;
    const uint32_t c = __shfl(cls, (int)leader);
    const uint64_t same = __ballot(cls == c) & pending;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(&offsets[c], (uint32_t)__popcll(same));
    base = __shfl(base, (int)leader);
    if (cls == c && i < n) order[base + __popcll(same & ((1ull << lane) - 1))] = (uint32_t)i;
    pending &= ~same;
  }
}
}  // namespace

void launch_ht_order(const HtDecTask* d_tasks, uint64_t n, uint32_t* d_counts, uint32_t* d_order, hipStream_t stream) {
  if (n == 0) return;
  launch_fill32(d_counts, 0, 4 * (kHtClasses + 1), stream);
  const uint32_t blocks = (uint32_t)std::min<uint64_t>(ceil_div(n, 256), 1024);
  hipLaunchKernelGGL(ht_class_count, dim3(blocks), dim3(256), 0, stream, d_tasks, n, d_counts);
  hipLaunchKernelGGL(ht_class_scan, dim3(1), dim3(1), 0, stream, d_counts);
  hipLaunchKernelGGL(ht_class_scatter, dim3((uint32_t)ceil_div(n, 64)), dim3(64), 0, stream, d_tasks, n, d_counts,
                     d_order);
  NDIC_HIP(hipGetLastError());
}

void launch_ht_decode(const HtDecTask* d_tasks, const uint32_t* d_order, uint64_t n, hipStream_t stream) {
  if (n == 0) return;
  static int cus = [] {
    int dev = 0, c = 0;
    NDIC_HIP(hipGetDevice(&dev));
    NDIC_HIP(hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev));
    return std::max(1, c);
  }();
  int64_t waves = tune(kTuneHtLaneWaves);
  if (waves == 0) waves = ceil_div(n, 64) >= 4 * (uint64_t)cus ? 4 : 1;
  if (waves == 4)
    hipLaunchKernelGGL(ht_decode_kernel<4>, dim3((uint32_t)ceil_div(n, 256)), dim3(256), 0, stream, d_tasks, d_order, n);
  else
    hipLaunchKernelGGL(ht_decode_kernel<1>, dim3((uint32_t)ceil_div(n, 64)), dim3(64), 0, stream, d_tasks, d_order, n);
  NDIC_HIP(hipGetLastError());
}

namespace {
// Workgroups of a persistent (or grid-stride) launch of a 64-thread kernel:
// enough to fill every CU at the kernel's occupancy, fewer when there are
// fewer grains.  The resident count is measured once per kernel.
template <class K> uint32_t persistent_grid(K kernel, uint64_t grains) {
  static std::mutex mu;
  static std::vector<std::pair<const void*, int>> caps;
  int cap = 0;
  {
    std::lock_guard<std::mutex> lock(mu);
    for (auto& c : caps)
      if (c.first == (const void*)kernel) cap = c.second;
    if (!cap) {
      int dev = 0, cus = 0, per = 0;
      NDIC_HIP(hipGetDevice(&dev));
      NDIC_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
      NDIC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, 64, 0));
      cap = std::max(1, cus * std::max(1, per));
      caps.emplace_back((const void*)kernel, cap);
    }
  }
  return (uint32_t)std::min<uint64_t>(grains, (uint64_t)cap);
}
}  // namespace

void launch_ht_vlc(const HtDecTask* d_tasks, uint64_t n, uint32_t* d_rec, hipStream_t stream) {
  if (n == 0) return;
  hipLaunchKernelGGL(ht_vlc_kernel<4>, dim3((uint32_t)ceil_div(n, 256)), dim3(256), 0, stream, d_tasks, n, d_rec);
  NDIC_HIP(hipGetLastError());
}

void launch_ht_magsgn(const HtDecTask* d_tasks, uint64_t n, const uint32_t* d_rec, hipStream_t stream) {
  if (n == 0) return;
  const uint32_t gx = (uint32_t)std::min<uint64_t>(n, 1u << 20), gy = (uint32_t)ceil_div(n, gx);
  hipLaunchKernelGGL(ht_magsgn_kernel, dim3(gx, gy), dim3(64), 0, stream, d_tasks, n, d_rec);
  NDIC_HIP(hipGetLastError());
}

void launch_ht_decode_split(const HtDecTask* d_tasks, uint64_t n, uint32_t* d_rec, hipStream_t stream) {
  launch_ht_vlc(d_tasks, n, d_rec, stream);
  launch_ht_magsgn(d_tasks, n, d_rec, stream);
}

void launch_ht_decode_wave(const HtDecTask* d_tasks, uint64_t n, uint64_t* queue, hipStream_t stream) {
  if (n == 0) return;
  const uint32_t grain = (uint32_t)std::max<int64_t>(1, tune(kTuneHtWaveGrain));
  hipLaunchKernelGGL(ht_decode_wave_kernel, dim3(persistent_grid(ht_decode_wave_kernel, ceil_div(n, grain))), dim3(64),
                     0, stream, d_tasks, n, queue, grain);
  NDIC_HIP(hipGetLastError());
}

void launch_ht_refine(const HtDecTask* d_tasks, uint64_t n, hipStream_t stream) {
  if (n == 0) return;
  hipLaunchKernelGGL(ht_refine_kernel, dim3(persistent_grid(ht_refine_kernel, ceil_div(n, 64))), dim3(64), 0, stream,
                     d_tasks, n);
  NDIC_HIP(hipGetLastError());
}

void launch_ht_encode_wave(const HtEncTask* d_tasks, uint64_t n, uint64_t* queue, hipStream_t stream) {
  if (n == 0) return;
  const uint32_t grain = (uint32_t)std::max<int64_t>(1, tune(kTuneHtWaveGrain));
  hipLaunchKernelGGL(ht_encode_wave_kernel, dim3(persistent_grid(ht_encode_wave_kernel, ceil_div(n, grain))), dim3(64),
                     0, stream, d_tasks, n, queue, grain);
  NDIC_HIP(hipGetLastError());
}

void launch_ht_encode(const HtEncTask* d_tasks, uint64_t n, hipStream_t stream) {
  if (n == 0) return;
  hipLaunchKernelGGL(ht_encode_kernel, dim3((uint32_t)ceil_div(n, 64)), dim3(64), 0, stream, d_tasks, n);
  NDIC_HIP(hipGetLastError());
}

}  // namespace ndic
