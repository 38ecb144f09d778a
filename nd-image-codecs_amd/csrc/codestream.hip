// Part 1/15 codestreams and the ndht container, parsed and written on the
// device, one wavefront, thread or workgroup per plane.
//
// Decode: each plane's parser locates its codestream (container index
// entry, crates/ndic-codestream/src/container.rs:287-340), parses the main
// and tile-part headers (crates/ndic-codestream/src/reader.rs:120-291,
// markers.rs:130-521, quant.rs:159-202), then every packet header in
// resolution order with its tag trees (packet.rs:199-362, tagtree.rs,
// bitio.rs:104-161), and writes one HT decode task per code-block.  The
// wavefront-per-plane form stages the bytes it parses and the tag trees in
// LDS first; the thread-per-plane form reads them from global memory.
//
// Encode: per-plane depth and reversible ranging (crates/ndic-zarr/src/
// htj2k.rs:261-276, quant.rs:51-85), packet headers with tag trees
// (packet.rs:69-168), markers and PLT/TLM (writer.rs:198-322), the container
// header and coefficient-plane index (container.rs:80-100, 249-280), then a
// workgroup per plane copies every piece into the packed output.
#include <algorithm>

#include "codestream.hpp"
#include "tune.hpp"

namespace ndic {

namespace {

constexpr uint32_t kMaxPackets = 33;  // one per resolution, 32 levels at most
constexpr uint16_t kSOC = 0xFF4F, kSIZ = 0xFF51, kCOD = 0xFF52, kCOC = 0xFF53, kQCD = 0xFF5C, kQCC = 0xFF5D,
                   kCAP = 0xFF50, kCOM = 0xFF64, kTLM = 0xFF55, kPLT = 0xFF58, kSOT = 0xFF90, kSOD = 0xFF93,
                   kEOC = 0xFFD9;

__device__ __forceinline__ uint32_t rd16(const uint8_t* p) { return (uint32_t(p[0]) << 8) | p[1]; }
__device__ __forceinline__ uint32_t rd32(const uint8_t* p) {
  return (uint32_t(p[0]) << 24) | (uint32_t(p[1]) << 16) | (uint32_t(p[2]) << 8) | p[3];
}
__device__ __forceinline__ uint32_t le32(const uint8_t* p) {
  return uint32_t(p[0]) | (uint32_t(p[1]) << 8) | (uint32_t(p[2]) << 16) | (uint32_t(p[3]) << 24);
}
__device__ __forceinline__ uint64_t le64(const uint8_t* p) { return uint64_t(le32(p)) | (uint64_t(le32(p + 4)) << 32); }

__device__ __forceinline__ uint32_t find_chunk(const DevChunk* c, uint32_t n, uint64_t plane) {
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) / 2;
    if (c[mid].plane_base <= plane) lo = mid;
    else hi = mid;
  }
  return lo;
}

__device__ void raise(const DevChunk& ch, int32_t code, uint64_t offset, uint64_t plane) {
  if (atomicCAS(ch.err, 0, code) == 0) {
    ch.err[1] = (int32_t)(uint32_t)offset;
    ch.err[2] = (int32_t)(uint32_t)(offset >> 32);
    ch.err[3] = (int32_t)plane;
  }
}

__device__ __forceinline__ uint32_t log2_ceil(uint32_t x) {
  uint32_t t = 31 - __clz(x);
  return t + ((x & (x - 1)) != 0 ? 1 : 0);
}

// Tag trees over an nx x ny grid (tagtree.rs:13-123): level 0 holds the
// leaves, each level halves the grid, one extra level reads 0.  Values and
// flags of both trees live in the plane's scratch.
constexpr uint32_t kTreeOffs = 34;  // level offsets a tag tree keeps (levels + 1 <= 34)
struct Tree {
  uint8_t* val;
  uint8_t* flag;
  uint32_t nx, ny, levels;
  uint32_t* off;  // kTreeOffs words the caller provides (LDS in the parsers)
  __device__ uint32_t init(uint8_t* mem, uint32_t* offs, uint32_t w, uint32_t h, uint8_t init_val) {
    off = offs;
    nx = w;
    ny = h;
    levels = 1 + max(log2_ceil(w), log2_ceil(h));
    uint32_t total = 0;
    for (uint32_t l = 0; l <= levels; ++l) {
      off[l] = total;
      uint32_t lw = max((w + (1u << l) - 1) >> l, 1u), lh = max((h + (1u << l) - 1) >> l, 1u);
      total += lw * lh;
    }
    val = mem;
    flag = mem + total;
    for (uint32_t i = 0; i < total; ++i) {
      val[i] = i < off[1] ? init_val : 0;
      flag[i] = 0;
    }
    return 2 * total;
  }
  __device__ __forceinline__ uint32_t idx(uint32_t x, uint32_t y, uint32_t l) const {
    uint32_t lw = max((nx + (1u << l) - 1) >> l, 1u);
    return off[l] + y * lw + x;
  }
  __device__ void reduce() {
    for (uint32_t l = 1; l <= levels; ++l) {
      uint32_t lw = max((nx + (1u << l) - 1) >> l, 1u), lh = max((ny + (1u << l) - 1) >> l, 1u);
      uint32_t clw = max((nx + (1u << (l - 1)) - 1) >> (l - 1), 1u), clh = max((ny + (1u << (l - 1)) - 1) >> (l - 1), 1u);
      for (uint32_t y = 0; y < lh; ++y)
        for (uint32_t x = 0; x < lw; ++x) {
          uint32_t m = 255;
          for (uint32_t d = 0; d < 4; ++d) {
            uint32_t cx = 2 * x + (d & 1), cy = 2 * y + (d >> 1);
            if (cx < clw && cy < clh) m = min(m, uint32_t(val[off[l - 1] + cy * clw + cx]));
          }
          val[off[l] + y * lw + x] = l == levels ? 0 : (uint8_t)m;
        }
    }
  }
};

uint32_t tree_bytes_host(uint32_t w, uint32_t h) {
  uint32_t lv = 1, t = 0;
  auto l2c = [](uint32_t x) {
    uint32_t r = 0;
    while ((1u << r) < x) ++r;
    return r;
  };
  lv = 1 + std::max(l2c(w), l2c(h));
  for (uint32_t l = 0; l <= lv; ++l)
    t += std::max((w + (1u << l) - 1) >> l, 1u) * std::max((h + (1u << l) - 1) >> l, 1u);
  return 2 * t;
}

// Bytes of one plane codestream: a window [base, base + n) copied to LDS,
// everything else read from global memory (the window is only a cache of
// the same bytes).
struct Bytes {
  const uint8_t* g;
  const uint8_t* s;
  uint64_t base;
  uint32_t n;
  __device__ __forceinline__ uint32_t operator[](uint64_t i) const {
    const uint64_t k = i - base;
    return k < n ? s[k] : g[i];
  }
};
__device__ __forceinline__ uint32_t rd16(const Bytes& b, uint64_t i) { return (b[i] << 8) | b[i + 1]; }
__device__ __forceinline__ uint32_t rd32(const Bytes& b, uint64_t i) {
  return (b[i] << 24) | (b[i + 1] << 16) | (b[i + 2] << 8) | b[i + 3];
}

// bitio.rs:104-161, over bytes [start, start + len) of a codestream
struct BitReader {
  Bytes src;
  uint64_t start, len, pos;
  uint32_t tmp, avail;
  bool unstuff, bad;
  __device__ void init(const Bytes& b, uint64_t st, uint64_t n) {
    src = b;
    start = st;
    len = n;
    pos = 0;
    tmp = 0;
    avail = 0;
    unstuff = false;
    bad = false;
  }
  // the next byte (7 bits after 0xFF); false, and `bad`, past the end
  __device__ bool fetch() {
    if (pos >= len) {
      bad = true;
      return false;
    }
    uint32_t t = src[start + pos];
    pos += 1;
    tmp = t;
    avail = 8 - (unstuff ? 1 : 0);
    unstuff = t == 0xFF;
    return true;
  }
  __device__ uint32_t bit() {
    if (avail == 0 && !fetch()) return 0;
    avail -= 1;
    return (tmp >> avail) & 1;
  }
  // n bits, most significant first, a byte's worth at a time (past the end
  // they read as 0, as bit() by bit())
  __device__ uint32_t bits(uint32_t n) {
    uint32_t v = 0;
    while (n) {
      if (avail == 0 && !fetch()) return v << n;
      const uint32_t k = n < avail ? n : avail;
      avail -= k;
      v = (v << k) | ((tmp >> avail) & ((1u << k) - 1));
      n -= k;
    }
    return v;
  }
  // the number of 0 bits before the next 1 bit, which is consumed (a run
  // that reaches the end stops there, `bad` set)
  __device__ uint32_t zeros_to_one() {
    uint32_t z = 0;
    for (;;) {
      if (avail == 0 && !fetch()) return z;
      const uint32_t b = tmp & ((1u << avail) - 1);
      if (b == 0) {
        z += avail;
        avail = 0;
        continue;
      }
      const uint32_t hb = 31 - __clz(b);  // position of the first 1 in the byte's remaining bits
      z += avail - 1 - hb;
      avail = hb;
      return z;
    }
  }
  // the number of 1 bits before the next 0 bit, which is consumed (a run
  // that reaches the end stops there, `bad` set)
  __device__ uint32_t ones_to_zero() {
    uint32_t c = 0;
    for (;;) {
      if (avail == 0 && !fetch()) return c;
      const uint32_t b = ~tmp & ((1u << avail) - 1);
      if (b == 0) {
        c += avail;
        avail = 0;
        continue;
      }
      const uint32_t hb = 31 - __clz(b);
      c += avail - 1 - hb;
      avail = hb;
      return c;
    }
  }
  __device__ uint64_t terminate() {
    if (unstuff) pos += pos < len ? 1 : 0;
    return pos;
  }
};

// bitio.rs:16-90
struct BitWriter {
  uint8_t* out;
  uint32_t len, cap, tmp, avail;
  __device__ void init(uint8_t* o, uint32_t c) {
    out = o;
    cap = c;
    len = 0;
    tmp = 0;
    avail = 8;
  }
  __device__ void bit(uint32_t b) {
    avail -= 1;
    tmp |= (b & 1) << avail;
    if (avail == 0) {
      if (len < cap) out[len] = (uint8_t)tmp;
      len += 1;
      avail = tmp == 0xFF ? 7 : 8;
      tmp = 0;
    }
  }
  __device__ void bits(uint32_t d, uint32_t n) {
    for (uint32_t i = n; i-- > 0;) bit(d >> i);
  }
  __device__ void terminate() {
    if (avail < 8) {
      if (len < cap) out[len] = (uint8_t)tmp;
      len += 1;
    }
  }
};

// ---------------------------------------------------------------------------
// decode: plane parse
// ---------------------------------------------------------------------------
struct Hdr {
  uint32_t xsiz, ysiz, depth, is_signed;
  uint32_t scod, decomps, cbw, cbh, cb_style, wavelet, layers, mct;
  uint32_t guard, style, nvals;
  uint16_t vals[100];  // style 0: exponents; style 2: (exponent << 11) | mantissa
};

// K_max of band index kidx (quant.rs:88-107): guard bits + exponent - 1
__device__ uint32_t kmax_of(const Hdr& h, uint32_t kidx) {
  uint32_t i = kidx < h.nvals ? kidx : h.nvals - 1;
  uint32_t eps = h.style == 0 ? h.vals[i] : (uint32_t)(h.vals[i] >> 11);
  return (eps == 0 ? 0 : eps - 1) + h.guard;
}

// The dequantisation step of band index kidx of a 9/7 plane decoded to
// resolution max_res: Quant::irrev_delta (quant.rs:204-221: (1 + mu / 2^11)
// x band gain / 2^eps) x 2^depth (T.800 E-3's 2^R_b) x the K factor the
// unscaled lifting (dwt/mod.rs:7-9) leaves in the band relative to the output
// LL -- K^2 per 2-D level of low-pass: LL of level L holds K^2L, HL / LH of
// level d K^2(d-1), HH of level d K^2(d-2), all over K^2(L - max_res) -- over
// 2^(31 - K_max), the alignment of the block decoder's magnitudes.  Double
// precision, rounded once to f32.
__device__ float irrev_scale(const Hdr& h, uint32_t kidx, uint32_t levels, uint32_t max_res) {
  constexpr double kK = 1.230174104914001;  // T.800 Table F.4
  constexpr double kK2 = kK * kK;
  const uint32_t res = kidx == 0 ? 0 : (kidx - 1) / 3 + 1, band = kidx == 0 ? 0 : (kidx - 1) % 3 + 1;
  const uint32_t i = kidx < h.nvals ? kidx : h.nvals - 1;
  const uint32_t v = h.vals[i], eps = v >> 11;
  const float gain = band == 0 ? 1.0f : band == 3 ? 4.0f : 2.0f;
  const float delta = (float)((v & 0x7FF) | 0x800) * gain / 2048.0f / __uint_as_float((127 + eps) << 23);
  const int d = res == 0 ? (int)levels : (int)(levels - res + 1);
  int e = res == 0 ? (int)levels : band == 3 ? d - 2 : d - 1;
  e -= (int)(levels - max_res);
  double f = e < 0 ? 1.0 / kK2 : 1.0;
  for (int k = 0; k < e; ++k) f *= kK2;
  return (float)ldexp((double)delta * f, (int)h.depth - 31 + (int)kmax_of(h, kidx));
}

// Locates plane p's codestream inside the chunk (container.rs:287-340,
// htj2k.rs:354-384); returns false after raising an error.
__device__ bool locate(const DevChunk& ch, uint64_t p, uint64_t gid, DevPlane& pl) {
  if (ch.no_container) {
    pl.cs_off = 0;
    pl.cs_len = ch.in_len;
    return true;
  }
  const uint8_t* b = ch.in;
  uint64_t n = ch.in_len;
  if (n < 8) return raise(ch, kErrChunkTruncated, n, gid), false;
  if (b[0] != 'n' || b[1] != 'd' || b[2] != 'h' || b[3] != 't') return raise(ch, kErrChunkMagic, 0, gid), false;
  if (b[4] != 1) return raise(ch, kErrChunkVersion, 4, gid), false;
  uint32_t flags = b[5];
  if (flags & ~1u) return raise(ch, kErrChunkFlags, 5, gid), false;
  uint32_t lv = b[6];
  if (lv > 32) return raise(ch, kErrChunkLevels, 6, gid), false;
  uint32_t ndim = b[7];
  if (ndim < 2 || ndim > 32) return raise(ch, kErrChunkNdim, 7, gid), false;
  uint64_t fixed = 8 + 4ull * ndim;
  if (n < fixed) return raise(ch, kErrChunkTruncated, n, gid), false;
  for (uint32_t i = 0; i < ndim; ++i)
    if (le32(b + 8 + 4 * i) == 0) return raise(ch, kErrChunkZeroDim, 8 + 4 * i, gid), false;
  if (ndim != ch.ndim) return raise(ch, kErrChunkShape, 0, gid), false;
  for (uint32_t i = 0; i < ndim; ++i)
    if (le32(b + 8 + 4 * i) != ch.dims[i]) return raise(ch, kErrChunkShape, 0, gid), false;
  if (lv != ch.L) return raise(ch, kErrCsLevels, 6, gid), false;
  if (!(flags & 1)) {
    // offsets come from the walk kernel
    if (pl.cs_len == 0) return raise(ch, kErrChunkTruncated, n, gid), false;
    return true;
  }
  // the container plane this decoded plane is (a plane-selective decode
  // reads only the index entries of the planes it decodes)
  const uint64_t cp = ch.plane_sel ? ch.plane_sel[p] : p;
  uint64_t entry = 12 + 4ull * (lv + 1);
  uint64_t need = fixed + (uint64_t)ch.P_all * entry;
  if (n < need) return raise(ch, kErrChunkTruncated, n, gid), false;
  const uint8_t* e = b + fixed + cp * entry;
  uint64_t off = le64(e);
  uint32_t len = le32(e + 8);
  uint64_t min_off = need;
  if (cp > 0) {
    const uint8_t* pe = e - entry;
    min_off = le64(pe) + le32(pe + 8);
  }
  if (off < min_off) return raise(ch, kErrChunkIndex, fixed + cp * entry, gid), false;
  for (uint32_t r = 0; r + 1 <= lv; ++r)
    if (le32(e + 12 + 4 * r) > le32(e + 16 + 4 * r)) return raise(ch, kErrChunkIndex, fixed + cp * entry, gid), false;
  if (le32(e + 12 + 4 * lv) > len) return raise(ch, kErrChunkIndex, fixed + cp * entry, gid), false;
  if (off + len > n) return raise(ch, kErrPlaneRange, off, gid), false;
  pl.cs_off = off;
  pl.cs_len = len;
  return true;
}

// Main header (reader.rs:120-176, markers.rs Siz/Cod/Cap/TLM parsers,
// quant.rs:159-202) and the device path's scope checks.
__device__ bool parse_main(const DevChunk& ch, const Bytes& cs, uint64_t len, uint64_t base, uint64_t gid, Hdr& h,
                           uint64_t* pos_out) {
  if (len < 2 || rd16(cs, 0) != kSOC) return raise(ch, kErrCsMarker, base, gid), false;
  uint64_t pos = 2;
  bool has_siz = false, has_cod = false, has_qcd = false;
  for (;;) {
    if (pos + 2 > len) return raise(ch, kErrCsMarker, base + pos, gid), false;
    uint32_t m = rd16(cs, pos);
    if (m == kSOT) break;
    if (m < 0xFF00 || pos + 4 > len) return raise(ch, kErrCsMarker, base + pos, gid), false;
    uint32_t L = rd16(cs, pos + 2);
    if (L < 2 || pos + 2 + L > len) return raise(ch, kErrCsMarker, base + pos, gid), false;
    const uint64_t q = pos + 4;
    uint32_t n = L - 2;
    if (m == kSIZ) {
      if (n < 36) return raise(ch, kErrCsMarker, base + pos, gid), false;
      uint32_t csiz = rd16(cs, q + 34);
      if (csiz == 0 || n < 36 + 3 * csiz) return raise(ch, kErrCsMarker, base + pos, gid), false;
      uint32_t xs = rd32(cs, q + 2), ys = rd32(cs, q + 6), xo = rd32(cs, q + 10), yo = rd32(cs, q + 14);
      uint32_t xt = rd32(cs, q + 18), yt = rd32(cs, q + 22), xto = rd32(cs, q + 26), yto = rd32(cs, q + 30);
      if (xt == 0 || yt == 0 || xs <= xo || ys <= yo || xto > xo || yto > yo)
        return raise(ch, kErrCsMarker, base + pos, gid), false;
      uint64_t tx = (uint64_t(xs - xto) + xt - 1) / xt, ty = (uint64_t(ys - yto) + yt - 1) / yt;
      if (tx != 1 || ty != 1 || xo || yo || xto || yto) return raise(ch, kErrCsUnsupported, base + pos, gid), false;
      if (csiz != 1) return raise(ch, kErrCsGeometry, base + pos, gid), false;
      if (cs[q + 37] != 1 || cs[q + 38] != 1) return raise(ch, kErrCsUnsupported, base + pos, gid), false;
      h.xsiz = xs;
      h.ysiz = ys;
      h.depth = (cs[q + 36] & 0x7F) + 1;
      h.is_signed = (cs[q + 36] & 0x80) != 0;
      if (h.depth > 38) return raise(ch, kErrCsMarker, base + pos, gid), false;
      has_siz = true;
    } else if (m == kCOD) {
      if (n < 10) return raise(ch, kErrCsMarker, base + pos, gid), false;
      h.scod = cs[q + 0];
      h.decomps = cs[q + 5];
      if (h.decomps > 32) return raise(ch, kErrCsMarker, base + pos, gid), false;
      if (cs[q + 6] > 8 || cs[q + 7] > 8 || cs[q + 6] + cs[q + 7] > 8) return raise(ch, kErrCsMarker, base + pos, gid), false;
      if (h.scod & 1) {
        if (n < 10 + h.decomps + 1) return raise(ch, kErrCsMarker, base + pos, gid), false;
        for (uint32_t i = 1; i <= h.decomps; ++i)
          if ((cs[q + 10 + i] & 0xF) == 0 || (cs[q + 10 + i] >> 4) == 0) return raise(ch, kErrCsMarker, base + pos, gid), false;
      }
      h.layers = rd16(cs, q + 2);
      h.cbw = cs[q + 6];
      h.cbh = cs[q + 7];
      h.cb_style = cs[q + 8];
      h.wavelet = cs[q + 9];
      h.mct = cs[q + 4];
      has_cod = true;
    } else if (m == kQCD) {
      if (n == 0) return raise(ch, kErrCsMarker, base + pos, gid), false;
      h.style = cs[q + 0] & 0x1F;
      h.guard = cs[q + 0] >> 5;
      if (h.style == 0) {
        h.nvals = n - 1;
        if (h.nvals > 100) return raise(ch, kErrCsUnsupported, base + pos, gid), false;
        for (uint32_t i = 0; i < h.nvals; ++i) h.vals[i] = cs[q + 1 + i] >> 3;
      } else if (h.style == 2) {
        if ((n - 1) % 2) return raise(ch, kErrCsMarker, base + pos, gid), false;
        h.nvals = (n - 1) / 2;  // irreversible: decoded only where the call allows it
        if (h.nvals > 100) return raise(ch, kErrCsUnsupported, base + pos, gid), false;
        for (uint32_t i = 0; i < h.nvals; ++i) h.vals[i] = (uint16_t)rd16(cs, q + 1 + 2 * i);
      } else {
        return raise(ch, kErrCsMarker, base + pos, gid), false;
      }
      if (h.nvals == 0) return raise(ch, kErrCsMarker, base + pos, gid), false;
      for (uint32_t i = 0; i < h.nvals; ++i)  // quant.rs parse: K_max within the 31-bit datapath
        if (kmax_of(h, i) > 31) return raise(ch, kErrCsMarker, base + pos, gid), false;
      has_qcd = true;
    } else if (m == kCAP) {
      if (n < 6 || !(rd32(cs, q) & 0x00020000u)) return raise(ch, kErrCsMarker, base + pos, gid), false;
    } else if (m == kTLM) {
      if (n < 2) return raise(ch, kErrCsMarker, base + pos, gid), false;
      uint32_t st = (cs[q + 1] >> 4) & 3, sp = (cs[q + 1] & 0x40) ? 4 : 2;
      if (st == 3 || (n - 2) % (st + sp)) return raise(ch, kErrCsMarker, base + pos, gid), false;
    } else if (m == kCOC || m == kQCC) {
      return raise(ch, kErrCsUnsupported, base + pos, gid), false;
    }
    pos += 2 + L;
  }
  if (!has_siz || !has_cod || !has_qcd) return raise(ch, kErrCsMarker, base + pos, gid), false;
  // check_supported (reader.rs:304-341) and this path's scope: one HT
  // component, reversible 5/3 (or, where the call allows it, 9/7 with an
  // expounded QCD, no component transform and at most 30-bit samples), one
  // layer, maximal precincts, 64x64 blocks.
  const bool rev = h.wavelet == 1 && h.style == 0;
  const bool irrev = ch.irrev && h.wavelet == 0 && h.style == 2 && h.mct == 0 && h.depth <= 30;
  if (!(h.cb_style & 0x40) || !(rev || irrev) || h.layers != 1) return raise(ch, kErrCsUnsupported, base, gid), false;
  if ((h.scod & 1) || h.cbw != 4 || h.cbh != 4)
    return raise(ch, kErrCsUnsupported, base, gid), false;
  if (h.decomps != ch.L) return raise(ch, kErrCsLevels, base, gid), false;
  if (h.xsiz != ch.W || h.ysiz != ch.H) return raise(ch, kErrCsGeometry, base, gid), false;
  *pos_out = pos;
  return true;
}


// Tile-part (reader.rs:178-291): SOT, PLT..., SOD; returns the body range.
__device__ bool parse_tile(const DevChunk& ch, const Bytes& cs, uint64_t len, uint64_t pos, uint64_t base,
                           uint64_t gid, bool strict, uint64_t* body_start, uint64_t* body_end, uint32_t* plt = nullptr,
                           uint32_t* nplt = nullptr) {
  uint64_t sot = pos;
  if (pos + 12 > len || rd16(cs, pos + 2) != 10) return raise(ch, kErrCsMarker, base + pos, gid), false;
  uint32_t psot = rd32(cs, pos + 6);
  pos += 12;
  for (;;) {
    if (pos + 2 > len) return raise(ch, kErrCsMarker, base + pos, gid), false;
    uint32_t m = rd16(cs, pos);
    if (m == kSOD) {
      pos += 2;
      break;
    }
    if (pos + 4 > len) return raise(ch, kErrCsMarker, base + pos, gid), false;
    uint32_t L = rd16(cs, pos + 2);
    if (L < 2 || pos + 2 + L > len) return raise(ch, kErrCsMarker, base + pos, gid), false;
    if (m == kPLT) {
      const uint64_t q = pos + 4;
      bool pending = false;
      uint32_t acc = 0;
      for (uint32_t i = 1; i < L - 2; ++i) {
        const uint32_t v = cs[q + i];
        pending = (v & 0x80) != 0;
        acc = (acc << 7) | (v & 0x7F);
        if (!pending && plt) {
          // packet lengths (markers.rs plt_payload): a cache hint for staging
          if (*nplt < kMaxPackets) plt[*nplt] = acc;
          *nplt += 1;
          acc = 0;
        }
      }
      if (L - 2 == 0 || pending) return raise(ch, kErrCsMarker, base + pos, gid), false;
    } else if (m == kCOD || m == kQCD || m == kCOC || m == kQCC) {
      return raise(ch, kErrCsUnsupported, base + pos, gid), false;
    }
    pos += 2 + L;
  }
  uint64_t end;
  if (psot == 0) {
    bool ends_eoc = len >= 2 && cs[len - 2] == 0xFF && cs[len - 1] == 0xD9;
    end = (strict || ends_eoc) ? (len >= 2 ? len - 2 : 0) : len;
  } else {
    end = sot + psot;
  }
  if (end < pos || (strict && end > len)) return raise(ch, kErrCsMarker, base + sot, gid), false;
  if (strict) {
    // one tile-part, then EOC
    if (end + 2 > len || rd16(cs, end) != kEOC) {
      if (end + 2 <= len && rd16(cs, end) == kSOT) return raise(ch, kErrCsUnsupported, base + end, gid), false;
      return raise(ch, kErrCsMarker, base + end, gid), false;
    }
  }
  *body_start = pos;
  *body_end = end < len ? end : len;
  return true;
}

__device__ __forceinline__ uint32_t find_chunk_task(const DevChunk* c, uint32_t n, uint64_t task) {
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) / 2;
    if (c[mid].task_base <= task) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Every code-block task of a decode window starts as "skip" (resolutions
// past max_res: w = h = 0) or "zero fill" (not included, until the packet
// header parse marks it), one thread per task.
__global__ __launch_bounds__(256) void init_tasks_kernel(const DevChunk* __restrict__ chunks, uint32_t nchunks,
                                                         uint64_t ntasks, HtDecTask* tasks,
                                                         const GeoPacket* __restrict__ gp,
                                                         const GeoBand* __restrict__ gb) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (i >= ntasks) return;
  const DevChunk& ch = chunks[find_chunk_task(chunks, nchunks, i)];
  const uint64_t local = i - ch.task_base;
  if (!ch.blocks_per_plane || local >= (uint64_t)ch.P * ch.blocks_per_plane) return;
  const uint64_t p = local / ch.blocks_per_plane;
  const uint32_t k = (uint32_t)(local % ch.blocks_per_plane);
  for (uint32_t r = 0; r <= ch.L; ++r) {
    const GeoPacket& pk = gp[ch.geo_packet + r];
    for (uint32_t b = 0; b < pk.nbands; ++b) {
      const GeoBand& bd = gb[ch.geo_band + pk.band_first + b];
      if (k < bd.block_first || k >= bd.block_first + bd.nx * bd.ny) continue;
      const uint32_t by = (k - bd.block_first) / bd.nx, bx = (k - bd.block_first) % bd.nx;
      HtDecTask t;
      t.seg = nullptr;
      t.out = ch.coef + p * (uint64_t)ch.W * ch.H + (uint64_t)(bd.y0 + by * 64) * ch.W + bd.x0 + bx * 64;
      t.err = ch.err;
      t.stride = ch.W;
      t.len1 = t.len2 = 0;
      const bool used = (int32_t)r <= ch.max_res;
      t.w = used ? (uint16_t)min(64u, bd.w - bx * 64) : 0;
      t.h = used ? (uint16_t)min(64u, bd.h - by * 64) : 0;
      t.mmsb = 0;
      t.passes = 0;
      t.kmax = 0;
      t.flags = 0;
      t.err_offset = 0;
      tasks[i] = t;
      return;
    }
  }
}

// Parses one plane: locate, main and tile-part headers, then every packet
// header in resolution order with its tag trees (tag: scratch of
// ch.tag_stride bytes), writing the plane's HT tasks.  win(r, off) gives the
// byte source for packet r starting at codestream offset off.
template <class Win>
__device__ void parse_plane(const DevChunk& ch, uint64_t p, uint64_t gid, DevPlane& pl, HtDecTask* pt,
                            const Bytes& cs, const Hdr& h, uint64_t bs, uint64_t be, uint8_t* tag, uint32_t* toffs,
                            const GeoPacket* __restrict__ gp, const GeoBand* __restrict__ gb, Win win) {
  uint64_t blen = be - bs;
  bool sop = (h.scod & 2) != 0, eph = (h.scod & 4) != 0;
  uint64_t cursor = 0;
  for (uint32_t r = 0; r <= ch.L && (int32_t)r <= ch.max_res; ++r) {
    const GeoPacket& pk = gp[ch.geo_packet + r];
    if (cursor > blen) return raise(ch, kErrPacket, pl.cs_off + bs + cursor, gid);
    uint64_t start = 0;
    const uint64_t d = bs + cursor;  // packet start (codestream offset)
    uint64_t dn = blen - cursor;
    const Bytes pw = win(r, d);
    if (sop && dn >= 6 && pw[d] == 0xFF && pw[d + 1] == 0x91) start = 6;
    BitReader bb;
    bb.init(pw, d + start, dn - start);
    bool empty = true, any_band = false;
    // first pass over the header: inclusion, missing MSBs, passes, lengths
    uint32_t nincl = 0;
    for (uint32_t b = 0; b < pk.nbands; ++b) {
      const GeoBand& bd = gb[ch.geo_band + pk.band_first + b];
      if (bd.nx == 0 || bd.ny == 0) continue;
      any_band = true;
      if (empty) {
        if (bb.bit() == 0) break;
        empty = false;
      }
      Tree inc, mm;
      uint32_t used = inc.init(tag, toffs, bd.nx, bd.ny, 0);
      mm.init(tag + used, toffs + kTreeOffs, bd.nx, bd.ny, 0);
      uint32_t nl = inc.levels, kmax = pl.kmax[bd.kidx];
      for (uint32_t y = 0; y < bd.ny; ++y)
        for (uint32_t x = 0; x < bd.nx; ++x) {
          bool empty_cb = false;
          for (uint32_t cl = nl; cl >= 1; --cl) {
            uint32_t cur = cl - 1;
            uint32_t ii = inc.idx(x >> cur, y >> cur, cur);
            empty_cb = inc.val[ii] == 1;
            if (empty_cb) break;
            if (inc.flag[ii] == 0) {
              uint32_t bit = bb.bit();
              empty_cb = bit == 0;
              inc.val[ii] = (uint8_t)(1 - bit);
              inc.flag[ii] = 1;
            }
            if (empty_cb) break;
          }
          if (empty_cb) continue;
          uint32_t mmsbs = 0;
          for (uint32_t lp1 = nl; lp1 >= 1; --lp1) {
            uint32_t cur = lp1 - 1;
            mmsbs = mm.val[mm.idx(x >> lp1, y >> lp1, lp1)];
            uint32_t ci = mm.idx(x >> cur, y >> cur, cur);
            if (mm.flag[ci] == 0) {
              mmsbs += bb.zeros_to_one();
              mm.val[ci] = (uint8_t)min(mmsbs, 255u);
              mm.flag[ci] = 1;
            }
          }
          uint32_t np = 1;
          if (bb.bit() == 1) {
            np = 2;
            if (bb.bit() == 1) {
              uint32_t v = bb.bits(2);
              np = 3 + v;
              if (v == 3) {
                v = bb.bits(5);
                np = 6 + v;
                if (v == 31) np = 37 + bb.bits(7);
              }
            }
          }
          uint32_t phld = (np - 1) / 3;
          mmsbs += phld;
          np -= phld * 3;
          if (mmsbs > kmax) return raise(ch, kErrPacket, pl.cs_off + bs + cursor, gid);
          const uint32_t lblock = 3 + bb.ones_to_zero();
          if (lblock > 24) return raise(ch, kErrPacket, pl.cs_off + bs + cursor, gid);
          uint32_t nb = lblock + 31 - __clz(phld * 3 + 1);
          uint32_t len1 = bb.bits(nb);
          if (len1 < 2 || len1 >= 65535) return raise(ch, kErrPacket, pl.cs_off + bs + cursor, gid);
          uint32_t len2 = 0;
          if (np > 1) {
            len2 = bb.bits(lblock + (np > 2 ? 1 : 0));
            if (len2 >= 2047) return raise(ch, kErrPacket, pl.cs_off + bs + cursor, gid);
          }
          if (bb.bad) return raise(ch, kErrPacket, pl.cs_off + bs + cursor, gid);
          HtDecTask& t = pt[bd.block_first + y * bd.nx + x];
          t.len1 = len1;
          t.len2 = len2;
          t.mmsb = (uint8_t)min(mmsbs, 255u);
          t.passes = (uint8_t)min(np, 255u);  // more than 3 is refused by the block decoder
          t.kmax = (uint8_t)(pl.irrev ? 31 : kmax);  // 9/7: the aligned magnitude itself, dequantised later
          t.flags = kTaskHeader | ((h.cb_style & 0x08) ? kTaskCausal : 0);
          nincl += 1;
        }
    }
    // a precinct without code-blocks still carries the one-bit empty header
    if (!any_band) (void)bb.bit();
    if (bb.bad) return raise(ch, kErrPacket, pl.cs_off + bs + cursor, gid);
    uint64_t hl = start + bb.terminate();
    if (eph) {
      if (dn >= hl + 2 && pw[d + hl] == 0xFF && pw[d + hl + 1] == 0x92) hl += 2;
      else return raise(ch, kErrPacket, pl.cs_off + bs + cursor + hl, gid);
    }
    // bodies follow in band / raster order
    uint64_t q = cursor + hl;
    for (uint32_t b = 0; b < pk.nbands && nincl; ++b) {
      const GeoBand& bd = gb[ch.geo_band + pk.band_first + b];
      for (uint32_t i = 0; i < bd.nx * bd.ny; ++i) {
        HtDecTask& t = pt[bd.block_first + i];
        if (!(t.flags & kTaskHeader)) continue;
        uint64_t sl = (uint64_t)t.len1 + t.len2;
        if (q + sl > blen) return raise(ch, kErrPacket, pl.cs_off + bs + q, gid);
        t.seg = cs.g + bs + q;
        t.err_offset = (uint32_t)(pl.cs_off + bs + q);
        t.flags = (uint8_t)((t.flags & ~kTaskHeader) | kTaskIncluded);
        q += sl;
      }
    }
    cursor = q;
  }
  pl.ok = 1;
}

__device__ bool plane_headers(const DevChunk& ch, uint64_t p, uint64_t gid, DevPlane& pl, const Bytes& cs,
                              Hdr& h, uint64_t* bs, uint64_t* be, uint32_t* plt, uint32_t* nplt) {
  const uint64_t len = pl.cs_len;
  uint64_t pos;
  if (!parse_main(ch, cs, len, pl.cs_off, gid, h, &pos)) return false;
  pl.depth = h.depth;
  pl.is_signed = h.is_signed;
  for (uint32_t i = 0; i < 1 + 3 * ch.L && i < 100; ++i) pl.kmax[i] = (uint8_t)kmax_of(h, i);
  pl.irrev = h.wavelet == 0;
  if (ch.irrev) {
    IrrevPlane& ip = ch.irrev[p];
    ip.on = pl.irrev;
    if (pl.irrev)
      for (uint32_t i = 0; i < 1 + 3 * ch.L && i < 100; ++i) ip.scale[i] = irrev_scale(h, i, ch.L, ch.max_res);
  }
  return parse_tile(ch, cs, len, pos, pl.cs_off, gid, !(ch.no_container & 2), bs, be, plt, nplt);
}

// Thread per plane, everything read from global memory (tag trees in the
// plane's global scratch).
__global__ __launch_bounds__(64) void parse_planes_kernel(const DevChunk* __restrict__ chunks, uint32_t nchunks,
                                                          uint64_t nplanes, DevPlane* planes, HtDecTask* tasks,
                                                          const GeoPacket* __restrict__ gp,
                                                          const GeoBand* __restrict__ gb, uint32_t tag_lds) {
  // per lane in LDS: the tag trees' level offsets and, when the window's
  // trees fit tag_lds bytes, their nodes (the serial header walk reads and
  // writes them at every code-block)
  extern __shared__ __align__(16) uint8_t lane_lds[];
  uint8_t* mine = lane_lds + threadIdx.x * (tag_lds + 8 * kTreeOffs);
  uint32_t* toffs = (uint32_t*)(mine + tag_lds);
  uint64_t gid = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (gid >= nplanes) return;
  const DevChunk& ch = chunks[find_chunk(chunks, nchunks, gid)];
  uint64_t p = gid - ch.plane_base;
  DevPlane& pl = planes[gid];
  pl.ok = 0;
  // the plane's tasks were set to "skip" / "zero fill" by init_tasks_kernel
  HtDecTask* pt = tasks + ch.task_base + p * ch.blocks_per_plane;
  if (!locate(ch, p, gid, pl)) return;
  const Bytes cs{ch.in + pl.cs_off, nullptr, 0, 0};
  Hdr h;
  uint64_t bs, be;
  if (!plane_headers(ch, p, gid, pl, cs, h, &bs, &be, nullptr, nullptr)) return;
  parse_plane(ch, p, gid, pl, pt, cs, h, bs, be, tag_lds ? mine : ch.tag + gid * ch.tag_stride, toffs, gp, gb,
              [&](uint32_t, uint64_t) { return cs; });
}

// The wave copies n bytes g[from, from + n) to LDS dst with aligned dword
// loads (a dword holding a byte of the codestream lies in its allocation).
__device__ void stage_bytes(uint8_t* dst, const uint8_t* g, uint64_t from, uint32_t n) {
  const uint32_t lane = threadIdx.x & 63;
  const uintptr_t a0 = (uintptr_t)(g + from) & ~(uintptr_t)3;
  const uint32_t skew = (uint32_t)((uintptr_t)(g + from) - a0);
  const uint32_t words = (skew + n + 3) / 4;
  for (uint32_t w0 = 0; w0 < words; w0 += 256) {
    uint32_t v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint32_t w = w0 + lane + 64 * k;
      v[k] = w < words ? *(const __attribute__((address_space(1))) uint32_t*)(a0 + 4 * (uintptr_t)w) : 0u;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint32_t w = w0 + lane + 64 * k;
      for (uint32_t b = 0; b < 4; ++b) {
        const int64_t i = (int64_t)(4 * w + b) - skew;
        if (w < words && i >= 0 && i < (int64_t)n) dst[i] = (uint8_t)(v[k] >> (8 * b));
      }
    }
  }
}

// Wavefront per plane: the wave copies the codestream's first kStageBytes
// (or its headers up to SOD and the start of every packet that PLT locates)
// into LDS, then lane 0 parses from there with the tag trees in LDS.
constexpr uint32_t kStageBytes = 8192, kTagLds = 4096;
__global__ __launch_bounds__(64) void parse_planes_wave_kernel(const DevChunk* __restrict__ chunks, uint32_t nchunks,
                                                               uint64_t nplanes, DevPlane* planes, HtDecTask* tasks,
                                                               const GeoPacket* __restrict__ gp,
                                                               const GeoBand* __restrict__ gb) {
  __shared__ __align__(16) uint8_t stage[kStageBytes];
  __shared__ __align__(16) uint8_t tagl[kTagLds];
  __shared__ uint32_t toffl[2 * kTreeOffs];
  __shared__ uint32_t plt[kMaxPackets];
  __shared__ uint64_t wbase[kMaxPackets];
  __shared__ uint32_t woff[kMaxPackets], wlen[kMaxPackets];
  __shared__ uint32_t s_ok, s_nplt;
  __shared__ uint64_t s_bs, s_be, s_len, s_off;
  const uint64_t gid = blockIdx.x;
  const uint32_t lane = threadIdx.x;
  if (gid >= nplanes) return;
  const DevChunk& ch = chunks[find_chunk(chunks, nchunks, gid)];
  const uint64_t p = gid - ch.plane_base;
  DevPlane& pl = planes[gid];
  HtDecTask* pt = tasks + ch.task_base + p * ch.blocks_per_plane;
  Hdr h;
  if (lane == 0) {
    pl.ok = 0;
    s_ok = locate(ch, p, gid, pl) ? 1u : 0u;
    s_len = pl.cs_len;
    s_off = pl.cs_off;
  }
  __syncthreads();
  if (!s_ok) return;
  const uint8_t* g = ch.in + s_off;
  const uint64_t len = s_len;
  // the first bytes: every header up to SOD, and all of a small codestream
  const uint32_t n0 = (uint32_t)min<uint64_t>(len, kStageBytes);
  stage_bytes(stage, g, 0, n0);
  __syncthreads();
  if (lane == 0) {
    const Bytes cs{g, stage, 0, n0};
    s_nplt = 0;
    s_ok = plane_headers(ch, p, gid, pl, cs, h, &s_bs, &s_be, plt, &s_nplt) ? 1u : 0u;
  }
  __syncthreads();
  if (!s_ok) return;
  const uint64_t bs = s_bs;
  // packet windows from the PLT lengths, when the first bytes do not cover
  // the packets: the staging after the headers is shared out among them
  const uint32_t npk = min<uint32_t>((uint32_t)ch.L + 1, (uint32_t)max(ch.max_res, 0) + 1);
  const bool use_plt = len > kStageBytes && s_nplt == ch.L + 1 && bs < kStageBytes;
  if (lane == 0) {
    uint64_t at = bs;
    const uint32_t room = use_plt ? (uint32_t)(kStageBytes - ((bs + 15) & ~15ull)) : 0u;
    const uint32_t each = npk ? (room / npk) & ~15u : 0u;
    for (uint32_t r = 0; r < npk && r < kMaxPackets; ++r) {
      wbase[r] = at;
      woff[r] = (uint32_t)((bs + 15) & ~15ull) + r * each;
      wlen[r] = use_plt ? (uint32_t)min<uint64_t>(each, plt[r]) : 0u;
      if (use_plt) at += plt[r];
    }
  }
  __syncthreads();
  if (use_plt)
    for (uint32_t r = 0; r < npk; ++r) stage_bytes(stage + woff[r], g, wbase[r], (uint32_t)min<uint64_t>(wlen[r], len - min(len, wbase[r])));
  __syncthreads();
  if (lane != 0) return;
  // with packet windows the staging after the headers holds them, so the
  // first window shrinks to the headers
  const Bytes cs{g, stage, 0, use_plt ? (uint32_t)((bs + 15) & ~15ull) : n0};
  uint8_t* tag = ch.tag_stride <= kTagLds ? tagl : ch.tag + gid * ch.tag_stride;
  parse_plane(ch, p, gid, pl, pt, cs, h, bs, s_be, tag, toffl, gp, gb, [&](uint32_t r, uint64_t off) {
    if (use_plt && r < npk && off == wbase[r]) return Bytes{g, stage + woff[r], wbase[r], wlen[r]};
    return cs;
  });
}

// Walks concatenated codestreams of index-less chunks (htj2k.rs:367-381).
__global__ void walk_chunks_kernel(const DevChunk* __restrict__ chunks, uint32_t nchunks, DevPlane* planes) {
  uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nchunks) return;
  const DevChunk& ch = chunks[c];
  if (ch.no_container || ch.in_len < 8 || (ch.in[5] & 1)) return;
  uint32_t ndim = ch.in[7];
  uint64_t off = 8 + 4ull * ndim;
  // every container plane is walked (lengths are only known in order); the
  // decoded planes (all, or the selected ones in order) take their offsets
  uint64_t j = 0;
  for (uint64_t cp = 0; cp < ch.P_all && j < ch.P; ++cp) {
    DevPlane* pl = nullptr;
    if (!ch.plane_sel || ch.plane_sel[j] == cp) {
      pl = &planes[ch.plane_base + j];
      ++j;
      pl->cs_off = off;
      pl->cs_len = 0;
    }
    const uint8_t* cs = ch.in + off;
    if (off > ch.in_len) return;
    uint64_t len = ch.in_len - off, pos = 2;
    if (len < 2 || rd16(cs) != kSOC) return;
    for (;;) {  // main header
      if (pos + 4 > len) return;
      uint32_t m = rd16(cs + pos);
      if (m == kSOT) break;
      uint32_t L = rd16(cs + pos + 2);
      if (m < 0xFF00 || L < 2 || pos + 2 + L > len) return;
      pos += 2 + L;
    }
    for (;;) {  // tile-parts until EOC
      if (pos + 2 > len) return;
      uint32_t m = rd16(cs + pos);
      if (m == kEOC) {
        pos += 2;
        break;
      }
      if (m != kSOT || pos + 12 > len) return;
      uint32_t psot = rd32(cs + pos + 6);
      if (psot == 0) {
        pos = len;
        break;
      }
      pos += psot;
    }
    if (pl) pl->cs_len = pos;
    off += pos;
  }
}

// DC un-shift with clamping for unsigned planes (reader.rs:626-638).
__global__ __launch_bounds__(256) void unshift_kernel(const DevChunk* __restrict__ chunks, uint32_t nchunks,
                                                      const DevPlane* __restrict__ planes, int32_t* const* data,
                                                      const uint64_t* dims) {
  uint64_t gid = blockIdx.x;
  const DevPlane& pl = planes[gid];
  if (pl.is_signed || !pl.ok) return;
  uint32_t c = find_chunk(chunks, nchunks, gid);
  if (!data[c]) return;  // un-shifted by a later fused pass
  const DevChunk& ch = chunks[c];
  uint64_t ow = dims[2 * c], oh = dims[2 * c + 1];
  int32_t* d = data[c] + (gid - ch.plane_base) * (uint64_t)ch.W * ch.H;
  int32_t half = 1 << (pl.depth - 1);
  int64_t mx = ((int64_t)1 << pl.depth) - 1;
  for (uint64_t i = threadIdx.x; i < ow * oh; i += blockDim.x) {
    uint64_t y = i / ow, x = i % ow;
    int64_t v = (int64_t)d[y * ch.W + x] + half;
    d[y * ch.W + x] = (int32_t)(v < 0 ? 0 : v > mx ? mx : v);
  }
}

// Irreversible planes: aligned sign-magnitude words (the HT decoder's output
// with K_max taken as 31) to f32 coefficients, scaled by the step of the
// coefficient's band.  Grid: (plane, 64 strides); the band of (x, y) is found
// from the level sizes: d = min(levels column x and row y lie inside), LL at
// d = L, else HL / LH / HH of level d + 1 (resolution L - d).
__global__ __launch_bounds__(256) void dequantize_kernel(const DevChunk* __restrict__ chunks, uint32_t nchunks) {
  const uint64_t gid = blockIdx.x;
  const DevChunk& ch = chunks[find_chunk(chunks, nchunks, gid)];
  if (!ch.irrev) return;
  const uint64_t p = gid - ch.plane_base;
  const IrrevPlane& ip = ch.irrev[p];
  if (!ip.on) return;
  __shared__ uint32_t lw[34], lh[34];
  __shared__ float sc[100];
  const uint32_t L = ch.L;
  if (threadIdx.x == 0) {
    uint32_t w = ch.W, h = ch.H;
    for (uint32_t l = 0; l <= L; ++l) {
      lw[l] = w;
      lh[l] = h;
      w = (w + 1) / 2;
      h = (h + 1) / 2;
    }
  }
  for (uint32_t i = threadIdx.x; i < 1 + 3 * L && i < 100; i += blockDim.x) sc[i] = ip.scale[i];
  __syncthreads();
  int32_t* plane = ch.coef + p * (uint64_t)ch.W * ch.H;
  const uint64_t n = (uint64_t)ch.W * ch.H;
  for (uint64_t i = (uint64_t)blockIdx.y * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.y * blockDim.x) {
    const int32_t v = plane[i];
    if (v == 0) continue;  // 0 and +0.0f share their bits
    const uint32_t x = (uint32_t)(i % ch.W), y = (uint32_t)(i / ch.W);
    uint32_t dx = 0, dy = 0;
    while (dx < L && x < lw[dx + 1]) ++dx;
    while (dy < L && y < lh[dy + 1]) ++dy;
    const uint32_t d = dx < dy ? dx : dy;
    uint32_t k = 0;
    if (d < L) {
      const uint32_t band = (dx == d && dy == d) ? 3 : (dx == d ? 1 : 2);
      k = 3 * (L - d - 1) + band;
    }
    const float c = __fmul_rn((float)(uint32_t)(v < 0 ? -(int64_t)v : v), sc[k]);
    plane[i] = __float_as_int(v < 0 ? -c : c);
  }
}

// Irreversible planes: f32 region -> nearest int32, clamped to the
// level-shifted range of the sample depth.
__global__ __launch_bounds__(256) void irrev_round_kernel(const DevChunk* __restrict__ chunks, uint32_t nchunks,
                                                          const DevPlane* __restrict__ planes,
                                                          int32_t* const* data, const uint64_t* dims) {
  const uint64_t gid = blockIdx.x;
  const DevPlane& pl = planes[gid];
  if (!pl.ok || !pl.irrev) return;
  const uint32_t c = find_chunk(chunks, nchunks, gid);
  if (!data[c]) return;
  const DevChunk& ch = chunks[c];
  const uint64_t ow = dims[2 * c], oh = dims[2 * c + 1];
  int32_t* d = data[c] + (gid - ch.plane_base) * (uint64_t)ch.W * ch.H;
  const float lo = -(float)(1 << (pl.depth - 1)), hi = (float)((1 << (pl.depth - 1)) - 1);
  for (uint64_t i = (uint64_t)blockIdx.y * blockDim.x + threadIdx.x; i < ow * oh;
       i += (uint64_t)gridDim.y * blockDim.x) {
    const uint64_t y = i / ow, x = i % ow;
    const float f = __int_as_float(d[y * ch.W + x]);
    d[y * ch.W + x] = (int32_t)fminf(fmaxf(rintf(f), lo), hi);
  }
}

// Window copy of decoded planes with narrowing (htj2k_codec.rs:393-407: the
// (y, x) window of each selected plane, narrowed like htj2k.rs
// narrow_samples); one workgroup per output row.
__device__ __forceinline__ uint32_t find_window_job(const WindowJob* jobs, uint32_t njobs, uint64_t item) {
  uint32_t lo = 0, hi = njobs;
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) / 2;
    if (jobs[mid].item_begin <= item) lo = mid;
    else hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(256) void window_kernel(const WindowJob* __restrict__ jobs, uint32_t njobs) {
  const uint64_t item = blockIdx.x + (uint64_t)blockIdx.y * gridDim.x;
  const WindowJob& j = jobs[find_window_job(jobs, njobs, item)];
  const uint64_t local = item - j.item_begin;
  if (local >= j.items) return;
  const uint64_t plane = local / j.h, row = local % j.h;
  const int32_t* src = j.src + plane * j.plane_elems + (uint64_t)(j.y0 + row) * j.W + j.x0;
  const uint64_t o = (plane * j.h + row) * j.w;
  bool ok = true;
  int32_t lo = INT32_MAX, hi = INT32_MIN;
  for (uint32_t x = threadIdx.x; x < j.w; x += blockDim.x) {
    const int32_t v = src[x];
    if (!j.dst) {
      lo = v < lo ? v : lo;
      hi = v > hi ? v : hi;
      continue;
    }
    uint8_t* d = j.dst + (o + x) * j.dsize;
    switch (j.dtype) {
      case NDIC_U8: *d = (uint8_t)v; ok &= v >= 0 && v <= 255; break;
      case NDIC_I8: *d = (uint8_t)v; ok &= v >= -128 && v <= 127; break;
      case NDIC_U16: *(uint16_t*)d = (uint16_t)v; ok &= v >= 0 && v <= 65535; break;
      case NDIC_I16: *(int16_t*)d = (int16_t)v; ok &= v >= -32768 && v <= 32767; break;
      case NDIC_U32: *(uint32_t*)d = (uint32_t)v; ok &= v >= 0; break;
      default: *(int32_t*)d = v; break;
    }
  }
  if (!ok) atomicOr(j.flag, 1);
  if (!j.dst) {
    for (int off = 32; off > 0; off >>= 1) {
      const int32_t a = __shfl_xor(lo, off), b = __shfl_xor(hi, off);
      lo = a < lo ? a : lo;
      hi = b > hi ? b : hi;
    }
    if ((threadIdx.x & 63) == 0 && lo <= hi) {
      atomicMin((long long*)&j.minmax[0], (long long)lo);
      atomicMax((long long*)&j.minmax[1], (long long)hi);
    }
  }
}

// ---------------------------------------------------------------------------
// encode
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void plane_minmax_kernel(const DevChunk* __restrict__ chunks, uint32_t nchunks,
                                                           DevPlane* planes) {
  uint64_t gid = blockIdx.x;
  const DevChunk& ch = chunks[find_chunk(chunks, nchunks, gid)];
  uint64_t n = (uint64_t)ch.W * ch.H;
  const int32_t* d = ch.planes_in + (gid - ch.plane_base) * n;
  int32_t lo = 0, hi = 0;  // depth_for folds from (0, 0): htj2k.rs:262-264
  for (uint64_t i = threadIdx.x; i < n; i += 256) {
    int32_t v = d[i];
    lo = v < lo ? v : lo;
    hi = v > hi ? v : hi;
  }
  __shared__ int32_t slo[256], shi[256];
  slo[threadIdx.x] = lo;
  shi[threadIdx.x] = hi;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      slo[threadIdx.x] = min(slo[threadIdx.x], slo[threadIdx.x + s]);
      shi[threadIdx.x] = max(shi[threadIdx.x], shi[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    planes[gid].lo = slo[0];
    planes[gid].hi = shi[0];
  }
}

// Depth (htj2k.rs:261-276), reversible ranging (quant.rs:51-85) and the
// K_max <= 30 datapath check (writer.rs:132-143), per plane.
__constant__ uint8_t c_xll[34] = {0, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2,
                                  2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2};
__constant__ uint8_t c_xlh[34] = {0, 2, 3, 3, 3, 3, 3, 3, 3, 3, 3, 3, 3, 3, 3, 3, 3,
                                  3, 3, 3, 3, 3, 3, 3, 3, 3, 3, 3, 3, 3, 3, 3, 3, 3};
__constant__ uint8_t c_xhh[34] = {0, 2, 3, 3, 3, 3, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4,
                                  4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4};

__device__ uint32_t raw_exp(uint32_t L, uint32_t b, uint32_t i) {
  if (i == 0) return b + (L == 0 ? 1 : c_xll[L]);
  uint32_t d = L - (i - 1) / 3, k = (i - 1) % 3;
  return b + (k == 2 ? c_xhh[d] : c_xlh[d]);
}

__global__ void plane_setup_kernel(const DevChunk* __restrict__ chunks, uint32_t nchunks, uint64_t nplanes,
                                   DevPlane* planes) {
  uint64_t gid = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (gid >= nplanes) return;
  const DevChunk& ch = chunks[find_chunk(chunks, nchunks, gid)];
  DevPlane& pl = planes[gid];
  bool sg = ch.sample_signed != 0;
  int64_t lo = pl.lo, hi = pl.hi;
  uint32_t depth = 32;
  for (uint32_t b = 1; b <= 32; ++b) {
    bool fits = sg ? (lo >= -((int64_t)1 << (b - 1)) && hi < ((int64_t)1 << (b - 1))) : hi < ((int64_t)1 << b);
    if (fits) {
      depth = b;
      break;
    }
  }
  pl.depth = (int32_t)depth;
  pl.is_signed = sg;
  uint32_t n = 1 + 3 * ch.L, max_raw = 0;
  for (uint32_t i = 0; i < n; ++i) max_raw = max(max_raw, raw_exp(ch.L, depth, i));
  uint32_t guard = max(max_raw > 31 ? max_raw - 31 : 0u, 1u);
  uint32_t kmm = 0;
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t eps = raw_exp(ch.L, depth, i) - guard;
    uint32_t km = (eps == 0 ? 0 : eps - 1) + guard;
    pl.kmax[i] = (uint8_t)km;
    kmm = max(kmm, km);
  }
  pl.ok = 1;
  if (max_raw > 38) {
    pl.ok = 0;
    raise(ch, kErrEncodeKmax, 0, gid);
  } else if (kmm > 30) {
    pl.ok = 0;
    raise(ch, kErrEncodeKmax, 0, gid);
  }
}

// Unsigned planes are DC level shifted before the DWT (writer.rs:337-346).
__global__ __launch_bounds__(256) void dc_shift_kernel(const DevChunk* __restrict__ chunks, uint32_t nchunks,
                                                       const DevPlane* __restrict__ planes, int32_t* data_base_unused) {
  uint64_t gid = blockIdx.x;
  const DevPlane& pl = planes[gid];
  if (pl.is_signed) return;
  const DevChunk& ch = chunks[find_chunk(chunks, nchunks, gid)];
  uint64_t n = (uint64_t)ch.W * ch.H;
  int32_t* d = ch.planes_in + (gid - ch.plane_base) * n;
  int32_t half = 1 << (pl.depth - 1);
  for (uint64_t i = threadIdx.x; i < n; i += 256) d[i] -= half;
}

__global__ void build_enc_tasks_kernel(const DevChunk* __restrict__ chunks, uint32_t nchunks, uint64_t nplanes,
                                       const DevPlane* __restrict__ planes, const GeoPacket* __restrict__ gp,
                                       const GeoBand* __restrict__ gb, HtEncTask* tasks, uint32_t* lens,
                                       uint64_t slot_stride) {
  uint64_t gid = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (gid >= nplanes) return;
  const DevChunk& ch = chunks[find_chunk(chunks, nchunks, gid)];
  const DevPlane& pl = planes[gid];
  uint64_t p = gid - ch.plane_base;
  uint64_t tb = ch.task_base + p * ch.blocks_per_plane;
  const int32_t* plane = ch.coef + p * (uint64_t)ch.W * ch.H;
  for (uint32_t r = 0; r <= ch.L; ++r) {
    const GeoPacket& pk = gp[ch.geo_packet + r];
    for (uint32_t b = 0; b < pk.nbands; ++b) {
      const GeoBand& bd = gb[ch.geo_band + pk.band_first + b];
      for (uint32_t by = 0; by < bd.ny; ++by)
        for (uint32_t bx = 0; bx < bd.nx; ++bx) {
          uint64_t ti = tb + bd.block_first + by * bd.nx + bx;
          HtEncTask& t = tasks[ti];
          t.in = plane + (uint64_t)(bd.y0 + by * 64) * ch.W + bd.x0 + bx * 64;
          t.slot = ch.slots + ti * slot_stride;
          t.lens = lens + 4 * ti;
          t.err = ch.err;
          t.stride = ch.W;
          t.w = (uint16_t)min(64u, bd.w - bx * 64);
          t.h = (uint16_t)min(64u, bd.h - by * 64);
          t.kmax = pl.ok ? pl.kmax[bd.kidx] : 0;
          t.ms_cap = ht_ms_cap(t.w, t.h, t.kmax);
          t.flags = 0;
          if (!pl.ok) t.w = 0;  // a refused plane codes nothing
        }
    }
  }
}

__device__ __forceinline__ void put16(uint8_t* o, uint32_t& n, uint32_t v) {
  o[n++] = uint8_t(v >> 8);
  o[n++] = uint8_t(v);
}
__device__ __forceinline__ void put32(uint8_t* o, uint32_t& n, uint32_t v) {
  put16(o, n, v >> 16);
  put16(o, n, v & 0xFFFF);
}

// Packet headers (packet.rs:69-168) and the plane's main + tile-part header
// (writer.rs:198-322).  Sizes are bounded by the planner's strides.
constexpr uint32_t kWriteTagLds = 32768;  // LDS tag-tree scratch of a 64-thread workgroup
__global__ __launch_bounds__(64) void write_planes_kernel(const DevChunk* __restrict__ chunks, uint32_t nchunks,
                                                          uint64_t nplanes, DevPlane* planes,
                                                          const HtEncTask* __restrict__ tasks, const uint32_t* lens,
                                                          const GeoPacket* __restrict__ gp,
                                                          const GeoBand* __restrict__ gb) {
  __shared__ __align__(16) uint8_t ltag[kWriteTagLds];
  uint64_t gid = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (gid >= nplanes) return;
  const DevChunk& ch = chunks[find_chunk(chunks, nchunks, gid)];
  DevPlane& pl = planes[gid];
  if (!pl.ok) return;
  uint64_t p = gid - ch.plane_base;
  uint64_t tb = ch.task_base + p * ch.blocks_per_plane;
  // the tag trees in LDS when every thread's share fits, else in the plane's global scratch
  uint8_t* tag = 64ull * ch.tag_stride <= kWriteTagLds ? ltag + threadIdx.x * ch.tag_stride
                                                       : ch.tag + gid * ch.tag_stride;
  uint64_t body = 0;
  for (uint32_t r = 0; r <= ch.L; ++r) {
    const GeoPacket& pk = gp[ch.geo_packet + r];
    uint8_t* hdr = ch.pkt_hdr + (gid * (uint64_t)ch.pkt_slots + r) * (uint64_t)ch.pkt_stride;
    BitWriter bw;
    bw.init(hdr, ch.pkt_stride);
    bool any = false;
    uint64_t bodies = 0;
    for (uint32_t i = 0; i < pk.nblocks; ++i) {
      const uint32_t* l = lens + 4 * (tb + pk.block_first + i);
      if (l[3]) {
        any = true;
        bodies += (uint64_t)l[0] + l[1] + l[2];
      }
    }
    if (!any) {
      hdr[0] = 0;
      pl.packet_hdr[r] = 1;
      pl.packet_len[r] = 1;
      body += 1;
      continue;
    }
    bw.bit(1);
    for (uint32_t b = 0; b < pk.nbands; ++b) {
      const GeoBand& bd = gb[ch.geo_band + pk.band_first + b];
      if (bd.nx == 0 || bd.ny == 0) continue;
      Tree inc, mm;
      uint32_t toffs[2 * kTreeOffs];
      uint32_t used = inc.init(tag, toffs, bd.nx, bd.ny, 255);
      mm.init(tag + used, toffs + kTreeOffs, bd.nx, bd.ny, 255);
      uint32_t kmm = pl.kmax[bd.kidx] - 1;  // missing MSBs of every coded block
      for (uint32_t y = 0; y < bd.ny; ++y)
        for (uint32_t x = 0; x < bd.nx; ++x) {
          bool present = lens[4 * (tb + bd.block_first + y * bd.nx + x) + 3] != 0;
          inc.val[inc.idx(x, y, 0)] = present ? 0 : 1;
          mm.val[mm.idx(x, y, 0)] = present ? (uint8_t)kmm : 0;
        }
      inc.reduce();
      mm.reduce();
      uint32_t nl = inc.levels;
      if (inc.val[inc.idx(0, 0, nl - 1)] != 0) {
        bw.bit(0);
        continue;
      }
      for (uint32_t y = 0; y < bd.ny; ++y)
        for (uint32_t x = 0; x < bd.nx; ++x) {
          for (uint32_t cur = nl; cur >= 1; --cur) {
            uint32_t lm = cur - 1;
            uint32_t ii = inc.idx(x >> lm, y >> lm, lm);
            if (inc.flag[ii] == 0) {
              uint32_t skipped = inc.val[ii] - inc.val[inc.idx(x >> cur, y >> cur, cur)];
              bw.bits(1 - min(skipped, 1u), 1);
              inc.flag[ii] = 1;
            }
            if (inc.val[ii] > 0) break;
          }
          const uint32_t* l = lens + 4 * (tb + bd.block_first + y * bd.nx + x);
          if (!l[3]) continue;
          for (uint32_t cur = nl; cur >= 1; --cur) {
            uint32_t lm = cur - 1;
            uint32_t ii = mm.idx(x >> lm, y >> lm, lm);
            if (mm.flag[ii] == 0) {
              uint32_t zeros = mm.val[ii] - mm.val[mm.idx(x >> cur, y >> cur, cur)];
              for (uint32_t z = 0; z < zeros; ++z) bw.bit(0);
              bw.bit(1);
              mm.flag[ii] = 1;
            }
          }
          bw.bits(0, 1);  // one pass
          uint32_t len1 = l[0] + l[1] + l[2];
          uint32_t bits1 = 32 - __clz(len1);
          uint32_t nb = bits1 > 3 ? bits1 - 3 : 0;
          bw.bits(0xFFFFFFFEu, nb + 1);
          bw.bits(len1, nb + 3);
        }
    }
    bw.terminate();
    pl.packet_hdr[r] = bw.len;
    pl.packet_len[r] = (uint32_t)(bw.len + bodies);
    body += bw.len + bodies;
    if (bw.len > ch.pkt_stride) raise(ch, kErrEncodeScup, 0, gid);
  }
  // ---- headers
  uint8_t* o = ch.head + gid * (uint64_t)ch.head_stride;
  uint32_t n = 0;
  // tile-part header first (its length goes into TLM): SOT, PLT..., SOD
  uint8_t plt[1024];
  uint32_t pn = 0;
  if (ch.index) {
    uint32_t zplt = 0, start = 0, npk = ch.L + 1;
    for (;;) {
      uint32_t end = start, payload = 1;
      while (end < npk) {
        uint32_t bits = 32 - __clz(pl.packet_len[end] | 1);
        uint32_t vl = (bits + 6) / 7;
        if (payload + vl > 65000) break;
        payload += vl;
        end += 1;
      }
      put16(plt, pn, kPLT);
      put16(plt, pn, payload + 2);
      plt[pn++] = (uint8_t)zplt;
      for (uint32_t k = start; k < end; ++k) {
        uint32_t v = pl.packet_len[k];
        bool started = false;
        for (int sh = 28; sh >= 7; sh -= 7) {
          uint32_t b = (v >> sh) & 0x7F;
          if (b != 0 || started) {
            plt[pn++] = uint8_t(0x80 | b);
            started = true;
          }
        }
        plt[pn++] = uint8_t(v & 0x7F);
      }
      start = end;
      if (start >= npk) break;
      zplt += 1;
    }
  }
  uint64_t psot = 12 + pn + 2 + body;
  uint32_t tile_part = (uint32_t)psot;
  put16(o, n, kSOC);
  // SIZ (markers.rs:105-125)
  put16(o, n, kSIZ);
  put16(o, n, 41);
  put16(o, n, 0x4000);
  put32(o, n, ch.W);
  put32(o, n, ch.H);
  put32(o, n, 0);
  put32(o, n, 0);
  put32(o, n, ch.W);
  put32(o, n, ch.H);
  put32(o, n, 0);
  put32(o, n, 0);
  put16(o, n, 1);
  o[n++] = uint8_t((pl.depth - 1) | (pl.is_signed ? 0x80 : 0));
  o[n++] = 1;
  o[n++] = 1;
  // COD (markers.rs:233-251)
  put16(o, n, kCOD);
  put16(o, n, 12);
  o[n++] = 0;
  o[n++] = (uint8_t)ch.progression;
  put16(o, n, 1);
  o[n++] = 0;
  o[n++] = (uint8_t)ch.L;
  o[n++] = 4;
  o[n++] = 4;
  o[n++] = 0x40;
  o[n++] = 1;
  // QCD (quant.rs:133-145)
  uint32_t nv = 1 + 3 * ch.L, guard = 1;
  {
    uint32_t max_raw = 0;
    for (uint32_t i = 0; i < nv; ++i) max_raw = max(max_raw, raw_exp(ch.L, pl.depth, i));
    guard = max(max_raw > 31 ? max_raw - 31 : 0u, 1u);
  }
  put16(o, n, kQCD);
  put16(o, n, 3 + nv);
  o[n++] = uint8_t(guard << 5);
  uint32_t magb = 0;
  for (uint32_t i = 0; i < nv; ++i) {
    uint32_t eps = raw_exp(ch.L, pl.depth, i) - guard;
    o[n++] = uint8_t(eps << 3);
    uint32_t s = eps + guard;
    magb = max(magb, s > 0 ? s - 1 : 0u);
  }
  // CAP (markers.rs:274-300)
  uint32_t bp = magb <= 8 ? 0 : (magb < 28 ? magb - 8 : 13 + (magb >> 2));
  put16(o, n, kCAP);
  put16(o, n, 8);
  put32(o, n, 0x00020000u);
  put16(o, n, bp & 0x1F);
  // COM
  const char* com = "nd-image-codecs ndic-codestream";
  put16(o, n, kCOM);
  put16(o, n, 2 + 2 + 31);
  put16(o, n, 1);
  for (int i = 0; i < 31; ++i) o[n++] = (uint8_t)com[i];
  // TLM
  if (ch.index) {
    put16(o, n, kTLM);
    put16(o, n, 8);
    o[n++] = 0;
    o[n++] = 0x40;
    put32(o, n, tile_part);
  }
  // SOT, PLT, SOD
  put16(o, n, kSOT);
  put16(o, n, 10);
  put16(o, n, 0);
  put32(o, n, (uint32_t)psot);
  o[n++] = 0;
  o[n++] = 1;
  for (uint32_t i = 0; i < pn; ++i) o[n++] = plt[i];
  put16(o, n, kSOD);
  pl.head_len = n;
  pl.total_len = n + body + 2;
}

// Container header and coefficient-plane index (container.rs:249-280,
// PlaneEntry::from_codestream container.rs:80-100).
// Container header and coefficient-plane index of each chunk
// (container.rs:80-100, 249-280), one 256-thread workgroup per chunk: the
// planes' codestream offsets are a workgroup prefix sum of their lengths,
// and each thread writes the index entries of its planes.
__global__ __launch_bounds__(256) void write_chunks_kernel(const DevChunk* __restrict__ chunks, uint32_t nchunks,
                                                           DevPlane* planes, uint64_t* chunk_len) {
  __shared__ uint64_t part[256];
  __shared__ uint64_t carry;
  const uint32_t c = blockIdx.x, tid = threadIdx.x;
  if (c >= nchunks) return;
  const DevChunk& ch = chunks[c];
  uint8_t* o = ch.chunk_head;
  const uint64_t fixed = 8 + 4ull * ch.ndim;
  if (tid == 0) {
    o[0] = 'n';
    o[1] = 'd';
    o[2] = 'h';
    o[3] = 't';
    o[4] = 1;
    o[5] = ch.index ? 1 : 0;
    o[6] = (uint8_t)ch.L;
    o[7] = (uint8_t)ch.ndim;
    carry = 0;
  }
  for (uint32_t i = tid; i < 4 * ch.ndim; i += 256) o[8 + i] = uint8_t(ch.dims[i / 4] >> (8 * (i % 4)));
  const uint64_t entry = 12 + 4ull * (ch.L + 1);
  const uint64_t first = fixed + (ch.index ? (uint64_t)ch.P * entry : 0);
  for (uint64_t p0 = 0; p0 < ch.P; p0 += 256) {
    const uint64_t p = p0 + tid;
    const uint64_t len = p < ch.P ? planes[ch.plane_base + p].total_len : 0;
    // inclusive scan of the lengths over the workgroup
    part[tid] = len;
    __syncthreads();
    for (uint32_t d = 1; d < 256; d <<= 1) {
      const uint64_t v = tid >= d ? part[tid - d] : 0;
      __syncthreads();
      part[tid] += v;
      __syncthreads();
    }
    const uint64_t off = first + carry + part[tid] - len;
    if (p < ch.P) {
      DevPlane& pl = planes[ch.plane_base + p];
      pl.cs_off = off;
      if (ch.index) {
        uint8_t* e = o + fixed + p * entry;
        for (int k = 0; k < 8; ++k) e[k] = uint8_t(off >> (8 * k));
        const uint32_t l32 = (uint32_t)len;
        for (int k = 0; k < 4; ++k) e[8 + k] = uint8_t(l32 >> (8 * k));
        uint32_t pre = pl.head_len;
        for (uint32_t r = 0; r <= ch.L; ++r) {
          pre += pl.packet_len[r];
          for (int k = 0; k < 4; ++k) e[12 + 4 * r + k] = uint8_t(pre >> (8 * k));
        }
      }
    }
    __syncthreads();
    if (tid == 255) carry += part[255];
    __syncthreads();
  }
  if (tid == 0) chunk_len[c] = first + carry;
}

// The workgroup copies n bytes (src read backward when rev), four loads in
// flight per thread.
__device__ __forceinline__ void wg_copy(uint8_t* dst, const uint8_t* src, uint32_t n, bool rev) {
  for (uint32_t i0 = threadIdx.x; i0 < n; i0 += 1024) {
    uint8_t v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint32_t i = i0 + 256 * k;
      v[k] = i < n ? (rev ? src[n - 1 - i] : src[i]) : 0;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint32_t i = i0 + 256 * k;
      if (i < n) dst[i] = v[k];
    }
  }
}

// One workgroup per plane copies [chunk header (plane 0)] plane header,
// packets (header + segments) and EOC into the packed output.
__global__ __launch_bounds__(256) void assemble_kernel(const DevChunk* __restrict__ chunks, uint32_t nchunks,
                                                       const DevPlane* __restrict__ planes,
                                                       const HtEncTask* __restrict__ tasks, const uint32_t* lens,
                                                       const GeoPacket* __restrict__ gp, uint8_t* out) {
  uint64_t gid = blockIdx.x;
  const DevChunk& ch = chunks[find_chunk(chunks, nchunks, gid)];
  if (ch.out_len == 0) return;  // refused chunk: nothing to write
  const DevPlane& pl = planes[gid];
  uint64_t p = gid - ch.plane_base;
  uint8_t* dst = out + ch.out_off;
  if (p == 0) wg_copy(dst, ch.chunk_head, (uint32_t)ch.chunk_head_len, false);
  dst += pl.cs_off;
  const uint8_t* head = ch.head + gid * (uint64_t)ch.head_stride;
  wg_copy(dst, head, pl.head_len, false);
  uint64_t pos = pl.head_len;
  uint64_t tb = ch.task_base + p * ch.blocks_per_plane;
  for (uint32_t r = 0; r <= ch.L; ++r) {
    const uint8_t* hdr = ch.pkt_hdr + (gid * (uint64_t)ch.pkt_slots + r) * (uint64_t)ch.pkt_stride;
    uint32_t hl = pl.packet_hdr[r];
    wg_copy(dst + pos, hdr, hl, false);
    pos += hl;
    const GeoPacket& pk = gp[ch.geo_packet + r];
    if (hl == 1 && pl.packet_len[r] == 1) continue;  // empty packet
    for (uint32_t b = 0; b < pk.nblocks; ++b) {
      uint64_t ti = tb + pk.block_first + b;
      const uint32_t* l = lens + 4 * ti;
      if (!l[3]) continue;
      // [MagSgn][MEL][VLC reversed] (the layout ht_seg_byte reads)
      const HtEncTask& t = tasks[ti];
      wg_copy(dst + pos, t.slot, l[0], false);
      wg_copy(dst + pos + l[0], t.slot + t.ms_cap, l[1], false);
      wg_copy(dst + pos + l[0] + l[1], t.slot + t.ms_cap + kMelCap, l[2], true);
      pos += (uint64_t)l[0] + l[1] + l[2];
    }
  }
  if (threadIdx.x == 0) {
    dst[pos] = 0xFF;
    dst[pos + 1] = 0xD9;
  }
}

}  // namespace

uint32_t tag_scratch_bytes(uint32_t nx, uint32_t ny) { return tree_bytes_host(nx, ny) * 2; }

void launch_parse_planes(const DevChunk* d_chunks, uint32_t nchunks, uint64_t nplanes, DevPlane* planes,
                         HtDecTask* tasks, const GeoPacket* gp, const GeoBand* gb, uint32_t tag_stride,
                         hipStream_t stream) {
  if (!nplanes) return;
  if ((int64_t)nplanes <= tune(kTuneParseWaveMaxPlanes) && nplanes <= 0xFFFFFFFFull) {
    hipLaunchKernelGGL(parse_planes_wave_kernel, dim3((uint32_t)nplanes), dim3(64), 0, stream, d_chunks, nchunks,
                       nplanes, planes, tasks, gp, gb);
  } else {
    // tag trees of up to 512 bytes per plane (64x64 blocks of planes up to
    // 2048 x 2048) live in LDS with the level offsets; larger ones in the
    // window's global scratch
    const uint32_t tl = tag_stride <= 512 ? tag_stride : 0;
    hipLaunchKernelGGL(parse_planes_kernel, dim3((uint32_t)ceil_div(nplanes, 64)), dim3(64),
                       64 * (tl + 8 * kTreeOffs), stream, d_chunks, nchunks, nplanes, planes, tasks, gp, gb, tl);
  }
  NDIC_HIP(hipGetLastError());
}

void launch_init_tasks(const DevChunk* d_chunks, uint32_t nchunks, uint64_t ntasks, HtDecTask* tasks,
                       const GeoPacket* gp, const GeoBand* gb, hipStream_t stream) {
  if (!ntasks) return;
  hipLaunchKernelGGL(init_tasks_kernel, dim3((uint32_t)ceil_div(ntasks, 256)), dim3(256), 0, stream, d_chunks,
                     nchunks, ntasks, tasks, gp, gb);
  NDIC_HIP(hipGetLastError());
}

void launch_walk_chunks(const DevChunk* d_chunks, uint32_t nchunks, DevPlane* planes, hipStream_t stream) {
  if (!nchunks) return;
  hipLaunchKernelGGL(walk_chunks_kernel, dim3((uint32_t)ceil_div(nchunks, 64)), dim3(64), 0, stream, d_chunks,
                     nchunks, planes);
  NDIC_HIP(hipGetLastError());
}

__global__ void fill32_kernel(uint32_t* p, uint32_t value, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = value;
}

void launch_fill32(void* p, uint32_t value, uint64_t bytes, hipStream_t stream) {
  const uint64_t n = bytes / 4;
  if (!n) return;
  const uint32_t blocks = (uint32_t)std::min<uint64_t>(ceil_div(n, 256), 1024);
  hipLaunchKernelGGL(fill32_kernel, dim3(blocks), dim3(256), 0, stream, (uint32_t*)p, value, n);
  NDIC_HIP(hipGetLastError());
}

void launch_window(const WindowJob* d_jobs, uint32_t njobs, uint64_t rows, hipStream_t stream) {
  if (!rows) return;
  uint32_t gx = (uint32_t)(rows < 65535 ? rows : 65535);
  uint32_t gy = (uint32_t)ceil_div(rows, gx);
  hipLaunchKernelGGL(window_kernel, dim3(gx, gy), dim3(256), 0, stream, d_jobs, njobs);
  NDIC_HIP(hipGetLastError());
}

void launch_unshift(const DevChunk* d_chunks, uint32_t nchunks, uint64_t nplanes, const DevPlane* planes,
                    int32_t* const* d_data, const uint64_t* d_dims, hipStream_t stream) {
  if (!nplanes) return;
  hipLaunchKernelGGL(unshift_kernel, dim3((uint32_t)nplanes), dim3(256), 0, stream, d_chunks, nchunks, planes,
                     d_data, d_dims);
  NDIC_HIP(hipGetLastError());
}

void launch_dequantize(const DevChunk* d_chunks, uint32_t nchunks, uint64_t nplanes, hipStream_t stream) {
  if (!nplanes) return;
  hipLaunchKernelGGL(dequantize_kernel, dim3((uint32_t)nplanes, 64), dim3(256), 0, stream, d_chunks, nchunks);
  NDIC_HIP(hipGetLastError());
}

void launch_irrev_round(const DevChunk* d_chunks, uint32_t nchunks, uint64_t nplanes, const DevPlane* planes,
                        int32_t* const* d_data, const uint64_t* d_dims, hipStream_t stream) {
  if (!nplanes) return;
  hipLaunchKernelGGL(irrev_round_kernel, dim3((uint32_t)nplanes, 16), dim3(256), 0, stream, d_chunks, nchunks, planes,
                     d_data, d_dims);
  NDIC_HIP(hipGetLastError());
}

void launch_plane_minmax(const DevChunk* d_chunks, uint32_t nchunks, uint64_t nplanes, DevPlane* planes,
                         hipStream_t stream) {
  if (!nplanes) return;
  hipLaunchKernelGGL(plane_minmax_kernel, dim3((uint32_t)nplanes), dim3(256), 0, stream, d_chunks, nchunks, planes);
  NDIC_HIP(hipGetLastError());
}

void launch_plane_setup(const DevChunk* d_chunks, uint32_t nchunks, uint64_t nplanes, DevPlane* planes,
                        hipStream_t stream) {
  if (!nplanes) return;
  hipLaunchKernelGGL(plane_setup_kernel, dim3((uint32_t)ceil_div(nplanes, 64)), dim3(64), 0, stream, d_chunks,
                     nchunks, nplanes, planes);
  NDIC_HIP(hipGetLastError());
}

void launch_dc_shift(const DevChunk* d_chunks, uint32_t nchunks, uint64_t nplanes, const DevPlane* planes,
                     hipStream_t stream) {
  if (!nplanes) return;
  hipLaunchKernelGGL(dc_shift_kernel, dim3((uint32_t)nplanes), dim3(256), 0, stream, d_chunks, nchunks, planes,
                     nullptr);
  NDIC_HIP(hipGetLastError());
}

void launch_build_enc_tasks(const DevChunk* d_chunks, uint32_t nchunks, uint64_t nplanes, const DevPlane* planes,
                            const GeoPacket* gp, const GeoBand* gb, HtEncTask* tasks, uint32_t* lens,
                            uint64_t slot_stride, hipStream_t stream) {
  if (!nplanes) return;
  hipLaunchKernelGGL(build_enc_tasks_kernel, dim3((uint32_t)ceil_div(nplanes, 64)), dim3(64), 0, stream, d_chunks,
                     nchunks, nplanes, planes, gp, gb, tasks, lens, slot_stride);
  NDIC_HIP(hipGetLastError());
}

void launch_write_planes(const DevChunk* d_chunks, uint32_t nchunks, uint64_t nplanes, DevPlane* planes,
                         const HtEncTask* tasks, const uint32_t* lens, const GeoPacket* gp, const GeoBand* gb,
                         hipStream_t stream) {
  if (!nplanes) return;
  hipLaunchKernelGGL(write_planes_kernel, dim3((uint32_t)ceil_div(nplanes, 64)), dim3(64), 0, stream, d_chunks,
                     nchunks, nplanes, planes, tasks, lens, gp, gb);
  NDIC_HIP(hipGetLastError());
}

void launch_write_chunks(const DevChunk* d_chunks, uint32_t nchunks, DevPlane* planes, uint64_t* chunk_len,
                         hipStream_t stream) {
  if (!nchunks) return;
  hipLaunchKernelGGL(write_chunks_kernel, dim3(nchunks), dim3(256), 0, stream, d_chunks, nchunks, planes, chunk_len);
  NDIC_HIP(hipGetLastError());
}

void launch_assemble(const DevChunk* d_chunks, uint32_t nchunks, uint64_t nplanes, const DevPlane* planes,
                     const HtEncTask* tasks, const uint32_t* lens, const GeoPacket* gp, uint8_t* out,
                     hipStream_t stream) {
  if (!nplanes) return;
  hipLaunchKernelGGL(assemble_kernel, dim3((uint32_t)nplanes), dim3(256), 0, stream, d_chunks, nchunks, planes,
                     tasks, lens, gp, out);
  NDIC_HIP(hipGetLastError());
}

}  // namespace ndic
