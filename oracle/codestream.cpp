// Oracle (test infrastructure only): Part 1/15 codestream, ndht container and
// the htj2k chunk cores, restated from the reference.
//   bit I/O:    crates/ndic-codestream/src/bitio.rs:16-161
//   tag trees:  crates/ndic-codestream/src/tagtree.rs:13-123
//   packets:    crates/ndic-codestream/src/packet.rs:69-362
//   markers:    crates/ndic-codestream/src/markers.rs:55-521
//   quant:      crates/ndic-codestream/src/quant.rs:18-202
//   geometry:   crates/ndic-codestream/src/geometry.rs:10-127
//   writer:     crates/ndic-codestream/src/writer.rs:55-390
//   reader:     crates/ndic-codestream/src/reader.rs:99-680
//   container:  crates/ndic-codestream/src/container.rs:80-340
//   chunk core: crates/ndic-zarr/src/htj2k.rs:113-441
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <sstream>

#include "oracle.hpp"

namespace orc {

namespace {

// ---------------------------------------------------------------------------
// bitio.rs
// ---------------------------------------------------------------------------
struct HeaderBitWriter {
  std::vector<uint8_t> out;
  uint32_t tmp = 0, avail = 8;
  void put_bit(uint32_t bit) {
    avail -= 1;
    tmp |= (bit & 1) << avail;
    if (avail == 0) {
      out.push_back((uint8_t)tmp);
      avail = tmp == 0xFF ? 7 : 8;
      tmp = 0;
    }
  }
  void put_bits(uint32_t data, uint32_t n) {
    for (uint32_t i = n; i-- > 0;) put_bit(data >> i);
  }
  void put_zeros(uint32_t n) {
    for (uint32_t i = 0; i < n; ++i) put_bit(0);
  }
  std::vector<uint8_t> terminate() {
    if (avail < 8) out.push_back((uint8_t)tmp);
    return out;
  }
};

struct HeaderBitReader {
  const uint8_t* data;
  size_t len, pos = 0, base;
  uint32_t tmp = 0, avail = 0;
  bool unstuff = false;
  HeaderBitReader(const uint8_t* d, size_t l, size_t b) : data(d), len(l), base(b) {}
  void fill() {
    if (pos >= len) fail(kCodestream, "packet header truncated", base + pos);
    uint8_t t = data[pos++];
    tmp = t;
    avail = 8 - (unstuff ? 1 : 0);
    unstuff = t == 0xFF;
  }
  uint32_t read_bit() {
    if (avail == 0) fill();
    avail -= 1;
    return (tmp >> avail) & 1;
  }
  uint32_t read_bits(uint32_t n) {
    uint32_t v = 0;
    for (uint32_t i = 0; i < n; ++i) v = (v << 1) | read_bit();
    return v;
  }
  size_t terminate() {
    if (unstuff) pos += pos < len ? 1 : 0;
    return pos;
  }
};

// ---------------------------------------------------------------------------
// tagtree.rs
// ---------------------------------------------------------------------------
uint32_t log2_ceil(size_t x) {
  uint32_t t = 63 - __builtin_clzll((unsigned long long)x);
  return t + ((x & (x - 1)) != 0 ? 1 : 0);
}

struct TagTree {
  std::vector<std::vector<uint8_t>> vals, flags;
  size_t w;
  static uint32_t num_levels(size_t w, size_t h) { return 1 + std::max(log2_ceil(w), log2_ceil(h)); }
  TagTree(size_t w_, size_t h, uint8_t init) : w(w_) {
    size_t nl = num_levels(w_, h);
    for (size_t lev = 0; lev <= nl; ++lev) {
      size_t lw = (w_ + ((size_t)1 << lev) - 1) >> lev, lh = (h + ((size_t)1 << lev) - 1) >> lev;
      vals.emplace_back(std::max<size_t>(lw, 1) * std::max<size_t>(lh, 1), lev == 0 ? init : 0);
      flags.emplace_back(std::max<size_t>(lw, 1) * std::max<size_t>(lh, 1), 0);
    }
  }
  size_t levels() const { return vals.size() - 1; }
  size_t idx(size_t x, size_t y, size_t lev) const {
    size_t lw = std::max<size_t>((w + ((size_t)1 << lev) - 1) >> lev, 1);
    return y * lw + x;
  }
  uint8_t val(size_t x, size_t y, size_t lev) const { return vals[lev][idx(x, y, lev)]; }
  void set_val(size_t x, size_t y, size_t lev, uint8_t v) { vals[lev][idx(x, y, lev)] = v; }
  uint8_t flag(size_t x, size_t y, size_t lev) const { return flags[lev][idx(x, y, lev)]; }
  void set_flag(size_t x, size_t y, size_t lev) { flags[lev][idx(x, y, lev)] = 1; }
  void reduce() {
    for (size_t lev = 1; lev <= levels(); ++lev) {
      size_t lw = std::max<size_t>((w + ((size_t)1 << lev) - 1) >> lev, 1);
      size_t n = vals[lev].size();
      for (size_t i = 0; i < n; ++i) {
        size_t x = i % lw, y = i / lw;
        uint8_t m = 255;
        size_t clw = std::max<size_t>((w + ((size_t)1 << (lev - 1)) - 1) >> (lev - 1), 1);
        size_t ch = (vals[lev - 1].size() + clw - 1) / clw;
        const size_t dd[4][2] = {{0, 0}, {1, 0}, {0, 1}, {1, 1}};
        for (auto& d : dd) {
          size_t cx = 2 * x + d[0], cy = 2 * y + d[1];
          if (cx < clw && cy < ch) m = std::min(m, vals[lev - 1][cy * clw + cx]);
        }
        vals[lev][i] = lev == levels() ? 0 : m;
      }
    }
    for (auto& v : vals[levels()]) v = 0;
  }
};

// ---------------------------------------------------------------------------
// packet.rs
// ---------------------------------------------------------------------------
struct CodedBlock {
  bool present = false;
  uint8_t missing_msbs = 0, num_passes = 1;
  uint32_t len_cleanup = 0, len_refinement = 0;
  std::vector<uint8_t> data;
};
struct BandBlocks {
  size_t nx = 0, ny = 0;
  std::vector<CodedBlock> blocks;
};

std::vector<uint8_t> write_packet(const std::vector<BandBlocks>& bands) {
  bool any = false;
  for (auto& b : bands)
    for (auto& c : b.blocks) any |= c.present;
  if (!any) return {0x00};
  HeaderBitWriter bb;
  bb.put_bit(1);
  for (auto& band : bands) {
    if (band.nx == 0 || band.ny == 0) continue;
    size_t nx = band.nx, ny = band.ny;
    size_t num_levels = TagTree::num_levels(nx, ny);
    TagTree inc(nx, ny, 255), mmsb(nx, ny, 255);
    for (size_t y = 0; y < ny; ++y)
      for (size_t x = 0; x < nx; ++x) {
        const CodedBlock& b = band.blocks[y * nx + x];
        inc.set_val(x, y, 0, b.present ? 0 : 1);
        mmsb.set_val(x, y, 0, b.present ? b.missing_msbs : 0);
      }
    inc.reduce();
    mmsb.reduce();
    if (inc.val(0, 0, num_levels - 1) != 0) {
      bb.put_bit(0);
      continue;
    }
    for (size_t y = 0; y < ny; ++y)
      for (size_t x = 0; x < nx; ++x) {
        for (size_t cur = num_levels; cur >= 1; --cur) {
          size_t lm = cur - 1;
          if (inc.flag(x >> lm, y >> lm, lm) == 0) {
            uint8_t skipped = inc.val(x >> lm, y >> lm, lm) - inc.val(x >> cur, y >> cur, cur);
            bb.put_bits(1 - std::min<uint32_t>(skipped, 1), 1);
            inc.set_flag(x >> lm, y >> lm, lm);
          }
          if (inc.val(x >> lm, y >> lm, lm) > 0) break;
        }
        const CodedBlock& cb = band.blocks[y * nx + x];
        if (!cb.present) continue;
        for (size_t cur = num_levels; cur >= 1; --cur) {
          size_t lm = cur - 1;
          if (mmsb.flag(x >> lm, y >> lm, lm) == 0) {
            uint8_t zeros = mmsb.val(x >> lm, y >> lm, lm) - mmsb.val(x >> cur, y >> cur, cur);
            bb.put_zeros(zeros);
            bb.put_bit(1);
            mmsb.set_flag(x >> lm, y >> lm, lm);
          }
        }
        switch (cb.num_passes) {
          case 1: bb.put_bits(0, 1); break;
          case 2: bb.put_bits(2, 2); break;
          default: bb.put_bits(12, 4); break;
        }
        uint32_t bits1 = 32 - (cb.len_cleanup ? __builtin_clz(cb.len_cleanup) : 32);
        uint32_t extra_bit = cb.num_passes > 2 ? 1 : 0;
        uint32_t bits2 = cb.num_passes > 1 ? 32 - (cb.len_refinement ? __builtin_clz(cb.len_refinement) : 32) : 0;
        uint32_t b2 = bits2 > extra_bit ? bits2 - extra_bit : 0;
        uint32_t mx = std::max(bits1, b2);
        uint32_t bits = mx > 3 ? mx - 3 : 0;
        bb.put_bits(0xFFFFFFFEu, bits + 1);
        bb.put_bits(cb.len_cleanup, bits + 3);
        if (cb.num_passes > 1) bb.put_bits(cb.len_refinement, bits + 3 + extra_bit);
      }
  }
  std::vector<uint8_t> out = bb.terminate();
  for (auto& band : bands)
    for (auto& cb : band.blocks)
      if (cb.present) out.insert(out.end(), cb.data.begin(), cb.data.end());
  return out;
}

struct ParseBand { size_t nx, ny; uint8_t k_max; };
struct ParsedBlock {
  size_t band, bx, by;
  uint8_t missing_msbs, num_passes;
  uint32_t len_cleanup, len_refinement;
};

size_t consume_eph(const uint8_t* data, size_t len, size_t pos, size_t offset) {
  if (len >= pos + 2 && data[pos] == 0xFF && data[pos + 1] == 0x92) return pos + 2;
  fail(kCodestream, "expected EPH after packet header", offset + pos);
}

std::vector<ParsedBlock> parse_packet_header(const uint8_t* data, size_t len, const std::vector<ParseBand>& bands,
                                             bool uses_sop, bool uses_eph, size_t offset, size_t* header_len) {
  size_t start = 0;
  if (uses_sop && len >= 6 && data[0] == 0xFF && data[1] == 0x91) start = 6;
  HeaderBitReader bb(data + start, len - start, offset + start);
  std::vector<ParsedBlock> blocks;
  bool empty_packet = true;
  for (size_t bi = 0; bi < bands.size(); ++bi) {
    const ParseBand& band = bands[bi];
    if (band.nx == 0 || band.ny == 0) continue;
    if (empty_packet) {
      if (bb.read_bit() == 0) {
        size_t hl = start + bb.terminate();
        if (uses_eph) hl = consume_eph(data, len, hl, offset);
        *header_len = hl;
        return blocks;
      }
      empty_packet = false;
    }
    size_t nx = band.nx, ny = band.ny;
    size_t num_levels = TagTree::num_levels(nx, ny);
    TagTree inc(nx, ny, 0), mmsb(nx, ny, 0);
    for (size_t y = 0; y < ny; ++y)
      for (size_t x = 0; x < nx; ++x) {
        bool empty_cb = false;
        for (size_t cl = num_levels; cl >= 1; --cl) {
          size_t cur = cl - 1;
          empty_cb = inc.val(x >> cur, y >> cur, cur) == 1;
          if (empty_cb) break;
          if (inc.flag(x >> cur, y >> cur, cur) == 0) {
            uint32_t bit = bb.read_bit();
            empty_cb = bit == 0;
            inc.set_val(x >> cur, y >> cur, cur, (uint8_t)(1 - bit));
            inc.set_flag(x >> cur, y >> cur, cur);
          }
          if (empty_cb) break;
        }
        if (empty_cb) continue;
        uint32_t mmsbs = 0;
        for (size_t lp1 = num_levels; lp1 >= 1; --lp1) {
          size_t cur = lp1 - 1;
          mmsbs = mmsb.val(x >> lp1, y >> lp1, lp1);
          if (mmsb.flag(x >> cur, y >> cur, cur) == 0) {
            while (bb.read_bit() != 1) mmsbs += 1;
            mmsb.set_val(x >> cur, y >> cur, cur, (uint8_t)std::min<uint32_t>(mmsbs, 255));
            mmsb.set_flag(x >> cur, y >> cur, cur);
          }
        }
        uint32_t num_passes = 1;
        if (bb.read_bit() == 1) {
          num_passes = 2;
          if (bb.read_bit() == 1) {
            uint32_t v = bb.read_bits(2);
            num_passes = 3 + v;
            if (v == 3) {
              v = bb.read_bits(5);
              num_passes = 6 + v;
              if (v == 31) num_passes = 37 + bb.read_bits(7);
            }
          }
        }
        uint32_t num_phld = (num_passes - 1) / 3;
        mmsbs += num_phld;
        num_passes -= num_phld * 3;
        if (mmsbs > band.k_max) fail(kCodestream, "missing MSBs exceed the QCD K_max bound", offset);
        uint32_t lblock = 3;
        while (bb.read_bit() == 1) {
          lblock += 1;
          if (lblock > 24) fail(kCodestream, "packet header Lblock exceeds the length bound", offset);
        }
        uint32_t bits = lblock + 31 - __builtin_clz(num_phld * 3 + 1);
        uint32_t len1 = bb.read_bits(bits);
        if (len1 < 2 || len1 >= 65535) fail(kCodestream, "HT cleanup segment length out of range", offset);
        uint32_t len2 = 0;
        if (num_passes > 1) {
          len2 = bb.read_bits(lblock + (num_passes > 2 ? 1 : 0));
          if (len2 >= 2047) fail(kCodestream, "HT refinement segment length out of range", offset);
        }
        blocks.push_back({bi, x, y, (uint8_t)std::min<uint32_t>(mmsbs, 255), (uint8_t)std::min<uint32_t>(num_passes, 3),
                          len1, len2});
      }
  }
  if (empty_packet) (void)bb.read_bit();
  size_t hl = start + bb.terminate();
  if (uses_eph) hl = consume_eph(data, len, hl, offset);
  *header_len = hl;
  return blocks;
}

// ---------------------------------------------------------------------------
// quant.rs
// ---------------------------------------------------------------------------
const uint32_t X_LL[34] = {0, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2};
const uint32_t X_LH[34] = {0, 2, 3, 3, 3, 3, 3, 3, 3, 3, 3, 3, 3, 3, 3, 3, 3, 3, 3, 3, 3, 3, 3, 3, 3, 3, 3, 3, 3, 3, 3, 3, 3, 3};
const uint32_t X_HH[34] = {0, 2, 3, 3, 3, 3, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4};

struct Quant {
  uint8_t guard_bits = 1, style = 0;
  std::vector<uint16_t> values;
  static Quant reversible(unsigned nd, unsigned bit_depth) {
    uint32_t b = bit_depth;
    if (nd >= 34) fail(kInvalidArgument, "more than 33 decomposition levels");
    std::vector<uint32_t> raw;
    raw.push_back(b + (nd == 0 ? 1 : X_LL[nd]));
    for (unsigned d = nd; d >= 1; --d) {
      raw.push_back(b + X_LH[d]);
      raw.push_back(b + X_LH[d]);
      raw.push_back(b + X_HH[d]);
    }
    uint32_t max_raw = *std::max_element(raw.begin(), raw.end());
    if (max_raw > 38) fail(kInvalidArgument, "bit depth + wavelet gain exceeds the 38-bit J2K bound");
    uint32_t g = std::max<uint32_t>(max_raw > 31 ? max_raw - 31 : 0, 1);
    Quant q;
    q.guard_bits = (uint8_t)g;
    q.style = 0;
    for (uint32_t r : raw) q.values.push_back((uint16_t)(r - g));
    return q;
  }
  uint8_t k_max(unsigned res, unsigned band) const {
    size_t idx = res == 0 ? 0 : (res - 1) * 3 + band;
    idx = std::min(idx, values.empty() ? 0 : values.size() - 1);
    uint32_t nb;
    if (style == 0) {
      uint32_t eps = values[idx];
      nb = eps == 0 ? 0 : eps - 1;
    } else {
      uint32_t e = values[idx] >> 11;
      nb = e > 0 ? e - 1 : 0;
    }
    return (uint8_t)(nb + guard_bits);
  }
  // quant.rs:204-221: the irreversible step of (res, band) with the band gain
  // (T.800 E-3, relative to the samples' nominal range 2^depth)
  float irrev_delta(unsigned res, unsigned band) const {
    const float gains[4] = {1.0f, 2.0f, 2.0f, 4.0f};
    size_t idx = res == 0 ? 0 : (res - 1) * 3 + band;
    idx = std::min(idx, values.empty() ? 0 : values.size() - 1);
    uint16_t v = values[idx];
    uint32_t eps = v >> 11;
    float mantissa = (float)((v & 0x7FF) | 0x800) * gains[std::min(band, 3u)] / (float)(1u << 11);
    float two_eps;
    uint32_t bits = (127 + eps) << 23;
    std::memcpy(&two_eps, &bits, 4);
    return mantissa / two_eps;
  }
  uint32_t magb() const {
    uint32_t b = 0;
    for (size_t idx = 0; idx < values.size(); ++idx) {
      uint32_t t;
      if (style == 0) {
        uint32_t s = uint32_t(values[idx]) + guard_bits;
        t = s > 0 ? s - 1 : 0;
      } else {
        size_t nd = (values.size() - 1) / 3;
        size_t nb = nd - (idx > 0 ? (idx - 1) / 3 : 0);
        uint32_t s = uint32_t(values[idx] >> 11) + guard_bits;
        t = s > nb ? s - (uint32_t)nb : 0;
      }
      b = std::max(b, t);
    }
    return b;
  }
  std::vector<uint8_t> to_bytes() const {
    std::vector<uint8_t> out{(uint8_t)((guard_bits << 5) | style)};
    for (uint16_t v : values) {
      if (style == 0) out.push_back((uint8_t)(v << 3));
      else { out.push_back(v >> 8); out.push_back(v & 0xFF); }
    }
    return out;
  }
  static Quant parse(const uint8_t* p, size_t n, size_t offset) {
    if (n == 0) fail(kCodestream, "empty QCD payload", offset);
    Quant q;
    q.style = p[0] & 0x1F;
    q.guard_bits = p[0] >> 5;
    if (q.style == 0) {
      for (size_t i = 1; i < n; ++i) q.values.push_back(p[i] >> 3);
    } else if (q.style == 2) {
      if ((n - 1) % 2 != 0) fail(kCodestream, "odd-length expounded QCD payload", offset);
      for (size_t i = 1; i + 1 < n; i += 2) q.values.push_back(uint16_t(p[i] << 8 | p[i + 1]));
    } else {
      fail(kCodestream, "unsupported quantization style (derived)", offset);
    }
    if (q.values.empty()) fail(kCodestream, "QCD with no subband step sizes", offset);
    size_t max_band = (q.values.size() + 2) / 3;
    for (size_t r = 0; r <= std::min<size_t>(max_band, 255); ++r)
      for (unsigned b = r > 0 ? 1 : 0; b <= (r > 0 ? 3u : 0u); ++b)
        if (q.k_max((unsigned)r, b) > 31) fail(kCodestream, "QCD K_max exceeds the 31-bit magnitude datapath", offset);
    return q;
  }
};

// ---------------------------------------------------------------------------
// geometry.rs
// ---------------------------------------------------------------------------
struct Band { unsigned res, band; size_t x0, y0, w, h; };

std::vector<Band> bands_of_resolution(size_t width, size_t height, unsigned levels, unsigned res) {
  std::vector<Band> out;
  if (res == 0) {
    size_t w, h;
    level_dims(width, height, levels, &w, &h);
    out.push_back({0, 0, 0, 0, w, h});
    return out;
  }
  unsigned l = levels - res + 1;
  size_t wp, hp, wl, hl;
  level_dims(width, height, l - 1, &wp, &hp);
  level_dims(width, height, l, &wl, &hl);
  size_t wh = wp - wl, hh = hp - hl;
  out.push_back({res, 1, wl, 0, wh, hl});
  out.push_back({res, 2, 0, hl, wl, hh});
  out.push_back({res, 3, wl, hl, wh, hh});
  return out;
}
void band_grid(const Band& b, size_t cbw, size_t cbh, size_t* nx, size_t* ny) {
  if (b.w == 0 || b.h == 0) { *nx = 0; *ny = 0; return; }
  *nx = (b.w + cbw - 1) / cbw;
  *ny = (b.h + cbh - 1) / cbh;
}
void effective_cb(size_t cbw, size_t cbh, unsigned ppx, unsigned ppy, unsigned res, size_t* ew, size_t* eh) {
  unsigned shift = res > 0 ? 1 : 0;
  size_t pw = (size_t)1 << std::min<unsigned>(ppx - shift, 31);
  size_t ph = (size_t)1 << std::min<unsigned>(ppy - shift, 31);
  *ew = std::min(cbw, pw);
  *eh = std::min(cbh, ph);
}

// ---------------------------------------------------------------------------
// markers.rs
// ---------------------------------------------------------------------------
const uint16_t SOC = 0xFF4F, SIZ = 0xFF51, COD = 0xFF52, COC = 0xFF53, QCD = 0xFF5C, QCC = 0xFF5D, CAP = 0xFF50,
               COM = 0xFF64, TLM = 0xFF55, PLT = 0xFF58, SOT = 0xFF90, SOD = 0xFF93, EOC = 0xFFD9;

void be16(std::vector<uint8_t>& o, uint32_t v) { o.push_back(uint8_t(v >> 8)); o.push_back(uint8_t(v)); }
void be32(std::vector<uint8_t>& o, uint32_t v) { be16(o, v >> 16); be16(o, v & 0xFFFF); }
uint16_t rd16(const uint8_t* p) { return uint16_t(p[0] << 8 | p[1]); }
uint32_t rd32(const uint8_t* p) { return uint32_t(p[0]) << 24 | uint32_t(p[1]) << 16 | uint32_t(p[2]) << 8 | p[3]; }

void push_segment(std::vector<uint8_t>& out, uint16_t marker, const std::vector<uint8_t>& payload) {
  if (payload.size() + 2 > 0xFFFF) fail(kInvalidArgument, "marker segment payload exceeds 65533 bytes");
  be16(out, marker);
  be16(out, uint32_t(payload.size() + 2));
  out.insert(out.end(), payload.begin(), payload.end());
}

std::vector<uint8_t> plt_payload(uint8_t zplt, const uint32_t* lens, size_t n) {
  std::vector<uint8_t> p{zplt};
  for (size_t i = 0; i < n; ++i) {
    uint32_t len = lens[i];
    bool started = false;
    for (uint32_t shift : {28u, 21u, 14u, 7u}) {
      uint32_t b = (len >> shift) & 0x7F;
      if (b != 0 || started) {
        p.push_back(uint8_t(0x80 | b));
        started = true;
      }
    }
    p.push_back(uint8_t(len & 0x7F));
  }
  return p;
}

struct Siz {
  uint16_t rsiz = 0;
  uint32_t xsiz = 0, ysiz = 0, xosiz = 0, yosiz = 0, xtsiz = 0, ytsiz = 0, xtosiz = 0, ytosiz = 0;
  struct Comp { uint8_t depth; bool is_signed; uint8_t xr, yr; };
  std::vector<Comp> comps;
};
struct Cod {
  uint8_t scod = 0, progression = 0, mct = 0, decomps = 0, cbw_exp = 0, cbh_exp = 0, cb_style = 0, wavelet = 0;
  uint16_t layers = 0;
  std::vector<std::pair<uint8_t, uint8_t>> precincts;
  std::pair<unsigned, unsigned> precinct_exp(unsigned r) const {
    if (precincts.empty()) return {15, 15};
    auto pp = precincts[std::min<size_t>(r, precincts.size() - 1)];
    return {pp.first, pp.second};
  }
};

Siz parse_siz(const uint8_t* p, size_t n, size_t off) {
  if (n < 36) fail(kCodestream, "SIZ too short", off);
  size_t csiz = rd16(p + 34);
  if (n < 36 + 3 * csiz || csiz == 0) fail(kCodestream, "SIZ component list truncated", off);
  Siz s;
  for (size_t c = 0; c < csiz; ++c) {
    uint8_t ss = p[36 + 3 * c];
    uint8_t depth = (ss & 0x7F) + 1;
    if (depth < 1 || depth > 38) fail(kCodestream, "SIZ component depth outside 1..=38", off);
    s.comps.push_back({depth, (ss & 0x80) != 0, p[37 + 3 * c], p[38 + 3 * c]});
  }
  s.rsiz = rd16(p);
  s.xsiz = rd32(p + 2); s.ysiz = rd32(p + 6); s.xosiz = rd32(p + 10); s.yosiz = rd32(p + 14);
  s.xtsiz = rd32(p + 18); s.ytsiz = rd32(p + 22); s.xtosiz = rd32(p + 26); s.ytosiz = rd32(p + 30);
  if (s.xtsiz == 0 || s.ytsiz == 0) fail(kCodestream, "SIZ with a zero tile dimension", off);
  if (s.xsiz <= s.xosiz || s.ysiz <= s.yosiz) fail(kCodestream, "SIZ canvas origin at or past its size", off);
  if (s.xtosiz > s.xosiz || s.ytosiz > s.yosiz) fail(kCodestream, "SIZ tile origin past the canvas origin", off);
  return s;
}

Cod parse_cod(const uint8_t* p, size_t n, size_t off) {
  if (n < 10) fail(kCodestream, "COD too short", off);
  Cod c;
  c.scod = p[0];
  c.decomps = p[5];
  if (c.decomps > 32) fail(kCodestream, "COD with more than 32 decomposition levels", off);
  if (p[6] > 8 || p[7] > 8 || p[6] + p[7] > 8) fail(kCodestream, "COD code-block exponents out of range", off);
  if (c.scod & 1) {
    size_t need = size_t(c.decomps) + 1;
    if (n < 10 + need) fail(kCodestream, "COD precinct list truncated", off);
    for (size_t i = 0; i < need; ++i) {
      uint8_t b = p[10 + i];
      std::pair<uint8_t, uint8_t> pp{uint8_t(b & 0xF), uint8_t(b >> 4)};
      if (i > 0 && (pp.first == 0 || pp.second == 0)) fail(kCodestream, "COD precinct exponent 0 above resolution 0", off);
      c.precincts.push_back(pp);
    }
  }
  c.progression = p[1];
  c.layers = rd16(p + 2);
  c.mct = p[4];
  c.cbw_exp = p[6];
  c.cbh_exp = p[7];
  c.cb_style = p[8];
  c.wavelet = p[9];
  return c;
}

void parse_tlm(const uint8_t* p, size_t n, size_t off, std::vector<uint32_t>& tlm) {
  if (n < 2) fail(kCodestream, "TLM too short", off);
  uint8_t stlm = p[1];
  size_t st = (stlm >> 4) & 0x3, sp = (stlm & 0x40) ? 4 : 2;
  if (st == 3) fail(kCodestream, "invalid TLM ST field", off);
  size_t entry = st + sp, body = n - 2;
  if (body % entry != 0) fail(kCodestream, "TLM entries truncated", off);
  for (size_t i = 0; i < body; i += entry) {
    const uint8_t* e = p + 2 + i;
    tlm.push_back(sp == 4 ? rd32(e + st) : rd16(e + st));
  }
}

void parse_plt(const uint8_t* p, size_t n, size_t off, std::vector<uint32_t>& out) {
  if (n == 0) fail(kCodestream, "PLT too short", off);
  uint32_t acc = 0;
  bool pending = false;
  for (size_t i = 1; i < n; ++i) {
    uint8_t b = p[i];
    acc = (acc << 7) | (b & 0x7F);
    if ((b & 0x80) == 0) {
      out.push_back(acc);
      acc = 0;
      pending = false;
    } else {
      pending = true;
    }
  }
  if (pending) fail(kCodestream, "PLT ends mid-varint", off);
}

// ---------------------------------------------------------------------------
// writer.rs
// ---------------------------------------------------------------------------
BandBlocks code_band(const std::vector<int32_t>& plane, size_t plane_width, const Band& band, size_t cbw, size_t cbh,
                     const Quant& quant) {
  BandBlocks bb;
  band_grid(band, cbw, cbh, &bb.nx, &bb.ny);
  uint8_t k_max = quant.k_max(band.res, band.band);
  uint32_t shift = 31 - k_max;
  for (size_t by = 0; by < bb.ny; ++by)
    for (size_t bx = 0; bx < bb.nx; ++bx) {
      size_t rx = bx * cbw, ry = by * cbh;
      size_t rw = std::min(cbw, band.w - rx), rh = std::min(cbh, band.h - ry);
      std::vector<uint32_t> buf(rw * rh);
      uint32_t mv = 0;
      for (size_t y = 0; y < rh; ++y)
        for (size_t x = 0; x < rw; ++x) {
          int32_t src = plane[(band.y0 + ry + y) * plane_width + band.x0 + rx + x];
          uint32_t sm = (src < 0 ? 0x80000000u : 0u) | ((uint32_t)(src < 0 ? -(int64_t)src : src) << shift);
          mv |= sm & 0x7FFFFFFF;
          buf[y * rw + x] = sm;
        }
      CodedBlock cb;
      if (mv < (1u << shift)) {
        bb.blocks.push_back(cb);
        continue;
      }
      cb.data = ht_encode_block(buf.data(), rw, rh, rw, uint32_t(k_max) - 1);
      cb.present = true;
      cb.missing_msbs = k_max - 1;
      cb.num_passes = 1;
      cb.len_cleanup = (uint32_t)cb.data.size();
      bb.blocks.push_back(std::move(cb));
    }
  return bb;
}

// ---------------------------------------------------------------------------
// reader.rs
// ---------------------------------------------------------------------------
struct TilePart { uint32_t psot; size_t offset, body_start, body_end; std::vector<uint32_t> plt; };
struct Codestream {
  const uint8_t* data;
  size_t len, total_len, first_tile_offset;
  Siz siz;
  Cod cod;
  Quant quant;
  std::vector<uint32_t> tlm;
  std::vector<TilePart> tile_parts;
};

Codestream parse_codestream(const uint8_t* data, size_t len, bool strict) {
  auto r16 = [&](size_t pos) -> uint16_t {
    if (pos + 2 > len) fail(kCodestream, "unexpected end of codestream", pos);
    return rd16(data + pos);
  };
  if (r16(0) != SOC) fail(kCodestream, "missing SOC", 0);
  size_t pos = 2;
  bool has_siz = false, has_cod = false, has_qcd = false;
  Codestream cs;
  cs.data = data;
  cs.len = len;
  for (;;) {
    uint16_t marker = r16(pos);
    if (marker == SOT) break;
    if (marker < 0xFF00) fail(kCodestream, "expected a marker in the main header", pos);
    size_t l = r16(pos + 2);
    if (l < 2 || pos + 2 + l > len) fail(kCodestream, "marker segment length out of bounds", pos);
    const uint8_t* pl = data + pos + 4;
    size_t pn = l - 2;
    switch (marker) {
      case SIZ: cs.siz = parse_siz(pl, pn, pos); has_siz = true; break;
      case COD: cs.cod = parse_cod(pl, pn, pos); has_cod = true; break;
      case QCD: cs.quant = Quant::parse(pl, pn, pos); has_qcd = true; break;
      case CAP: {
        if (pn < 6) fail(kCodestream, "CAP too short", pos);
        if ((rd32(pl) & 0x00020000u) == 0) fail(kCodestream, "CAP without the Part-15 (Pcap^15) bit", pos);
        break;
      }
      case TLM: parse_tlm(pl, pn, pos, cs.tlm); break;
      case COC: case QCC: fail(kUnsupported, "per-component COC/QCC overrides");
      default: break;
    }
    pos += 2 + l;
  }
  if (!has_siz) fail(kCodestream, "missing SIZ", pos);
  if (!has_cod) fail(kCodestream, "missing COD", pos);
  if (!has_qcd) fail(kCodestream, "missing QCD", pos);
  cs.first_tile_offset = pos;
  auto declared_len = [&]() {
    size_t m = 0;
    bool any = false;
    for (auto& tp : cs.tile_parts) {
      m = std::max(m, tp.psot == 0 ? tp.body_end : tp.body_end + 2);
      any = true;
    }
    return any ? m : len;
  };
  for (;;) {
    if (!strict && pos >= len) { cs.total_len = declared_len(); break; }
    if (pos + 2 > len) {
      if (!strict) { cs.total_len = declared_len(); break; }
      fail(kCodestream, "unexpected end of codestream", pos);
    }
    uint16_t marker = rd16(data + pos);
    if (marker == EOC) { cs.total_len = pos + 2; break; }
    if (marker != SOT) fail(kCodestream, "expected SOT or EOC", pos);
    size_t sot_off = pos;
    if (r16(pos + 2) != 10) fail(kCodestream, "SOT with unexpected length", pos);
    if (pos + 12 > len) fail(kCodestream, "truncated SOT segment", pos);
    uint32_t psot = rd32(data + pos + 6);
    pos += 12;
    TilePart tp;
    tp.psot = psot;
    tp.offset = sot_off;
    for (;;) {
      uint16_t m = r16(pos);
      if (m == SOD) { pos += 2; break; }
      size_t l = r16(pos + 2);
      if (l < 2 || pos + 2 + l > len) fail(kCodestream, "tile-part marker length out of bounds", pos);
      if (m == PLT) parse_plt(data + pos + 4, l - 2, pos, tp.plt);
      else if (m == COD || m == QCD || m == COC || m == QCC) fail(kUnsupported, "tile-part coding-style overrides");
      pos += 2 + l;
    }
    size_t body_end;
    if (psot == 0) {
      bool ends_eoc = len >= 2 && data[len - 2] == 0xFF && data[len - 1] == 0xD9;
      body_end = (strict || ends_eoc) ? (len >= 2 ? len - 2 : 0) : len;
    } else {
      body_end = sot_off + psot;
    }
    if (body_end < pos || (strict && body_end > len)) fail(kCodestream, "Psot out of bounds", sot_off);
    tp.body_start = pos;
    tp.body_end = body_end;
    cs.tile_parts.push_back(tp);
    pos = body_end;
  }
  return cs;
}

bool irreversible(const Codestream& cs) { return cs.cod.wavelet == 0 && cs.quant.style == 2; }

void check_supported(const Codestream& cs, bool allow_irreversible = false) {
  const Siz& s = cs.siz;
  size_t tx = (s.xsiz - s.xtosiz + s.xtsiz - 1) / s.xtsiz, ty = (s.ysiz - s.ytosiz + s.ytsiz - 1) / s.ytsiz;
  if (tx != 1 || ty != 1 || s.xosiz || s.yosiz || s.xtosiz || s.ytosiz)
    fail(kUnsupported, "multi-tile or offset-canvas codestreams");
  for (auto& c : s.comps)
    if (c.xr != 1 || c.yr != 1) fail(kUnsupported, "component subsampling");
  if (!(cs.cod.cb_style & 0x40)) fail(kUnsupported, "non-HT (legacy J2K-1) block coding");
  if (allow_irreversible && irreversible(cs)) {
    // the decode-only 9/7 path: one component, no ICT, depths the int32
    // level-shifted samples hold
    if (cs.siz.comps.size() != 1 || cs.cod.mct) fail(kUnsupported, "irreversible codestreams with several components");
    if (cs.siz.comps[0].depth > 30) fail(kUnsupported, "irreversible codestreams deeper than 30 bits");
  } else if (cs.cod.wavelet != 1 || cs.quant.style != 0) {
    fail(kUnsupported, "irreversible (9/7) codestreams");
  }
  if (cs.cod.layers != 1) fail(kUnsupported, "multiple quality layers");
  if ((cs.cod.progression == 3 || cs.cod.progression == 4) && (cs.cod.scod & 1))
    fail(kUnsupported, "PCRL/CPRL with explicit precinct sizes");
}

void precinct_grid(const Codestream& cs, unsigned r, size_t* npx, size_t* npy) {
  size_t rw, rh;
  level_dims(cs.siz.xsiz, cs.siz.ysiz, cs.cod.decomps - r, &rw, &rh);
  auto pp = cs.cod.precinct_exp(r);
  if (rw == 0 || rh == 0) { *npx = 0; *npy = 0; return; }
  size_t pw = (size_t)1 << std::min(pp.first, 31u), ph = (size_t)1 << std::min(pp.second, 31u);
  *npx = (rw + pw - 1) / pw;
  *npy = (rh + ph - 1) / ph;
}

struct Seq { size_t comp; unsigned res; size_t px, py; };
std::vector<Seq> packet_sequence(const Codestream& cs) {
  size_t comps = cs.siz.comps.size();
  std::vector<Seq> order;
  if (cs.cod.progression == 2) {
    for (unsigned r = 0; r <= cs.cod.decomps; ++r) {
      size_t npx, npy;
      precinct_grid(cs, r, &npx, &npy);
      for (size_t py = 0; py < npy; ++py)
        for (size_t px = 0; px < npx; ++px)
          for (size_t c = 0; c < comps; ++c) order.push_back({c, r, px, py});
    }
  } else if (cs.cod.progression == 3 || cs.cod.progression == 4) {
    for (size_t c = 0; c < comps; ++c)
      for (unsigned r = 0; r <= cs.cod.decomps; ++r) {
        size_t npx, npy;
        precinct_grid(cs, r, &npx, &npy);
        for (size_t py = 0; py < npy; ++py)
          for (size_t px = 0; px < npx; ++px) order.push_back({c, r, px, py});
      }
  } else {
    for (unsigned r = 0; r <= cs.cod.decomps; ++r)
      for (size_t c = 0; c < comps; ++c) {
        size_t npx, npy;
        precinct_grid(cs, r, &npx, &npy);
        for (size_t py = 0; py < npy; ++py)
          for (size_t px = 0; px < npx; ++px) order.push_back({c, r, px, py});
      }
  }
  return order;
}

struct PacketSpan { size_t comp; unsigned res; size_t offset, len; };
std::vector<PacketSpan> packet_index(const Codestream& cs) {
  check_supported(cs);
  for (auto& tp : cs.tile_parts)
    if (tp.plt.empty()) fail(kUnsupported, "codestream has no PLT index");
  if (!cs.tlm.empty()) {
    size_t off = cs.first_tile_offset;
    for (size_t i = 0; i < cs.tlm.size(); ++i) {
      if (i >= cs.tile_parts.size()) fail(kCodestream, "TLM lists more tile-parts than found", off);
      if (cs.tile_parts[i].offset != off) fail(kCodestream, "TLM offsets disagree with SOT positions", off);
      off += cs.tlm[i];
    }
  }
  std::vector<Seq> order = packet_sequence(cs);
  std::vector<PacketSpan> spans;
  size_t oi = 0;
  for (auto& tp : cs.tile_parts) {
    size_t off = tp.body_start;
    for (uint32_t l : tp.plt) {
      if (oi >= order.size()) fail(kCodestream, "PLT lists more packets than expected", off);
      spans.push_back({order[oi].comp, order[oi].res, off, l});
      off += l;
      oi += 1;
    }
    if (off != tp.body_end) fail(kCodestream, "PLT lengths do not cover the tile-part body", off);
  }
  if (oi != order.size()) fail(kCodestream, "PLT covers fewer packets than expected", cs.len);
  return spans;
}

void precinct_block_range(size_t bw, size_t bh, unsigned ppx_log, unsigned ppy_log, size_t px, size_t py, size_t cbw,
                          size_t cbh, size_t r[4]) {
  r[0] = r[1] = r[2] = r[3] = 0;
  if (bw == 0 || bh == 0) return;
  size_t ppw = (size_t)1 << std::min(ppx_log, 31u), pph = (size_t)1 << std::min(ppy_log, 31u);
  size_t x0 = px * ppw, y0 = py * pph;
  if (x0 >= bw || y0 >= bh) return;
  size_t x1 = std::min(x0 + ppw, bw), y1 = std::min(y0 + pph, bh);
  size_t bx0 = x0 / cbw, by0 = y0 / cbh;
  r[0] = bx0;
  r[1] = by0;
  r[2] = (x1 + cbw - 1) / cbw - bx0;
  r[3] = (y1 + cbh - 1) / cbh - by0;
}

// reader.rs:478-645 (single component is all the htj2k chunk core accepts;
// multi-component streams decode each component the same way)
// The step that turns the block decoder's aligned magnitude (bit 30 down,
// with the half-bin bit below the last decoded plane: decode.rs:5-6) into a
// coefficient of the unscaled 9/7 lifting (dwt/mod.rs:7-9) for a decode to
// resolution max_res: irrev_delta x 2^depth (T.800 E-3's 2^R_b) x the K
// factor that lifting leaves in the band relative to the output LL
// (K^2 per 2-D level of low-pass: LL of level L holds K^2L, HL/LH of level d
// K^2(d-1), HH of level d K^2(d-2), all over K^2(L-max_res)) / 2^(31-K_max).
// Double precision, rounded once to f32.
float irrev_scale(const Quant& q, unsigned res, unsigned band, unsigned levels, unsigned max_res, unsigned depth) {
  constexpr double kK = 1.230174104914001;  // T.800 Table F.4
  constexpr double kK2 = kK * kK;
  unsigned d = res == 0 ? levels : levels - res + 1;
  int e = res == 0 ? (int)levels : band == 3 ? (int)d - 2 : (int)d - 1;
  e -= (int)(levels - max_res);
  double f = 1.0;
  if (e < 0) f = 1.0 / kK2;
  for (int i = 0; i < e; ++i) f *= kK2;
  double s = (double)q.irrev_delta(res, band) * f;
  return (float)std::ldexp(s, (int)depth - 31 + (int)q.k_max(res, band));
}

std::vector<std::vector<int32_t>> decode_to_resolution(const Codestream& cs, unsigned max_res, size_t* ow, size_t* oh,
                                                       bool allow_irreversible = false) {
  check_supported(cs, allow_irreversible);
  const bool irrev = irreversible(cs);
  max_res = std::min<unsigned>(max_res, cs.cod.decomps);
  unsigned levels = cs.cod.decomps;
  size_t width = cs.siz.xsiz, height = cs.siz.ysiz, ncomps = cs.siz.comps.size();
  std::vector<uint8_t> body;
  for (auto& tp : cs.tile_parts) {
    size_t s = std::min(tp.body_start, cs.len), e = std::min(tp.body_end, cs.len);
    body.insert(body.end(), cs.data + s, cs.data + e);
  }
  std::vector<std::vector<int32_t>> planes(ncomps, std::vector<int32_t>(width * height, 0));
  std::vector<float> fplane(irrev ? width * height : 0, 0.0f);  // irreversible: dequantised coefficients
  size_t cbw_n = (size_t)1 << (cs.cod.cbw_exp + 2), cbh_n = (size_t)1 << (cs.cod.cbh_exp + 2);
  bool uses_sop = cs.cod.scod & 2, uses_eph = cs.cod.scod & 4, stripe_causal = cs.cod.cb_style & 0x08;
  std::vector<Seq> seq = packet_sequence(cs);
  long last_needed = -1;
  for (size_t i = 0; i < seq.size(); ++i)
    if (seq[i].res <= max_res) last_needed = (long)i;
  size_t cursor = 0;
  for (size_t i = 0; i < seq.size(); ++i) {
    if (last_needed < 0 || (long)i > last_needed) break;
    if (cursor > body.size()) fail(kCodestream, "packet body truncated", cursor);
    unsigned res = seq[i].res;
    std::vector<Band> bands = bands_of_resolution(width, height, levels, res);
    auto pp = cs.cod.precinct_exp(res);
    size_t ecbw, ecbh;
    effective_cb(cbw_n, cbh_n, pp.first, pp.second, res, &ecbw, &ecbh);
    unsigned shift = res > 0 ? 1 : 0;
    std::vector<std::array<size_t, 4>> sub(bands.size());
    std::vector<ParseBand> pbands;
    for (size_t b = 0; b < bands.size(); ++b) {
      size_t r[4];
      precinct_block_range(bands[b].w, bands[b].h, pp.first > shift ? pp.first - shift : 0,
                           pp.second > shift ? pp.second - shift : 0, seq[i].px, seq[i].py, ecbw, ecbh, r);
      sub[b] = {r[0], r[1], r[2], r[3]};
      pbands.push_back({r[2], r[3], cs.quant.k_max(bands[b].res, bands[b].band)});
    }
    size_t hl;
    std::vector<ParsedBlock> blocks =
        parse_packet_header(body.data() + cursor, body.size() - cursor, pbands, uses_sop, uses_eph, cursor, &hl);
    size_t p = cursor + hl;
    for (auto& blk : blocks) {
      const Band& band = bands[blk.band];
      size_t bx = sub[blk.band][0] + blk.bx, by = sub[blk.band][1] + blk.by;
      size_t rx = bx * ecbw, ry = by * ecbh;
      size_t rw = std::min(ecbw, band.w - rx), rh = std::min(ecbh, band.h - ry);
      size_t seg_len = size_t(blk.len_cleanup) + blk.len_refinement;
      if (p + seg_len > body.size()) fail(kCodestream, "packet body truncated", p);
      const uint8_t* seg = body.data() + p;
      p += seg_len;
      if (res > max_res) continue;
      std::vector<uint32_t> out(rw * rh, 0);
      ht_decode_block(seg, seg_len, out.data(), rw, rh, rw, blk.missing_msbs, blk.num_passes, blk.len_cleanup,
                      blk.len_refinement, stripe_causal);
      if (irrev) {
        // T.800 E.1.1.2 with the decoder's mid-point reconstruction
        const float scale = irrev_scale(cs.quant, band.res, band.band, levels, max_res, cs.siz.comps[0].depth);
        for (size_t y = 0; y < rh; ++y)
          for (size_t x = 0; x < rw; ++x) {
            uint32_t v = out[y * rw + x];
            float c = (float)(v & 0x7FFFFFFF) * scale;
            fplane[(band.y0 + ry + y) * width + band.x0 + rx + x] = (v & 0x80000000u) ? -c : c;
          }
        continue;
      }
      uint8_t k_max = cs.quant.k_max(band.res, band.band);
      uint32_t sh = 31 - k_max;
      std::vector<int32_t>& plane = planes[seq[i].comp];
      for (size_t y = 0; y < rh; ++y)
        for (size_t x = 0; x < rw; ++x) {
          uint32_t v = out[y * rw + x];
          int32_t mag = (int32_t)((v & 0x7FFFFFFF) >> sh);
          plane[(band.y0 + ry + y) * width + band.x0 + rx + x] = (v & 0x80000000u) ? -mag : mag;
        }
    }
    cursor = p;
  }
  size_t out_w, out_h;
  level_dims(width, height, levels - max_res, &out_w, &out_h);
  std::vector<std::vector<int32_t>> comps_out;
  if (irrev) {
    // inverse 9/7 over the kept levels, then to the nearest integer (ties
    // to even) inside the level-shifted sample range
    std::vector<float> region(out_w * out_h);
    for (size_t y = 0; y < out_h; ++y)
      std::copy(fplane.begin() + y * width, fplane.begin() + y * width + out_w, region.begin() + y * out_w);
    dwt97_inverse(region.data(), out_w, out_h, max_res);
    const unsigned depth = cs.siz.comps[0].depth;
    const float lo = -(float)(1 << (depth - 1)), hi = (float)((1 << (depth - 1)) - 1);
    std::vector<int32_t> ints(out_w * out_h);
    for (size_t i = 0; i < ints.size(); ++i)
      ints[i] = (int32_t)std::fmin(std::fmax(std::nearbyint(region[i]), lo), hi);
    comps_out.push_back(std::move(ints));
    planes.clear();
  }
  for (auto& plane : planes) {
    std::vector<int32_t> region(out_w * out_h);
    for (size_t y = 0; y < out_h; ++y)
      std::copy(plane.begin() + y * width, plane.begin() + y * width + out_w, region.begin() + y * out_w);
    dwt53_inverse(region.data(), out_w, out_h, max_res);
    comps_out.push_back(std::move(region));
  }
  if (cs.cod.mct == 1 && ncomps >= 3) {
    for (size_t i = 0; i < out_w * out_h; ++i) {
      int32_t y = comps_out[0][i], cb = comps_out[1][i], cr = comps_out[2][i];
      int32_t g = y - ((cb + cr) >> 2);
      comps_out[0][i] = cr + g;
      comps_out[1][i] = g;
      comps_out[2][i] = cb + g;
    }
  }
  for (size_t ci = 0; ci < comps_out.size(); ++ci) {
    auto c = cs.siz.comps[std::min(ci, cs.siz.comps.size() - 1)];
    if (!c.is_signed) {
      int32_t half = 1 << (c.depth - 1);
      int64_t mx = ((int64_t)1 << c.depth) - 1;
      for (auto& v : comps_out[ci]) v = (int32_t)std::min<int64_t>(std::max<int64_t>((int64_t)v + half, 0), mx);
    }
  }
  *ow = out_w;
  *oh = out_h;
  return comps_out;
}

// ---------------------------------------------------------------------------
// container.rs
// ---------------------------------------------------------------------------
struct PlaneEntry { uint64_t offset; uint32_t len; std::vector<uint32_t> prefix; };
struct ChunkHeader {
  std::vector<uint32_t> dims;
  unsigned xy_levels = 0;
  bool has_index = false;
  std::vector<PlaneEntry> planes;
  size_t num_planes() const {
    size_t n = 1;
    for (size_t i = 0; i + 2 < dims.size(); ++i) n *= dims[i];
    return n;
  }
  static size_t fixed_len(size_t ndim) { return 8 + 4 * ndim; }
  static size_t entry_len(unsigned lv) { return 12 + 4 * (size_t(lv) + 1); }
  size_t header_len() const { return fixed_len(dims.size()) + (has_index ? planes.size() * entry_len(xy_levels) : 0); }
};

PlaneEntry plane_entry(const Codestream& cs, uint64_t offset) {
  std::vector<PacketSpan> spans = packet_index(cs);
  if (cs.total_len > 0xFFFFFFFFull) fail(kUnsupported, "codestream exceeds the 4 GiB plane index bound");
  PlaneEntry e{offset, (uint32_t)cs.total_len, {}};
  for (unsigned r = 0; r <= cs.cod.decomps; ++r) {
    uint64_t m = 0;
    for (auto& s : spans)
      if (s.res <= r) m = std::max<uint64_t>(m, s.offset + s.len);
    if (m > 0xFFFFFFFFull) fail(kUnsupported, "codestream exceeds the 4 GiB plane index bound");
    e.prefix.push_back((uint32_t)m);
  }
  return e;
}

std::vector<uint8_t> header_bytes(const ChunkHeader& h) {
  std::vector<uint8_t> out{'n', 'd', 'h', 't', 1, uint8_t(h.has_index ? 1 : 0), uint8_t(h.xy_levels), uint8_t(h.dims.size())};
  auto le32 = [&](uint32_t v) { for (int i = 0; i < 4; ++i) out.push_back(uint8_t(v >> (8 * i))); };
  for (uint32_t d : h.dims) le32(d);
  if (h.has_index)
    for (auto& p : h.planes) {
      for (int i = 0; i < 8; ++i) out.push_back(uint8_t(p.offset >> (8 * i)));
      le32(p.len);
      for (uint32_t l : p.prefix) le32(l);
    }
  return out;
}

ChunkHeader parse_header(const uint8_t* b, size_t n) {
  if (n < 8) fail(kCodestream, "chunk header prefix too short", 0);
  if (std::memcmp(b, "ndht", 4) != 0) fail(kCodestream, "not an htj2k chunk (bad magic)", 0);
  if (b[4] != 1) fail(kCodestream, "unsupported htj2k chunk container version", 4);
  uint8_t flags = b[5];
  if (flags & ~1) fail(kCodestream, "unknown htj2k chunk container flags", 5);
  unsigned lv = b[6];
  if (lv > 32) fail(kCodestream, "xy_levels exceeds the 32-level J2K bound (T.800 SPcod)", 6);
  size_t ndim = b[7];
  if (ndim < 2 || ndim > 32) fail(kCodestream, "chunk ndim outside 2..=32", 7);
  size_t fixed = ChunkHeader::fixed_len(ndim);
  if (n < fixed) fail(kCodestream, "truncated chunk header", n);
  ChunkHeader h;
  h.xy_levels = lv;
  for (size_t i = 0; i < ndim; ++i) {
    size_t at = 8 + 4 * i;
    uint32_t d = uint32_t(b[at]) | uint32_t(b[at + 1]) << 8 | uint32_t(b[at + 2]) << 16 | uint32_t(b[at + 3]) << 24;
    if (d == 0) fail(kCodestream, "zero chunk dimension", at);
    h.dims.push_back(d);
  }
  uint64_t np = 1;
  for (size_t i = 0; i + 2 < ndim; ++i) {
    np *= h.dims[i];
    if (np > (1ull << 24)) fail(kCodestream, "chunk declares more planes than supported", 8);
  }
  if (flags & 1) {
    h.has_index = true;
    size_t entry = ChunkHeader::entry_len(lv), need = fixed + np * entry;
    if (n < need) fail(kCodestream, "truncated coefficient-plane index", n);
    uint64_t min_offset = need;
    for (size_t i = 0; i < np; ++i) {
      const uint8_t* e = b + fixed + i * entry;
      PlaneEntry pe;
      pe.offset = 0;
      for (int k = 0; k < 8; ++k) pe.offset |= uint64_t(e[k]) << (8 * k);
      pe.len = uint32_t(e[8]) | uint32_t(e[9]) << 8 | uint32_t(e[10]) << 16 | uint32_t(e[11]) << 24;
      for (size_t k = 12; k < entry; k += 4)
        pe.prefix.push_back(uint32_t(e[k]) | uint32_t(e[k + 1]) << 8 | uint32_t(e[k + 2]) << 16 | uint32_t(e[k + 3]) << 24);
      if (pe.offset < min_offset) fail(kCodestream, "plane offsets overlap or precede the index", fixed + i * entry);
      for (size_t k = 1; k < pe.prefix.size(); ++k)
        if (pe.prefix[k - 1] > pe.prefix[k]) fail(kCodestream, "plane prefix lengths are inconsistent", fixed + i * entry);
      if (!pe.prefix.empty() && pe.prefix.back() > pe.len) fail(kCodestream, "plane prefix lengths are inconsistent", fixed + i * entry);
      min_offset = pe.offset + pe.len;
      h.planes.push_back(pe);
    }
  }
  return h;
}

int32_t read_sample(const uint8_t* p, int dtype) {
  switch (dtype) {
    case kU8: return p[0];
    case kI8: return (int8_t)p[0];
    case kU16: return uint16_t(p[0] | p[1] << 8);
    case kI16: return int16_t(p[0] | p[1] << 8);
    default: {
      uint32_t v = uint32_t(p[0]) | uint32_t(p[1]) << 8 | uint32_t(p[2]) << 16 | uint32_t(p[3]) << 24;
      return (int32_t)v;
    }
  }
}

}  // namespace

// writer.rs:55-323
std::vector<uint8_t> write_codestream(const std::vector<int32_t>& plane, size_t width, size_t height,
                                      unsigned bit_depth, bool is_signed, unsigned levels, int progression,
                                      bool emit_tlm_plt, const Quant& quant, uint8_t wavelet);

std::vector<uint8_t> encode_image_with_depth(const int32_t* samples, size_t width, size_t height, unsigned bit_depth,
                                             bool is_signed, unsigned levels, int progression, bool emit_tlm_plt) {
  if (width == 0 || height == 0) fail(kInvalidArgument, "empty image");
  if (bit_depth < 1 || bit_depth > 38) fail(kInvalidArgument, "declared bit depth must be in 1..=38");
  int64_t lo = is_signed ? -((int64_t)1 << (bit_depth - 1)) : 0;
  int64_t hi = is_signed ? ((int64_t)1 << (bit_depth - 1)) - 1 : ((int64_t)1 << bit_depth) - 1;
  for (size_t i = 0; i < width * height; ++i)
    if (samples[i] < lo || samples[i] > hi) fail(kInvalidArgument, "sample outside the declared dtype range");
  Quant quant = Quant::reversible(levels, bit_depth);
  unsigned kmm = 0;
  for (unsigned r = 0; r <= levels; ++r)
    for (unsigned b = r > 0 ? 1 : 0; b <= (r > 0 ? 3u : 0u); ++b) kmm = std::max<unsigned>(kmm, quant.k_max(r, b));
  if (kmm > 30) fail(kUnsupported, "bit depth exceeds the 32-bit HT datapath (K_max > 30)");

  std::vector<int32_t> plane(samples, samples + width * height);
  if (!is_signed) {
    int32_t half = 1 << (bit_depth - 1);
    for (auto& s : plane) s -= half;
  }
  dwt53_forward(plane.data(), width, height, levels);
  return write_codestream(plane, width, height, bit_depth, is_signed, levels, progression, emit_tlm_plt, quant, 1);
}

// The codestream of one transformed (and, on the irreversible path,
// quantised) plane: packets of every resolution, PLT/TLM, SIZ/COD/QCD/CAP/COM
// (the tail of htj2k.rs encode_image / writer.rs); wavelet 1 = 5/3, 0 = 9/7.
std::vector<uint8_t> write_codestream(const std::vector<int32_t>& plane, size_t width, size_t height,
                                      unsigned bit_depth, bool is_signed, unsigned levels, int progression,
                                      bool emit_tlm_plt, const Quant& quant, uint8_t wavelet) {
  const size_t cbw = 64, cbh = 64;
  std::vector<std::vector<BandBlocks>> per_res;
  for (unsigned r = 0; r <= levels; ++r) {
    std::vector<Band> bands = bands_of_resolution(width, height, levels, r);
    size_t ecbw, ecbh;
    effective_cb(cbw, cbh, 15, 15, r, &ecbw, &ecbh);
    std::vector<BandBlocks> bb;
    for (auto& b : bands) bb.push_back(code_band(plane, width, b, ecbw, ecbh, quant));
    per_res.push_back(std::move(bb));
  }
  std::vector<uint8_t> body;
  std::vector<uint32_t> packet_lengths;
  for (unsigned r = 0; r <= levels; ++r) {
    std::vector<uint8_t> bytes = write_packet(per_res[r]);
    packet_lengths.push_back((uint32_t)bytes.size());
    body.insert(body.end(), bytes.begin(), bytes.end());
  }
  std::vector<uint8_t> plt;
  if (emit_tlm_plt) {
    unsigned zplt = 0;
    size_t start = 0;
    for (;;) {
      size_t end = start, payload_len = 1;
      while (end < packet_lengths.size()) {
        size_t bits = 32 - __builtin_clz(packet_lengths[end] | 1);
        size_t vlen = (bits + 6) / 7;
        if (payload_len + vlen > 65000) break;
        payload_len += vlen;
        end += 1;
      }
      push_segment(plt, PLT, plt_payload((uint8_t)zplt, packet_lengths.data() + start, end - start));
      start = end;
      if (start >= packet_lengths.size()) break;
      if (zplt == 255) fail(kInvalidArgument, "more than 256 PLT segments required");
      zplt += 1;
    }
  }
  size_t psot = 12 + plt.size() + 2 + body.size();
  std::vector<uint8_t> tile_part;
  {
    std::vector<uint8_t> sot;
    be16(sot, 0);
    be32(sot, (uint32_t)psot);
    sot.push_back(0);
    sot.push_back(1);
    push_segment(tile_part, SOT, sot);
  }
  tile_part.insert(tile_part.end(), plt.begin(), plt.end());
  be16(tile_part, SOD);
  tile_part.insert(tile_part.end(), body.begin(), body.end());

  std::vector<uint8_t> out;
  be16(out, SOC);
  {
    std::vector<uint8_t> p;
    be16(p, 0x4000);
    for (uint32_t v : {(uint32_t)width, (uint32_t)height, 0u, 0u, (uint32_t)width, (uint32_t)height, 0u, 0u}) be32(p, v);
    be16(p, 1);
    p.push_back(uint8_t((bit_depth - 1) | (is_signed ? 0x80 : 0)));
    p.push_back(1);
    p.push_back(1);
    push_segment(out, SIZ, p);
  }
  {
    std::vector<uint8_t> p{0, (uint8_t)progression};
    be16(p, 1);
    p.push_back(0);
    p.push_back((uint8_t)levels);
    p.push_back(4);  // log2(64) - 2
    p.push_back(4);
    p.push_back(0x40);
    p.push_back(wavelet);
    push_segment(out, COD, p);
  }
  push_segment(out, QCD, quant.to_bytes());
  {
    uint32_t magb = quant.magb();
    uint32_t bp = magb <= 8 ? 0 : (magb < 28 ? magb - 8 : 13 + (magb >> 2));
    std::vector<uint8_t> p;
    be32(p, 0x00020000u);
    be16(p, (wavelet == 0 ? 0x20 : 0) | (bp & 0x1F));  // Ccap15 (markers.rs:325-339)
    push_segment(out, CAP, p);
  }
  {
    std::vector<uint8_t> p;
    be16(p, 1);
    const char* text = "nd-image-codecs ndic-codestream";
    p.insert(p.end(), text, text + std::strlen(text));
    push_segment(out, COM, p);
  }
  if (emit_tlm_plt) {
    std::vector<uint8_t> p{0, 0x40};
    be32(p, (uint32_t)tile_part.size());
    push_segment(out, TLM, p);
  }
  out.insert(out.end(), tile_part.begin(), tile_part.end());
  be16(out, EOC);
  return out;
}

// A lossy codestream for the tests of the decode-only 9/7 path (the
// reference refuses lossy encode, htj2k.rs:293-297): level shift, forward
// 9/7 (dwt/mod.rs:205-240), a dead-zone quantiser q = sign(c) floor(|c| /
// step) with every band at exponent eps and mantissa mu (QCD style 2,
// guard bits `guard`) and step = irrev_scale x 2^(31 - K_max), then the
// reversible path's block coder and writer.  Fails when a band's indices do
// not fit K_max = guard + eps - 1 bits.
std::vector<uint8_t> encode_image_irreversible(const int32_t* samples, size_t width, size_t height,
                                               unsigned bit_depth, bool is_signed, unsigned levels, unsigned eps,
                                               unsigned mu, unsigned guard) {
  if (width == 0 || height == 0) fail(kInvalidArgument, "empty image");
  if (bit_depth < 1 || bit_depth > 30) fail(kInvalidArgument, "bit depth outside 1..=30");
  if (eps > 31 || mu > 0x7FF || guard > 7) fail(kInvalidArgument, "quantiser parameters out of range");
  Quant q;
  q.guard_bits = (uint8_t)guard;
  q.style = 2;
  q.values.assign(1 + 3 * levels, (uint16_t)(eps << 11 | mu));
  int32_t half = is_signed ? 0 : 1 << (bit_depth - 1);
  std::vector<float> f(width * height);
  for (size_t i = 0; i < f.size(); ++i) f[i] = (float)(samples[i] - half);
  dwt97_forward(f.data(), width, height, levels);
  std::vector<int32_t> plane(width * height, 0);
  for (unsigned r = 0; r <= levels; ++r) {
    for (const Band& b : bands_of_resolution(width, height, levels, r)) {
      unsigned km = q.k_max(b.res, b.band);
      if (km > 30) fail(kInvalidArgument, "K_max above 30");
      double step = std::ldexp((double)irrev_scale(q, b.res, b.band, levels, levels, bit_depth), 31 - (int)km);
      for (size_t y = 0; y < b.h; ++y)
        for (size_t x = 0; x < b.w; ++x) {
          size_t at = (b.y0 + y) * width + b.x0 + x;
          double m = std::floor(std::fabs((double)f[at]) / step);
          if (m >= std::ldexp(1.0, (int)km)) fail(kInvalidArgument, "quantisation index exceeds K_max bits");
          plane[at] = f[at] < 0 ? -(int32_t)m : (int32_t)m;
        }
    }
  }
  return write_codestream(plane, width, height, bit_depth, is_signed, levels, 2, true, q, 0);
}

Decoded decode_codestream(const uint8_t* data, size_t len, int max_res, bool prefix, bool allow_irreversible) {
  Codestream cs = parse_codestream(data, len, !prefix);
  Decoded d;
  std::vector<std::vector<int32_t>> comps = decode_to_resolution(
      cs, max_res < 0 ? cs.cod.decomps : (unsigned)max_res, &d.width, &d.height, allow_irreversible);
  if (comps.size() != 1) fail(kUnsupported, "oracle decodes single-component codestreams");
  d.samples = std::move(comps[0]);
  return d;
}

// htj2k.rs:286-346
std::vector<uint8_t> htj2k_encode_chunk(const uint8_t* bytes, size_t len, const std::vector<size_t>& shape, int dtype,
                                        const Htj2kConfig& cfg) {
  if (cfg.xy_levels > 32) fail(kInvalidArgument, "htj2k: xy_levels exceeds the 32-level J2K bound (T.800 SPcod)");
  if (cfg.progression < 0 || cfg.progression > 4) fail(kUnsupported, "htj2k: unknown progression order");
  if (!cfg.reversible) fail(kUnsupported, "htj2k: lossy (reversible: false) coding lands with a later phase");
  if (shape.size() < 2) fail(kInvalidArgument, "htj2k needs a chunk with trailing 2D (y, x) planes");
  if (shape.size() > 32) fail(kInvalidArgument, "htj2k chunks carry at most 32 dimensions");
  for (size_t d : shape)
    if (d == 0) fail(kInvalidArgument, "htj2k needs non-zero chunk extents");
  if (!(dtype == kU8 || dtype == kI8 || dtype == kU16 || dtype == kI16 || dtype == kU32 || dtype == kI32))
    fail(kUnsupported, "htj2k does not support this dtype");
  size_t height = shape[shape.size() - 2], width = shape[shape.size() - 1], np = 1, elems = 1;
  for (size_t i = 0; i + 2 < shape.size(); ++i) np *= shape[i];
  for (size_t d : shape) elems *= d;
  size_t isz = dtype_size(dtype);
  if (len != elems * isz) fail(kInvalidArgument, "chunk byte length does not match its shape and dtype");
  bool is_signed = dtype_signed(dtype);
  size_t plane_bytes = width * height * isz;
  std::vector<std::vector<uint8_t>> streams;
  std::vector<int32_t> samples(width * height);
  for (size_t p = 0; p < np; ++p) {
    const uint8_t* pb = bytes + p * plane_bytes;
    for (size_t i = 0; i < width * height; ++i) {
      if (dtype == kU32) {
        uint32_t v = uint32_t(pb[4 * i]) | uint32_t(pb[4 * i + 1]) << 8 | uint32_t(pb[4 * i + 2]) << 16 | uint32_t(pb[4 * i + 3]) << 24;
        if (v > 0x7FFFFFFFu) fail(kInvalidArgument, "uint32 sample exceeds the 32-bit HT datapath; decorrelate or reduce the dynamic range");
      }
      samples[i] = read_sample(pb + i * isz, dtype);
    }
    int64_t lo = 0, hi = 0;
    for (int32_t v : samples) { lo = std::min<int64_t>(lo, v); hi = std::max<int64_t>(hi, v); }
    unsigned depth = 32;
    for (unsigned b = 1; b <= 32; ++b) {
      bool fits = is_signed ? (lo >= -((int64_t)1 << (b - 1)) && hi < ((int64_t)1 << (b - 1))) : hi < ((int64_t)1 << b);
      if (fits) { depth = b; break; }
    }
    try {
      streams.push_back(encode_image_with_depth(samples.data(), width, height, depth, is_signed, cfg.xy_levels,
                                                cfg.progression, cfg.index));
    } catch (Error& e) {
      if (e.kind == kUnsupported) e.message = "htj2k plane " + std::to_string(p) + ": " + e.message;
      throw;
    }
  }
  ChunkHeader h;
  for (size_t d : shape) {
    if (d > 0xFFFFFFFFull) fail(kInvalidArgument, "chunk extent exceeds u32");
    h.dims.push_back((uint32_t)d);
  }
  h.xy_levels = cfg.xy_levels;
  h.has_index = cfg.index;
  if (cfg.index) {
    uint64_t offset = ChunkHeader::fixed_len(shape.size()) + np * ChunkHeader::entry_len(cfg.xy_levels);
    for (auto& s : streams) {
      Codestream cs = parse_codestream(s.data(), s.size(), true);
      PlaneEntry e = plane_entry(cs, offset);
      offset += e.len;
      h.planes.push_back(e);
    }
  }
  std::vector<uint8_t> out = header_bytes(h);
  for (auto& s : streams) out.insert(out.end(), s.begin(), s.end());
  return out;
}

// htj2k.rs:354-441
std::vector<uint8_t> htj2k_decode_chunk(const uint8_t* chunk, size_t len, const std::vector<size_t>& shape, int dtype) {
  ChunkHeader h = parse_header(chunk, len);
  size_t np = h.num_planes();
  std::vector<std::pair<size_t, size_t>> ranges;
  if (h.has_index) {
    for (auto& p : h.planes) {
      size_t start = (size_t)p.offset;
      if (start + p.len > len) fail(kCodestream, "plane byte range exceeds the chunk", start);
      ranges.push_back({start, p.len});
    }
  } else {
    size_t off = h.header_len();
    for (size_t i = 0; i < np; ++i) {
      if (off > len) fail(kCodestream, "chunk ends before all planes", off);
      Codestream cs = parse_codestream(chunk + off, len - off, true);
      ranges.push_back({off, cs.total_len});
      off += cs.total_len;
    }
  }
  bool same = h.dims.size() == shape.size();
  for (size_t i = 0; same && i < shape.size(); ++i) same = h.dims[i] == shape[i];
  if (!same) fail(kCodestream, "chunk header declares a different shape than decode expects", 0);
  std::vector<uint8_t> out;
  size_t isz = dtype_size(dtype);
  for (auto& r : ranges) {
    Codestream cs = parse_codestream(chunk + r.first, r.second, true);
    size_t w, hh;
    std::vector<std::vector<int32_t>> comps = decode_to_resolution(cs, cs.cod.decomps, &w, &hh);
    if (w != h.dims.back() || hh != h.dims[h.dims.size() - 2] || comps.size() != 1)
      fail(kCodestream, "plane decodes to a different geometry than the chunk header declares", 0);
    const std::vector<int32_t>& s = comps[0];
    int64_t lo = 0, hi = 0;
    switch (dtype) {
      case kU8: lo = 0; hi = 0xFF; break;
      case kI8: lo = -128; hi = 127; break;
      case kU16: lo = 0; hi = 0xFFFF; break;
      case kI16: lo = -32768; hi = 32767; break;
      case kU32: lo = 0; hi = 0xFFFFFFFFll; break;
      case kI32: lo = INT64_MIN; hi = INT64_MAX; break;
      default: fail(kUnsupported, "htj2k does not support this dtype");
    }
    if (!s.empty()) {
      int64_t mn = *std::min_element(s.begin(), s.end()), mx = *std::max_element(s.begin(), s.end());
      if (mn < lo || mx > hi) fail(kCodestream, "decoded sample range does not fit the dtype (corrupt or mismatched chunk)", 0);
    }
    for (int32_t v : s) {
      uint32_t u = (uint32_t)v;
      for (size_t k = 0; k < isz; ++k) out.push_back(uint8_t(u >> (8 * k)));
    }
  }
  return out;
}

}  // namespace orc
