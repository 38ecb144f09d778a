// nd_zfp / zfp chunk codec: configuration and stream-header handling on the
// host, block coding on the device (zfp.hip).
//
// Chunk semantics follow crates/ndic-zfp/src/chunk.rs:246-417 (effective
// shape, narrow-integer promotion, length checks) and the stream wrapper
// crates/ndic-zfp/src/lib.rs:88-384 (mode validation, expert parameters,
// full header, exact header match and fixed-rate length on decode, maximum
// stream size), with the brick reader of lib.rs:386-599.
#include <cmath>
#include <cstring>
#include <sstream>

#include "engine.hpp"
#include "zfp.hpp"

namespace ndic {

namespace {

std::string shape_str(const std::vector<size_t>& s) {
  std::ostringstream o;
  o << "[";
  for (size_t i = 0; i < s.size(); ++i) o << (i ? ", " : "") << s[i];
  o << "]";
  return o.str();
}

std::string num_str(double v) {
  std::ostringstream o;
  o << v;
  return o.str();
}

const char* mode_name(int m) {
  switch (m) {
    case NDIC_ZFP_FIXED_RATE: return "fixed_rate";
    case NDIC_ZFP_FIXED_ACCURACY: return "fixed_accuracy";
    case NDIC_ZFP_FIXED_PRECISION: return "fixed_precision";
    case NDIC_ZFP_REVERSIBLE: return "reversible";
  }
  return "?";
}

// NdZfpConfig::zfp_mode + ZfpMode::validate (chunk.rs:162-206, lib.rs:94-109)
void validate_config(const ndic_zfp_config& c) {
  if (c.dims != 0 && (c.dims < 1 || c.dims > 4))
    fail(NDIC_INVALID_ARGUMENT, "zfp: dims must be 1..=4, got " + std::to_string(c.dims));
  switch (c.mode) {
    case NDIC_ZFP_FIXED_RATE:
      if (!(std::isfinite(c.rate) && c.rate > 0.0))
        fail(NDIC_INVALID_ARGUMENT, "nd_zfp: fixed-rate mode needs a positive finite rate, got " + num_str(c.rate));
      break;
    case NDIC_ZFP_FIXED_ACCURACY:
      if (!(std::isfinite(c.tolerance) && c.tolerance >= 0.0))
        fail(NDIC_INVALID_ARGUMENT,
             "nd_zfp: fixed-accuracy mode needs a non-negative finite tolerance, got " + num_str(c.tolerance));
      break;
    case NDIC_ZFP_FIXED_PRECISION:
      if (c.precision < 1 || c.precision > 64)
        fail(NDIC_INVALID_ARGUMENT,
             "nd_zfp: fixed-precision mode needs 1..=64 bit planes, got " + std::to_string(c.precision));
      break;
    case NDIC_ZFP_REVERSIBLE: break;
    default: fail(NDIC_UNSUPPORTED, "zfp: unknown mode code " + std::to_string(c.mode));
  }
}

// chunk.rs:246-278
std::vector<size_t> effective_shape(const std::vector<size_t>& shape, int32_t dims) {
  if (shape.empty()) fail(NDIC_INVALID_ARGUMENT, "zfp: chunk shape has no dimensions");
  for (size_t d : shape)
    if (d == 0) fail(NDIC_INVALID_ARGUMENT, "zfp: chunk shape " + shape_str(shape) + " has a zero extent");
  if (dims == 0) {
    if (shape.size() > 4)
      fail(NDIC_INVALID_ARGUMENT, "zfp: chunk shape " + shape_str(shape) + " has " + std::to_string(shape.size()) +
                                      " dimensions but ZFP fields are 1-4 dimensional; collapse singleton "
                                      "dimensions with a reshape codec upstream");
    return shape;
  }
  std::vector<size_t> sq;
  for (size_t d : shape)
    if (d > 1) sq.push_back(d);
  if (sq.size() > (size_t)dims)
    fail(NDIC_INVALID_ARGUMENT, "zfp: chunk shape " + shape_str(shape) + " has " + std::to_string(sq.size()) +
                                    " non-singleton dimensions but the configuration declares dims=" +
                                    std::to_string(dims) + "; reduce chunking or raise dims");
  std::vector<size_t> eff((size_t)dims - sq.size(), 1);
  eff.insert(eff.end(), sq.begin(), sq.end());
  return eff;
}

uint32_t scalar_of(int dtype) {
  switch (dtype) {
    case NDIC_U8: case NDIC_I8: case NDIC_U16: case NDIC_I16: case NDIC_I32: return kZfpInt32;
    case NDIC_I64: return kZfpInt64;
    case NDIC_F32: return kZfpFloat;
    case NDIC_F64: return kZfpDouble;
  }
  fail(NDIC_UNSUPPORTED, std::string("zfp has no path for dtype ") + dtype_name(dtype));
}

const char* scalar_name(uint32_t s) {
  switch (s) {
    case kZfpInt32: return "int32";
    case kZfpInt64: return "int64";
    case kZfpFloat: return "float";
    case kZfpDouble: return "double";
  }
  return "none";
}

// Everything the host knows about one chunk's stream.
struct Plan {
  std::vector<size_t> eff;
  uint32_t dims = 0, scalar = 0;
  int32_t dtype = 0;
  int32_t mode = 0;
  ZfpParams prm{};
  uint32_t n[4] = {1, 1, 1, 1}, nb[4] = {1, 1, 1, 1};
  uint64_t nblocks = 0, elems = 0;
  uint64_t meta = 0, mode_word = 0;
  uint32_t mode_bits = 12;
  uint64_t header_bits = 0;
  bool fixed = false;
  uint64_t max_bytes = 0;     // zfp_stream_maximum_size
  uint64_t fixed_bytes = 0;   // fixed-rate: the exact stream length
  uint32_t slot_bits = 0;     // bound on one block's bits
};

// ZfpConfig::{fixed_rate, fixed_accuracy, fixed_precision, reversible}
ZfpParams params_of(const ndic_zfp_config& c, uint32_t scalar, uint32_t dims) {
  ZfpParams p{1, kZfpMaxBits, 64, kZfpMinExp};
  switch (c.mode) {
    case NDIC_ZFP_FIXED_RATE: {
      uint32_t n = 1u << (2 * dims);
      double b = std::floor(n * c.rate + 0.5);
      uint32_t bits = b >= 4294967295.0 ? 0xffffffffu : (uint32_t)b;
      if (scalar == kZfpFloat && bits < 9) bits = 9;
      if (scalar == kZfpDouble && bits < 12) bits = 12;
      // the header's long mode form holds 1..32768 bits per block
      if (bits < 1 || bits > 0x8000)
        fail(NDIC_INVALID_ARGUMENT, "nd_zfp: fixed-rate mode needs 1..=32768 bits per block");
      p.minbits = p.maxbits = bits;
      break;
    }
    case NDIC_ZFP_FIXED_PRECISION: p.maxprec = c.precision; break;
    case NDIC_ZFP_FIXED_ACCURACY: {
      int emin = kZfpMinExp;
      if (c.tolerance > 0) {
        int e;
        (void)std::frexp(c.tolerance, &e);
        emin = e - 1;
      }
      p.minexp = emin;
      break;
    }
    default: p.minexp = kZfpMinExp - 1; break;
  }
  return p;
}

// zfp_stream_mode: short 12-bit forms for the common modes, else the 64-bit
// expert form (the all-default parameter set is always expert)
uint64_t mode_word(int mode, const ZfpParams& p, uint32_t* nbits) {
  *nbits = 12;
  bool defaults = p.minbits == 1 && p.maxbits == kZfpMaxBits && p.maxprec == 64 && p.minexp == kZfpMinExp;
  if (!defaults) {
    switch (mode) {
      case NDIC_ZFP_FIXED_RATE:
        if (p.maxbits <= 2048) return p.maxbits - 1;
        break;
      case NDIC_ZFP_FIXED_PRECISION:
        if (p.maxprec <= 128) return (p.maxprec - 1) + 2048;
        break;
      case NDIC_ZFP_FIXED_ACCURACY:
        if (p.minexp <= 843) return (uint64_t)(p.minexp - kZfpMinExp) + (2048 + 128 + 1);
        break;
      default: return 2048 + 128;
    }
  }
  *nbits = 64;
  auto clampu = [](uint64_t v, uint64_t hi) { return std::max<uint64_t>(1, std::min<uint64_t>(v, hi)) - 1; };
  uint64_t minexp = (uint64_t)std::max<int64_t>(0, std::min<int64_t>((int64_t)p.minexp + 16495, 0x7fff));
  uint64_t w = minexp;
  w = (w << 7) + clampu(p.maxprec, 0x80);
  w = (w << 15) + clampu(p.maxbits, 0x8000);
  w = (w << 15) + clampu(p.minbits, 0x8000);
  w = (w << 12) + 0xfff;
  return w;
}

Plan make_plan(const std::vector<size_t>& shape, int dtype, const ndic_zfp_config& cfg) {
  validate_config(cfg);
  Plan pl;
  pl.eff = effective_shape(shape, cfg.dims);
  pl.dims = (uint32_t)pl.eff.size();
  pl.scalar = scalar_of(dtype);
  pl.dtype = dtype;
  pl.mode = cfg.mode;
  pl.prm = params_of(cfg, pl.scalar, pl.dims);
  pl.elems = 1;
  for (size_t a = 0; a < pl.dims; ++a) {
    size_t e = pl.eff[pl.dims - 1 - a];
    if (e > 0xffffffffull) fail(NDIC_INVALID_ARGUMENT, "nd_zfp: shape " + shape_str(pl.eff) + " exceeds the ZFP header metadata bounds");
    pl.n[a] = (uint32_t)e;
    pl.nb[a] = (uint32_t)((e + 3) / 4);
  }
  pl.nblocks = 1;
  for (size_t a = 0; a < 4; ++a) pl.nblocks *= pl.nb[a];
  for (size_t d : pl.eff) pl.elems *= d;
  // field metadata: extents (48/24/16/12 bits per axis), dims - 1, type - 1
  uint32_t fb = pl.dims == 1 ? 48 : pl.dims == 2 ? 24 : pl.dims == 3 ? 16 : 12;
  uint64_t meta = 0;
  for (int a = (int)pl.dims - 1; a >= 0; --a) {
    uint64_t v = (uint64_t)pl.n[a] - 1;
    if (fb < 48 && (v >> fb))
      fail(NDIC_INVALID_ARGUMENT, "nd_zfp: shape " + shape_str(pl.eff) + " exceeds the ZFP header metadata bounds");
    meta = (meta << fb) + v;
  }
  meta = (meta << 2) + (pl.dims - 1);
  meta = (meta << 2) + (pl.scalar - 1);
  pl.meta = meta;
  pl.mode_word = mode_word(cfg.mode, pl.prm, &pl.mode_bits);
  pl.header_bits = 32 + 52 + pl.mode_bits;
  pl.fixed = cfg.mode == NDIC_ZFP_FIXED_RATE;
  // zfp_stream_maximum_size (148 = the largest header)
  uint32_t values = 1u << (2 * pl.dims);
  uint32_t prec = (pl.scalar == kZfpInt32 || pl.scalar == kZfpFloat) ? 32 : 64;
  uint64_t mb = 1;
  if (pl.scalar == kZfpFloat) mb += 8;
  if (pl.scalar == kZfpDouble) mb += 11;
  if (pl.prm.reversible() && (pl.scalar == kZfpFloat || pl.scalar == kZfpDouble)) mb += prec == 32 ? 5 : 6;
  mb += values - 1 + (uint64_t)values * std::min(pl.prm.maxprec, prec);
  mb = std::min<uint64_t>(mb, pl.prm.maxbits);
  mb = std::max<uint64_t>(mb, pl.prm.minbits);
  pl.max_bytes = ((148 + pl.nblocks * mb + 63) & ~uint64_t(63)) / 8;
  pl.fixed_bytes = pl.fixed ? ((pl.header_bits + pl.nblocks * (uint64_t)pl.prm.maxbits + 63) / 64) * 8 : 0;
  // a block's bits: header + value bits + group tests, capped by maxbits
  uint64_t bound = 2 + 11 + 6 + (uint64_t)values * prec + 2ull * values + prec;
  pl.slot_bits = (uint32_t)std::min<uint64_t>(bound, pl.prm.maxbits);
  return pl;
}

// 32-bit magic, 52-bit metadata, mode word: up to 3 LSB-first words
void header_words(const Plan& pl, uint64_t* w, uint32_t* nwords) {
  uint64_t bits[3] = {0, 0, 0};
  uint64_t pos = 0;
  auto put = [&](uint64_t v, uint32_t n) {
    for (uint32_t i = 0; i < n; ++i, ++pos)
      if ((v >> i) & 1) bits[pos >> 6] |= 1ull << (pos & 63);
  };
  put('z', 8);
  put('f', 8);
  put('p', 8);
  put(5, 8);
  put(pl.meta, 52);
  put(pl.mode_word, pl.mode_bits);
  *nwords = (uint32_t)((pos + 63) / 64);
  for (uint32_t i = 0; i < 3; ++i) w[i] = bits[i];
}

struct HeaderBits {
  const uint8_t* p;
  size_t n;
  uint64_t pos = 0;
  uint64_t get(uint32_t k) {
    uint64_t v = 0;
    for (uint32_t i = 0; i < k; ++i, ++pos) {
      uint64_t byte = pos >> 3;
      uint64_t b = byte < n ? (p[byte] >> (pos & 7)) & 1 : 0;
      v |= b << i;
    }
    return v;
  }
};

// decompress: maximum size, then the exact header, then the fixed-rate length
// (lib.rs:315-384).  head: the first bytes of the stream (zero-filled past len).
void check_stream(const Plan& pl, const uint8_t* head, size_t head_len, uint64_t len) {
  if (len > pl.max_bytes)
    fail(NDIC_CODESTREAM, "nd_zfp: stream is " + std::to_string(len) +
                              " bytes but this configuration compresses to at most " + std::to_string(pl.max_bytes));
  HeaderBits hb{head, (size_t)std::min<uint64_t>(head_len, len)};
  if (hb.get(8) != 'z' || hb.get(8) != 'f' || hb.get(8) != 'p')
    fail(NDIC_CODESTREAM, "nd_zfp: invalid ZFP header: bad magic");
  uint64_t codec = hb.get(8);
  if (codec != 5) fail(NDIC_CODESTREAM, "nd_zfp: invalid ZFP header: unsupported codec version " + std::to_string(codec));
  uint64_t meta = hb.get(52);
  uint32_t type = (uint32_t)(meta & 3) + 1, sdims = (uint32_t)((meta >> 2) & 3) + 1;
  if (type != pl.scalar)
    fail(NDIC_CODESTREAM, std::string("nd_zfp: stream holds ") + scalar_name(type) + " samples, expected " +
                              scalar_name(pl.scalar));
  uint32_t fb = sdims == 1 ? 48 : sdims == 2 ? 24 : sdims == 3 ? 16 : 12;
  uint64_t m = meta >> 4;
  std::vector<uint64_t> got(4, 0), want(4, 0);
  for (uint32_t a = 0; a < sdims; ++a) {
    got[a] = (m & ((1ull << fb) - 1)) + 1;
    m >>= fb;
  }
  for (uint32_t a = 0; a < pl.dims; ++a) want[a] = pl.n[a];
  if (got != want) {
    std::ostringstream o;
    o << "nd_zfp: stream declares dimensions [" << got[0] << ", " << got[1] << ", " << got[2] << ", " << got[3]
      << "], expected [" << want[0] << ", " << want[1] << ", " << want[2] << ", " << want[3] << "] (shape "
      << shape_str(pl.eff) << ")";
    fail(NDIC_CODESTREAM, o.str());
  }
  uint64_t mw = hb.get(12);
  uint32_t mbits = 12;
  if (mw == 0xfff) {
    mw |= hb.get(52) << 12;
    mbits = 64;
  }
  if (mw != pl.mode_word || mbits != pl.mode_bits)
    fail(NDIC_CODESTREAM, "nd_zfp: stream compression parameters do not match the configuration");
  if (pl.fixed && len != pl.fixed_bytes)
    fail(NDIC_CODESTREAM, "nd_zfp: fixed-rate stream is " + std::to_string(len) + " bytes, computed size is " +
                              std::to_string(pl.fixed_bytes));
}

ZfpChunk chunk_of(const Plan& pl) {
  ZfpChunk c{};
  c.dims = pl.dims;
  for (int a = 0; a < 4; ++a) {
    c.n[a] = pl.n[a];
    c.nb[a] = pl.nb[a];
  }
  c.nblocks = pl.nblocks;
  c.scalar = pl.scalar;
  c.dtype = pl.dtype;
  c.minbits = pl.prm.minbits;
  c.maxbits = pl.prm.maxbits;
  c.maxprec = pl.prm.maxprec;
  c.minexp = pl.prm.minexp;
  c.fixed = pl.fixed ? 1 : 0;
  c.slot_words = (pl.slot_bits + 63) / 64 + 1;
  c.header_bits = pl.header_bits;
  return c;
}

uint64_t al256(uint64_t x) { return (x + 255) & ~uint64_t(255); }

}  // namespace

// One chunk of a ZFP batch (device pointers).
struct ZfpItem {
  const uint8_t* d_in = nullptr;
  uint64_t in_len = 0;
  uint8_t* d_out = nullptr;  // decode: samples
  std::vector<size_t> shape;
  int32_t dtype = 0;
  ndic_zfp_config cfg{};
  int status = NDIC_OK;
  std::string message;
  uint64_t out_off = 0, out_len = 0;
  Plan plan;
  const uint8_t* host_head = nullptr;  // decode: first bytes on the host, if there
  // Block index of a variable-size stream (nblocks + 1 uint32 bit offsets):
  // encode writes it at index_off in the index arena when one is given;
  // decode seeks with it when index_len > 0.
  uint8_t* d_index = nullptr;
  uint64_t index_off = 0, index_len = 0;
};

namespace {

void set_fail(ZfpItem& it, const Failure& f) {
  it.status = f.status;
  it.message = f.message;
}

// class key: dims * 8 + scalar
uint32_t class_of(const Plan& p) { return p.dims * 8 + p.scalar; }

}  // namespace

void zfp_encode_items(Context& ctx, std::vector<ZfpItem>& items, uint8_t* d_out, uint64_t capacity, hipStream_t s,
                      uint8_t* d_index, uint64_t index_capacity) {
  ctx.release_kept();  // earlier calls on this context ended synchronised
  for (auto& it : items) {
    if (it.status != NDIC_OK) continue;
    try {
      it.plan = make_plan(it.shape, it.dtype, it.cfg);
      uint64_t want = it.plan.elems * dtype_size(it.dtype);
      if (it.in_len != want)
        fail(NDIC_INVALID_ARGUMENT, "zfp: chunk shape " + shape_str(it.shape) + " (" + dtype_name(it.dtype) +
                                        ") needs " + std::to_string(want) + " bytes, got " +
                                        std::to_string(it.in_len));
    } catch (const Failure& f) {
      set_fail(it, f);
    }
  }
  uint64_t cursor = 0, icursor = 0;
  for (uint32_t cls = 8; cls < 40; ++cls) {
    std::vector<ZfpItem*> group;
    for (auto& it : items)
      if (it.status == NDIC_OK && class_of(it.plan) == cls) group.push_back(&it);
    if (group.empty()) continue;
    std::vector<ZfpChunk> ch;
    uint64_t total = 0, vblocks = 0;
    bool any_index = false;
    uint32_t slot_words = 1;
    for (ZfpItem* it : group) {
      ZfpChunk c = chunk_of(it->plan);
      c.block_begin = total;
      c.src = it->d_in;
      total += c.nblocks;
      if (!c.fixed) vblocks += c.nblocks;
      slot_words = std::max(slot_words, c.slot_words);
      ch.push_back(c);
    }
    // fixed-rate streams have known lengths: place them now
    std::vector<ZfpHeader> hdr;
    for (size_t i = 0; i < group.size(); ++i) {
      if (!ch[i].fixed) continue;
      uint64_t len = group[i]->plan.fixed_bytes;
      if (cursor + len > capacity) {
        group[i]->status = NDIC_INVALID_ARGUMENT;
        group[i]->message = "zfp encode output capacity exceeded";
        continue;  // dst stays null: the kernels skip this chunk
      }
      group[i]->out_off = cursor;
      group[i]->out_len = len;
      ch[i].dst = d_out + cursor;
      cursor += len;
      ZfpHeader h{};
      h.dst = (uint64_t*)ch[i].dst;
      header_words(group[i]->plan, h.w, &h.nwords);
      hdr.push_back(h);
      NDIC_HIP(hipMemsetAsync(ch[i].dst, 0, len, s));
    }
    // scratch for the variable-size modes, indexed by the class-wide block
    uint64_t* offs = nullptr;
    uint64_t* scan = nullptr;
    uint64_t* slots = nullptr;
    uint64_t* totals = nullptr;
    void* scan_tmp = nullptr;
    size_t scan_bytes = 0;
    uint8_t* meta = nullptr;
    uint64_t sz_ch = al256(ch.size() * sizeof(ZfpChunk)), sz_hdr = al256(group.size() * sizeof(ZfpHeader));
    if (vblocks) {
      scan_bytes = zfp_scan_temp_bytes(total);
      uint64_t sz_offs = al256(total * 8), sz_tot = al256(group.size() * 8), sz_tmp = al256(scan_bytes);
      meta = (uint8_t*)ctx.work.ensure(sz_ch + sz_hdr + 2 * sz_offs + sz_tot + sz_tmp);
      offs = (uint64_t*)(meta + sz_ch + sz_hdr);
      scan = (uint64_t*)((uint8_t*)offs + sz_offs);
      totals = (uint64_t*)((uint8_t*)scan + sz_offs);
      scan_tmp = (uint8_t*)totals + sz_tot;
      slots = (uint64_t*)ctx.tmp.ensure(total * slot_words * 8 + 64);
      NDIC_HIP(hipMemsetAsync(offs, 0, total * 8, s));
    } else {
      meta = (uint8_t*)ctx.work.ensure(sz_ch + sz_hdr);
    }
    for (auto& c : ch) {
      c.offs = offs;
      c.slots = slots;
      c.slot_words = slot_words;
    }
    ZfpChunk* d_ch = (ZfpChunk*)meta;
    ZfpHeader* d_hdr = (ZfpHeader*)(meta + sz_ch);
    ctx.up(d_ch, ch.data(), ch.size() * sizeof(ZfpChunk), s);
    if (!hdr.empty()) {
      ctx.up(d_hdr, hdr.data(), hdr.size() * sizeof(ZfpHeader), s);
      launch_zfp_headers(d_hdr, (uint32_t)hdr.size(), s);
    }
    ZfpClock clock(s);
    clock.begin(ZfpClock::kEncode);
    launch_zfp_encode(d_ch, (uint32_t)ch.size(), total, cls / 8, cls % 8, s);
    clock.end();
    if (vblocks) {
      zfp_exclusive_scan(scan_tmp, scan_bytes, offs, scan, total, s);
      launch_zfp_totals(d_ch, (uint32_t)ch.size(), scan, totals, s);
      const uint64_t* p_tot = (const uint64_t*)ctx.down(totals, group.size() * 8, s);
      NDIC_HIP(hipStreamSynchronize(s));
      std::vector<uint64_t> tot(p_tot, p_tot + group.size());
      hdr.clear();
      for (size_t i = 0; i < group.size(); ++i) {
        if (ch[i].fixed || group[i]->status != NDIC_OK) continue;
        uint64_t len = ((ch[i].header_bits + tot[i] + 63) / 64) * 8;
        // its block index: 32-bit offsets, so only streams of fewer than
        // 2^32 block bits have one
        const bool with_index = d_index && tot[i] < (1ull << 32);
        const uint64_t ilen = with_index ? 4 * (ch[i].nblocks + 1) : 0;
        if (cursor + len > capacity || icursor + ilen > index_capacity) {
          group[i]->status = NDIC_INVALID_ARGUMENT;
          group[i]->message = cursor + len > capacity ? "zfp encode output capacity exceeded"
                                                      : "zfp encode block-index capacity exceeded";
          ch[i].offs = nullptr;
          continue;
        }
        group[i]->out_off = cursor;
        group[i]->out_len = len;
        ch[i].dst = d_out + cursor;
        cursor += len;
        ZfpHeader h{};
        h.dst = (uint64_t*)ch[i].dst;
        header_words(group[i]->plan, h.w, &h.nwords);
        hdr.push_back(h);
        NDIC_HIP(hipMemsetAsync(ch[i].dst, 0, len, s));
        if (with_index) {
          group[i]->index_off = icursor;
          group[i]->index_len = ilen;
          ch[i].index = (uint32_t*)(d_index + icursor);
          icursor += (ilen + 7) & ~uint64_t(7);
          any_index = true;
        }
      }
      // fixed-rate chunks of the class take no part in placement
      for (size_t i = 0; i < ch.size(); ++i)
        if (ch[i].fixed) ch[i].offs = nullptr;
      ctx.up(d_ch, ch.data(), ch.size() * sizeof(ZfpChunk), s);
      if (!hdr.empty()) {
        ctx.up(d_hdr, hdr.data(), hdr.size() * sizeof(ZfpHeader), s);
        launch_zfp_headers(d_hdr, (uint32_t)hdr.size(), s);
      }
      clock.begin(ZfpClock::kPlace);
      launch_zfp_place(d_ch, (uint32_t)ch.size(), total, scan, s);
      if (any_index) launch_zfp_index_write(d_ch, (uint32_t)ch.size(), total, scan, s);
      clock.end();
    }
    NDIC_HIP(hipStreamSynchronize(s));
    clock.collect();
  }
}

void zfp_decode_items(Context& ctx, std::vector<ZfpItem>& items, hipStream_t s) {
  ctx.release_kept();  // earlier calls on this context ended synchronised
  // header bytes of every stream, for the host-side checks
  std::vector<uint8_t> heads(items.size() * kZfpHeadBytes, 0);
  {
    std::vector<ZfpHead> hv;
    std::vector<size_t> which;
    for (size_t i = 0; i < items.size(); ++i) {
      if (items[i].host_head) {
        std::memcpy(&heads[i * kZfpHeadBytes], items[i].host_head,
                    (size_t)std::min<uint64_t>(items[i].in_len, kZfpHeadBytes));
      } else if (items[i].status == NDIC_OK) {
        hv.push_back({items[i].d_in, items[i].in_len});
        which.push_back(i);
      }
    }
    if (!hv.empty()) {
      uint64_t sz_h = al256(hv.size() * sizeof(ZfpHead));
      uint8_t* m = (uint8_t*)ctx.small.ensure(sz_h + hv.size() * kZfpHeadBytes);
      ctx.up(m, hv.data(), hv.size() * sizeof(ZfpHead), s);
      launch_zfp_heads((ZfpHead*)m, (uint32_t)hv.size(), m + sz_h, s);
      const uint8_t* got = (const uint8_t*)ctx.down(m + sz_h, hv.size() * kZfpHeadBytes, s);
      NDIC_HIP(hipStreamSynchronize(s));
      for (size_t k = 0; k < which.size(); ++k)
        std::memcpy(&heads[which[k] * kZfpHeadBytes], &got[k * kZfpHeadBytes], kZfpHeadBytes);
    }
  }
  for (size_t i = 0; i < items.size(); ++i) {
    ZfpItem& it = items[i];
    if (it.status != NDIC_OK) continue;
    try {
      it.plan = make_plan(it.shape, it.dtype, it.cfg);
      check_stream(it.plan, &heads[i * kZfpHeadBytes], kZfpHeadBytes, it.in_len);
      if (it.index_len && !it.plan.fixed && it.index_len != 4 * (it.plan.nblocks + 1))
        fail(NDIC_INVALID_ARGUMENT, "nd_zfp: the block index has " + std::to_string(it.index_len) +
                                        " bytes, the stream's " + std::to_string(it.plan.nblocks) +
                                        " blocks need " + std::to_string(4 * (it.plan.nblocks + 1)));
    } catch (const Failure& f) {
      set_fail(it, f);
    }
  }
  for (uint32_t cls = 8; cls < 40; ++cls) {
    std::vector<ZfpItem*> group;
    for (auto& it : items)
      if (it.status == NDIC_OK && class_of(it.plan) == cls) group.push_back(&it);
    if (group.empty()) continue;
    std::vector<ZfpChunk> ch;
    uint64_t total = 0;
    bool walk = false, indexed = false;
    for (ZfpItem* it : group) {
      ZfpChunk c = chunk_of(it->plan);
      c.block_begin = total;
      c.src = it->d_in;
      c.src_len = it->in_len;
      c.dst = it->d_out;
      total += c.nblocks;
      if (!c.fixed && it->index_len) {
        c.index = (uint32_t*)it->d_index;
        indexed = true;
      } else {
        walk |= !c.fixed;
      }
      ch.push_back(c);
    }
    const uint64_t sz_ch = al256(ch.size() * sizeof(ZfpChunk)), sz_offs = walk ? al256(total * 8) : 0;
    uint8_t* meta = (uint8_t*)ctx.work.ensure(sz_ch + sz_offs + (indexed ? al256(ch.size() * 4) : 0));
    uint64_t* offs = walk ? (uint64_t*)(meta + sz_ch) : nullptr;
    uint32_t* bad = indexed ? (uint32_t*)(meta + sz_ch + sz_offs) : nullptr;
    for (size_t j = 0; j < ch.size(); ++j) {
      ch[j].offs = offs;
      if (ch[j].index) ch[j].bad = bad + j;
    }
    ZfpChunk* d_ch = (ZfpChunk*)meta;
    ctx.up(d_ch, ch.data(), ch.size() * sizeof(ZfpChunk), s);
    if (indexed) NDIC_HIP(hipMemsetAsync(bad, 0, ch.size() * 4, s));
    ZfpClock clock(s);
    if (walk) {
      clock.begin(ZfpClock::kWalk);
      launch_zfp_index(d_ch, (uint32_t)ch.size(), cls / 8, cls % 8, s);
      clock.end();
    }
    clock.begin(ZfpClock::kDecode);
    launch_zfp_decode(d_ch, (uint32_t)ch.size(), total, cls / 8, cls % 8, nullptr, zfp_fixed_maxbits(ch), s);
    clock.end();
    const uint32_t* flags = indexed ? (const uint32_t*)ctx.down(bad, ch.size() * 4, s) : nullptr;
    NDIC_HIP(hipStreamSynchronize(s));
    clock.collect();
    for (size_t j = 0; flags && j < ch.size(); ++j)
      if (flags[j]) {
        group[j]->status = NDIC_CODESTREAM;
        group[j]->message = "nd_zfp: the block index disagrees with the stream";
      }
  }
}

}  // namespace ndic

using namespace ndic;

namespace {

std::vector<size_t> shape_vec(const size_t* shape, size_t ndim) {
  if (ndim && !shape) fail(NDIC_INVALID_ARGUMENT, "null shape");
  return std::vector<size_t>(shape, shape + ndim);
}

const ndic_zfp_config& cfg_ref(const ndic_zfp_config* c) {
  if (!c) fail(NDIC_INVALID_ARGUMENT, "null zfp configuration");
  return *c;
}

void raise_item(const ZfpItem& it) {
  if (it.status != NDIC_OK) fail(it.status, it.message);
}

}  // namespace

extern "C" ndic_status ndic_zfp_encode_chunk(const uint8_t* bytes, size_t len, const size_t* shape, size_t ndim,
                                             ndic_dtype dtype, const ndic_zfp_config* config, ndic_buffer* out,
                                             ndic_error* err) {
  return guard_call(err, [&] {
    if (!out) fail(NDIC_INVALID_ARGUMENT, "null output buffer");
    out->data = nullptr;
    out->len = 0;
    ZfpItem it;
    it.shape = shape_vec(shape, ndim);
    it.dtype = dtype;
    it.cfg = cfg_ref(config);
    it.in_len = len;
    // validate before touching the device
    it.plan = make_plan(it.shape, it.dtype, it.cfg);
    uint64_t want = it.plan.elems * dtype_size(dtype);
    if (len != want)
      fail(NDIC_INVALID_ARGUMENT, "zfp: chunk shape " + shape_str(it.shape) + " (" + dtype_name(dtype) + ") needs " +
                                      std::to_string(want) + " bytes, got " + std::to_string(len));
    if (len && !bytes) fail(NDIC_INVALID_ARGUMENT, "null input");
    Context& ctx = Context::get();
    std::lock_guard<std::recursive_mutex> lock(ctx.mu);
    hipStream_t s = ctx.stream;
    uint8_t* d_in = (uint8_t*)ctx.in.ensure(len + 64);
    NDIC_HIP(hipMemcpyAsync(d_in, bytes, len, hipMemcpyHostToDevice, s));
    uint64_t cap = it.plan.max_bytes + 64;
    uint8_t* d_out = (uint8_t*)ctx.out.ensure(cap);
    it.d_in = d_in;
    std::vector<ZfpItem> items{std::move(it)};
    zfp_encode_items(ctx, items, d_out, cap, s, nullptr, 0);
    raise_item(items[0]);
    emit_host(d_out + items[0].out_off, items[0].out_len, out, s);
  });
}

extern "C" ndic_status ndic_zfp_decode_chunk(const uint8_t* stream, size_t len, const size_t* shape, size_t ndim,
                                             ndic_dtype dtype, const ndic_zfp_config* config, ndic_buffer* out,
                                             ndic_error* err) {
  return guard_call(err, [&] {
    if (!out) fail(NDIC_INVALID_ARGUMENT, "null output buffer");
    out->data = nullptr;
    out->len = 0;
    ZfpItem it;
    it.shape = shape_vec(shape, ndim);
    it.dtype = dtype;
    it.cfg = cfg_ref(config);
    it.in_len = len;
    it.plan = make_plan(it.shape, it.dtype, it.cfg);
    uint8_t head[kZfpHeadBytes] = {};
    if (len && !stream) fail(NDIC_INVALID_ARGUMENT, "null input");
    std::memcpy(head, stream, std::min<size_t>(len, kZfpHeadBytes));
    check_stream(it.plan, head, kZfpHeadBytes, len);
    Context& ctx = Context::get();
    std::lock_guard<std::recursive_mutex> lock(ctx.mu);
    hipStream_t s = ctx.stream;
    uint8_t* d_in = (uint8_t*)ctx.in.ensure(len + 64);
    if (len) NDIC_HIP(hipMemcpyAsync(d_in, stream, len, hipMemcpyHostToDevice, s));
    uint64_t nout = it.plan.elems * dtype_size(dtype);
    uint8_t* d_out = (uint8_t*)ctx.out.ensure(nout + 64);
    it.d_in = d_in;
    it.d_out = d_out;
    it.host_head = head;
    std::vector<ZfpItem> items{std::move(it)};
    zfp_decode_items(ctx, items, s);
    raise_item(items[0]);
    emit_host(d_out, nout, out, s);
  });
}

extern "C" ndic_status ndic_zfp_decode_brick(const uint8_t* stream, size_t len, const size_t* shape, size_t ndim,
                                             ndic_dtype dtype, const ndic_zfp_config* config, const size_t* brick,
                                             size_t nbrick, ndic_buffer* out, size_t* brick_shape, ndic_error* err) {
  return guard_call(err, [&] {
    if (!out) fail(NDIC_INVALID_ARGUMENT, "null output buffer");
    out->data = nullptr;
    out->len = 0;
    const ndic_zfp_config& cfg = cfg_ref(config);
    validate_config(cfg);
    if (cfg.mode != NDIC_ZFP_FIXED_RATE)
      fail(NDIC_INVALID_ARGUMENT, std::string("zfp: brick decode needs fixed_rate mode, configuration says \"") +
                                      mode_name(cfg.mode) + "\"");
    Plan pl = make_plan(shape_vec(shape, ndim), dtype, cfg);
    if (len != pl.fixed_bytes)
      fail(NDIC_CODESTREAM, "nd_zfp: fixed-rate stream is " + std::to_string(len) + " bytes, computed size is " +
                                std::to_string(pl.fixed_bytes));
    uint8_t head[kZfpHeadBytes] = {};
    if (len && !stream) fail(NDIC_INVALID_ARGUMENT, "null input");
    std::memcpy(head, stream, std::min<size_t>(len, kZfpHeadBytes));
    check_stream(pl, head, kZfpHeadBytes, len);
    // BrickIndex::linear (lib.rs:452-472): row-major over the effective shape
    std::vector<size_t> grid(pl.dims), coords(brick ? brick : nullptr, brick ? brick + nbrick : nullptr);
    for (uint32_t a = 0; a < pl.dims; ++a) grid[a] = pl.nb[pl.dims - 1 - a];
    std::ostringstream cs;
    cs << shape_str(coords);
    if (coords.size() != grid.size())
      fail(NDIC_INVALID_ARGUMENT, "nd_zfp: brick coordinates " + cs.str() + " do not match the " +
                                      std::to_string(grid.size()) + "-axis grid");
    uint64_t k = 0;
    for (size_t a = 0; a < grid.size(); ++a) {
      if (coords[a] >= grid[a])
        fail(NDIC_INVALID_ARGUMENT, "nd_zfp: brick coordinates " + cs.str() + " exceed the grid " + shape_str(grid));
      k = k * grid[a] + coords[a];
    }
    uint64_t cnt = 1;
    for (uint32_t a = 0; a < pl.dims; ++a) {
      size_t e = std::min<size_t>(pl.eff[a] - coords[a] * 4, 4);
      if (brick_shape) brick_shape[a] = e;
      cnt *= e;
    }
    Context& ctx = Context::get();
    std::lock_guard<std::recursive_mutex> lock(ctx.mu);
      hipStream_t s = ctx.stream;
    uint64_t nout = cnt * dtype_size(dtype);
    uint64_t sz_in = al256(len + 8), sz_out = al256(nout), sz_ch = al256(sizeof(ZfpChunk));
    uint8_t* m = (uint8_t*)ctx.in.ensure(sz_in + sz_out + sz_ch + sizeof(ZfpBrick));
    uint8_t* d_in = m;
    uint8_t* d_out = m + sz_in;
    ZfpChunk* d_ch = (ZfpChunk*)(m + sz_in + sz_out);
    ZfpBrick* d_br = (ZfpBrick*)(m + sz_in + sz_out + sz_ch);
    NDIC_HIP(hipMemcpyAsync(d_in, stream, len, hipMemcpyHostToDevice, s));
    ZfpChunk c = chunk_of(pl);
    c.src = d_in;
    c.src_len = len;
    ZfpBrick b{};
    b.chunk = 0;
    b.block = k;
    b.out = d_out;
    NDIC_HIP(hipMemcpyAsync(d_ch, &c, sizeof(c), hipMemcpyHostToDevice, s));
    NDIC_HIP(hipMemcpyAsync(d_br, &b, sizeof(b), hipMemcpyHostToDevice, s));
    launch_zfp_decode(d_ch, 1, 1, pl.dims, pl.scalar, d_br, zfp_fixed_maxbits({c}), s);
    emit_host(d_out, nout, out, s);
  });
}

namespace {

ZfpItem item_of(const ndic_zfp_desc& d, const uint8_t* d_in, uint8_t* d_out, bool encode) {
  ZfpItem it;
  if (d.ndim > NDIC_MAX_NDIM) fail(NDIC_INVALID_ARGUMENT, "descriptor ndim exceeds NDIC_MAX_NDIM");
  it.shape.assign(d.shape, d.shape + d.ndim);
  it.dtype = d.dtype;
  it.cfg = d.config;
  it.in_len = d.in_len;
  it.d_in = d_in + d.in_offset;
  if (encode) {
    size_t es = dtype_size(d.dtype);
    if (d.in_offset % es) fail(NDIC_INVALID_ARGUMENT, "zfp batch: sample offset is not element-aligned");
  } else {
    if (d.in_offset % 8) fail(NDIC_INVALID_ARGUMENT, "zfp batch: stream offset is not 8-byte aligned");
    it.d_out = d_out + d.out_offset;
    if (d.out_offset % dtype_size(d.dtype)) fail(NDIC_INVALID_ARGUMENT, "zfp batch: output offset is not element-aligned");
  }
  return it;
}

ndic_status batch_status(ndic_zfp_desc* descs, std::vector<ZfpItem>& items, ndic_error* err) {
  ndic_status first = NDIC_OK;
  for (size_t i = 0; i < items.size(); ++i) {
    descs[i].status = items[i].status;
    if (items[i].status != NDIC_OK && first == NDIC_OK) {
      first = (ndic_status)items[i].status;
      if (err) {
        err->status = items[i].status;
        err->offset = 0;
        std::snprintf(err->message, sizeof(err->message), "chunk %zu: %s", i, items[i].message.c_str());
      }
    }
  }
  return first;
}

}  // namespace

namespace {

ndic_status encode_batch_call(const uint8_t* d_in, uint8_t* d_out, size_t capacity, uint8_t* d_index,
                              size_t index_capacity, ndic_zfp_desc* descs, size_t n, void* stream, ndic_error* err) {
  std::vector<ZfpItem> items(n);
  ndic_status st = guard_call(err, [&] {
    if (n && !descs) fail(NDIC_INVALID_ARGUMENT, "null descriptors");
    for (size_t i = 0; i < n; ++i) {
      try {
        items[i] = item_of(descs[i], d_in, d_out, true);
      } catch (const Failure& f) {
        items[i].status = f.status;
        items[i].message = f.message;
      }
    }
    Context& ctx = Context::get();
    std::lock_guard<std::recursive_mutex> lock(ctx.mu);
    hipStream_t s = stream ? (hipStream_t)stream : ctx.stream;
    zfp_encode_items(ctx, items, d_out, capacity, s, d_index, index_capacity);
    for (size_t i = 0; i < n; ++i) {
      const bool ok = items[i].status == NDIC_OK;
      descs[i].out_offset = items[i].out_off;
      descs[i].out_len = ok ? items[i].out_len : 0;
      descs[i].index_offset = ok ? items[i].index_off : 0;
      descs[i].index_len = ok ? items[i].index_len : 0;
    }
  });
  if (st != NDIC_OK) return st;
  return batch_status(descs, items, err);
}

ndic_status decode_batch_call(const uint8_t* d_in, const uint8_t* d_index, uint8_t* d_out, ndic_zfp_desc* descs,
                              size_t n, void* stream, ndic_error* err) {
  std::vector<ZfpItem> items(n);
  ndic_status st = guard_call(err, [&] {
    if (n && !descs) fail(NDIC_INVALID_ARGUMENT, "null descriptors");
    for (size_t i = 0; i < n; ++i) {
      try {
        items[i] = item_of(descs[i], d_in, d_out, false);
        if (d_index && descs[i].index_len) {
          if (descs[i].index_offset % 4) fail(NDIC_INVALID_ARGUMENT, "zfp batch: block index offset is not 4-byte aligned");
          items[i].d_index = const_cast<uint8_t*>(d_index) + descs[i].index_offset;
          items[i].index_len = descs[i].index_len;
        }
      } catch (const Failure& f) {
        items[i].status = f.status;
        items[i].message = f.message;
      }
    }
    Context& ctx = Context::get();
    std::lock_guard<std::recursive_mutex> lock(ctx.mu);
    hipStream_t s = stream ? (hipStream_t)stream : ctx.stream;
    zfp_decode_items(ctx, items, s);
  });
  if (st != NDIC_OK) return st;
  return batch_status(descs, items, err);
}

}  // namespace

extern "C" ndic_status ndic_zfp_encode_batch(const uint8_t* d_in, uint8_t* d_out, size_t capacity,
                                             ndic_zfp_desc* descs, size_t n, void* stream, ndic_error* err) {
  return encode_batch_call(d_in, d_out, capacity, nullptr, 0, descs, n, stream, err);
}

extern "C" ndic_status ndic_zfp_decode_batch(const uint8_t* d_in, uint8_t* d_out, ndic_zfp_desc* descs, size_t n,
                                             void* stream, ndic_error* err) {
  return decode_batch_call(d_in, nullptr, d_out, descs, n, stream, err);
}

extern "C" ndic_status ndic_zfp_encode_batch_indexed(const uint8_t* d_in, uint8_t* d_out, size_t capacity,
                                                     uint8_t* d_index, size_t index_capacity, ndic_zfp_desc* descs,
                                                     size_t n, void* stream, ndic_error* err) {
  return encode_batch_call(d_in, d_out, capacity, d_index, index_capacity, descs, n, stream, err);
}

extern "C" ndic_status ndic_zfp_decode_batch_indexed(const uint8_t* d_in, const uint8_t* d_index, uint8_t* d_out,
                                                     ndic_zfp_desc* descs, size_t n, void* stream, ndic_error* err) {
  return decode_batch_call(d_in, d_index, d_out, descs, n, stream, err);
}

namespace ndic {

void plan_bricks(const uint8_t* d_in, uint8_t* d_out, const ndic_zfp_brick_desc* descs, size_t n,
                 const uint8_t* heads, BrickPlan& bp) {
  bp.status.assign(n, NDIC_OK);
  bp.msg.assign(n, std::string());
  bp.classes.clear();
  std::vector<Plan> plans(n);
  std::vector<uint64_t> block(n, 0);
  for (size_t i = 0; i < n; ++i) {
    const ndic_zfp_brick_desc& d = descs[i];
    try {
      if (d.ndim < 1 || d.ndim > 4) fail(NDIC_INVALID_ARGUMENT, "nd_zfp: brick descriptor ndim must be 1..=4");
      if (d.in_offset % 8) fail(NDIC_INVALID_ARGUMENT, "zfp bricks: stream offset is not 8-byte aligned");
      ndic_zfp_config cfg{};
      cfg.mode = NDIC_ZFP_FIXED_RATE;
      cfg.rate = d.rate;
      std::vector<size_t> shape(d.shape, d.shape + d.ndim);
      Plan pl = make_plan(shape, d.dtype, cfg);
      if (d.in_len != pl.fixed_bytes)
        fail(NDIC_CODESTREAM, "nd_zfp: fixed-rate stream is " + std::to_string(d.in_len) +
                                  " bytes, computed size is " + std::to_string(pl.fixed_bytes));
      check_stream(pl, &heads[i * kZfpHeadBytes], kZfpHeadBytes, d.in_len);
      uint64_t k = 0;
      for (uint32_t a = 0; a < pl.dims; ++a) {
        uint32_t g = pl.nb[pl.dims - 1 - a];
        if (d.brick[a] >= g) fail(NDIC_INVALID_ARGUMENT, "nd_zfp: brick coordinates exceed the grid");
        k = k * g + d.brick[a];
      }
      if (d.out_offset % dtype_size(d.dtype)) fail(NDIC_INVALID_ARGUMENT, "zfp bricks: output offset is not aligned");
      plans[i] = pl;
      block[i] = k;
    } catch (const Failure& f) {
      bp.status[i] = f.status;
      bp.msg[i] = f.message;
    }
  }
  for (uint32_t cls = 8; cls < 40; ++cls) {
    BrickClass bc;
    bc.dims = cls / 8;
    bc.scalar = cls % 8;
    for (size_t i = 0; i < n; ++i) {
      if (bp.status[i] != NDIC_OK || class_of(plans[i]) != cls) continue;
      ZfpChunk c = chunk_of(plans[i]);
      c.src = d_in + descs[i].in_offset;
      c.src_len = descs[i].in_len;
      ZfpBrick b{};
      b.chunk = (uint32_t)bc.chunks.size();
      b.block = block[i];
      b.out = d_out + descs[i].out_offset;
      bc.chunks.push_back(c);
      bc.bricks.push_back(b);
    }
    if (!bc.chunks.empty()) bp.classes.push_back(std::move(bc));
  }
}

ndic_status report_bricks(ndic_zfp_brick_desc* descs, size_t n, const BrickPlan& bp, ndic_error* err) {
  ndic_status first = NDIC_OK;
  for (size_t i = 0; i < n; ++i) {
    descs[i].status = bp.status[i];
    if (bp.status[i] != NDIC_OK && first == NDIC_OK) {
      first = (ndic_status)bp.status[i];
      if (err) {
        err->status = bp.status[i];
        err->offset = 0;
        std::snprintf(err->message, sizeof(err->message), "brick %zu: %s", i, bp.msg[i].c_str());
      }
    }
  }
  return first;
}

}  // namespace ndic

extern "C" ndic_status ndic_zfp_decode_bricks(const uint8_t* d_in, uint8_t* d_out, ndic_zfp_brick_desc* descs,
                                              size_t n, void* stream, ndic_error* err) {
  BrickPlan bp;
  ndic_status st = guard_call(err, [&] {
    if (n && !descs) fail(NDIC_INVALID_ARGUMENT, "null descriptors");
    Context& ctx = Context::get();
    std::lock_guard<std::recursive_mutex> lock(ctx.mu);
      ctx.release_kept();
    hipStream_t s = stream ? (hipStream_t)stream : ctx.stream;
    // header bytes of every stream for the checks of BrickReader::fixed_rate
    std::vector<ZfpHead> hv(n);
    for (size_t i = 0; i < n; ++i) hv[i] = {d_in + descs[i].in_offset, descs[i].in_len};
    uint64_t sz_h = al256(n * sizeof(ZfpHead));
    uint8_t* m = (uint8_t*)ctx.small.ensure(sz_h + n * kZfpHeadBytes + 64);
    std::vector<uint8_t> heads(n * kZfpHeadBytes);
    if (n) {
      ctx.up(m, hv.data(), n * sizeof(ZfpHead), s);
      launch_zfp_heads((ZfpHead*)m, (uint32_t)n, m + sz_h, s);
      const uint8_t* got = (const uint8_t*)ctx.down(m + sz_h, heads.size(), s);
      NDIC_HIP(hipStreamSynchronize(s));
      std::memcpy(heads.data(), got, heads.size());
    }
    plan_bricks(d_in, d_out, descs, n, heads.data(), bp);
    for (const BrickClass& bc : bp.classes) {
      uint64_t sz_ch = al256(bc.chunks.size() * sizeof(ZfpChunk));
      uint8_t* w = (uint8_t*)ctx.work.ensure(sz_ch + bc.bricks.size() * sizeof(ZfpBrick));
      ctx.up(w, bc.chunks.data(), bc.chunks.size() * sizeof(ZfpChunk), s);
      ctx.up(w + sz_ch, bc.bricks.data(), bc.bricks.size() * sizeof(ZfpBrick), s);
      ZfpClock clock(s);
      clock.begin(ZfpClock::kDecode);
      launch_zfp_decode((ZfpChunk*)w, (uint32_t)bc.chunks.size(), bc.bricks.size(), bc.dims, bc.scalar,
                        (ZfpBrick*)(w + sz_ch), zfp_fixed_maxbits(bc.chunks), s);
      clock.end();
      NDIC_HIP(hipStreamSynchronize(s));
      clock.collect();
    }
  });
  if (st != NDIC_OK) return st;
  return report_bricks(descs, n, bp, err);
}
