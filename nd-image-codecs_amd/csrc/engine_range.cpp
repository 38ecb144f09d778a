// Byte-range plans over an ndht chunk's coefficient-plane index and
// plane-selective decodes of device-resident chunks.
//   planner:  crates/ndic-codestream/src/range.rs:26-369 (RangeIndex),
//             container.rs:173-340 (ChunkHeader::required_len / parse),
//             crates/ndic-lift/src/lib.rs:195-225 (approx_len,
//             low_pass_indices)
//   decoder:  crates/ndic-zarr/src/htj2k_codec.rs:319-411 (plane-selective
//             partial decode), range.rs:308-369 (3D preview planes)
//
// Which planes a request needs and at which resolution follow from the
// chunk shape, the lift steps and max_px alone, so the host computes them;
// the device reads each selected plane's codestream location from the
// chunk's own index in HBM (the parse kernels' locate step) and decodes
// only those planes, to that resolution.
#include <algorithm>
#include <cstring>
#include <sstream>

#include "engine.hpp"
#include "items.hpp"

namespace ndic {

namespace {

constexpr uint64_t kMaxPlanes = 1ull << 24;

uint32_t le32(const uint8_t* p) { return uint32_t(p[0]) | uint32_t(p[1]) << 8 | uint32_t(p[2]) << 16 | uint32_t(p[3]) << 24; }
uint64_t le64(const uint8_t* p) { return uint64_t(le32(p)) | uint64_t(le32(p + 4)) << 32; }

struct PlaneEntry {
  uint64_t offset;
  uint32_t len;
  std::vector<uint32_t> prefix;
};

struct ChunkIndex {
  std::vector<uint32_t> dims;
  uint32_t xy_levels = 0;
  uint64_t header_len = 0;
  bool has_index = false;
  std::vector<PlaneEntry> planes;
};

// ChunkHeader::parse (container.rs:199-340), errors included.
ChunkIndex parse_chunk_prefix(const uint8_t* b, size_t n) {
  if (n < 8) fail(NDIC_CODESTREAM, "chunk header prefix too short", 0);
  if (std::memcmp(b, "ndht", 4) != 0) fail(NDIC_CODESTREAM, "not an htj2k chunk (bad magic)", 0);
  if (b[4] != 1) fail(NDIC_CODESTREAM, "unsupported htj2k chunk container version", 4);
  const uint32_t flags = b[5];
  if (flags & ~1u) fail(NDIC_CODESTREAM, "unknown htj2k chunk container flags", 5);
  const uint32_t lv = b[6];
  if (lv > 32) fail(NDIC_CODESTREAM, "xy_levels exceeds the 32-level J2K bound (T.800 SPcod)", 6);
  const uint32_t ndim = b[7];
  if (ndim < 2 || ndim > 32) fail(NDIC_CODESTREAM, "chunk ndim outside 2..=32", 7);
  const size_t fixed = 8 + 4 * (size_t)ndim;
  if (n < fixed) fail(NDIC_CODESTREAM, "truncated chunk header", n);
  ChunkIndex c;
  c.xy_levels = lv;
  for (uint32_t i = 0; i < ndim; ++i) {
    const uint32_t d = le32(b + 8 + 4 * i);
    if (d == 0) fail(NDIC_CODESTREAM, "zero chunk dimension", 8 + 4 * i);
    c.dims.push_back(d);
  }
  uint64_t np = 1;
  for (uint32_t i = 0; i + 2 < ndim; ++i) {
    np *= c.dims[i];
    if (np > kMaxPlanes) fail(NDIC_CODESTREAM, "chunk declares more planes than supported", 8);
  }
  c.header_len = fixed;
  if (!(flags & 1)) return c;
  const size_t entry = 12 + 4 * ((size_t)lv + 1);
  const size_t need = fixed + (size_t)np * entry;
  if (n < need) fail(NDIC_CODESTREAM, "truncated coefficient-plane index", n);
  c.has_index = true;
  c.header_len = need;
  uint64_t min_off = need;
  for (uint64_t i = 0; i < np; ++i) {
    const size_t at = fixed + (size_t)i * entry;
    const uint8_t* e = b + at;
    PlaneEntry p;
    p.offset = le64(e);
    p.len = le32(e + 8);
    for (uint32_t r = 0; r <= lv; ++r) p.prefix.push_back(le32(e + 12 + 4 * r));
    if (p.offset < min_off) fail(NDIC_CODESTREAM, "plane offsets overlap or precede the index", at);
    for (uint32_t r = 0; r < lv; ++r)
      if (p.prefix[r] > p.prefix[r + 1]) fail(NDIC_CODESTREAM, "plane prefix lengths are inconsistent", at);
    if (p.prefix[lv] > p.len) fail(NDIC_CODESTREAM, "plane prefix lengths are inconsistent", at);
    if (p.offset + p.len < p.offset) fail(NDIC_CODESTREAM, "plane byte range overflows", at);
    min_off = p.offset + p.len;
    c.planes.push_back(std::move(p));
  }
  return c;
}

// RangeIndex::resolution_for (range.rs:227-240): the largest resolution
// whose decoded long side fits max_px (0 when none does) and its size.
void resolution_for(uint64_t W, uint64_t H, uint32_t L, uint64_t max_px, uint32_t* res, uint64_t* w, uint64_t* h) {
  uint64_t bw, bh;
  level_dims(W, H, L, &bw, &bh);
  uint32_t best = 0;
  for (uint32_t r = 0; r <= L; ++r) {
    uint64_t dw, dh;
    level_dims(W, H, L - r, &dw, &dh);
    if (std::max(dw, dh) <= max_px) {
      best = r;
      bw = dw;
      bh = dh;
    }
  }
  *res = best;
  *w = bw;
  *h = bh;
}

uint32_t band_count(uint64_t n, uint32_t levels) {
  uint32_t c = 0;
  for (uint64_t m = n; c < levels && m >= 2; m = (m + 1) / 2) ++c;
  return c;
}

// approx_len / low_pass_indices (ndic-lift lib.rs:195-225).
std::vector<uint64_t> low_pass_indices(const ndic_lift_step& st, uint64_t extent) {
  std::vector<uint64_t> out;
  if (extent == 0) return out;
  const uint64_t group = st.group == 0 ? extent : std::min<uint64_t>(st.group, extent);
  for (uint64_t base = 0; base < extent; base += group) {
    const uint64_t len = std::min(group, extent - base);
    uint64_t keep;
    if (st.kind == 0) keep = std::min<uint64_t>(len, 1);
    else keep = (len + (1ull << band_count(len, st.levels)) - 1) >> band_count(len, st.levels);
    for (uint64_t i = 0; i < keep; ++i) out.push_back(base + i);
  }
  return out;
}

// Per leading axis, which indices hold low-pass planes (range.rs:322-338),
// and the selected planes in linear order (range.rs:342-353).
std::vector<uint32_t> thumbnail3d_planes(const std::vector<uint64_t>& dims, const ndic_lift_step* steps, size_t nsteps,
                                        std::vector<uint64_t>* counts) {
  const size_t lead = dims.size() - 2;
  for (size_t k = 0; k < nsteps; ++k)
    if (steps[k].dimension >= lead) {
      std::ostringstream m;
      m << "nd_lift step on dimension " << steps[k].dimension << " is not a leading chunk axis (chunk has " << lead
        << " leading dimensions)";
      fail(NDIC_INVALID_ARGUMENT, m.str());
    }
  std::vector<std::vector<char>> masks(lead);
  for (size_t d = 0; d < lead; ++d) {
    masks[d].assign(dims[d], 0);
    const ndic_lift_step* st = nullptr;
    for (size_t k = 0; k < nsteps && !st; ++k)
      if (steps[k].dimension == d) st = &steps[k];
    if (!st) std::fill(masks[d].begin(), masks[d].end(), 1);
    else
      for (uint64_t i : low_pass_indices(*st, dims[d])) masks[d][i] = 1;
  }
  uint64_t np = 1;
  for (size_t d = 0; d < lead; ++d) np *= dims[d];
  std::vector<uint32_t> sel;
  for (uint64_t lin = 0; lin < np; ++lin) {
    uint64_t rem = lin;
    bool keep = true;
    for (size_t d = lead; d-- > 0;) {
      keep &= masks[d][rem % dims[d]] != 0;
      rem /= dims[d];
    }
    if (keep) sel.push_back((uint32_t)lin);
  }
  counts->clear();
  for (size_t d = 0; d < lead; ++d) counts->push_back((uint64_t)std::count(masks[d].begin(), masks[d].end(), 1));
  return sel;
}

struct Range {
  uint64_t start, end;  // end inclusive
};

std::vector<Range> coalesce(std::vector<Range> r) {
  std::sort(r.begin(), r.end(), [](const Range& a, const Range& b) { return a.start < b.start; });
  std::vector<Range> out;
  for (auto& x : r) {
    if (!out.empty() && x.start <= out.back().end + 1) out.back().end = std::max(out.back().end, x.end);
    else out.push_back(x);
  }
  return out;
}

Range prefix_range(const ChunkIndex& c, uint64_t plane, uint32_t max_res) {
  if (plane >= c.planes.size()) {
    std::ostringstream m;
    m << "plane " << plane << " out of range for a " << c.planes.size() << "-plane chunk";
    fail(NDIC_INVALID_ARGUMENT, m.str());
  }
  const PlaneEntry& p = c.planes[plane];
  const uint64_t len = p.prefix[std::min(max_res, c.xy_levels)];
  if (len == 0) fail(NDIC_INVALID_ARGUMENT, "plane declares an empty resolution prefix");
  return {p.offset, p.offset + len - 1};
}

void to_buffer(const std::vector<uint64_t>& v, ndic_buffer* out) {
  out->len = v.size() * 8;
  out->data = (uint8_t*)std::malloc(out->len + 8);
  if (!out->data) fail(NDIC_IO, "out of host memory");
  if (!v.empty()) std::memcpy(out->data, v.data(), out->len);
}

}  // namespace

}  // namespace ndic

using namespace ndic;

extern "C" ndic_status ndic_range_plan_chunk(const uint8_t* prefix, size_t len, int32_t target, uint64_t arg,
                                             const ndic_lift_step* steps, size_t nsteps, ndic_range_plan* out,
                                             ndic_error* err) {
  return guard_call(err, [&] {
    if (!out || (!prefix && len)) fail(NDIC_INVALID_ARGUMENT, "range plan: null argument");
    std::memset(out, 0, sizeof(*out));
    const ChunkIndex c = parse_chunk_prefix(prefix, len);
    if (!c.has_index) fail(NDIC_UNSUPPORTED, "chunk was written without a coefficient-plane index (index: false)");
    const uint64_t W = c.dims.back(), H = c.dims[c.dims.size() - 2];
    std::vector<Range> ranges;
    if (c.header_len) ranges.push_back({0, c.header_len - 1});
    std::vector<uint64_t> planes, size;
    uint32_t max_res = 0;
    if (target == 0 || target == 2) {
      uint64_t w, h;
      resolution_for(W, H, c.xy_levels, arg, &max_res, &w, &h);
      if (target == 0) {
        ranges.push_back(prefix_range(c, 0, max_res));
        planes.push_back(0);
      } else {
        std::vector<uint64_t> dims(c.dims.begin(), c.dims.end()), counts;
        for (uint32_t p : thumbnail3d_planes(dims, steps, nsteps, &counts)) {
          ranges.push_back(prefix_range(c, p, max_res));
          planes.push_back(p);
        }
        size = counts;
      }
      size.push_back(h);
      size.push_back(w);
    } else if (target == 1) {
      if (arg >= c.planes.size()) {
        std::ostringstream m;
        m << "plane " << arg << " out of range for a " << c.planes.size() << "-plane chunk";
        fail(NDIC_INVALID_ARGUMENT, m.str());
      }
      const PlaneEntry& p = c.planes[arg];
      if (p.len) ranges.push_back({p.offset, p.offset + p.len - 1});
      planes.push_back(arg);
      size = {H, W};
      max_res = c.xy_levels;
    } else {
      fail(NDIC_INVALID_ARGUMENT, "range plan: target is 0 (thumbnail), 1 (plane) or 2 (thumbnail_3d)");
    }
    ranges = coalesce(ranges);
    std::vector<uint64_t> flat;
    uint64_t total = 0;
    for (auto& r : ranges) {
      flat.push_back(r.start);
      flat.push_back(r.end);
      total += r.end - r.start + 1;
    }
    out->target = target;
    out->max_res = max_res;
    out->ndims = (uint32_t)std::min<size_t>(size.size(), 34);
    for (uint32_t i = 0; i < out->ndims; ++i) out->decoded_size[i] = size[i];
    out->total_bytes = total;
    out->header_len = c.header_len;
    to_buffer(planes, &out->planes);
    to_buffer(flat, &out->ranges);
  });
}

extern "C" ndic_status ndic_htj2k_decode_select_batch(const uint8_t* d_in, uint8_t* d_out, ndic_select_desc* descs,
                                                      size_t n, void* stream, ndic_error* err) {
  return guard_call(err, [&] {
    Context& ctx = Context::get();
    std::lock_guard<std::recursive_mutex> lock(ctx.mu);
    hipStream_t s = stream ? (hipStream_t)stream : ctx.stream;
    std::vector<DecodeItem> items;
    std::vector<size_t> item_of(n, SIZE_MAX);
    for (size_t i = 0; i < n; ++i) {
      ndic_select_desc& d = descs[i];
      d.status = NDIC_OK;
      d.max_res = 0;
      d.out_ndim = 0;
      d.out_bytes = 0;
      if (d.ndim < 2 || d.ndim > NDIC_MAX_NDIM || d.xy_levels > 32 || d.nsteps > NDIC_MAX_STEPS)
        fail(NDIC_INVALID_ARGUMENT, "select descriptor out of range");
      std::vector<uint64_t> dims(d.shape, d.shape + d.ndim);
      for (uint64_t x : dims)
        if (x == 0) fail(NDIC_INVALID_ARGUMENT, "select descriptor: zero chunk dimension");
      const uint64_t W = dims.back(), H = dims[d.ndim - 2];
      DecodeItem it;
      it.d_in = d_in + d.in_offset;
      it.len = d.in_len;
      it.shape.assign(dims.begin(), dims.end());
      it.L = d.xy_levels;
      it.window = true;
      std::vector<uint64_t> out_shape;
      if (d.target == 0) {
        // the planes an axis-aligned subset touches, in C order of the
        // leading dims, each decoded in full; the (y, x) window copied and
        // narrowed to the dtype (htj2k_codec.rs:360-407)
        if (!(d.dtype == NDIC_U8 || d.dtype == NDIC_I8 || d.dtype == NDIC_U16 || d.dtype == NDIC_I16 ||
              d.dtype == NDIC_U32 || d.dtype == NDIC_I32))
          fail(NDIC_UNSUPPORTED, std::string("htj2k does not support dtype ") + dtype_name(d.dtype));
        for (uint32_t k = 0; k < d.ndim; ++k)
          if ((uint64_t)d.start[k] + d.count[k] > dims[k])
            fail(NDIC_INVALID_ARGUMENT, "htj2k partial decode: subset exceeds the chunk");
        std::vector<uint32_t> planes{0};
        for (uint32_t k = 0; k + 2 < d.ndim; ++k) {
          uint64_t extent = 1;
          for (uint32_t q = k + 1; q + 2 < d.ndim; ++q) extent *= dims[q];
          std::vector<uint32_t> next;
          for (uint32_t base : planes)
            for (uint32_t j = 0; j < d.count[k]; ++j) next.push_back((uint32_t)(base + (d.start[k] + j) * extent));
          planes.swap(next);
        }
        for (uint32_t k = 0; k < d.ndim; ++k) out_shape.push_back(d.count[k]);
        it.ht_dtype = d.dtype;
        it.max_res = -1;
        it.win[0] = d.start[d.ndim - 1];
        it.win[1] = d.start[d.ndim - 2];
        it.win[2] = d.count[d.ndim - 1];
        it.win[3] = d.count[d.ndim - 2];
        it.plane_sel = planes;
        d.max_res = d.xy_levels;
      } else if (d.target == 1) {
        // the low-pass planes of every lifted leading axis (all planes of an
        // untransformed one) at the largest resolution that fits max_px
        // (range.rs:308-369), as int32 coefficient samples
        lift_validate("0.1", d.steps, d.nsteps, d.ndim);
        std::vector<uint64_t> counts;
        it.plane_sel = thumbnail3d_planes(dims, d.steps, d.nsteps, &counts);
        uint32_t res;
        uint64_t w, h;
        resolution_for(W, H, d.xy_levels, d.max_px, &res, &w, &h);
        it.ht_dtype = NDIC_I32;
        it.max_res = (int32_t)res;
        it.win[2] = (uint32_t)w;
        it.win[3] = (uint32_t)h;
        out_shape = counts;
        out_shape.push_back(h);
        out_shape.push_back(w);
        d.max_res = res;
      } else {
        fail(NDIC_INVALID_ARGUMENT, "select descriptor: target is 0 (subset) or 1 (thumbnail_3d)");
      }
      uint64_t elems = 1;
      for (uint64_t x : out_shape) elems *= x;
      d.out_ndim = (uint32_t)out_shape.size();
      for (uint32_t k = 0; k < d.out_ndim; ++k) d.out_shape[k] = out_shape[k];
      d.out_bytes = elems * dtype_size(it.ht_dtype);
      if (!d_out || elems == 0) continue;  // planning only, or nothing to decode
      if (d.out_offset % dtype_size(it.ht_dtype)) fail(NDIC_INVALID_ARGUMENT, "select output offsets must be aligned to the dtype");
      it.d_out = d_out + d.out_offset;
      item_of[i] = items.size();
      items.push_back(std::move(it));
    }
    if (items.empty()) return;
    decode_items(ctx, items, s, window_elems_knob());
    int first = -1;
    for (size_t i = 0; i < n; ++i) {
      if (item_of[i] == SIZE_MAX) continue;
      descs[i].status = items[item_of[i]].status;
      if (descs[i].status != NDIC_OK && first < 0) first = (int)i;
    }
    if (first >= 0) {
      const DecodeItem& it = items[item_of[first]];
      std::ostringstream m;
      m << "chunk " << first << ": " << it.message;
      fail(it.status, m.str(), it.offset);
    }
  });
}
