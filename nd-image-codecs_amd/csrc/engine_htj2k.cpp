// htj2k / nd-lift-ht orchestration and the C-ABI entry points.
//   chunk cores: crates/ndic-zarr/src/htj2k.rs:113-441
//   writer:      crates/ndic-codestream/src/writer.rs:55-323
//   reader:      crates/ndic-codestream/src/reader.rs:469-645
//
// A batch of chunks runs as one launch sequence: parse (thread per plane) ->
// HT decode (lane per code-block) -> inverse DWT (per level, all planes) ->
// [DC un-shift] -> inverse nd_lift with the dtype narrowing fused into its
// last step.  Encode mirrors it and ends with the device-side assembly of
// every codestream and container.  Large batches are cut into windows that
// bound the device memory of the intermediate planes.
#include <algorithm>
#include <functional>
#include <memory>
#include <cstring>
#include <map>
#include <sstream>
#include <tuple>

#include "codestream.hpp"
#include "dwt.hpp"
#include "engine.hpp"
#include "ht.hpp"
#include "items.hpp"
#include "tune.hpp"

namespace ndic {

namespace {

struct Geometry {
  uint32_t packet_first, band_first;
  uint32_t blocks;
  uint32_t tag_bytes;
};

// Geometry tables for a batch: one entry per distinct (W, H, L).
struct GeoTables {
  std::vector<GeoPacket> packets;
  std::vector<GeoBand> bands;
  std::map<std::tuple<uint32_t, uint32_t, uint32_t>, Geometry> cache;

  // bands_of_resolution + code-block grids (geometry.rs:77-127)
  const Geometry& get(uint32_t W, uint32_t H, uint32_t L) {
    auto key = std::make_tuple(W, H, L);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    Geometry g{(uint32_t)packets.size(), (uint32_t)bands.size(), 0, 0};
    uint32_t block = 0;
    for (uint32_t r = 0; r <= L; ++r) {
      GeoPacket pk{(uint32_t)bands.size() - g.band_first, 0, block, 0};
      auto add = [&](uint32_t x0, uint32_t y0, uint32_t w, uint32_t h, uint32_t kidx) {
        GeoBand b{};
        b.x0 = x0;
        b.y0 = y0;
        b.w = w;
        b.h = h;
        b.nx = (w == 0 || h == 0) ? 0 : (uint32_t)ceil_div(w, 64);
        b.ny = (w == 0 || h == 0) ? 0 : (uint32_t)ceil_div(h, 64);
        b.kidx = kidx;
        b.block_first = block;
        block += b.nx * b.ny;
        if (b.nx) g.tag_bytes = std::max(g.tag_bytes, tag_scratch_bytes(b.nx, b.ny));
        bands.push_back(b);
        pk.nbands += 1;
      };
      if (r == 0) {
        uint64_t w, h;
        level_dims(W, H, L, &w, &h);
        add(0, 0, (uint32_t)w, (uint32_t)h, 0);
      } else {
        uint32_t l = L - r + 1;
        uint64_t wp, hp, wl, hl;
        level_dims(W, H, l - 1, &wp, &hp);
        level_dims(W, H, l, &wl, &hl);
        uint32_t k = (r - 1) * 3;
        add((uint32_t)wl, 0, (uint32_t)(wp - wl), (uint32_t)hl, k + 1);
        add(0, (uint32_t)hl, (uint32_t)wl, (uint32_t)(hp - hl), k + 2);
        add((uint32_t)wl, (uint32_t)hl, (uint32_t)(wp - wl), (uint32_t)(hp - hl), k + 3);
      }
      pk.nblocks = block - pk.block_first;
      packets.push_back(pk);
    }
    g.blocks = block;
    return cache.emplace(key, g).first->second;
  }
};

const char* err_text(int32_t code) {
  switch (code) {
    case kErrNarrow: return "nd_lift decode: coefficients do not narrow back to the dtype (corrupt or mismatched chunk)";
    case kErrHtNarrow: return "decoded sample range does not fit the dtype (corrupt or mismatched chunk)";
    case kErrChunkMagic: return "not an htj2k chunk (bad magic)";
    case kErrChunkVersion: return "unsupported htj2k chunk container version";
    case kErrChunkFlags: return "unknown htj2k chunk container flags";
    case kErrChunkLevels: return "xy_levels exceeds the 32-level J2K bound (T.800 SPcod)";
    case kErrChunkNdim: return "chunk ndim outside 2..=32";
    case kErrChunkTruncated: return "truncated chunk header or coefficient-plane index";
    case kErrChunkZeroDim: return "zero chunk dimension";
    case kErrChunkShape: return "chunk header declares a different shape than decode expects";
    case kErrChunkIndex: return "plane offsets overlap or prefixes are inconsistent";
    case kErrPlaneRange: return "plane byte range exceeds the chunk";
    case kErrCsMarker: return "malformed codestream marker segment";
    case kErrCsUnsupported:
      return "codestream outside the device decoder's scope (one HT component, reversible 5/3, one layer, maximal "
             "precincts, 64x64 code-blocks, one tile-part)";
    case kErrCsGeometry: return "plane decodes to a different geometry than the chunk header declares";
    case kErrCsPlt: return "PLT lengths do not cover the tile-part body";
    case kErrPacket: return "malformed packet header or truncated packet body";
    case kErrBlock: return "malformed HT code-block segment";
    case kErrCsLevels: return "codestream decomposition levels differ from the chunk's xy_levels";
    case kErrEncodeScup: return "cleanup segment suffix out of range (Scup)";
    case kErrEncodeKmax: return "bit depth exceeds the 32-bit HT datapath (K_max > 30)";
  }
  return "device error";
}

int err_status(int32_t code) {
  switch (code) {
    case kErrNarrow: case kErrEncodeScup: return NDIC_INVALID_ARGUMENT;
    case kErrCsUnsupported: case kErrEncodeKmax: return NDIC_UNSUPPORTED;
    default: return NDIC_CODESTREAM;
  }
}

}  // namespace

// Per-stage device times of the decode pipeline (ndic_dev_profile).
struct StageProfile {
  bool on = false;
  double ms[8] = {0}, bytes[8] = {0};
  uint64_t calls = 0;
  double zfp_ms[ZfpClock::kGroups] = {0};
  uint64_t zfp_spans[ZfpClock::kGroups] = {0};
};
StageProfile g_prof;

ZfpClock::ZfpClock(hipStream_t s) : s_(s), on_(g_prof.on) {}

ZfpClock::~ZfpClock() {
  for (auto& sp : spans_) {
    (void)hipEventDestroy(sp.second.first);
    if (sp.second.second) (void)hipEventDestroy(sp.second.second);
  }
}

void ZfpClock::begin(Group g) {
  if (!on_) return;
  hipEvent_t a = nullptr;
  NDIC_HIP(hipEventCreate(&a));
  spans_.push_back({(int)g, {a, nullptr}});
  NDIC_HIP(hipEventRecord(a, s_));
}

void ZfpClock::end() {
  if (!on_ || spans_.empty() || spans_.back().second.second) return;
  hipEvent_t b = nullptr;
  NDIC_HIP(hipEventCreate(&b));
  spans_.back().second.second = b;
  NDIC_HIP(hipEventRecord(b, s_));
}

void ZfpClock::collect() {
  if (!on_) return;
  for (auto& sp : spans_) {
    if (!sp.second.second) continue;
    float ms = 0;
    NDIC_HIP(hipEventElapsedTime(&ms, sp.second.first, sp.second.second));
    g_prof.zfp_ms[sp.first] += ms;
    g_prof.zfp_spans[sp.first] += 1;
  }
  for (auto& sp : spans_) {
    (void)hipEventDestroy(sp.second.first);
    if (sp.second.second) (void)hipEventDestroy(sp.second.second);
  }
  spans_.clear();
}

namespace {

struct StageTimer {
  hipEvent_t ev[8];
  int n = 0, nsub = 0;
  hipStream_t s;
  explicit StageTimer(hipStream_t st) : s(st) {
    if (!g_prof.on) return;
    for (auto& e : ev) NDIC_HIP(hipEventCreate(&e));
  }
  void mark() {
    if (g_prof.on && n < 6) NDIC_HIP(hipEventRecord(ev[n++], s));
  }
  // inside the HT stage (marks 1..2): the ends of its VLC and MagSgn passes
  void sub() {
    if (g_prof.on && nsub < 2) NDIC_HIP(hipEventRecord(ev[6 + nsub++], s));
  }
  // after the stream is synchronised: stage i spans marks i..i+1; with the
  // two sub-marks, entries 5..7 split the HT stage into its VLC pass, its
  // MagSgn pass and the refinement passes
  void collect(const double* bytes) {
    if (!g_prof.on) return;
    auto add = [](int i, hipEvent_t a, hipEvent_t b, double by) {
      float ms = 0;
      NDIC_HIP(hipEventElapsedTime(&ms, a, b));
      g_prof.ms[i] += ms;
      g_prof.bytes[i] += by;
    };
    for (int i = 0; i + 1 < n; ++i) add(i, ev[i], ev[i + 1], bytes[i]);
    if (nsub == 2 && n > 2) {
      add(5, ev[1], ev[6], 0);
      add(6, ev[6], ev[7], 0);
      add(7, ev[7], ev[2], 0);
    }
    g_prof.calls += 1;
  }
  ~StageTimer() {
    if (!g_prof.on) return;
    for (auto& e : ev) (void)hipEventDestroy(e);
  }
};

uint64_t planes_of(const std::vector<size_t>& shape) {
  uint64_t p = 1;
  for (size_t i = 0; i + 2 < shape.size(); ++i) p *= shape[i];
  return p;
}

// Whether a chunk's finest synthesis level, its inverse nd_lift step and the
// narrowing store can run as one pass (dwt_lift_kernel): a lift-ht chunk
// decoded at full resolution with at least one DWT level, one lift step along
// a plane axis (not y or x) with at most two synthesis levels, and a dtype of
// at most 32 bits (int32 planes).
bool fusable(const DecodeItem& it, const DevChunk& c) {
  if (it.bare || it.irreversible || !it.lift || !it.d_out || it.steps.size() != 1 || c.P == 0) return false;
  if (c.L == 0 || c.max_res != (int32_t)c.L || (c.W == 1 && c.H == 1)) return false;
  if (lift_wide(it.out_dtype)) return false;
  const ndic_lift_step& st = it.steps[0];
  if ((size_t)st.dimension + 2 >= it.shape.size() || st.kind > 2) return false;
  const uint64_t len = it.shape[st.dimension];
  const uint64_t group = st.group == 0 ? len : std::min<uint64_t>(st.group, len);
  return st.kind == 0 || st.levels <= 2 || group <= 4;
}

// Device memory for one window of a decode batch.  Up to
// ht_decode_wave_max_blocks code-blocks a window decodes one block per
// wavefront (latency); above, one block per lane (throughput).

// Queues every pass of one window on s and returns the step that waits for
// it and fills in the items' results (status, message, decoded region).
std::function<void()> decode_window_begin(Context& ctx, const std::vector<DecodeItem*>& items, hipStream_t s) {
  ctx.release_kept();  // every earlier window of this context ended synchronised
  GeoTables geo;
  std::vector<DevChunk> dc(items.size());
  uint64_t nplanes = 0, ntasks = 0, coef = 0;
  uint32_t tag_stride = 16;
  for (size_t i = 0; i < items.size(); ++i) {
    DecodeItem& it = *items[i];
    DevChunk& c = dc[i];
    std::memset(&c, 0, sizeof(c));
    uint32_t W = (uint32_t)it.shape.back(), H = (uint32_t)it.shape[it.shape.size() - 2];
    const Geometry& g = geo.get(W, H, it.L);
    c.in = it.d_in;
    c.in_len = it.len;
    c.ndim = (uint32_t)it.shape.size();
    for (size_t d = 0; d < it.shape.size() && d < 32; ++d) c.dims[d] = (uint32_t)it.shape[d];
    c.W = W;
    c.H = H;
    c.P_all = (uint32_t)planes_of(it.shape);
    c.P = it.plane_sel.empty() ? c.P_all : (uint32_t)it.plane_sel.size();
    c.L = it.L;
    c.max_res = it.max_res < 0 ? (int32_t)it.L : std::min<int32_t>(it.max_res, it.L);
    c.geo_packet = g.packet_first;
    c.geo_band = g.band_first;
    c.blocks_per_plane = g.blocks;
    c.task_base = ntasks;
    c.plane_base = nplanes;
    c.no_container = (it.bare ? 1u : 0u) | (it.bare && it.prefix ? 2u : 0u);
    tag_stride = std::max(tag_stride, (g.tag_bytes + 15) & ~15u);
    ntasks += (uint64_t)c.P * g.blocks;
    nplanes += c.P;
    coef += (uint64_t)c.P * W * H;
  }
  size_t n = dc.size();
  // one allocation, carved
  auto al = [](uint64_t x) { return (x + 255) & ~uint64_t(255); };
  uint64_t sz_chunks = al(n * sizeof(DevChunk)), sz_gp = al(geo.packets.size() * sizeof(GeoPacket)),
           sz_gb = al(geo.bands.size() * sizeof(GeoBand)), sz_planes = al(nplanes * sizeof(DevPlane)),
           sz_tasks = al(ntasks * sizeof(HtDecTask)), sz_tag = al(nplanes * tag_stride), sz_err = al(n * 16),
           sz_ptr = al(n * sizeof(void*)), sz_dims = al(n * 16), sz_flag = al(n * 4);
  uint64_t nsel = 0;
  for (auto* it : items) nsel += it->plane_sel.size();
  const uint64_t sz_sel = al(nsel * 4);
  const bool lane_decode = (int64_t)ntasks > tune(kTuneHtDecodeWaveMaxBlocks);
  const bool split = lane_decode && tune(kTuneHtDecodeSplit) != 0;
  const bool sort = lane_decode && !split && tune(kTuneHtDecodeSort) != 0;
  // per-quad records of the two-pass decoder (allocated before any capture)
  uint32_t* d_rec = split ? (uint32_t*)ctx.htrec.ensure(ntasks * kHtRecBytes + 256) : nullptr;
  const uint64_t sz_order = sort ? al(ntasks * 4) + al(4 * (kHtClasses + 1)) : 0;
  uint64_t meta = sz_chunks + sz_gp + sz_gb + sz_planes + sz_tasks + sz_tag + sz_err + sz_ptr + sz_dims + sz_flag + sz_sel +
                  sz_order;
  uint8_t* m = (uint8_t*)ctx.work.ensure(meta);
  DevChunk* d_chunks = (DevChunk*)m;
  GeoPacket* d_gp = (GeoPacket*)(m + sz_chunks);
  GeoBand* d_gb = (GeoBand*)(m + sz_chunks + sz_gp);
  DevPlane* d_planes = (DevPlane*)(m + sz_chunks + sz_gp + sz_gb);
  HtDecTask* d_tasks = (HtDecTask*)(m + sz_chunks + sz_gp + sz_gb + sz_planes);
  uint8_t* d_tag = m + sz_chunks + sz_gp + sz_gb + sz_planes + sz_tasks;
  int32_t* d_err = (int32_t*)(d_tag + sz_tag);
  int32_t** d_ptr = (int32_t**)((uint8_t*)d_err + sz_err);
  uint64_t* d_dims = (uint64_t*)((uint8_t*)d_ptr + sz_ptr);
  int32_t* d_flag = (int32_t*)((uint8_t*)d_dims + sz_dims);
  uint32_t* d_sel = (uint32_t*)((uint8_t*)d_flag + sz_flag);
  uint32_t* d_order = sort ? (uint32_t*)((uint8_t*)d_sel + sz_sel) : nullptr;
  uint32_t* d_counts = sort ? d_order + al(ntasks * 4) / 4 : nullptr;
  if (nsel) {
    std::vector<uint32_t> sel;
    sel.reserve(nsel);
    for (size_t i = 0; i < n; ++i) {
      dc[i].plane_sel = items[i]->plane_sel.empty() ? nullptr : d_sel + sel.size();
      sel.insert(sel.end(), items[i]->plane_sel.begin(), items[i]->plane_sel.end());
    }
    ctx.up(d_sel, sel.data(), nsel * 4, s);
  }
  int32_t* A = (int32_t*)ctx.coef.ensure(al(3 * coef * 4 + 64));
  int32_t* T = A + coef;
  int32_t* O = A + 2 * coef;
  // items that may hold 9/7 codestreams: one IrrevPlane per plane, and a
  // fourth coefficient buffer for the 9/7 row pass
  bool any_irrev = false;
  for (auto* it : items) any_irrev |= it->irreversible;
  IrrevPlane* d_irrev = any_irrev ? (IrrevPlane*)ctx.irrev.ensure(al(nplanes * sizeof(IrrevPlane))) : nullptr;
  int32_t* S97 = any_irrev ? (int32_t*)ctx.coef97.ensure(al(coef * 4 + 64)) : nullptr;
  uint64_t co = 0;
  for (size_t i = 0; i < n; ++i) {
    dc[i].coef = A + co;
    dc[i].err = d_err + 4 * i;
    dc[i].tag = d_tag;
    dc[i].tag_stride = tag_stride;
    dc[i].irrev = items[i]->irreversible ? d_irrev + dc[i].plane_base : nullptr;
    co += (uint64_t)dc[i].P * dc[i].W * dc[i].H;
  }
  // With graph_windows the window's whole launch sequence, from here to the
  // status downloads, is captured into a graph and launched as one.
  const bool graph = tune(kTuneGraphWindows) != 0 && !g_prof.on;
  struct Capture {
    hipStream_t s = nullptr;
    ~Capture() {  // an exception mid-capture: end it, drop the partial graph
      if (!s) return;
      g_parked_frees = nullptr;
      hipGraph_t g = nullptr;
      if (hipStreamEndCapture(s, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
      (void)hipGetLastError();
    }
  } capture;
  if (graph) {
    if ((int64_t)ntasks <= tune(kTuneHtDecodeWaveMaxBlocks)) (void)ctx.queue(0);  // allocated outside the capture
    NDIC_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
    capture.s = s;
    g_parked_frees = &ctx.parked;
  }
  launch_fill32(d_planes, 0, sz_planes, s);
  if (d_irrev) launch_fill32(d_irrev, 0, al(nplanes * sizeof(IrrevPlane)), s);
  launch_fill32(d_err, 0, sz_err + sz_ptr + sz_dims + sz_flag, s);
  ctx.up(d_chunks, dc.data(), n * sizeof(DevChunk), s);
  ctx.up(d_gp, geo.packets.data(), geo.packets.size() * sizeof(GeoPacket), s);
  ctx.up(d_gb, geo.bands.data(), geo.bands.size() * sizeof(GeoBand), s);
  auto timer = std::make_shared<StageTimer>(s);
  timer->mark();
  launch_walk_chunks(d_chunks, (uint32_t)n, d_planes, s);
  launch_init_tasks(d_chunks, (uint32_t)n, ntasks, d_tasks, d_gp, d_gb, s);
  launch_parse_planes(d_chunks, (uint32_t)n, nplanes, d_planes, d_tasks, d_gp, d_gb, tag_stride, s);
  timer->mark();
  if (split) {
    launch_ht_vlc(d_tasks, ntasks, d_rec, s);
    timer->sub();
    launch_ht_magsgn(d_tasks, ntasks, d_rec, s);
    timer->sub();
  } else if (lane_decode) {
    if (sort) launch_ht_order(d_tasks, ntasks, d_counts, d_order, s);
    launch_ht_decode(d_tasks, d_order, ntasks, s);
  } else {
    launch_ht_decode_wave(d_tasks, ntasks, ctx.queue(0), s);
  }
  launch_ht_refine(d_tasks, ntasks, s);
  if (any_irrev) launch_dequantize(d_chunks, (uint32_t)n, nplanes, s);
  timer->mark();
  // fused only when the window has the work items to fill the GPU many
  // times over: each fused workgroup walks its tile's whole lift axis, so a
  // small window (a serving read of one chunk) runs faster as separate passes
  std::vector<char> fuse(n, 0);
  if (const int64_t min_items = tune(kTuneFuseDwtLift)) {
    uint64_t total = 0;
    for (size_t i = 0; i < n; ++i) {
      if (!(fuse[i] = fusable(*items[i], dc[i]))) continue;
      const ndic_lift_step& st = items[i]->steps[0];
      const uint64_t len = items[i]->shape[st.dimension];
      const uint64_t group = st.group == 0 ? len : std::min<uint64_t>(st.group, len);
      total += (uint64_t)dc[i].P / len * ceil_div(len, group) * ceil_div(dc[i].W, kDwtLiftTileW) *
               ceil_div(dc[i].H, kDwtLiftTileH);
    }
    if (total < (uint64_t)min_items) std::fill(fuse.begin(), fuse.end(), 0);
  }
  std::vector<DwtChunk> dw(n);
  for (size_t i = 0; i < n; ++i) {
    DwtChunk& w = dw[i];
    w.skip_finest = fuse[i] != 0;
    w.A = dc[i].coef;
    uint64_t off = (uint64_t)(dc[i].coef - A);
    w.T = T + off;
    w.O = O + off;
    w.nplanes = dc[i].P;
    w.stride = dc[i].W;
    w.plane_elems = (uint64_t)dc[i].W * dc[i].H;
    uint32_t lv = (uint32_t)dc[i].max_res;
    level_dims(dc[i].W, dc[i].H, dc[i].L - lv, &w.W, &w.H);
    w.levels = lv;
    if (dc[i].irrev) {
      w.irrev = dc[i].irrev;
      w.S97 = S97 + off;
    }
    items[i]->out_w = w.W;
    items[i]->out_h = w.H;
  }
  run_dwt_inverse(ctx, dw, s);
  timer->mark();
  std::vector<int32_t*> res(n);
  std::vector<uint64_t> dims(2 * n);
  for (size_t i = 0; i < n; ++i) {
    res[i] = fuse[i] ? nullptr : dw[i].result;  // fused chunks un-shift in dwt_lift_kernel
    items[i]->d_result = dw[i].result;
    dims[2 * i] = dw[i].W;
    dims[2 * i + 1] = dw[i].H;
  }
  ctx.up(d_ptr, res.data(), n * sizeof(void*), s);
  ctx.up(d_dims, dims.data(), n * 16, s);
  if (any_irrev) launch_irrev_round(d_chunks, (uint32_t)n, nplanes, d_planes, d_ptr, d_dims, s);
  launch_unshift(d_chunks, (uint32_t)n, nplanes, d_planes, d_ptr, d_dims, s);
  timer->mark();
  // outputs
  std::vector<LiftChunk> lifts;
  std::vector<DwtLiftJob> fused;
  std::vector<size_t> fused_of(n, SIZE_MAX);
  uint64_t fused_items = 0;
  std::vector<WindowJob> windows;
  std::vector<size_t> window_of(n, SIZE_MAX);
  uint64_t window_rows = 0;
  for (size_t i = 0; i < n; ++i) {
    DecodeItem& it = *items[i];
    uint64_t total = (uint64_t)dc[i].P * dc[i].W * dc[i].H;
    if (fuse[i]) {
      const ndic_lift_step& st = it.steps[0];
      DwtLiftJob f{};
      f.A = dc[i].coef;
      f.LL = dw[i].result;  // the reconstruction of level 1 (A when there is one level)
      f.planes = d_planes + dc[i].plane_base;
      f.dst = it.d_out;
      f.flag = d_flag + i;
      f.w = dc[i].W;
      f.h = dc[i].H;
      f.plane_elems = (uint64_t)f.w * f.h;
      f.tiles_x = (uint32_t)ceil_div(f.w, kDwtLiftTileW);
      f.tiles = f.tiles_x * (uint32_t)ceil_div(f.h, kDwtLiftTileH);
      f.len = it.shape[st.dimension];
      f.outer = 1;
      for (size_t d = 0; d < st.dimension; ++d) f.outer *= it.shape[d];
      f.R = 1;
      for (size_t d = st.dimension + 1; d + 2 < it.shape.size(); ++d) f.R *= it.shape[d];
      f.group = st.group == 0 ? f.len : std::min<uint64_t>(st.group, f.len);
      f.segs = ceil_div(f.len, f.group);
      f.kind = st.kind;
      f.levels = st.levels;
      f.dtype = (uint32_t)it.out_dtype;
      f.dsize = (uint32_t)dtype_size(it.out_dtype);
      f.items = f.outer * f.segs * f.R * f.tiles;
      f.item_begin = fused_items;
      fused_items += f.items;
      if (f.items) {
        fused_of[i] = fused.size();
        fused.push_back(f);
      }
      continue;
    }
    if (it.window) {
      WindowJob wj{};
      wj.src = res[i];
      wj.dst = it.d_out;
      wj.flag = d_flag + i;
      wj.plane_elems = (uint64_t)dc[i].W * dc[i].H;
      wj.W = dc[i].W;
      wj.x0 = it.win[0];
      wj.y0 = it.win[1];
      wj.w = it.win[2];
      wj.h = it.win[3];
      wj.planes = dc[i].P;
      wj.dtype = (uint32_t)it.ht_dtype;
      wj.dsize = (uint32_t)dtype_size(it.ht_dtype);
      wj.items = (uint64_t)wj.planes * wj.h;
      wj.item_begin = window_rows;
      window_rows += wj.items;
      window_of[i] = windows.size();
      if (wj.items && wj.w) windows.push_back(wj);
      continue;
    }
    if (it.bare) {
      // plane decodes hand back the int32 region: into the caller's buffer
      // when one is given (the window's planes are reused afterwards)
      if (it.d_out && it.out_w && it.out_h)
        NDIC_HIP(hipMemcpy2DAsync(it.d_out, it.out_w * 4, res[i], (size_t)dc[i].W * 4, it.out_w * 4, it.out_h,
                                  hipMemcpyDeviceToDevice, s));
      continue;
    }
    if (it.lift) {
      LiftChunk lc;
      lc.plane = res[i];
      lc.tmp = (res[i] == O + (dc[i].coef - A)) ? (void*)(T + (dc[i].coef - A)) : (void*)(O + (dc[i].coef - A));
      lc.dst = it.d_out;
      lc.flag = d_flag + i;
      lc.shape = it.shape;
      lc.dtype = it.out_dtype;
      lc.steps = it.steps;
      lifts.push_back(lc);
    } else {
      launch_convert(nullptr, 0, res[i], false, it.d_out, it.ht_dtype, total, nullptr, d_flag + i, s);
    }
  }
  if (!fused.empty()) {
    // (the host table stays alive until the window's synchronisation below)
    DwtLiftJob* d_fused = (DwtLiftJob*)ctx.fusedjobs.ensure(fused.size() * sizeof(DwtLiftJob));
    ctx.up(d_fused, fused.data(), fused.size() * sizeof(DwtLiftJob), s);
    launch_dwt_lift(d_fused, (uint32_t)fused.size(), fused_items, s);
  }
  if (!lifts.empty()) run_lift(ctx, lifts, false, s);
  if (!windows.empty()) {
    WindowJob* d_win = (WindowJob*)ctx.winjobs.ensure(windows.size() * sizeof(WindowJob));
    ctx.up(d_win, windows.data(), windows.size() * sizeof(WindowJob), s);
    launch_window(d_win, (uint32_t)windows.size(), window_rows, s);
  }
  timer->mark();
  const int32_t* p_err = (const int32_t*)ctx.down(d_err, n * 16, s);
  const int32_t* p_flag = (const int32_t*)ctx.down(d_flag, n * 4, s);
  if (graph) {
    g_parked_frees = nullptr;
    capture.s = nullptr;
    hipGraph_t g = nullptr;
    NDIC_HIP(hipStreamEndCapture(s, &g));
    ctx.launch_graph(g, s);
  }
  return [&ctx, s, items, n, p_err, p_flag, coef, timer, dc = std::move(dc), dw = std::move(dw),
          fuse = std::move(fuse), fused = std::move(fused), fused_of = std::move(fused_of),
          windows = std::move(windows), window_of = std::move(window_of),
          lifts = std::move(lifts), res = std::move(res)]() {
  NDIC_HIP(hipStreamSynchronize(s));
  std::vector<int32_t> herr(p_err, p_err + 4 * n), hflag(p_flag, p_flag + n);
  if (g_prof.on) {
    // algorithmic bytes per stage: parse reads the streams; HT decode reads
    // the segments and writes 4 B per coefficient; the DWT reads and writes
    // 4 B per sample; the lift reads 4 B and writes the dtype per sample
    // (a fused chunk's finest level counts in the last stage, which then reads
    // 4 B per sample once and writes the dtype)
    double in = 0, dwt = 0, lift = 0;
    for (size_t i = 0; i < n; ++i) {
      in += (double)items[i]->len;
      double sm = (double)dc[i].P * dw[i].W * dw[i].H;
      for (uint32_t l = fuse[i] ? 1 : 0; l < dw[i].levels; ++l) {
        uint64_t lw, lh;
        level_dims(dw[i].W, dw[i].H, l, &lw, &lh);
        if (lw * lh > 1) dwt += 8.0 * (double)dc[i].P * (double)(lw * lh);
      }
      if (!items[i]->bare) lift += sm * (4 + dtype_size(items[i]->lift ? items[i]->out_dtype : items[i]->ht_dtype));
    }
    double bytes[5] = {in, in + 4 * (double)coef, dwt, 0, lift};
    timer->collect(bytes);
  }
  for (size_t i = 0; i < n; ++i) {
    DecodeItem& it = *items[i];
    if (herr[4 * i]) {
      it.status = err_status(herr[4 * i]);
      it.offset = (uint64_t)(uint32_t)herr[4 * i + 1] | ((uint64_t)(uint32_t)herr[4 * i + 2] << 32);
      std::ostringstream msg;
      msg << err_text(herr[4 * i]);
      if (herr[4 * i] != kErrCsUnsupported) msg << " (plane " << (herr[4 * i + 3] - (int64_t)dc[i].plane_base) << ")";
      it.message = msg.str();
    } else if (hflag[i]) {
      // error path: the actual range, for the reference's message
      int64_t* mm = (int64_t*)ctx.small.ensure(64);
      int64_t init[2] = {INT64_MAX, INT64_MIN}, h[2];
      uint64_t total = (uint64_t)dc[i].P * dc[i].W * dc[i].H;
      int32_t* src = it.lift ? nullptr : res[i];
      if (it.lift && fused_of[i] == SIZE_MAX) {
        // the narrowing happened inside the lift: recompute on the lifted plane
        LiftChunk lc = lifts[0];
        for (auto& l : lifts)
          if (l.dst == it.d_out) lc = l;
        lift_complete_plane(ctx, lc, s);
        src = (int32_t*)lc.plane;
      }
      NDIC_HIP(hipMemcpyAsync(mm, init, 16, hipMemcpyHostToDevice, s));
      if (window_of[i] != SIZE_MAX) {
        // the window that was narrowed, again in range mode
        WindowJob wj = windows[window_of[i]];
        wj.dst = nullptr;
        wj.minmax = mm;
        wj.item_begin = 0;
        WindowJob* d_w = (WindowJob*)ctx.winjobs.ensure(sizeof(WindowJob));
        NDIC_HIP(hipMemcpyAsync(d_w, &wj, sizeof(wj), hipMemcpyHostToDevice, s));
        launch_window(d_w, 1, wj.items, s);
      } else if (fused_of[i] != SIZE_MAX) {
        // no lifted plane was written: the fused pass again, in range mode
        DwtLiftJob f = fused[fused_of[i]];
        f.dst = nullptr;
        f.minmax = mm;
        f.item_begin = 0;
        DwtLiftJob* d_f = (DwtLiftJob*)ctx.fusedjobs.ensure(sizeof(DwtLiftJob));
        NDIC_HIP(hipMemcpyAsync(d_f, &f, sizeof(f), hipMemcpyHostToDevice, s));
        launch_dwt_lift(d_f, 1, f.items, s);
      } else {
        launch_convert(nullptr, 0, src, false, nullptr, 0, total, mm, nullptr, s);
      }
      NDIC_HIP(hipMemcpyAsync(h, mm, 16, hipMemcpyDeviceToHost, s));
      NDIC_HIP(hipStreamSynchronize(s));
      std::ostringstream msg;
      if (it.lift) {
        it.status = NDIC_INVALID_ARGUMENT;
        msg << "nd_lift decode: coefficient range [" << h[0] << ", " << h[1] << "] does not narrow back to "
            << dtype_name(it.out_dtype) << " (corrupt or mismatched chunk)";
      } else {
        it.status = NDIC_CODESTREAM;
        msg << "decoded sample range [" << h[0] << ", " << h[1] << "] does not fit dtype " << dtype_name(it.ht_dtype)
            << " (corrupt or mismatched chunk)";
      }
      it.message = msg.str();
    }
  }
  };
}

}  // namespace

// Windows take turns over ctx (on s) and its twin contexts (own buffers and
// streams), so that one window's parse and entropy decode overlap another's
// synthesis and lift; a window's results are collected once the windows after
// it are queued.  Under the stage profiler, or with pipeline_windows = 0, windows run
// one after the other.
void decode_items(Context& ctx, std::vector<DecodeItem>& items, hipStream_t s, uint64_t window_elems) {
  ht_init_tables();
  std::vector<std::vector<DecodeItem*>> wins(1);
  uint64_t acc = 0;
  for (auto& it : items) {
    uint64_t e = 1;
    for (size_t d : it.shape) e *= d;
    if (!wins.back().empty() && acc + e > window_elems) {
      wins.emplace_back();
      acc = 0;
    }
    wins.back().push_back(&it);
    acc += e;
  }
  if (wins.back().empty()) return;
  if (wins.size() == 1 || g_prof.on || !tune(kTunePipelineWindows)) {
    for (auto& w : wins) decode_window_begin(ctx, w, s)();
    return;
  }
  // pipeline_windows = n keeps n + 1 windows in flight (at most four), each
  // on its own context: ctx, its twin, the twin's twin ...
  const size_t depth = 1 + std::min<size_t>(tune(kTunePipelineWindows), 3);
  std::vector<Context*> cs{&ctx};
  while (cs.size() < depth) cs.push_back(&cs.back()->twin());
  hipEvent_t ready;
  NDIC_HIP(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
  std::vector<std::function<void()>> pending(depth);
  auto close_all = [&] {
    for (size_t j = 1; j < depth; ++j) cs[j]->close_stream();
  };
  try {
    NDIC_HIP(hipEventRecord(ready, s));  // the other contexts start after what s holds
    for (size_t j = 1; j < depth; ++j) {
      cs[j]->open_stream();
      NDIC_HIP(hipStreamWaitEvent(cs[j]->stream, ready, 0));
    }
    for (size_t k = 0; k < wins.size() + depth; ++k) {
      auto& slot = pending[k % depth];
      if (slot) {
        auto f = std::move(slot);
        slot = nullptr;
        f();
      }
      if (k < wins.size()) {
        Context& c = *cs[k % depth];
        slot = decode_window_begin(c, wins[k], k % depth ? c.stream : s);
      }
    }
  } catch (...) {
    (void)hipStreamSynchronize(s);
    close_all();
    (void)hipEventDestroy(ready);
    throw;
  }
  close_all();
  NDIC_HIP(hipEventDestroy(ready));
}

namespace {

void encode_window(Context& ctx, std::vector<EncodeItem*>& items, uint8_t* d_out, uint64_t capacity,
                   uint64_t* out_cursor, hipStream_t s) {
  ctx.release_kept();  // every earlier window of this context ended synchronised
  GeoTables geo;
  size_t n = items.size();
  std::vector<DevChunk> dc(n);
  uint64_t nplanes = 0, ntasks = 0, coef = 0, maxL = 0;
  uint32_t tag_stride = 16;
  for (size_t i = 0; i < n; ++i) {
    EncodeItem& it = *items[i];
    DevChunk& c = dc[i];
    std::memset(&c, 0, sizeof(c));
    uint32_t W = (uint32_t)it.shape.back(), H = (uint32_t)it.shape[it.shape.size() - 2];
    const Geometry& g = geo.get(W, H, it.L);
    c.ndim = (uint32_t)it.shape.size();
    for (size_t d = 0; d < it.shape.size(); ++d) c.dims[d] = (uint32_t)it.shape[d];
    c.W = W;
    c.H = H;
    c.P = (uint32_t)planes_of(it.shape);
    c.L = it.L;
    c.max_res = (int32_t)it.L;
    c.geo_packet = g.packet_first;
    c.geo_band = g.band_first;
    c.blocks_per_plane = g.blocks;
    c.task_base = ntasks;
    c.plane_base = nplanes;
    c.progression = it.progression;
    c.index = it.index ? 1 : 0;
    c.sample_signed = it.lift ? 1 : (dtype_signed(it.in_dtype) ? 1 : 0);
    tag_stride = std::max(tag_stride, (g.tag_bytes + 15) & ~15u);
    ntasks += (uint64_t)c.P * g.blocks;
    nplanes += c.P;
    coef += (uint64_t)c.P * W * H;
    maxL = std::max<uint64_t>(maxL, it.L);
  }
  auto al = [](uint64_t x) { return (x + 255) & ~uint64_t(255); };
  const uint64_t slot_stride = al(ht_ms_cap(64, 64, 31) + kMelCap + kVlcCap);
  uint64_t pkt_stride_eff = 0;
  const uint32_t head_stride = (uint32_t)al(256 + 6 * (maxL + 1) + 16);
  uint64_t chead_stride = 0;
  for (auto& c : dc) chead_stride = std::max<uint64_t>(chead_stride, 8 + 4 * c.ndim + (uint64_t)c.P * (12 + 4 * (c.L + 1)));
  chead_stride = al(chead_stride);
  // per-packet header bound: every block of the largest packet
  uint64_t max_blocks_pkt = 1;
  for (auto& p : geo.packets) max_blocks_pkt = std::max<uint64_t>(max_blocks_pkt, p.nblocks);
  pkt_stride_eff = al(max_blocks_pkt * 16 + 64);
  uint64_t sz_chunks = al(n * sizeof(DevChunk)), sz_gp = al(geo.packets.size() * sizeof(GeoPacket)),
           sz_gb = al(geo.bands.size() * sizeof(GeoBand)), sz_planes = al(nplanes * sizeof(DevPlane)),
           sz_tasks = al(ntasks * sizeof(HtEncTask)), sz_lens = al(ntasks * 16), sz_tag = al(nplanes * tag_stride),
           sz_err = al(n * 16), sz_clen = al(n * 8), sz_head = al(nplanes * head_stride),
           sz_pkt = al(nplanes * (maxL + 1) * pkt_stride_eff), sz_chead = al(n * chead_stride),
           sz_flag = al(n * 4 + 64);
  uint64_t meta = sz_chunks + sz_gp + sz_gb + sz_planes + sz_tasks + sz_lens + sz_tag + sz_err + sz_clen + sz_head +
                  sz_pkt + sz_chead + sz_flag;
  uint8_t* m = (uint8_t*)ctx.work.ensure(meta);
  uint8_t* cur = m;
  auto take = [&](uint64_t sz) {
    uint8_t* p = cur;
    cur += sz;
    return p;
  };
  DevChunk* d_chunks = (DevChunk*)take(sz_chunks);
  GeoPacket* d_gp = (GeoPacket*)take(sz_gp);
  GeoBand* d_gb = (GeoBand*)take(sz_gb);
  DevPlane* d_planes = (DevPlane*)take(sz_planes);
  HtEncTask* d_tasks = (HtEncTask*)take(sz_tasks);
  uint32_t* d_lens = (uint32_t*)take(sz_lens);
  uint8_t* d_tag = take(sz_tag);
  int32_t* d_err = (int32_t*)take(sz_err);
  uint64_t* d_clen = (uint64_t*)take(sz_clen);
  uint8_t* d_head = take(sz_head);
  uint8_t* d_pkt = take(sz_pkt);
  uint8_t* d_chead = take(sz_chead);
  int32_t* d_flag = (int32_t*)take(sz_flag);
  uint8_t* d_slots = (uint8_t*)ctx.tmp.ensure(al(ntasks * slot_stride + 64));
  int32_t* X = (int32_t*)ctx.coef.ensure(al(3 * coef * 4 + 64));
  int32_t* T = X + coef;
  int32_t* A = X + 2 * coef;
  uint64_t co = 0;
  for (size_t i = 0; i < n; ++i) {
    dc[i].planes_in = X + co;
    dc[i].coef = A + co;
    dc[i].err = d_err + 4 * i;
    dc[i].tag = d_tag;
    dc[i].tag_stride = tag_stride;
    dc[i].slots = d_slots;
    dc[i].head = d_head;
    dc[i].head_stride = head_stride;
    dc[i].pkt_hdr = d_pkt;
    dc[i].pkt_stride = (uint32_t)pkt_stride_eff;
    dc[i].pkt_slots = (uint32_t)(maxL + 1);
    dc[i].chunk_head = d_chead + i * chead_stride;
    dc[i].chunk_head_len = 0;
    co += (uint64_t)dc[i].P * dc[i].W * dc[i].H;
  }
  NDIC_HIP(hipMemsetAsync(d_planes, 0, sz_planes, s));
  NDIC_HIP(hipMemsetAsync(d_err, 0, sz_err, s));
  NDIC_HIP(hipMemsetAsync(d_flag, 0, sz_flag, s));
  // ---- samples -> int32 planes (nd_lift forward, or widening)
  std::vector<int64_t> mm_h(2 * n);
  {
    // per-chunk min/max for the lift budget / u32 range check
    int64_t* mm = (int64_t*)ctx.small.ensure(al(n * 16 + 64));
    std::vector<int64_t> init(2 * n);
    for (size_t i = 0; i < n; ++i) {
      init[2 * i] = INT64_MAX;
      init[2 * i + 1] = INT64_MIN;
    }
    ctx.up(mm, init.data(), n * 16, s);
    for (size_t i = 0; i < n; ++i) {
      uint64_t total = (uint64_t)dc[i].P * dc[i].W * dc[i].H;
      launch_convert(items[i]->d_in, items[i]->in_dtype, dc[i].planes_in, false, nullptr, 0, total, mm + 2 * i,
                     d_flag + i, s);
    }
    const int64_t* p_mm = (const int64_t*)ctx.down(mm, n * 16, s);
    const int32_t* p_fl = (const int32_t*)ctx.down(d_flag, n * 4, s);
    NDIC_HIP(hipStreamSynchronize(s));
    std::copy(p_mm, p_mm + 2 * n, mm_h.begin());
    std::vector<int32_t> fl(p_fl, p_fl + n);
    std::vector<LiftChunk> lifts;
    for (size_t i = 0; i < n; ++i) {
      EncodeItem& it = *items[i];
      if (fl[i]) {
        it.status = NDIC_INVALID_ARGUMENT;
        it.message = it.lift ? "nd_lift overflow budget: input value does not fit the widened i32 coefficient plane"
                             : "uint32 sample exceeds the 32-bit HT datapath; decorrelate or reduce the dynamic range";
        continue;
      }
      if (!it.lift) continue;
      try {
        uint64_t total = (uint64_t)dc[i].P * dc[i].W * dc[i].H;
        if (total) lift_check_budget(mm_h[2 * i], mm_h[2 * i + 1], it.shape, it.steps, false);
      } catch (const Failure& f) {
        it.status = f.status;
        it.message = f.message;
        continue;
      }
      LiftChunk lc;
      lc.plane = dc[i].planes_in;
      lc.tmp = T + (dc[i].planes_in - X);
      lc.shape = it.shape;
      lc.dtype = it.in_dtype;
      lc.steps = it.steps;
      lifts.push_back(lc);
    }
    if (!lifts.empty()) run_lift(ctx, lifts, true, s);
  }
  ctx.up(d_chunks, dc.data(), n * sizeof(DevChunk), s);
  ctx.up(d_gp, geo.packets.data(), geo.packets.size() * sizeof(GeoPacket), s);
  ctx.up(d_gb, geo.bands.data(), geo.bands.size() * sizeof(GeoBand), s);
  launch_plane_minmax(d_chunks, (uint32_t)n, nplanes, d_planes, s);
  launch_plane_setup(d_chunks, (uint32_t)n, nplanes, d_planes, s);
  launch_dc_shift(d_chunks, (uint32_t)n, nplanes, d_planes, s);
  std::vector<DwtChunk> dw(n);
  for (size_t i = 0; i < n; ++i) {
    DwtChunk& w = dw[i];
    w.X = dc[i].planes_in;
    w.A = dc[i].coef;
    w.T = T + (dc[i].planes_in - X);
    w.nplanes = dc[i].P;
    w.W = dc[i].W;
    w.H = dc[i].H;
    w.stride = dc[i].W;
    w.plane_elems = (uint64_t)dc[i].W * dc[i].H;
    w.levels = dc[i].L;
  }
  run_dwt_forward(ctx, dw, s);
  // chunks whose DWT wrote nothing (no level applied) code X directly
  bool reupload = false;
  for (size_t i = 0; i < n; ++i)
    if (dw[i].result != dc[i].coef) {
      dc[i].coef = dw[i].result;
      reupload = true;
    }
  if (reupload) ctx.up(d_chunks, dc.data(), n * sizeof(DevChunk), s);
  launch_build_enc_tasks(d_chunks, (uint32_t)n, nplanes, d_planes, d_gp, d_gb, d_tasks, d_lens, slot_stride, s);
  if ((int64_t)ntasks > tune(kTuneHtEncodeWaveMaxBlocks)) launch_ht_encode(d_tasks, ntasks, s);
  else launch_ht_encode_wave(d_tasks, ntasks, ctx.queue(1), s);
  launch_write_planes(d_chunks, (uint32_t)n, nplanes, d_planes, d_tasks, d_lens, d_gp, d_gb, s);
  launch_write_chunks(d_chunks, (uint32_t)n, d_planes, d_clen, s);
  const uint64_t* p_clen = (const uint64_t*)ctx.down(d_clen, n * 8, s);
  const int32_t* p_err = (const int32_t*)ctx.down(d_err, n * 16, s);
  NDIC_HIP(hipStreamSynchronize(s));
  std::vector<uint64_t> clen(p_clen, p_clen + n);
  std::vector<int32_t> herr(p_err, p_err + 4 * n);
  for (size_t i = 0; i < n; ++i) {
    EncodeItem& it = *items[i];
    if (it.status == NDIC_OK && herr[4 * i]) {
      it.status = err_status(herr[4 * i]);
      std::ostringstream msg;
      msg << "htj2k plane " << (herr[4 * i + 3] - (int64_t)dc[i].plane_base) << ": " << err_text(herr[4 * i]);
      it.message = msg.str();
    }
    dc[i].chunk_head_len = 8 + 4ull * dc[i].ndim + (dc[i].index ? (uint64_t)dc[i].P * (12 + 4 * (dc[i].L + 1)) : 0);
    if (it.status != NDIC_OK) {
      dc[i].out_off = 0;
      it.out_len = 0;
      continue;
    }
    if (*out_cursor + clen[i] > capacity) {
      it.status = NDIC_INVALID_ARGUMENT;
      it.message = "encode output capacity exceeded";
      continue;
    }
    dc[i].out_off = *out_cursor;
    it.out_off = *out_cursor;
    it.out_len = clen[i];
    *out_cursor += clen[i];
  }
  // refused chunks assemble nothing (out_len 0)
  for (size_t i = 0; i < n; ++i) dc[i].out_len = items[i]->status == NDIC_OK ? items[i]->out_len : 0;
  ctx.up(d_chunks, dc.data(), n * sizeof(DevChunk), s);
  launch_assemble(d_chunks, (uint32_t)n, nplanes, d_planes, d_tasks, d_lens, d_gp, d_out, s);
  NDIC_HIP(hipStreamSynchronize(s));
}

}  // namespace

void encode_items(Context& ctx, std::vector<EncodeItem>& items, uint8_t* d_out, uint64_t capacity, hipStream_t s,
                  uint64_t window_elems) {
  ht_init_tables();
  uint64_t cursor = 0;
  std::vector<EncodeItem*> win;
  uint64_t acc = 0;
  for (auto& it : items) {
    uint64_t e = 1;
    for (size_t d : it.shape) e *= d;
    if (!win.empty() && acc + e > window_elems) {
      encode_window(ctx, win, d_out, capacity, &cursor, s);
      win.clear();
      acc = 0;
    }
    win.push_back(&it);
    acc += e;
  }
  if (!win.empty()) encode_window(ctx, win, d_out, capacity, &cursor, s);
}

}  // namespace ndic

extern "C" void ndic_dev_profile(int32_t enable) {
  ndic::g_prof = ndic::StageProfile();
  ndic::g_prof.on = enable != 0;
}

extern "C" size_t ndic_dev_zfp_times(double* ms, uint64_t* spans, size_t n) {
  size_t k = n < (size_t)ndic::ZfpClock::kGroups ? n : (size_t)ndic::ZfpClock::kGroups;
  for (size_t i = 0; i < k; ++i) {
    ms[i] = ndic::g_prof.zfp_ms[i];
    spans[i] = ndic::g_prof.zfp_spans[i];
  }
  return k;
}

extern "C" size_t ndic_dev_stage_times(double* ms, double* bytes, size_t n) {
  size_t k = n < 8 ? n : 8;
  for (size_t i = 0; i < k; ++i) {
    ms[i] = ndic::g_prof.ms[i];
    bytes[i] = ndic::g_prof.bytes[i];
  }
  return (size_t)ndic::g_prof.calls;
}
