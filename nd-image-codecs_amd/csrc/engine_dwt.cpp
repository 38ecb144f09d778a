// 5/3 DWT orchestration: per-level launches over every chunk of a batch.
// Level geometry: crates/ndic-htj2k/src/dwt/mod.rs:33-43, 133-198.
#include <algorithm>
#include <cstring>

#include "dwt.hpp"
#include "engine.hpp"

namespace ndic {

void level_dims(uint64_t w, uint64_t h, uint32_t level, uint64_t* ow, uint64_t* oh) {
  for (uint32_t i = 0; i < level; ++i) {
    w = (w + 1) / 2;
    h = (h + 1) / 2;
  }
  *ow = w;
  *oh = h;
}

namespace {

struct PassLaunch {
  int pass;
  size_t first, count;
  uint64_t items;
};

void submit(Context& ctx, std::vector<DwtJob>& all, const std::vector<PassLaunch>& launches, hipStream_t stream) {
  if (all.empty()) return;
  // the job table buffer is reused by the next submit of this context: the
  // caller synchronises before that (each window ends with a sync)
  DwtJob* d = (DwtJob*)ctx.dwtjobs.ensure(all.size() * sizeof(DwtJob));
  ctx.up(d, all.data(), all.size() * sizeof(DwtJob), stream);
  for (auto& l : launches) launch_dwt_pass(l.pass, d + l.first, (uint32_t)l.count, l.items, stream);
}

void add_launch(std::vector<DwtJob>& all, std::vector<PassLaunch>& launches, std::vector<DwtJob>& jobs, int pass) {
  if (jobs.empty()) return;
  uint64_t items = 0;
  for (auto& j : jobs) {
    j.item_begin = items;
    items += j.items;
  }
  launches.push_back({pass, all.size(), jobs.size(), items});
  all.insert(all.end(), jobs.begin(), jobs.end());
}

}  // namespace

// Synthesis, one fused launch per level (dwt_inv_level).  A level reads its
// LL rows from the previous level's reconstruction and writes the next one,
// so reconstructions alternate between O and T (the coarsest level reads LL
// from A); result names the buffer the finest level wrote.  Chunks that may
// hold 9/7 planes add the two 9/7 passes of each level (rows into S97,
// columns into the same destination), which take only those planes while
// dwt_inv_level skips them.
void run_dwt_inverse(Context& ctx, std::vector<DwtChunk>& chunks, hipStream_t stream) {
  uint32_t max_l = 0;
  for (auto& c : chunks) {
    max_l = std::max(max_l, c.levels);
    c.result = c.A;
  }
  std::vector<DwtJob> all;
  std::vector<PassLaunch> launches;
  for (uint32_t l = max_l; l-- > 0;) {
    std::vector<DwtJob> jobs, hjobs, vjobs;
    for (auto& c : chunks) {
      if (c.levels <= l || c.nplanes == 0 || (l == 0 && c.skip_finest)) continue;
      uint64_t w, h;
      level_dims(c.W, c.H, l, &w, &h);
      if (w == 1 && h == 1) continue;
      DwtJob j{};
      j.stride = c.stride ? c.stride : c.W;
      j.plane_elems = c.plane_elems ? c.plane_elems : c.W * c.H;
      j.w = (uint32_t)w;
      j.h = (uint32_t)h;
      j.hl = (uint32_t)((h + 1) / 2);
      j.src = c.A;
      j.src_ll = c.result;
      j.dst = c.result == c.O ? c.T : c.O;
      j.tiles_x = (uint32_t)ceil_div(w, 64);
      j.tiles = j.tiles_x * ceil_div(h, 64);
      j.items = c.nplanes * j.tiles;
      j.irrev = c.irrev;
      jobs.push_back(j);
      if (c.irrev) {
        // the chunk's 9/7 planes: rows A (+ LL) -> S97, columns S97 -> dst
        DwtJob hj = j, vj = j;
        hj.dst = c.S97;
        hj.tiles = ceil_div((w + 1) / 2, 64);
        hj.items = c.nplanes * h * hj.tiles;
        vj.src = c.S97;
        vj.tiles = ceil_div(w, 64);
        vj.items = c.nplanes * ((h + 1) / 2) * vj.tiles;
        hjobs.push_back(hj);
        vjobs.push_back(vj);
      }
      c.result = j.dst;
    }
    add_launch(all, launches, jobs, kDwtInvLevel);
    add_launch(all, launches, hjobs, kDwt97InvH);
    add_launch(all, launches, vjobs, kDwt97InvV);
  }
  submit(ctx, all, launches, stream);
}

void run_dwt_forward(Context& ctx, std::vector<DwtChunk>& chunks, hipStream_t stream) {
  uint32_t max_l = 0;
  std::vector<bool> done(chunks.size(), false);
  for (auto& c : chunks) {
    max_l = std::max(max_l, c.levels);
    c.result = c.X;
  }
  std::vector<DwtJob> all;
  std::vector<PassLaunch> launches;
  for (uint32_t l = 0; l < max_l; ++l) {
    std::vector<DwtJob> vj, hj;
    for (size_t ci = 0; ci < chunks.size(); ++ci) {
      auto& c = chunks[ci];
      if (c.levels <= l || c.nplanes == 0 || done[ci]) continue;
      uint64_t w, h;
      level_dims(c.W, c.H, l, &w, &h);
      if (w == 1 && h == 1) {
        done[ci] = true;
        continue;
      }
      DwtJob v{};
      v.stride = c.stride ? c.stride : c.W;
      v.plane_elems = c.plane_elems ? c.plane_elems : c.W * c.H;
      v.w = (uint32_t)w;
      v.h = (uint32_t)h;
      v.src = c.result;  // X at level 0, then the LL of A
      v.dst = c.T;
      v.tiles = ceil_div(w, 64);
      v.items = c.nplanes * ((h + 1) / 2) * v.tiles;
      vj.push_back(v);
      DwtJob j{};
      j.stride = c.stride ? c.stride : c.W;
      j.plane_elems = c.plane_elems ? c.plane_elems : c.W * c.H;
      j.w = (uint32_t)w;
      j.h = (uint32_t)h;
      j.src = c.T;
      j.dst = c.A;
      j.tiles = ceil_div((w + 1) / 2, 64);
      j.items = c.nplanes * h * j.tiles;
      hj.push_back(j);
      c.result = c.A;
    }
    add_launch(all, launches, vj, kDwtFwdV);
    add_launch(all, launches, hj, kDwtFwdH);
  }
  submit(ctx, all, launches, stream);
}

}  // namespace ndic

extern "C" ndic_status ndic_dev_dwt53(int32_t forward, int32_t* plane, size_t width, size_t height, uint32_t levels,
                                      ndic_error* err) {
  return ndic::guard_call(err, [&] {
    using namespace ndic;
    if (levels > 32) fail(NDIC_INVALID_ARGUMENT, "at most 32 decomposition levels");
    Context& ctx = Context::get();
    std::lock_guard<std::recursive_mutex> lock(ctx.mu);
    hipStream_t s = ctx.stream;
    size_t n = width * height;
    int32_t* base = (int32_t*)ctx.coef.ensure(3 * n * 4 + 64);
    DwtChunk c{};
    c.nplanes = n ? 1 : 0;
    c.W = width;
    c.H = height;
    c.levels = levels;
    c.A = base;
    c.T = base + n;
    c.O = base + 2 * n;
    c.X = base + 2 * n;
    NDIC_HIP(hipMemcpyAsync(forward ? c.X : c.A, plane, n * 4, hipMemcpyHostToDevice, s));
    std::vector<DwtChunk> cs{c};
    if (forward) run_dwt_forward(ctx, cs, s);
    else run_dwt_inverse(ctx, cs, s);
    NDIC_HIP(hipMemcpyAsync(plane, cs[0].result, n * 4, hipMemcpyDeviceToHost, s));
    NDIC_HIP(hipStreamSynchronize(s));
    ctx.release_kept();
  });
}
