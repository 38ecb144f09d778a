// nd_lift orchestration and its C-ABI entry points.
//   config/validation: crates/ndic-lift/src/lib.rs:122-242
//   budget:            crates/ndic-lift/src/chunk.rs:161-302
//   bytes cores:       crates/ndic-zarr/src/lift.rs:246-346
#include <algorithm>
#include <cstring>
#include <functional>
#include <sstream>
#include <thread>

#include "engine.hpp"

namespace ndic {

thread_local std::vector<void*>* g_parked_frees = nullptr;

void* DevBuf::ensure(size_t bytes) {
  if (bytes <= cap && ptr) return ptr;
  if (g_parked_frees && ptr) {
    g_parked_frees->push_back(ptr);
    ptr = nullptr;
    cap = 0;
  }
  release();
  // grow to the next size of the form m * 2^k (8 <= m < 16): at most 1/8
  // over, and a slowly growing sequence of requests reallocates (and
  // synchronises the device in hipFree) only a logarithmic number of times
  size_t want = bytes < 256 ? 256 : bytes;
  int k = 63 - __builtin_clzll((unsigned long long)want) - 3;
  if (k > 0) want = ((want + ((size_t)1 << k) - 1) >> k) << k;
  NDIC_HIP(hipMalloc(&ptr, want));
  cap = want;
  return ptr;
}

void DevBuf::release() {
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  cap = 0;
}

void* PinnedArena::take(size_t n) {
  n = (n + 63) & ~(size_t)63;
  if (off + n > cap) {
    if (base) retired.push_back(base);
    size_t want = std::max(cap * 2, n + ((size_t)1 << 20));
    base = nullptr;
    cap = 0;
    off = 0;
    NDIC_HIP(hipHostMalloc((void**)&base, want, hipHostMallocDefault));
    cap = want;
  }
  void* r = base + off;
  off += n;
  return r;
}

void PinnedArena::reset() {
  for (char* p : retired) (void)hipHostFree(p);
  retired.clear();
  off = 0;
}

void Context::release_kept() {
  pin.reset();
  for (void* p : parked) (void)hipFree(p);
  parked.clear();
}

void Context::launch_graph(hipGraph_t g, hipStream_t s) {
  size_t nn = 0;
  NDIC_HIP(hipGraphGetNodes(g, nullptr, &nn));
  std::vector<hipGraphNode_t> nodes(nn);
  if (nn) NDIC_HIP(hipGraphGetNodes(g, nodes.data(), &nn));
  uint64_t sig = 1469598103934665603ull ^ nn;
  for (hipGraphNode_t nd : nodes) {
    hipGraphNodeType t;
    NDIC_HIP(hipGraphNodeGetType(nd, &t));
    sig = (sig ^ (uint64_t)t) * 1099511628211ull;
    if (t == hipGraphNodeTypeKernel) {
      hipKernelNodeParams kp;
      NDIC_HIP(hipGraphKernelNodeGetParams(nd, &kp));
      sig = (sig ^ (uint64_t)(uintptr_t)kp.func) * 1099511628211ull;
    }
  }
  hipGraphExec_t exec = nullptr;
  for (size_t i = 0; i < graphs.size(); ++i) {
    if (graphs[i].sig != sig) continue;
    hipGraphNode_t err_node = nullptr;
    hipGraphExecUpdateResult res;
    if (hipGraphExecUpdate(graphs[i].exec, g, &err_node, &res) == hipSuccess && res == hipGraphExecUpdateSuccess) {
      exec = graphs[i].exec;
    } else {
      (void)hipGetLastError();
      (void)hipGraphExecDestroy(graphs[i].exec);
      graphs.erase(graphs.begin() + i);
    }
    break;
  }
  if (!exec) {
    NDIC_HIP(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0));
    if (graphs.size() >= 8) {
      (void)hipGraphExecDestroy(graphs.front().exec);
      graphs.erase(graphs.begin());
    }
    graphs.push_back({sig, exec});
  }
  (void)hipGraphDestroy(g);
  NDIC_HIP(hipGraphLaunch(exec, s));
}

void Context::up(void* d, const void* h, size_t n, hipStream_t s) {
  if (!n) return;
  const void* p = stage_keep(h, n);
  NDIC_HIP(hipMemcpyAsync(d, p, n, hipMemcpyHostToDevice, s));
}

void* Context::down(const void* d, size_t n, hipStream_t s) {
  void* p = pin.take(n);
  if (n) NDIC_HIP(hipMemcpyAsync(p, d, n, hipMemcpyDeviceToHost, s));
  return p;
}

// One context (stream + reusable device buffers) per host thread: calls
// from different threads run on their own streams and overlap on the GPU;
// calls from one thread are ordered on its stream.  A thread's context is
// released when the thread exits (its buffers are reused by the thread's
// later calls until then); the thread that loaded the library keeps its
// context until the process ends, when the runtime may already be gone.
namespace {
const std::thread::id g_loader_thread = std::this_thread::get_id();

void destroy_context(Context* c) {
  if (c->twin_ctx) destroy_context(c->twin_ctx);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto& g : c->graphs) (void)hipGraphExecDestroy(g.exec);
  for (void* p : c->parked) (void)hipFree(p);
  for (DevBuf* b : {&c->in, &c->out, &c->plane, &c->tmp, &c->coef, &c->work, &c->jobs, &c->dwtjobs, &c->fusedjobs,
                    &c->winjobs, &c->small, &c->htrec, &c->irrev, &c->coef97})
    b->release();
  if (c->queues) (void)hipFree(c->queues);
  c->pin.reset();
  if (c->pin.base) (void)hipHostFree(c->pin.base);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

Context* new_context() {
  auto* c = new Context();
  NDIC_HIP(hipGetDevice(&c->device));
  NDIC_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  return c;
}

struct ContextOwner {
  Context* c = nullptr;
  ~ContextOwner() {
    if (!c || std::this_thread::get_id() == g_loader_thread) return;
    destroy_context(c);
  }
};
}  // namespace

uint64_t* Context::queue(int k) {
  if (!queues) {
    // zeroed before any stream (the caller's may not be this one) uses them
    NDIC_HIP(hipMalloc((void**)&queues, 256));
    NDIC_HIP(hipMemsetAsync(queues, 0, 256, stream));
    NDIC_HIP(hipStreamSynchronize(stream));
  }
  return queues + 2 * k;
}

Context& Context::twin() {
  if (!twin_ctx) {
    twin_ctx = new Context();
    twin_ctx->device = device;
  }
  return *twin_ctx;
}

void Context::open_stream() {
  if (!stream) NDIC_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
}

void Context::close_stream() {
  if (!stream) return;
  (void)hipStreamSynchronize(stream);
  (void)hipStreamDestroy(stream);
  stream = nullptr;
}

Context& Context::get() {
  static thread_local ContextOwner owner;
  if (!owner.c) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
      fail(NDIC_IO, "no HIP device is visible: the nd-image-codecs MI355X path needs a GPU");
    owner.c = new_context();
  }
  return *owner.c;
}

bool lift_wide(int dtype) { return dtype == NDIC_U64 || dtype == NDIC_I64; }

namespace {

const char* kind_name(uint32_t k) { return k == 0 ? "delta" : k == 1 ? "haar" : k == 2 ? "lift53" : "?"; }

bool parse_u32(const std::string& s, uint32_t* v) {
  if (s.empty() || s.size() > 10) return false;
  uint64_t acc = 0;
  for (char c : s) {
    if (c < '0' || c > '9') return false;
    acc = acc * 10 + uint64_t(c - '0');
  }
  if (acc > 0xFFFFFFFFull) return false;
  *v = (uint32_t)acc;
  return true;
}

typedef __int128 i128;

std::string i128_str(i128 v) {
  if (v == 0) return "0";
  bool neg = v < 0;
  unsigned __int128 u = neg ? (unsigned __int128)(-(v + 1)) + 1 : (unsigned __int128)v;
  std::string s;
  while (u) {
    s.push_back(char('0' + int(u % 10)));
    u /= 10;
  }
  if (neg) s.push_back('-');
  std::reverse(s.begin(), s.end());
  return s;
}

uint32_t band_count(uint64_t n, uint32_t levels) {
  uint32_t c = 0;
  uint64_t m = n;
  for (uint32_t l = 0; l < levels && m >= 2; ++l) {
    ++c;
    m = (m + 1) / 2;
  }
  return c;
}

i128 floor_div(i128 a, i128 b) {
  i128 q = a / b;
  if (a % b < 0) q -= 1;
  return q;
}

struct Iv {
  i128 lo, hi;
};

}  // namespace

void lift_validate(const char* version, const ndic_lift_step* steps, size_t nsteps, size_t ndim) {
  std::string v = version ? version : "";
  size_t dot = v.find('.');
  uint32_t ma = 0, mi = 0;
  bool ok = dot != std::string::npos && parse_u32(v.substr(0, dot), &ma);
  if (ok) {
    size_t dot2 = v.find('.', dot + 1);
    ok = parse_u32(v.substr(dot + 1, dot2 == std::string::npos ? std::string::npos : dot2 - dot - 1), &mi);
  }
  if (!ok || ma != 0 || mi != 1)
    fail(NDIC_UNSUPPORTED, "nd_lift configuration version \"" + v +
                               "\" is not supported by this build (implements 0.1); refusing rather than mis-decoding");
  for (size_t i = 0; i < nsteps; ++i) {
    if (steps[i].kind > 2) fail(NDIC_INVALID_ARGUMENT, "nd_lift transform kind must be delta, haar or lift53");
    if ((steps[i].kind == 1 || steps[i].kind == 2) && steps[i].levels == 0)
      fail(NDIC_INVALID_ARGUMENT, std::string("nd_lift transform: kind \"") + kind_name(steps[i].kind) +
                                      "\" needs levels >= 1");
    if (steps[i].levels > 255) fail(NDIC_INVALID_ARGUMENT, "nd_lift transform levels must fit a u8");
  }
  for (size_t i = 0; i < nsteps; ++i)
    if (steps[i].dimension >= ndim) {
      std::ostringstream m;
      m << "nd_lift transform: dimension " << steps[i].dimension << " out of range for a " << ndim
        << "-dimensional chunk";
      fail(NDIC_INVALID_ARGUMENT, m.str());
    }
}

void lift_check_budget(int64_t lo64, int64_t hi64, const std::vector<size_t>& shape,
                       const std::vector<ndic_lift_step>& steps, bool wide) {
  i128 pmin = wide ? (i128)INT64_MIN : (i128)INT32_MIN, pmax = wide ? (i128)INT64_MAX : (i128)INT32_MAX;
  auto fits = [&](Iv v) { return v.lo >= pmin && v.hi <= pmax; };
  auto uni = [](Iv a, Iv b) { return Iv{std::min(a.lo, b.lo), std::max(a.hi, b.hi)}; };
  Iv all{lo64, hi64};
  for (auto& s : steps) {
    size_t len = s.dimension < shape.size() ? shape[s.dimension] : 0;
    if (len < 2) continue;
    size_t group = s.group == 0 ? len : std::min<size_t>(s.group, len);
    uint32_t levels = s.kind == 0 ? 1 : band_count(group, s.levels);
    Iv input = all, acc = all, band = all;
    bool ok = true;
    for (uint32_t l = 0; l < levels && ok; ++l) {
      i128 lo = band.lo, hi = band.hi;
      Iv next, level_all;
      if (s.kind == 0) {
        Iv r{lo - hi, hi - lo};
        next = input;
        level_all = uni(r, input);
      } else if (s.kind == 1) {
        Iv d{lo - hi, hi - lo};
        Iv sm{lo + floor_div(d.lo, 2), hi + floor_div(d.hi, 2)};
        next = sm;
        level_all = uni(sm, d);
      } else {
        Iv ps{2 * lo + 1, 2 * hi + 1};
        Iv d{lo - hi, hi - lo};
        Iv us{2 * d.lo + 2, 2 * d.hi + 2};
        if (!fits(ps) || !fits(us)) {
          ok = false;
          break;
        }
        Iv sm{lo + floor_div(us.lo, 4), hi + floor_div(us.hi, 4)};
        next = sm;
        level_all = uni(sm, d);
      }
      acc = uni(acc, level_all);
      if (!fits(acc)) ok = false;
      band = next;
    }
    if (!ok) {
      std::ostringstream m;
      m << "nd_lift overflow budget exceeded on axis " << s.dimension << " (" << std::max<uint32_t>(1, s.levels)
        << " levels of \"" << kind_name(s.kind) << "\" over input range [" << i128_str(all.lo) << ", "
        << i128_str(all.hi) << "]): coefficients would leave the " << (wide ? 64 : 32)
        << "-bit plane; reduce levels or shorten the group";
      fail(NDIC_INVALID_ARGUMENT, m.str());
    }
    all = acc;
  }
}

void run_lift(Context& ctx, std::vector<LiftChunk>& chunks, bool fwd, hipStream_t stream) {
  struct Launch {
    size_t first, count;
    uint64_t items;
    bool wide;
    uint32_t max_n;
    bool col;
  };
  std::vector<LiftJob> all;
  std::vector<Launch> launches;
  size_t rounds = 1;
  for (auto& c : chunks) rounds = std::max(rounds, c.steps.size());
  for (int wide = 0; wide < 2; ++wide) {
    for (size_t r = 0; r < rounds; ++r) {
      std::vector<LiftJob> lds_jobs, glob_jobs;
      uint32_t lds_max = 1, glob_max = 1;
      for (auto& c : chunks) {
        if (lift_wide(c.dtype) != (wide == 1)) continue;
        uint64_t total = 1;
        for (size_t d : c.shape) total *= d;
        if (total == 0) continue;
        size_t ns = c.steps.size();
        LiftJob j{};
        j.plane = c.plane;
        j.tmp = c.tmp;
        j.flag = c.flag;
        j.dtype = (uint32_t)c.dtype;
        j.dsize = (uint32_t)dtype_size(c.dtype);
        const ndic_lift_step* st = nullptr;
        bool active = ns == 0 ? r == 0 : r < ns;
        if (!active) continue;
        if (ns) st = &c.steps[fwd ? r : ns - 1 - r];
        bool first = fwd ? r == 0 : (ns == 0 || ns - 1 - r == 0);
        if (fwd && first) j.src = c.src;
        if (!fwd && first) j.dst = c.dst;
        size_t len = st ? c.shape[st->dimension] : 1;
        if (!st || len < 2) {
          // nothing to lift along this axis: widen/narrow only (or skip)
          if (!j.src && !j.dst) continue;
          j.kind = 3;
          j.outer = 1;
          j.len = 1;
          j.inner = total;
          j.group = 1;
        } else {
          uint64_t inner = 1;
          for (size_t d = st->dimension + 1; d < c.shape.size(); ++d) inner *= c.shape[d];
          j.kind = st->kind;
          j.levels = st->levels;
          j.len = len;
          j.inner = inner;
          j.outer = total / (len * inner);
          j.group = st->group == 0 ? len : std::min<uint64_t>(st->group, len);
        }
        j.tiles = ceil_div(j.inner, 64);
        j.segs = ceil_div(j.len, j.group);
        j.items = j.outer * j.segs * j.tiles;
        uint32_t n = (uint32_t)std::min<uint64_t>(j.group, 0xFFFFFFFFu);
        if (n <= lift_lds_limit(wide)) {
          lds_jobs.push_back(j);
          lds_max = std::max(lds_max, n);
        } else {
          if (!j.tmp && j.kind != 0 && j.kind != 3) fail(NDIC_IO, "internal: missing lift temporary");
          glob_jobs.push_back(j);
          glob_max = std::max(glob_max, n);
        }
      }
      // LDS jobs go out in segment-length classes so that short segments
      // are not launched with the LDS (and occupancy) of the longest one
      (void)lds_max;
      // (inverse int32 segments of at most two synthesis levels go to the
      // register-resident kernel, the rest to the LDS tile kernel)
      const uint32_t classes[] = {4, 8, 16, 32, 64, 128};
      uint32_t lo = 0;
      for (uint32_t cap : classes) {
        std::vector<LiftJob> jobs[2];
        for (auto& j : lds_jobs) {
          uint32_t n = (uint32_t)std::min<uint64_t>(j.group, 0xFFFFFFFFu);
          if (n <= lo || n > cap) continue;
          bool col = !fwd && !wide && cap <= lift_col_limit() &&
                     (j.kind == 0 || j.kind == 3 || j.levels <= 2 || n <= 4);
          jobs[col].push_back(j);
        }
        lo = cap;
        for (int col = 0; col < 2; ++col) {
          if (jobs[col].empty()) continue;
          uint64_t items = 0;
          for (auto& j : jobs[col]) {
            j.item_begin = items;
            items += j.items;
          }
          launches.push_back({all.size(), jobs[col].size(), items, wide == 1, cap, col == 1});
          all.insert(all.end(), jobs[col].begin(), jobs[col].end());
        }
      }
      if (!glob_jobs.empty()) {
        uint64_t items = 0;
        for (auto& j : glob_jobs) {
          j.item_begin = items;
          items += j.items;
        }
        launches.push_back({all.size(), glob_jobs.size(), items, wide == 1, glob_max, false});
        all.insert(all.end(), glob_jobs.begin(), glob_jobs.end());
      }
    }
  }
  if (all.empty()) return;
  // the host job table outlives the asynchronous upload in the context's
  // kept list; callers synchronise before the job buffer is reused
  LiftJob* d_jobs = (LiftJob*)ctx.jobs.ensure(all.size() * sizeof(LiftJob));
  ctx.up(d_jobs, all.data(), all.size() * sizeof(LiftJob), stream);
  for (auto& l : launches) launch_lift_step(d_jobs + l.first, (uint32_t)l.count, l.items, fwd, l.wide, l.max_n, l.col, stream);
}

void lift_complete_plane(Context& ctx, const LiftChunk& c, hipStream_t stream) {
  if (c.steps.empty()) return;  // the narrowing read the plane as it is
  LiftChunk last = c;
  last.dst = nullptr;
  last.steps.assign(1, c.steps[0]);
  std::vector<LiftChunk> cs{last};
  run_lift(ctx, cs, false, stream);
}

namespace {

void set_error(ndic_error* err, const Failure& f) {
  if (!err) return;
  err->status = f.status;
  err->offset = f.offset;
  std::strncpy(err->message, f.message.c_str(), sizeof(err->message) - 1);
  err->message[sizeof(err->message) - 1] = 0;
}

}  // namespace

ndic_status guard_call(ndic_error* err, const std::function<void()>& body) {
  try {
    body();
    if (err) {
      err->status = NDIC_OK;
      err->offset = 0;
      err->message[0] = 0;
    }
    return NDIC_OK;
  } catch (const Failure& f) {
    // work this call queued may still read its pinned staging: drain it
    (void)hipDeviceSynchronize();
    set_error(err, f);
    return (ndic_status)f.status;
  } catch (const std::bad_alloc&) {
    (void)hipDeviceSynchronize();
    set_error(err, Failure{NDIC_IO, 0, "host allocation failed"});
    return NDIC_IO;
  }
}

void emit_host(const void* d_src, size_t n, ndic_buffer* out, hipStream_t stream) {
  out->data = (uint8_t*)std::malloc(n ? n : 1);
  if (!out->data) fail(NDIC_IO, "host allocation failed");
  out->len = n;
  if (n) {
    NDIC_HIP(hipMemcpyAsync(out->data, d_src, n, hipMemcpyDeviceToHost, stream));
    NDIC_HIP(hipStreamSynchronize(stream));
  }
}

namespace {

ndic_status lift_chunk_call(bool fwd, const uint8_t* bytes, size_t len, const size_t* shape, size_t ndim,
                            ndic_dtype dtype, const char* version, const ndic_lift_step* steps, size_t nsteps,
                            ndic_buffer* out, ndic_error* err) {
  return guard_call(err, [&] {
    if (!out) fail(NDIC_INVALID_ARGUMENT, "nd_lift: null output buffer");
    out->data = nullptr;
    out->len = 0;
    if (dtype == NDIC_F32 || dtype == NDIC_F64 || dtype < 0 || dtype > NDIC_F64)
      fail(NDIC_UNSUPPORTED, "nd_lift: floating-point dtypes are not supported");
    lift_validate(version, steps, nsteps, ndim);
    std::vector<size_t> sh(shape, shape + ndim);
    uint64_t n = 1;
    for (size_t d : sh) {
      if (d != 0 && n > UINT64_MAX / 8 / d) fail(NDIC_INVALID_ARGUMENT, "nd_lift: chunk shape overflows usize");
      n *= d;
    }
    bool wide = lift_wide(dtype);
    size_t psz = wide ? 8 : 4, isz = dtype_size(dtype);
    if (fwd && len != n * isz) {
      std::ostringstream m;
      m << "nd_lift encode: got " << len << " bytes for " << n << " elements of " << isz << " bytes";
      fail(NDIC_INVALID_ARGUMENT, m.str());
    }
    if (!fwd && len != n * psz) {
      std::ostringstream m;
      m << "nd_lift decode: got " << len << " bytes for " << n << " coefficients of " << psz << " bytes";
      fail(NDIC_INVALID_ARGUMENT, m.str());
    }
    Context& ctx = Context::get();
    std::lock_guard<std::recursive_mutex> lock(ctx.mu);
    ctx.release_kept();  // earlier calls on this context ended synchronised
    hipStream_t s = ctx.stream;
    uint8_t* d_in = (uint8_t*)ctx.in.ensure(len + 16);
    void* d_plane = ctx.plane.ensure(n * psz + 16);
    uint8_t* d_out = fwd ? nullptr : (uint8_t*)ctx.out.ensure(n * isz + 16);
    int32_t* d_small = (int32_t*)ctx.small.ensure(64);
    NDIC_HIP(hipMemsetAsync(d_small, 0, 64, s));
    if (len) NDIC_HIP(hipMemcpyAsync(d_in, bytes, len, hipMemcpyHostToDevice, s));
    std::vector<ndic_lift_step> st(steps, steps + nsteps);
    bool need_tmp = false;
    for (auto& x : st) {
      size_t l = sh[x.dimension];
      size_t g = x.group == 0 ? l : std::min<size_t>(x.group, l);
      need_tmp |= g > lift_lds_limit(wide) && x.kind != 0;
    }
    void* d_tmp = need_tmp ? ctx.tmp.ensure(n * psz + 16) : nullptr;
    LiftChunk c;
    c.shape = sh;
    c.dtype = dtype;
    c.steps = st;
    c.tmp = d_tmp;
    c.flag = d_small;
    if (fwd) {
      // widen with min/max first: the budget needs the input range
      int64_t* mm = (int64_t*)(d_small + 4);
      int64_t init[2] = {INT64_MAX, INT64_MIN};
      NDIC_HIP(hipMemcpyAsync(mm, init, 16, hipMemcpyHostToDevice, s));
      launch_convert(d_in, dtype, d_plane, wide, nullptr, 0, n, mm, d_small, s);
      int64_t hmm[2];
      int32_t flag = 0;
      NDIC_HIP(hipMemcpyAsync(hmm, mm, 16, hipMemcpyDeviceToHost, s));
      NDIC_HIP(hipMemcpyAsync(&flag, d_small, 4, hipMemcpyDeviceToHost, s));
      NDIC_HIP(hipStreamSynchronize(s));
      if (flag) {
        uint64_t mx = 0;
        for (uint64_t i = 0; i < n; ++i) {
          uint64_t v = 0;
          std::memcpy(&v, bytes + i * isz, isz);
          mx = std::max(mx, v);
        }
        std::ostringstream m;
        m << "nd_lift overflow budget: input value " << mx << " does not fit the widened "
          << (wide ? "i64" : "i32") << " coefficient plane";
        fail(NDIC_INVALID_ARGUMENT, m.str());
      }
      if (n) lift_check_budget(hmm[0], hmm[1], sh, st, wide);
      c.plane = d_plane;
      std::vector<LiftChunk> cs{c};
      run_lift(ctx, cs, true, s);
      emit_host(d_plane, n * psz, out, s);
    } else {
      NDIC_HIP(hipMemcpyAsync(d_plane, d_in, len, hipMemcpyDeviceToDevice, s));
      c.plane = d_plane;
      c.dst = d_out;
      std::vector<LiftChunk> cs{c};
      run_lift(ctx, cs, false, s);
      int32_t flag = 0;
      NDIC_HIP(hipMemcpyAsync(&flag, d_small, 4, hipMemcpyDeviceToHost, s));
      NDIC_HIP(hipStreamSynchronize(s));
      if (flag) {
        // error path: the range of the reconstructed plane, for the message
        int64_t* mm = (int64_t*)(d_small + 4);
        int64_t init[2] = {INT64_MAX, INT64_MIN}, hmm[2];
        lift_complete_plane(ctx, c, s);
        NDIC_HIP(hipMemcpyAsync(mm, init, 16, hipMemcpyHostToDevice, s));
        launch_convert(nullptr, 0, d_plane, wide, nullptr, 0, n, mm, nullptr, s);
        NDIC_HIP(hipMemcpyAsync(hmm, mm, 16, hipMemcpyDeviceToHost, s));
        NDIC_HIP(hipStreamSynchronize(s));
        std::ostringstream m;
        m << "nd_lift decode: coefficient range [" << hmm[0] << ", " << hmm[1] << "] does not narrow back to "
          << dtype_name(dtype) << " (corrupt or mismatched chunk)";
        fail(NDIC_INVALID_ARGUMENT, m.str());
      }
      emit_host(d_out, n * isz, out, s);
    }
  });
}

}  // namespace

}  // namespace ndic

extern "C" {

void ndic_buffer_free(ndic_buffer* buf) {
  if (!buf) return;
  std::free(buf->data);
  buf->data = nullptr;
  buf->len = 0;
}

const char* ndic_version(void) { return "nd-image-codecs-amd 0.1.0 (gfx950)"; }

ndic_status ndic_lift_forward_chunk(const uint8_t* bytes, size_t len, const size_t* shape, size_t ndim,
                                    ndic_dtype dtype, const char* version, const ndic_lift_step* steps,
                                    size_t nsteps, ndic_buffer* out, ndic_error* err) {
  return ndic::lift_chunk_call(true, bytes, len, shape, ndim, dtype, version, steps, nsteps, out, err);
}

ndic_status ndic_lift_inverse_chunk(const uint8_t* bytes, size_t len, const size_t* shape, size_t ndim,
                                    ndic_dtype dtype, const char* version, const ndic_lift_step* steps,
                                    size_t nsteps, ndic_buffer* out, ndic_error* err) {
  return ndic::lift_chunk_call(false, bytes, len, shape, ndim, dtype, version, steps, nsteps, out, err);
}

}  // extern "C"
