// Host-side runtime of the library: the device context (stream + reusable
// device buffers) and the orchestration of the codec kernels.
#pragma once

#include <cstring>
#include <functional>
#include <mutex>
#include <utility>
#include <vector>

#include "common.hpp"
#include "lift.hpp"

namespace ndic {

// A device allocation that grows on demand and is reused across calls.
struct DevBuf {
  void* ptr = nullptr;
  size_t cap = 0;
  void* ensure(size_t bytes);
  template <class T> T* as() const { return (T*)ptr; }
  void release();
};

// While a window's launch sequence is being captured into a HIP graph a
// buffer that must grow cannot free its old allocation (a free synchronises
// the device): DevBuf::ensure parks it here instead, and the context frees it
// once the window has completed.
extern thread_local std::vector<void*>* g_parked_frees;

// Pinned host staging for the small tables a call uploads (descriptors, job
// lists) and the statuses it reads back: copies from pinned memory are
// queued without the host waiting, where a copy from pageable memory would
// wait for the stream to drain.  Space is handed out by a bump pointer and
// reclaimed by reset() at the start of the next window: every window ends
// with its stream synchronised, and a call that fails midway drains the
// device before it returns (guard_call).
struct PinnedArena {
  char* base = nullptr;
  size_t cap = 0, off = 0;
  std::vector<char*> retired;  // outgrown blocks, freed at reset()
  void* take(size_t n);
  void reset();
};

struct Context {
  std::recursive_mutex mu;
  hipStream_t stream = nullptr;
  int device = 0;
  DevBuf in, out, plane, tmp, coef, work, jobs, dwtjobs, fusedjobs, winjobs, small, htrec, irrev, coef97;
  uint64_t* queues = nullptr;  // work-queue counter pairs of the persistent kernels (zeroed)
  // Counter pair k (0: HT decode, 1: HT encode) of this context's queues.
  uint64_t* queue(int k);
  PinnedArena pin;
  // A pinned copy of p[0..n) for an upload queued on the stream; valid until
  // release_kept().
  const void* stage_keep(const void* p, size_t n) {
    void* h = pin.take(n);
    if (n) std::memcpy(h, p, n);
    return h;
  }
  // Queue an upload of host bytes through pinned staging.
  void up(void* d, const void* h, size_t n, hipStream_t s);
  // Queue a download into pinned staging; the bytes are there once s is
  // synchronised.
  void* down(const void* d, size_t n, hipStream_t s);
  // Reclaims the staging and the allocations parked during a capture; only
  // between windows (nothing of them in flight).
  std::vector<void*> parked;
  void release_kept();
  // Instantiated decode-window graphs, keyed by their node sequence (node
  // types and kernel functions); a new capture with the same sequence
  // updates one in place instead of instantiating again.
  struct CachedGraph {
    uint64_t sig;
    hipGraphExec_t exec;
  };
  std::vector<CachedGraph> graphs;
  // Launches graph g on s through the cache (g is consumed).
  void launch_graph(hipGraph_t g, hipStream_t s);
  // A second context of this one (same device, own buffers and staging) for
  // a window queued while one of this context is in flight; created on first
  // use, used under this context's lock.  Its stream exists only while a
  // batch uses it (open_stream / close_stream): HIP maps a process's streams
  // onto a few hardware queues, and an idle extra stream there changes which
  // of the other threads' streams share one.
  Context* twin_ctx = nullptr;
  Context& twin();
  void open_stream();
  void close_stream();
  static Context& get();
};

// One chunk of a lift launch sequence.
struct LiftChunk {
  void* plane = nullptr;   // coefficient plane (int32 or int64)
  void* tmp = nullptr;     // chunk-sized temporary, for segments too long for LDS
  const uint8_t* src = nullptr;
  uint8_t* dst = nullptr;
  int32_t* flag = nullptr;
  std::vector<size_t> shape;
  int dtype = NDIC_U8;
  std::vector<ndic_lift_step> steps;
};

// One chunk of a DWT launch sequence: nplanes planes of W x H, each buffer
// holding them back to back.
struct IrrevPlane;
struct DwtChunk {
  int32_t* X = nullptr;   // forward input (spatial planes)
  int32_t* A = nullptr;   // Mallat coefficients
  int32_t* T = nullptr;   // pass temporary
  int32_t* O = nullptr;   // inverse output
  int32_t* result = nullptr;  // set by the run: where the result planes are
  uint64_t nplanes = 0, W = 0, H = 0;  // region transformed (top-left of each plane)
  uint64_t stride = 0;                 // plane row pitch (0: W)
  uint64_t plane_elems = 0;            // plane spacing (0: W * H)
  uint32_t levels = 0;
  bool skip_finest = false;  // inverse: stop before the finest level (result = its LL source)
  // inverse: planes whose codestream is 9/7 (IrrevPlane::on, one entry per
  // plane) take the f32 9/7 synthesis instead, through S97 (a buffer laid
  // out like A); null: every plane is 5/3
  const IrrevPlane* irrev = nullptr;
  int32_t* S97 = nullptr;
};
void level_dims(uint64_t w, uint64_t h, uint32_t level, uint64_t* ow, uint64_t* oh);
void run_dwt_forward(Context& ctx, std::vector<DwtChunk>& chunks, hipStream_t stream);
void run_dwt_inverse(Context& ctx, std::vector<DwtChunk>& chunks, hipStream_t stream);

// Device time of the ZFP coders' kernels while the stage profiler is on
// (ndic_dev_profile): HIP events around each group of launches, summed per
// group once the stream has synchronised (collect).  Does nothing when the
// profiler is off.
class ZfpClock {
 public:
  enum Group { kEncode, kPlace, kDecode, kWalk, kGroups };
  explicit ZfpClock(hipStream_t s);
  ~ZfpClock();
  void begin(Group g);
  void end();
  void collect();

 private:
  hipStream_t s_;
  bool on_;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> spans_;
};

// Runs body, mapping a Failure to the C-ABI status and error record.
ndic_status guard_call(ndic_error* err, const std::function<void()>& body);
// Copies n device bytes into a fresh host ndic_buffer.
void emit_host(const void* d_src, size_t n, ndic_buffer* out, hipStream_t stream);

struct DecodeItem;
struct EncodeItem;
// Decode / encode batches of chunks on the device (engine_htj2k.cpp).
// window_elems bounds the samples whose intermediate planes are resident at
// once.
void decode_items(Context& ctx, std::vector<DecodeItem>& items, hipStream_t s, uint64_t window_elems);
// Samples per batch window: the window_elems knob, else NDIC_WINDOW_ELEMS,
// else 2^31.
uint64_t window_elems_knob();
void encode_items(Context& ctx, std::vector<EncodeItem>& items, uint8_t* d_out, uint64_t capacity, hipStream_t s,
                  uint64_t window_elems);

bool lift_wide(int dtype);
// Validates a configuration against a chunk dimensionality the way
// NdLiftConfig::validate does (crates/ndic-lift/src/lib.rs:122-184).
void lift_validate(const char* version, const ndic_lift_step* steps, size_t nsteps, size_t ndim);
// Budget check of the encode direction on the input's [lo, hi]
// (crates/ndic-lift/src/chunk.rs:188-302); throws the reference's refusal.
void lift_check_budget(int64_t lo, int64_t hi, const std::vector<size_t>& shape,
                       const std::vector<ndic_lift_step>& steps, bool wide);
// Runs every step of every chunk, forward (widening src into plane first)
// or inverse (narrowing plane into dst last).
void run_lift(Context& ctx, std::vector<LiftChunk>& chunks, bool fwd, hipStream_t stream);
// After an inverse run_lift whose last step narrowed straight into dst: runs
// that step again in place (no dst), so the plane holds the lifted samples the
// narrowing saw (the range an out-of-range error reports).
void lift_complete_plane(Context& ctx, const LiftChunk& c, hipStream_t stream);

}  // namespace ndic
