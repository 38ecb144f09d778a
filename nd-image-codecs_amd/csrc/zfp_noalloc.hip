// The allocation-free form of the ZFP brick batch (ndic_zfp_decode_bricks):
// for callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc).  The
// descriptor tables and stream-header bytes the batch needs on the device
// live in one file-scope array, `scratch`, of a fixed size; the call uploads
// into it from host memory with hipMemcpyAsync, launches the same kernels as
// ndic_zfp_decode_bricks and calls no hipMalloc, hipFree, pinned-host
// allocation or other allocating routine.
//
// Each GPU holds its own copy of the array (until the process ends) at its
// own address, which the library looks up once per device with
// hipGetSymbolAddress, with that device current, at the first call on it.
// Every call on a GPU shares that GPU's array: the library serialises the
// calls on one device (a call returns after its kernels complete), so calls
// made from several host threads never overlap in it; a caller that reaches
// the same device's array by other means would race with them.
#include <cstring>
#include <mutex>
#include <vector>

#include "engine.hpp"
#include "zfp.hpp"

namespace ndic {
namespace {

constexpr size_t al256c(size_t x) { return (x + 255) & ~size_t(255); }

// Layout: [ZfpHead x max][header bytes x max][per class: ZfpChunk x k | ZfpBrick x k]
constexpr size_t kHeadsOff = 0;
constexpr size_t kHeadBytesOff = al256c(NDIC_ZFP_NOALLOC_MAX_BRICKS * sizeof(ZfpHead));
constexpr size_t kClassesOff = kHeadBytesOff + al256c(NDIC_ZFP_NOALLOC_MAX_BRICKS * kZfpHeadBytes);
// 32 (dims, scalar) classes, two 256-byte aligned tables each
constexpr size_t kScratchBytes =
    kClassesOff + NDIC_ZFP_NOALLOC_MAX_BRICKS * (sizeof(ZfpChunk) + sizeof(ZfpBrick)) + 2 * 32 * 256;
static_assert(kScratchBytes <= (2u << 20), "the allocation-free scratch stays modest");

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

std::mutex g_mu;                      // serialises calls (per process; one array per device)
std::vector<uint8_t*> g_scratch_ptr;  // device address of `scratch`, per device ordinal

// The current device's address of `scratch`, looked up at the first call on it.
uint8_t* scratch_ptr() {
  int dev = 0;
  NDIC_HIP(hipGetDevice(&dev));
  if ((size_t)dev >= g_scratch_ptr.size()) g_scratch_ptr.resize(dev + 1, nullptr);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    NDIC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) fail(NDIC_IO, "hipGetSymbolAddress returned no address for the brick scratch");
    g_scratch_ptr[dev] = (uint8_t*)p;
  }
  return g_scratch_ptr[dev];
}

}  // namespace
}  // namespace ndic

using namespace ndic;

extern "C" ndic_status ndic_zfp_decode_bricks_noalloc(const uint8_t* d_in, uint8_t* d_out,
                                                      ndic_zfp_brick_desc* descs, size_t n, void* stream,
                                                      ndic_error* err) {
  BrickPlan bp;
  ndic_status st = guard_call(err, [&] {
    if (n && !descs) fail(NDIC_INVALID_ARGUMENT, "null descriptors");
    if (n > NDIC_ZFP_NOALLOC_MAX_BRICKS)
      fail(NDIC_INVALID_ARGUMENT, "zfp bricks (allocation-free): " + std::to_string(n) +
                                      " bricks exceed the per-call maximum of " +
                                      std::to_string(NDIC_ZFP_NOALLOC_MAX_BRICKS));
    std::lock_guard<std::mutex> lock(g_mu);
    hipStream_t s = stream ? (hipStream_t)stream : Context::get().stream;
      uint8_t* base = scratch_ptr();
    std::vector<uint8_t> heads(n * kZfpHeadBytes);
    if (n) {
      // header bytes of every stream for the checks of BrickReader::fixed_rate
      std::vector<ZfpHead> hv(n);
      for (size_t i = 0; i < n; ++i) hv[i] = {d_in + descs[i].in_offset, descs[i].in_len};
      NDIC_HIP(hipMemcpyAsync(base + kHeadsOff, hv.data(), n * sizeof(ZfpHead), hipMemcpyHostToDevice, s));
      launch_zfp_heads((const ZfpHead*)(base + kHeadsOff), (uint32_t)n, base + kHeadBytesOff, s);
      NDIC_HIP(hipMemcpyAsync(heads.data(), base + kHeadBytesOff, heads.size(), hipMemcpyDeviceToHost, s));
      NDIC_HIP(hipStreamSynchronize(s));
    }
    plan_bricks(d_in, d_out, descs, n, heads.data(), bp);
    // every class's tables at their own place, so the launches queue back to back
    size_t off = kClassesOff;
    for (const BrickClass& bc : bp.classes) {
      const size_t sz_ch = bc.chunks.size() * sizeof(ZfpChunk), sz_br = bc.bricks.size() * sizeof(ZfpBrick);
      uint8_t* ch = base + off;
      uint8_t* br = base + al256c(off + sz_ch);
      off = al256c(off + sz_ch) + sz_br;
      off = al256c(off);
      if (off > kScratchBytes) fail(NDIC_IO, "zfp bricks (allocation-free): scratch layout overflow");
      NDIC_HIP(hipMemcpyAsync(ch, bc.chunks.data(), sz_ch, hipMemcpyHostToDevice, s));
      NDIC_HIP(hipMemcpyAsync(br, bc.bricks.data(), sz_br, hipMemcpyHostToDevice, s));
      launch_zfp_decode((const ZfpChunk*)ch, (uint32_t)bc.chunks.size(), bc.bricks.size(), bc.dims, bc.scalar,
                        (const ZfpBrick*)br, zfp_fixed_maxbits(bc.chunks), s);
    }
    NDIC_HIP(hipStreamSynchronize(s));
  });
  if (st != NDIC_OK) return st;
  return report_bricks(descs, n, bp, err);
}
