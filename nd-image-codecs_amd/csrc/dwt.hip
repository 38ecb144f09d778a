// Reversible 5/3 DWT of (y, x) planes on the device, batched over chunks.
//
// Each decomposition level is two data-parallel passes, one thread per
// output sample pair: the lifting steps of T.800 F.4.8.2 need only the
// immediate neighbours, so a pair (2i, 2i+1) is computed straight from the
// subband samples i-1..i+1 (the predict of the right neighbour is recomputed
// rather than shared).  Rows are processed with lanes along x (a row segment
// is one coalesced access); columns with lanes along x as well, each lane
// walking its own column pair, so every pass is a coalesced streaming pass.
// Passes read one buffer and write another, which removes every in-place
// hazard between threads:
//   forward level l:  vertical   S(level input) -> T,  horizontal T -> A
//   inverse level l:  horizontal A (+ LL from O)  -> T,  vertical   T -> O
// A holds the Mallat coefficients, O the reconstruction.  Arithmetic and
// boundary handling follow crates/ndic-htj2k/src/dwt/mod.rs:51-101 (whole-
// sample symmetric extension; a length-1 signal passes through) and the
// level order of dwt/mod.rs:133-198 (vertical first on analysis, horizontal
// first on synthesis).
#include "dwt.hpp"

namespace ndic {

namespace {

__device__ __forceinline__ uint32_t find_job(const DwtJob* jobs, uint32_t njobs, uint64_t item) {
  uint32_t lo = 0, hi = njobs;
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) / 2;
    if (jobs[mid].item_begin <= item) lo = mid;
    else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ uint32_t umin(uint32_t a, uint32_t b) { return a < b ? a : b; }

// Inverse horizontal pass: row y of the level-l region.  Low half of the row
// comes from the LL source for y < hl (the reconstructed LL of the coarser
// level), else from A (the LH band); the high half always from A.
__global__ __launch_bounds__(256) void dwt_inv_h(const DwtJob* __restrict__ jobs, uint32_t njobs) {
  uint64_t item = (blockIdx.x + (uint64_t)blockIdx.y * gridDim.x) * 4 + (threadIdx.x >> 6);
  const DwtJob& j = jobs[find_job(jobs, njobs, item)];
  uint64_t local = item - j.item_begin;
  if (local >= j.items) return;
  uint64_t lines = j.h, t = local % j.tiles, r = local / j.tiles;
  uint64_t y = r % lines, plane = r / lines;
  uint32_t n = j.w, nl = (n + 1) / 2, nh = n / 2;
  uint32_t i = (uint32_t)t * 64 + (threadIdx.x & 63);
  if (i >= nl) return;
  uint64_t base = plane * j.plane_elems + y * j.stride;
  const int32_t* low = (y < j.hl ? j.src_ll : j.src) + base;
  const int32_t* high = j.src + base + nl;
  int32_t* out = j.dst + base;
  if (n == 1) {
    out[0] = low[0];
    return;
  }
  int32_t hm = high[umin(i == 0 ? 0 : i - 1, nh - 1)], h0 = high[umin(i, nh - 1)];
  int32_t x0 = low[i] - ((hm + h0 + 2) >> 2);
  out[2 * i] = x0;
  if (2 * i + 1 < n) {
    int32_t xr;
    if (2 * i + 2 < n) {
      int32_t h1 = high[umin(i + 1, nh - 1)];
      xr = low[i + 1] - ((h0 + h1 + 2) >> 2);
    } else {
      xr = x0;
    }
    out[2 * i + 1] = high[i] + ((x0 + xr) >> 1);
  }
}

// Inverse vertical pass: column pair (2i, 2i+1) of column x.
__global__ __launch_bounds__(256) void dwt_inv_v(const DwtJob* __restrict__ jobs, uint32_t njobs) {
  uint64_t item = (blockIdx.x + (uint64_t)blockIdx.y * gridDim.x) * 4 + (threadIdx.x >> 6);
  const DwtJob& j = jobs[find_job(jobs, njobs, item)];
  uint64_t local = item - j.item_begin;
  if (local >= j.items) return;
  uint32_t n = j.h, nl = (n + 1) / 2, nh = n / 2;
  uint64_t t = local % j.tiles, r = local / j.tiles;
  uint64_t i = r % nl, plane = r / nl;
  uint32_t x = (uint32_t)t * 64 + (threadIdx.x & 63);
  if (x >= j.w) return;
  uint64_t S = j.stride;
  const int32_t* col = j.src + plane * j.plane_elems + x;
  int32_t* out = j.dst + plane * j.plane_elems + x;
  if (n == 1) {
    out[0] = col[0];
    return;
  }
  const int32_t* high = col + (uint64_t)nl * S;
  uint32_t ii = (uint32_t)i;
  int32_t hm = high[umin(ii == 0 ? 0 : ii - 1, nh - 1) * S], h0 = high[umin(ii, nh - 1) * S];
  int32_t x0 = col[i * S] - ((hm + h0 + 2) >> 2);
  out[2 * i * S] = x0;
  if (2 * i + 1 < n) {
    int32_t xr;
    if (2 * i + 2 < n) {
      int32_t h1 = high[umin(ii + 1, nh - 1) * S];
      xr = col[(i + 1) * S] - ((h0 + h1 + 2) >> 2);
    } else {
      xr = x0;
    }
    out[(2 * i + 1) * S] = high[i * S] + ((x0 + xr) >> 1);
  }
}

// Analysis of one pair: high[i] and low[i] of x[0..n) (stride s).
__device__ __forceinline__ void ana_pair(const int32_t* x, uint64_t s, uint32_t n, uint32_t i, int32_t* lo,
                                         int32_t* hi, bool* has_hi) {
  uint32_t nh = n / 2;
  auto hv = [&](uint32_t k) -> int32_t {
    uint32_t r = 2 * k + 2 < n ? 2 * k + 2 : n - 2;
    return x[(2 * k + 1) * s] - ((x[2 * k * s] + x[r * s]) >> 1);
  };
  uint32_t kl = umin(i == 0 ? 0 : i - 1, nh - 1), kr = umin(i, nh - 1);
  int32_t dl = hv(kl), dr = kl == kr ? dl : hv(kr);
  *lo = x[2 * i * s] + ((dl + dr + 2) >> 2);
  *has_hi = i < nh;
  *hi = *has_hi ? (i == kr ? dr : hv(i)) : 0;
}

// Forward vertical pass: low rows [0, nl), high rows [nl, n) of column x.
__global__ __launch_bounds__(256) void dwt_fwd_v(const DwtJob* __restrict__ jobs, uint32_t njobs) {
  uint64_t item = (blockIdx.x + (uint64_t)blockIdx.y * gridDim.x) * 4 + (threadIdx.x >> 6);
  const DwtJob& j = jobs[find_job(jobs, njobs, item)];
  uint64_t local = item - j.item_begin;
  if (local >= j.items) return;
  uint32_t n = j.h, nl = (n + 1) / 2;
  uint64_t t = local % j.tiles, r = local / j.tiles;
  uint32_t i = (uint32_t)(r % nl);
  uint64_t plane = r / nl;
  uint32_t x = (uint32_t)t * 64 + (threadIdx.x & 63);
  if (x >= j.w) return;
  uint64_t S = j.stride;
  const int32_t* col = j.src + plane * j.plane_elems + x;
  int32_t* out = j.dst + plane * j.plane_elems + x;
  if (n == 1) {
    out[0] = col[0];
    return;
  }
  int32_t lo, hi;
  bool has_hi;
  ana_pair(col, S, n, i, &lo, &hi, &has_hi);
  out[(uint64_t)i * S] = lo;
  if (has_hi) out[(uint64_t)(nl + i) * S] = hi;
}

// Forward horizontal pass: row y, low columns [0, nl), high [nl, n).
__global__ __launch_bounds__(256) void dwt_fwd_h(const DwtJob* __restrict__ jobs, uint32_t njobs) {
  uint64_t item = (blockIdx.x + (uint64_t)blockIdx.y * gridDim.x) * 4 + (threadIdx.x >> 6);
  const DwtJob& j = jobs[find_job(jobs, njobs, item)];
  uint64_t local = item - j.item_begin;
  if (local >= j.items) return;
  uint64_t t = local % j.tiles, r = local / j.tiles;
  uint64_t y = r % j.h, plane = r / j.h;
  uint32_t n = j.w, nl = (n + 1) / 2;
  uint32_t i = (uint32_t)t * 64 + (threadIdx.x & 63);
  if (i >= nl) return;
  uint64_t base = plane * j.plane_elems + y * j.stride;
  const int32_t* row = j.src + base;
  int32_t* out = j.dst + base;
  if (n == 1) {
    out[0] = row[0];
    return;
  }
  int32_t lo, hi;
  bool has_hi;
  ana_pair(row, 1, n, i, &lo, &hi, &has_hi);
  out[i] = lo;
  if (has_hi) out[nl + i] = hi;
}

}  // namespace

void launch_dwt_pass(int pass, const DwtJob* d_jobs, uint32_t njobs, uint64_t items, hipStream_t stream) {
  if (items == 0) return;
  uint64_t blocks = ceil_div(items, 4);
  uint32_t gx = (uint32_t)(blocks < 65535 ? blocks : 65535);
  uint32_t gy = (uint32_t)ceil_div(blocks, gx);
  switch (pass) {
    case kDwtInvH: hipLaunchKernelGGL(dwt_inv_h, dim3(gx, gy), dim3(256), 0, stream, d_jobs, njobs); break;
    case kDwtInvV: hipLaunchKernelGGL(dwt_inv_v, dim3(gx, gy), dim3(256), 0, stream, d_jobs, njobs); break;
    case kDwtFwdV: hipLaunchKernelGGL(dwt_fwd_v, dim3(gx, gy), dim3(256), 0, stream, d_jobs, njobs); break;
    default: hipLaunchKernelGGL(dwt_fwd_h, dim3(gx, gy), dim3(256), 0, stream, d_jobs, njobs); break;
  }
  NDIC_HIP(hipGetLastError());
}

}  // namespace ndic
