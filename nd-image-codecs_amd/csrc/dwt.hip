// Reversible 5/3 DWT of (y, x) planes on the device, batched over chunks.
//
// Analysis runs each level as two data-parallel passes, one thread per
// output sample pair: the lifting steps of T.800 F.4.8.2 need only the
// immediate neighbours, so a pair (2i, 2i+1) is computed straight from the
// subband samples i-1..i+1 (the predict of the right neighbour is recomputed
// rather than shared).  Synthesis fuses both passes of a level in one
// LDS-tiled kernel (dwt_inv_level); the two-pass synthesis kernels remain for
// the single-plane test hook.  Rows are processed with lanes along x (a row segment
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

// Fused inverse level: one 256-thread workgroup reconstructs a 64x64 tile of
// the level region.  Horizontal synthesis of the 33 low and 34 high rows the
// tile's vertical synthesis needs goes to LDS (reading the Mallat bands, the
// LL rows from the coarser reconstruction), then vertical synthesis writes the
// tile's row pairs, a wave per row pair (256-byte stores).  Same arithmetic
// and boundary rules as dwt_inv_h + dwt_inv_v, one read and one write of the
// level instead of two of each.
constexpr int kTile = 64;
__global__ __launch_bounds__(256) void dwt_inv_level(const DwtJob* __restrict__ jobs, uint32_t njobs) {
  __shared__ int32_t TLo[kTile / 2 + 1][kTile];
  __shared__ int32_t THi[kTile / 2 + 2][kTile];
  uint64_t item = blockIdx.x + (uint64_t)blockIdx.y * gridDim.x;
  const DwtJob& j = jobs[find_job(jobs, njobs, item)];
  uint64_t local = item - j.item_begin;
  if (local >= j.items) return;
  const uint32_t tiles_x = j.tiles_x, tiles_y = (uint32_t)(j.tiles / tiles_x);
  const uint32_t tx = (uint32_t)(local % tiles_x);
  const uint64_t rest = local / tiles_x;
  const uint32_t ty = (uint32_t)(rest % tiles_y);
  const uint64_t plane = rest / tiles_y;
  const uint32_t w = j.w, h = j.h;
  const uint32_t nl = (w + 1) / 2, nh = w / 2;   // horizontal low / high widths
  const uint32_t hl = (h + 1) / 2, hh = h / 2;   // vertical low / high heights
  const uint32_t X0 = tx * kTile, Y0 = ty * kTile;
  const uint32_t i0 = Y0 / 2, jb = X0 / 2;
  const uint64_t base = plane * j.plane_elems;
  const uint64_t S = j.stride;
  const uint32_t tid = threadIdx.x;

  // ---- horizontal synthesis of T rows into LDS
  // rows 0..32: low rows i0 + m (m < 33, i0 + m < hl); rows 33..66: high rows
  // clamp(i0 - 1 + m, 0, hh - 1), m < 34
  for (uint32_t it = tid; it < (kTile + 3) * (kTile / 2); it += 256) {
    uint32_t r = it / (kTile / 2), jl = it % (kTile / 2);
    uint32_t y;       // T row (level-region row)
    int32_t* dstrow;
    if (r < kTile / 2 + 1) {
      if (i0 + r >= hl) continue;
      y = i0 + r;
      dstrow = TLo[r];
    } else {
      if (hh == 0) continue;
      int32_t k = (int32_t)i0 - 1 + (int32_t)(r - (kTile / 2 + 1));
      k = k < 0 ? 0 : (k > (int32_t)hh - 1 ? (int32_t)hh - 1 : k);
      y = hl + (uint32_t)k;
      dstrow = THi[r - (kTile / 2 + 1)];
    }
    uint32_t jj = jb + jl;
    if (jj >= nl) continue;
    const int32_t* low = (y < hl ? j.src_ll : j.src) + base + y * S;
    const int32_t* high = j.src + base + y * S + nl;
    uint32_t c = 2 * jl;  // tile column of the pair
    if (w == 1) {
      dstrow[0] = low[0];
      continue;
    }
    int32_t hm = high[umin(jj == 0 ? 0 : jj - 1, nh - 1)], h0 = high[umin(jj, nh - 1)];
    int32_t x0 = low[jj] - ((hm + h0 + 2) >> 2);
    dstrow[c] = x0;
    if (2 * jj + 1 < w) {
      int32_t xr;
      if (2 * jj + 2 < w) {
        int32_t h1 = high[umin(jj + 1, nh - 1)];
        xr = low[jj + 1] - ((h0 + h1 + 2) >> 2);
      } else {
        xr = x0;
      }
      dstrow[c + 1] = high[jj] + ((x0 + xr) >> 1);
    }
  }
  __syncthreads();

  // ---- vertical synthesis out of LDS: row pairs (2ii, 2ii+1), ii = i0 + iv
  int32_t* out = j.dst + base;
  for (uint32_t it = tid; it < (kTile / 2) * kTile; it += 256) {
    uint32_t iv = it / kTile, c = it % kTile;
    uint32_t ii = i0 + iv, x = X0 + c;
    if (ii >= hl || x >= w) continue;
    if (h == 1) {
      out[x] = TLo[0][c];
      continue;
    }
    // THi[m] holds high row clamp(i0 - 1 + m, 0, hh - 1), so the clamped
    // neighbours ii - 1, ii, ii + 1 are THi[iv], THi[iv + 1], THi[iv + 2]
    int32_t hm = THi[iv][c], h0 = THi[iv + 1][c];
    int32_t x0 = TLo[iv][c] - ((hm + h0 + 2) >> 2);
    out[(uint64_t)(2 * ii) * S + x] = x0;
    if (2 * ii + 1 < h) {
      int32_t xr = 2 * ii + 2 < h ? TLo[iv + 1][c] - ((h0 + THi[iv + 2][c] + 2) >> 2) : x0;
      out[(uint64_t)(2 * ii + 1) * S + x] = h0 + ((x0 + xr) >> 1);
    }
  }
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
    case kDwtInvLevel: {
      // one workgroup per tile
      uint32_t tx = (uint32_t)(items < 65535 ? items : 65535);
      uint32_t ty = (uint32_t)ceil_div(items, tx);
      hipLaunchKernelGGL(dwt_inv_level, dim3(tx, ty), dim3(256), 0, stream, d_jobs, njobs);
      break;
    }
    default: hipLaunchKernelGGL(dwt_fwd_h, dim3(gx, gy), dim3(256), 0, stream, d_jobs, njobs); break;
  }
  NDIC_HIP(hipGetLastError());
}

}  // namespace ndic
