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
#include "codestream.hpp"

namespace ndic {

namespace {

template <class Job>
__device__ __forceinline__ uint32_t find_job(const Job* jobs, uint32_t njobs, uint64_t item) {
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
// Eight waves per SIMD (at most 64 registers a lane, no spills) keep more
// tile loads in flight than the six of the unconstrained allocation.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void dwt_inv_level(
    const DwtJob* __restrict__ jobs, uint32_t njobs) {
  __shared__ int32_t TLo[kTile / 2 + 1][kTile];
  __shared__ int32_t THi[kTile / 2 + 2][kTile];
  const uint64_t item = xcd_item();
  if (item >= jobs[njobs - 1].item_begin + jobs[njobs - 1].items) return;
  const DwtJob& j = jobs[find_job(jobs, njobs, item)];
  uint64_t local = item - j.item_begin;
  if (local >= j.items) return;
  const uint32_t tiles_x = j.tiles_x, tiles_y = (uint32_t)(j.tiles / tiles_x);
  const uint32_t tx = (uint32_t)(local % tiles_x);
  const uint64_t rest = local / tiles_x;
  const uint32_t ty = (uint32_t)(rest % tiles_y);
  const uint64_t plane = rest / tiles_y;
  if (j.irrev && j.irrev[plane].on) return;  // a 9/7 plane
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
  // clamp(i0 - 1 + m, 0, hh - 1), m < 34.  Thread tid owns column pair
  // jl = tid % 32 of T rows tid / 32 + 8 k (k < 9); it takes them in two
  // batches and issues the loads of a whole batch before it uses any.
  constexpr uint32_t kTRows = kTile + 3, kIter = (kTRows * (kTile / 2) + 255) / 256, kBatch = 5;
  const uint32_t jl = tid & 31, rb = tid >> 5;
  const uint32_t jj = jb + jl;
  const uint32_t jm = umin(jj == 0 ? 0 : jj - 1, nh - 1), j0 = umin(jj, nh - 1), j1 = umin(jj + 1, nh - 1);
  const bool has_r = 2 * jj + 1 < w, has_r2 = 2 * jj + 2 < w;
#pragma unroll 1
  for (uint32_t k0 = 0; k0 < kIter; k0 += kBatch) {
    int32_t vl0[kBatch], vl1[kBatch], vhm[kBatch], vh0[kBatch], vh1[kBatch];
    bool ok[kBatch];
#pragma unroll
    for (uint32_t b = 0; b < kBatch; ++b) {
      const uint32_t r = rb + 8 * (k0 + b);
      uint32_t y = 0;
      ok[b] = r < kTRows && jj < nl;
      if (r < kTile / 2 + 1) {
        ok[b] = ok[b] && i0 + r < hl;
        y = i0 + r;
      } else {
        ok[b] = ok[b] && hh > 0;
        int32_t m = (int32_t)i0 - 1 + (int32_t)(r - (kTile / 2 + 1));
        m = m < 0 ? 0 : (m > (int32_t)hh - 1 ? (int32_t)hh - 1 : m);
        y = hl + (uint32_t)m;
      }
      vl0[b] = vl1[b] = vhm[b] = vh0[b] = vh1[b] = 0;
      if (ok[b]) {
        const int32_t* low = (y < hl ? j.src_ll : j.src) + base + y * S;
        const int32_t* high = j.src + base + y * S + nl;
        if (w == 1) {
          vl0[b] = low[0];
        } else {
          vl0[b] = low[jj];
          vhm[b] = high[jm];
          vh0[b] = high[j0];
          if (has_r2) {
            vl1[b] = low[jj + 1];
            vh1[b] = high[j1];
          }
        }
      }
    }
#pragma unroll
    for (uint32_t b = 0; b < kBatch; ++b) {
      if (!ok[b]) continue;
      const uint32_t r = rb + 8 * (k0 + b);
      int32_t* dstrow = r < kTile / 2 + 1 ? TLo[r] : THi[r - (kTile / 2 + 1)];
      if (w == 1) {
        dstrow[0] = vl0[b];
        continue;
      }
      const int32_t x0 = vl0[b] - ((vhm[b] + vh0[b] + 2) >> 2);
      dstrow[2 * jl] = x0;
      if (has_r) {
        const int32_t xr = has_r2 ? vl1[b] - ((vh0[b] + vh1[b] + 2) >> 2) : x0;
        dstrow[2 * jl + 1] = vh0[b] + ((x0 + xr) >> 1);
      }
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


// ---------------------------------------------------------------------------
// Irreversible 9/7 synthesis (dwt/mod.rs:243-325): the unscaled lifting
// (T.800 Table F.4 multipliers, K folded into the dequantisation step) on f32
// coefficients held as bits in the int32 buffers.  Each thread computes one
// output pair (2i, 2i+1) of a line straight from the subband samples: the
// four steps are evaluated at the neighbours each step reads, with the
// reference's index clamps applied at every step (its symmetric extension
// repeats the edge sample of each intermediate array), so the result equals
// the in-place sweep of the reference sample for sample.  Sums, products and
// updates are rounded separately (no contraction), in the reference's order.
// ---------------------------------------------------------------------------
constexpr float k97Alpha = -1.5861343f, k97Beta = -0.052980118f, k97Gamma = 0.8829111f, k97Delta = 0.44350685f;

__device__ __forceinline__ float lstep(float base, float c, float a, float b) {
#pragma clang fp contract(off)
  return base + c * (a + b);
}

// One line of a level: low half at lo[k * ls] (k < nl), high half at
// hi[k * hs] (k < nh), n >= 2.
struct Line97 {
  const int32_t* lo;
  const int32_t* hi;
  uint64_t ls, hs;
  int nl, nh;
  __device__ float L(int k) const { return __int_as_float(lo[(uint64_t)k * ls]); }
  __device__ float H(int k) const { return __int_as_float(hi[(uint64_t)k * hs]); }
  __device__ int ch(int k) const { return k < 0 ? 0 : (k > nh - 1 ? nh - 1 : k); }
  __device__ int cl(int k) const { return k > nl - 1 ? nl - 1 : k; }
  // after the -delta step (low from high), the -gamma step (high from low),
  // the -beta step and the -alpha step
  __device__ float L1(int k) const { return lstep(L(k), -k97Delta, H(ch(k - 1)), H(ch(k))); }
  __device__ float H1(int k) const { return lstep(H(k), -k97Gamma, L1(k), L1(cl(k + 1))); }
  __device__ float L2(int k) const { return lstep(L1(k), -k97Beta, H1(ch(k - 1)), H1(ch(k))); }
  __device__ float H2(int k) const { return lstep(H1(k), -k97Alpha, L2(k), L2(cl(k + 1))); }
};

// Inverse horizontal 9/7 pass: pair i of row y of the level region; the low
// half of rows y < hl comes from the LL source (the coarser level's
// reconstruction), else from A (LH); the high half from A.
__global__ __launch_bounds__(256) void dwt97_inv_h(const DwtJob* __restrict__ jobs, uint32_t njobs) {
  uint64_t item = (blockIdx.x + (uint64_t)blockIdx.y * gridDim.x) * 4 + (threadIdx.x >> 6);
  const DwtJob& j = jobs[find_job(jobs, njobs, item)];
  uint64_t local = item - j.item_begin;
  if (local >= j.items) return;
  uint64_t lines = j.h, t = local % j.tiles, r = local / j.tiles;
  uint64_t y = r % lines, plane = r / lines;
  if (!j.irrev || !j.irrev[plane].on) return;
  const uint32_t n = j.w, nl = (n + 1) / 2, nh = n / 2;
  const uint32_t i = (uint32_t)t * 64 + (threadIdx.x & 63);
  if (i >= nl) return;
  const uint64_t base = plane * j.plane_elems + y * j.stride;
  const int32_t* low = (y < j.hl ? j.src_ll : j.src) + base;
  int32_t* out = j.dst + base;
  if (n == 1) {
    out[0] = low[0];
    return;
  }
  const Line97 ln{low, j.src + base + nl, 1, 1, (int)nl, (int)nh};
  out[2 * i] = __float_as_int(ln.L2((int)i));
  if (i < nh) out[2 * i + 1] = __float_as_int(ln.H2((int)i));
}

// Inverse vertical 9/7 pass: pair (2i, 2i+1) of column x.
__global__ __launch_bounds__(256) void dwt97_inv_v(const DwtJob* __restrict__ jobs, uint32_t njobs) {
  uint64_t item = (blockIdx.x + (uint64_t)blockIdx.y * gridDim.x) * 4 + (threadIdx.x >> 6);
  const DwtJob& j = jobs[find_job(jobs, njobs, item)];
  uint64_t local = item - j.item_begin;
  if (local >= j.items) return;
  const uint32_t n = j.h, nl = (n + 1) / 2, nh = n / 2;
  uint64_t t = local % j.tiles, r = local / j.tiles;
  uint64_t i = r % nl, plane = r / nl;
  if (!j.irrev || !j.irrev[plane].on) return;
  const uint32_t x = (uint32_t)t * 64 + (threadIdx.x & 63);
  if (x >= j.w) return;
  const uint64_t S = j.stride;
  const int32_t* col = j.src + plane * j.plane_elems + x;
  int32_t* out = j.dst + plane * j.plane_elems + x;
  if (n == 1) {
    out[0] = col[0];
    return;
  }
  const Line97 ln{col, col + (uint64_t)nl * S, S, S, (int)nl, (int)nh};
  out[2 * i * S] = __float_as_int(ln.L2((int)i));
  if (i < nh) out[(2 * i + 1) * S] = __float_as_int(ln.H2((int)i));
}

// ---------------------------------------------------------------------------
// Finest synthesis level fused with the inverse nd_lift step along a plane
// axis and the narrowing store.
//
// A workgroup owns one 64 x 32 tile of the plane and one segment of the lift
// axis (z).  It synthesises the tile of one z-row's plane at a time -- the
// horizontal pass of the T rows it needs into LDS, a barrier, the vertical
// pass into registers: 8 samples per thread, column tid % 64, row pairs
// tid / 64 + 4 k -- and runs the lift's inverse levels along z as a stream
// over those register tiles: each level keeps its last even sample and detail
// (5/3) or nothing (haar, delta) and asks the level below for the next low
// sample when it needs one, so every plane is synthesised exactly once and
// each output plane is stored (narrowed to the dtype) as soon as it is known.
// Two LDS buffers alternate between planes, so one barrier per plane
// suffices.  Arithmetic: dwt_inv_level above (T.800 F.4.8.2), the DC
// un-shift of unsigned planes (reader.rs:626-638), kernel.rs:186-200 and
// 243-272 (inverse haar / 5-3 levels, wrapping) and the inverse delta.
// ---------------------------------------------------------------------------
constexpr uint32_t kFx = kDwtLiftTileW, kFy = kDwtLiftTileH;   // output tile
constexpr uint32_t kFLo = kFy / 2 + 1, kFHi = kFy / 2 + 2;     // T rows: low, high
constexpr uint32_t kFTRows = kFLo + kFHi;
// horizontal pass: kFx / 2 column pairs per T row, kFHRows T rows at a time
constexpr uint32_t kFPairs = kFx / 2, kFHRows = 256 / kFPairs, kFIter = (kFTRows + kFHRows - 1) / kFHRows;
// vertical pass: kFx columns, kFVGroups row-pair groups, kFRP row pairs per thread
constexpr uint32_t kFVGroups = 256 / kFx, kFRP = kDwtLiftRowPairs, kFS = 2 * kFRP;  // samples per thread
static_assert(kFVGroups * kFS == kFy, "a thread holds kFRP row pairs of one column");

struct V8 {
  int32_t v[kFS];
};
__device__ __forceinline__ int32_t wadd32(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
__device__ __forceinline__ int32_t wsub32(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }

struct FusedTile {
  const DwtLiftJob& j;
  int32_t (*TLo)[kFLo][kFx];
  int32_t (*THi)[kFHi][kFx];
  uint64_t zplane0;  // plane index of z-row 0 of the segment, z-rows R planes apart
  uint32_t X0, Y0, i0, jb, nl, nh, hl, hh;
  uint32_t buf;
  bool ok;
  int32_t lo, hi;  // range mode

  __device__ __forceinline__ uint64_t plane_of(uint32_t k) const { return zplane0 + (uint64_t)k * j.R; }

  // Synthesised tile of z-row k: sample e of this thread is column
  // X0 + tid % kFx, row Y0 + 2 (tid / kFx + kFVGroups (e / 2)) + e % 2.
  __device__ __forceinline__ V8 fetch(uint32_t k) {
    const uint32_t tid = threadIdx.x, w = j.w, h = j.h;
    const uint64_t p = plane_of(k);
    const int32_t* a = j.A + p * j.plane_elems;
    const int32_t* ll = j.LL + p * j.plane_elems;
    int32_t(*lo)[kFx] = TLo[buf];
    int32_t(*hi)[kFx] = THi[buf];
    buf ^= 1;
    // horizontal synthesis of the T rows (loads of all rows first)
    {
      const uint32_t jl = tid % kFPairs, rb = tid / kFPairs, jj = jb + jl;
      const uint32_t jm = umin(jj == 0 ? 0 : jj - 1, nh - 1), j0 = umin(jj, nh - 1), j1 = umin(jj + 1, nh - 1);
      const bool has_r = 2 * jj + 1 < w, has_r2 = 2 * jj + 2 < w;
      int32_t vl0[kFIter], vl1[kFIter], vhm[kFIter], vh0[kFIter], vh1[kFIter];
      bool okr[kFIter];
#pragma unroll
      for (uint32_t b = 0; b < kFIter; ++b) {
        const uint32_t r = rb + kFHRows * b;
        uint32_t y = 0;
        okr[b] = r < kFTRows && jj < nl;
        if (r < kFLo) {
          okr[b] = okr[b] && i0 + r < hl;
          y = i0 + r;
        } else {
          okr[b] = okr[b] && hh > 0;
          int32_t m = (int32_t)i0 - 1 + (int32_t)(r - kFLo);
          m = m < 0 ? 0 : (m > (int32_t)hh - 1 ? (int32_t)hh - 1 : m);
          y = hl + (uint32_t)m;
        }
        vl0[b] = vl1[b] = vhm[b] = vh0[b] = vh1[b] = 0;
        if (okr[b]) {
          const int32_t* low = (y < hl ? ll : a) + (uint64_t)y * w;
          const int32_t* high = a + (uint64_t)y * w + nl;
          if (w == 1) {
            vl0[b] = low[0];
          } else {
            vl0[b] = low[jj];
            vhm[b] = high[jm];
            vh0[b] = high[j0];
            if (has_r2) {
              vl1[b] = low[jj + 1];
              vh1[b] = high[j1];
            }
          }
        }
      }
#pragma unroll
      for (uint32_t b = 0; b < kFIter; ++b) {
        if (!okr[b]) continue;
        const uint32_t r = rb + kFHRows * b;
        int32_t* dstrow = r < kFLo ? lo[r] : hi[r - kFLo];
        if (w == 1) {
          dstrow[0] = vl0[b];
          continue;
        }
        const int32_t x0 = vl0[b] - ((vhm[b] + vh0[b] + 2) >> 2);
        dstrow[2 * jl] = x0;
        if (has_r) {
          const int32_t xr = has_r2 ? vl1[b] - ((vh0[b] + vh1[b] + 2) >> 2) : x0;
          dstrow[2 * jl + 1] = vh0[b] + ((x0 + xr) >> 1);
        }
      }
    }
    __syncthreads();
    // vertical synthesis into registers
    V8 out;
    const uint32_t c = tid % kFx, ivb = tid / kFx;
#pragma unroll
    for (uint32_t kk = 0; kk < kFRP; ++kk) {
      const uint32_t iv = ivb + kFVGroups * kk, ii = i0 + iv;
      int32_t x0 = 0, x1 = 0;
      if (h == 1) {
        x0 = lo[0][c];
      } else if (ii < hl) {
        const int32_t hm = hi[iv][c], h0 = hi[iv + 1][c];
        x0 = lo[iv][c] - ((hm + h0 + 2) >> 2);
        const int32_t xr = 2 * ii + 2 < h ? lo[iv + 1][c] - ((h0 + hi[iv + 2][c] + 2) >> 2) : x0;
        x1 = h0 + ((x0 + xr) >> 1);
      }
      out.v[2 * kk] = x0;
      out.v[2 * kk + 1] = x1;
    }
    // DC un-shift with clamping of unsigned planes
    const DevPlane& pl = j.planes[p];
    if (!pl.is_signed && pl.ok) {
      const int32_t half = 1 << (pl.depth - 1);
      const int64_t mx = ((int64_t)1 << pl.depth) - 1;
#pragma unroll
      for (int e = 0; e < (int)kFS; ++e) {
        const int64_t v = (int64_t)out.v[e] + half;
        out.v[e] = (int32_t)(v < 0 ? 0 : v > mx ? mx : v);
      }
    }
    return out;
  }

  template <class D>
  __device__ __forceinline__ void put_as(uint8_t* base, const V8& x, uint32_t c, uint32_t ivb) {
    D* d = (D*)base;
#pragma unroll
    for (uint32_t e = 0; e < kFS; ++e) {
      const uint32_t y = Y0 + 2 * (ivb + kFVGroups * (e / 2)) + (e & 1), xx = X0 + c;
      if (xx < j.w && y < j.h) {
        d[(uint64_t)y * j.w + xx] = (D)x.v[e];
        const int32_t v = x.v[e];
        bool f = true;
        if (sizeof(D) == 1) f = (D)-1 > 0 ? (v >= 0 && v <= 255) : (v >= -128 && v <= 127);
        else if (sizeof(D) == 2) f = (D)-1 > 0 ? (v >= 0 && v <= 65535) : (v >= -32768 && v <= 32767);
        else f = (D)-1 > 0 ? v >= 0 : true;
        ok &= f;
      }
    }
  }
  // Stores z-row k of the output, narrowed to the dtype (range mode: folds
  // it into lo / hi).
  __device__ __forceinline__ void put(uint32_t k, const V8& x) {
    const uint32_t c = threadIdx.x % kFx, ivb = threadIdx.x / kFx;
    if (!j.dst) {
#pragma unroll
      for (uint32_t e = 0; e < kFS; ++e) {
        const uint32_t y = Y0 + 2 * (ivb + kFVGroups * (e / 2)) + (e & 1), xx = X0 + c;
        if (xx < j.w && y < j.h) {
          lo = x.v[e] < lo ? x.v[e] : lo;
          hi = x.v[e] > hi ? x.v[e] : hi;
        }
      }
      return;
    }
    uint8_t* base = j.dst + plane_of(k) * j.plane_elems * j.dsize;
    switch (j.dtype) {
      case NDIC_U8: put_as<uint8_t>(base, x, c, ivb); break;
      case NDIC_I8: put_as<int8_t>(base, x, c, ivb); break;
      case NDIC_U16: put_as<uint16_t>(base, x, c, ivb); break;
      case NDIC_I16: put_as<int16_t>(base, x, c, ivb); break;
      case NDIC_U32: put_as<uint32_t>(base, x, c, ivb); break;
      default: put_as<int32_t>(base, x, c, ivb); break;
    }
  }
};

// Low samples read straight from z-rows base, base + 1, ...
struct RowSrc {
  FusedTile* t;
  uint32_t next_row;
  __device__ __forceinline__ V8 next() { return t->fetch(next_row++); }
};

// One inverse level along z: yields X[0], X[1], ... X[len - 1] from the low
// samples of Src (in order) and the detail z-rows dbase + m.
template <class Src>
struct ZLevel {
  FusedTile* t;
  Src* src;
  uint32_t kind, len, nd, dbase, m;
  bool has_odd;
  V8 xe, xo, dc;
  __device__ __forceinline__ void init(FusedTile* tile, Src* s, uint32_t k, uint32_t n, uint32_t db) {
    t = tile;
    src = s;
    kind = k;
    len = n;
    nd = n / 2;
    dbase = db;
    m = 0;
    has_odd = false;
    if (kind == 2) {  // X[0] needs S[0] and D[0] (len >= 2, so nd >= 1)
      dc = t->fetch(dbase);
      const V8 s0 = src->next();
#pragma unroll
      for (int e = 0; e < (int)kFS; ++e) xe.v[e] = wsub32(s0.v[e], wadd32(wadd32(dc.v[e], dc.v[e]), 2) >> 2);
    }
  }
  __device__ __forceinline__ V8 next() {
    if (has_odd) {
      has_odd = false;
      return xo;
    }
    if (kind == 1) {  // haar: X[2m] = S[m] - (D[m] >> 1), X[2m+1] = X[2m] + D[m]
      const V8 s = src->next();
      if (m >= nd) {  // odd length: the last low sample passes through
        ++m;
        return s;
      }
      const V8 d = t->fetch(dbase + m);
      V8 x0;
#pragma unroll
      for (int e = 0; e < (int)kFS; ++e) {
        x0.v[e] = wsub32(s.v[e], d.v[e] >> 1);
        xo.v[e] = wadd32(x0.v[e], d.v[e]);
      }
      ++m;
      has_odd = true;
      return x0;
    }
    // 5/3: xe = X[2m], dc = D[m]
    const V8 out = xe;
    if (m < nd) {
      V8 xn = xe, dn = dc;
      if (2 * m + 2 < len) {
        if (m + 1 < nd) dn = t->fetch(dbase + m + 1);
        const V8 s = src->next();
#pragma unroll
        for (int e = 0; e < (int)kFS; ++e) xn.v[e] = wsub32(s.v[e], wadd32(wadd32(dc.v[e], dn.v[e]), 2) >> 2);
      }
#pragma unroll
      for (int e = 0; e < (int)kFS; ++e) xo.v[e] = wadd32(dc.v[e], wadd32(wadd32(xe.v[e], xn.v[e]), 1) >> 1);
      has_odd = true;
      xe = xn;
      dc = dn;
    }
    ++m;
    return out;
  }
};

__device__ __forceinline__ uint32_t zbands(uint64_t n, uint32_t levels) {
  uint32_t c = 0;
  for (uint64_t m = n; c < levels && m >= 2; m = (m + 1) / 2) ++c;
  return c;
}

// Six waves per SIMD (at most 80 registers a lane) hide more of the z-row
// reads than the five the unconstrained allocation leaves room for.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void dwt_lift_kernel(
    const DwtLiftJob* __restrict__ jobs, uint32_t njobs) {
  __shared__ int32_t TLo[2][kFLo][kFx];
  __shared__ int32_t THi[2][kFHi][kFx];
  const uint64_t item = xcd_item();
  if (item >= jobs[njobs - 1].item_begin + jobs[njobs - 1].items) return;
  const DwtLiftJob& j = jobs[find_job(jobs, njobs, item)];
  const uint64_t local = item - j.item_begin;
  if (local >= j.items) return;
  uint64_t rest = local / j.tiles;
  const uint32_t t = (uint32_t)(local % j.tiles);
  const uint64_t r = rest % j.R;
  rest /= j.R;
  const uint64_t sg = rest % j.segs, o = rest / j.segs;
  const uint64_t g0 = sg * j.group;
  const uint32_t n = (uint32_t)(j.len - g0 < j.group ? j.len - g0 : j.group);
  FusedTile ft{j, TLo, THi};
  ft.zplane0 = (o * j.len + g0) * j.R + r;
  ft.X0 = (t % j.tiles_x) * kFx;
  ft.Y0 = (t / j.tiles_x) * kFy;
  ft.i0 = ft.Y0 / 2;
  ft.jb = ft.X0 / 2;
  ft.nl = (j.w + 1) / 2;
  ft.nh = j.w / 2;
  ft.hl = (j.h + 1) / 2;
  ft.hh = j.h / 2;
  ft.buf = 0;
  ft.ok = true;
  ft.lo = INT32_MAX;
  ft.hi = INT32_MIN;
  const uint32_t bc = j.kind == 1 || j.kind == 2 ? zbands(n, j.levels) : 0;
  if (bc == 0) {
    V8 acc = ft.fetch(0);
    ft.put(0, acc);
    for (uint32_t k = 1; k < n; ++k) {
      V8 v = ft.fetch(k);
      if (j.kind == 0) {
#pragma unroll
        for (int e = 0; e < (int)kFS; ++e) v.v[e] = wadd32(v.v[e], acc.v[e]);
      }
      ft.put(k, v);
      acc = v;
    }
  } else {
    const uint32_t len1 = (n + 1) / 2;
    if (bc == 1) {
      RowSrc s0{&ft, 0};
      ZLevel<RowSrc> l1;
      l1.init(&ft, &s0, j.kind, n, len1);
      for (uint32_t k = 0; k < n; ++k) ft.put(k, l1.next());
    } else {
      const uint32_t ns2 = (len1 + 1) / 2;
      RowSrc s0{&ft, 0};
      ZLevel<RowSrc> l2;
      l2.init(&ft, &s0, j.kind, len1, ns2);
      ZLevel<ZLevel<RowSrc>> l1;
      l1.init(&ft, &l2, j.kind, n, len1);
      for (uint32_t k = 0; k < n; ++k) ft.put(k, l1.next());
    }
  }
  if (!ft.ok) atomicOr(j.flag, 1);
  if (!j.dst) {  // range mode: one atomic pair per wave
    int32_t lo = ft.lo, hi = ft.hi;
    for (int off = 32; off > 0; off >>= 1) {
      const int32_t a = __shfl_xor(lo, off), b = __shfl_xor(hi, off);
      lo = a < lo ? a : lo;
      hi = b > hi ? b : hi;
    }
    if ((threadIdx.x & 63) == 0 && lo <= hi) {
      atomicMin((long long*)&j.minmax[0], (long long)lo);
      atomicMax((long long*)&j.minmax[1], (long long)hi);
    }
  }
}

}  // namespace

void launch_dwt_lift(const DwtLiftJob* d_jobs, uint32_t njobs, uint64_t items, hipStream_t stream) {
  if (items == 0) return;
  hipLaunchKernelGGL(dwt_lift_kernel, xcd_grid(items), dim3(256), 0, stream, d_jobs, njobs);
  NDIC_HIP(hipGetLastError());
}

void launch_dwt_pass(int pass, const DwtJob* d_jobs, uint32_t njobs, uint64_t items, hipStream_t stream) {
  if (items == 0) return;
  uint64_t blocks = ceil_div(items, 4);
  uint32_t gx = (uint32_t)(blocks < 65535 ? blocks : 65535);
  uint32_t gy = (uint32_t)ceil_div(blocks, gx);
  switch (pass) {
    case kDwtInvH: hipLaunchKernelGGL(dwt_inv_h, dim3(gx, gy), dim3(256), 0, stream, d_jobs, njobs); break;
    case kDwtInvV: hipLaunchKernelGGL(dwt_inv_v, dim3(gx, gy), dim3(256), 0, stream, d_jobs, njobs); break;
    case kDwtFwdV: hipLaunchKernelGGL(dwt_fwd_v, dim3(gx, gy), dim3(256), 0, stream, d_jobs, njobs); break;
    case kDwt97InvH: hipLaunchKernelGGL(dwt97_inv_h, dim3(gx, gy), dim3(256), 0, stream, d_jobs, njobs); break;
    case kDwt97InvV: hipLaunchKernelGGL(dwt97_inv_v, dim3(gx, gy), dim3(256), 0, stream, d_jobs, njobs); break;
    case kDwtInvLevel:  // one workgroup per tile, XCD-aware order
      hipLaunchKernelGGL(dwt_inv_level, xcd_grid(items), dim3(256), 0, stream, d_jobs, njobs);
      break;
    default: hipLaunchKernelGGL(dwt_fwd_h, dim3(gx, gy), dim3(256), 0, stream, d_jobs, njobs); break;
  }
  NDIC_HIP(hipGetLastError());
}

}  // namespace ndic
