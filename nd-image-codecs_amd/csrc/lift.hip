// nd_lift on the device: delta / haar / 5-3 lifting along one axis of many
// chunks per launch.
//
// Layout: a chunk's coefficient plane is row-major, viewed per step as
// (outer, len, inner) (crates/ndic-lift/src/chunk.rs:110-150).  The axis
// samples of one column j are `inner` elements apart, so a wave that takes
// 64 consecutive columns reads and writes 256-byte contiguous rows: the sweep
// is coalesced without any transpose.  Each lane owns one column of one group
// segment and runs every level of the step on it; the segment is staged in
// LDS (two buffers: samples and the level's subband output), or, for segments
// longer than the LDS budget, in global memory with a chunk-sized temporary.
// Lanes never touch each other's columns, so no barriers are needed.
//
// Arithmetic follows crates/ndic-lift/src/kernel.rs:66-128: plain arithmetic
// forward (the budget check has run), wrapping arithmetic inverse.  The first
// forward step can widen the dtype input on load; the last inverse step can
// narrow to the dtype on store and flags out-of-range coefficients.
#include "lift.hpp"

namespace ndic {

namespace {

template <class T> struct U_;
template <> struct U_<int32_t> { typedef uint32_t t; };
template <> struct U_<int64_t> { typedef uint64_t t; };
template <class T> __device__ __forceinline__ T wadd(T a, T b) { typedef typename U_<T>::t U; return (T)((U)a + (U)b); }
template <class T> __device__ __forceinline__ T wsub(T a, T b) { typedef typename U_<T>::t U; return (T)((U)a - (U)b); }

__device__ __forceinline__ int64_t load_dtype(const uint8_t* p, uint32_t dtype) {
  switch (dtype) {
    case NDIC_U8: return p[0];
    case NDIC_I8: return (int8_t)p[0];
    case NDIC_U16: return *(const uint16_t*)p;
    case NDIC_I16: return *(const int16_t*)p;
    case NDIC_U32: return *(const uint32_t*)p;
    case NDIC_I32: return *(const int32_t*)p;
    default: return *(const int64_t*)p;  // u64 / i64 (range-checked before)
  }
}

// Narrowing store; returns false when v does not fit the dtype.
template <class T> __device__ __forceinline__ bool store_dtype(uint8_t* p, uint32_t dtype, T v) {
  switch (dtype) {
    case NDIC_U8: *p = (uint8_t)v; return v >= 0 && v <= 255;
    case NDIC_I8: *p = (uint8_t)v; return v >= -128 && v <= 127;
    case NDIC_U16: *(uint16_t*)p = (uint16_t)v; return v >= 0 && v <= 65535;
    case NDIC_I16: *(int16_t*)p = (int16_t)v; return v >= -32768 && v <= 32767;
    case NDIC_U32: *(uint32_t*)p = (uint32_t)v; return v >= 0;
    case NDIC_I32: *(int32_t*)p = (int32_t)v; return true;
    case NDIC_U64: *(uint64_t*)p = (uint64_t)v; return v >= 0;
    default: *(int64_t*)p = (int64_t)v; return true;
  }
}

// A column of one segment: element k at base[k * stride].
template <class T> struct Col {
  T* p;
  uint64_t stride;
  __device__ __forceinline__ T& operator[](uint64_t k) const { return p[k * stride]; }
};

__device__ __forceinline__ uint32_t band_count(uint64_t n, uint32_t levels) {
  uint32_t c = 0;
  uint64_t m = n;
  for (uint32_t l = 0; l < levels && m >= 2; ++l) {
    ++c;
    m = (m + 1) / 2;
  }
  return c;
}

// kernel.rs:167-183, 214-240 (one forward level over x[0..n), subband order)
template <class T> __device__ void level_fwd(uint32_t kind, Col<T> x, Col<T> tmp, uint64_t n) {
  uint64_t ns = (n + 1) / 2, nd = n / 2;
  if (kind == 1) {
    for (uint64_t i = 0; i < nd; ++i) {
      T x0 = x[2 * i], x1 = x[2 * i + 1];
      T d = x1 - x0;
      tmp[ns + i] = d;
      tmp[i] = x0 + (d >> 1);
    }
    if (n & 1) tmp[ns - 1] = x[n - 1];
  } else {
    for (uint64_t i = 0; i < nd; ++i) {
      uint64_t r = 2 * i + 2 < n ? 2 * i + 2 : 2 * n - 2 - (2 * i + 2);
      tmp[ns + i] = x[2 * i + 1] - ((x[2 * i] + x[r] + 1) >> 1);
    }
    for (uint64_t i = 0; i < ns; ++i) {
      uint64_t il = i == 0 ? 0 : i - 1, ir = i < nd - 1 ? i : nd - 1;
      tmp[i] = x[2 * i] + ((tmp[ns + il] + tmp[ns + ir] + 2) >> 2);
    }
  }
  for (uint64_t i = 0; i < n; ++i) x[i] = tmp[i];
}

// kernel.rs:186-200, 243-272 (one inverse level, wrapping arithmetic)
template <class T> __device__ void level_inv(uint32_t kind, Col<T> x, Col<T> tmp, uint64_t n) {
  uint64_t ns = (n + 1) / 2, nd = n / 2;
  if (kind == 1) {
    for (uint64_t i = 0; i < nd; ++i) {
      T d = x[ns + i];
      T x0 = wsub(x[i], (T)(d >> 1));
      tmp[2 * i] = x0;
      tmp[2 * i + 1] = wadd(x0, d);
    }
    if (n & 1) tmp[n - 1] = x[ns - 1];
  } else {
    for (uint64_t i = 0; i < ns; ++i) {
      uint64_t il = i == 0 ? 0 : i - 1, ir = i < nd - 1 ? i : nd - 1;
      tmp[2 * i] = wsub(x[i], (T)(wadd(wadd(x[ns + il], x[ns + ir]), (T)2) >> 2));
    }
    for (uint64_t i = 0; i < nd; ++i) {
      uint64_t r = 2 * i + 2 < n ? 2 * i + 2 : 2 * n - 2 - (2 * i + 2);
      tmp[2 * i + 1] = wadd(x[ns + i], (T)(wadd(wadd(tmp[2 * i], tmp[r]), (T)1) >> 1));
    }
  }
  for (uint64_t i = 0; i < n; ++i) x[i] = tmp[i];
}

template <class T> __device__ void transform_column(const LiftJob& job, bool fwd, Col<T> x, Col<T> tmp, uint64_t n) {
  if (n < 2) return;
  if (job.kind == 0) {
    if (fwd) {
      for (uint64_t k = n - 1; k >= 1; --k) x[k] = x[k] - x[k - 1];
    } else {
      for (uint64_t k = 1; k < n; ++k) x[k] = wadd(x[k], x[k - 1]);
    }
    return;
  }
  uint32_t bc = band_count(n, job.levels);
  if (fwd) {
    for (uint32_t k = 0; k < bc; ++k) level_fwd<T>(job.kind, x, tmp, (n + (1ull << k) - 1) >> k);
  } else {
    for (uint32_t k = bc; k-- > 0;) level_inv<T>(job.kind, x, tmp, (n + (1ull << k) - 1) >> k);
  }
}

__device__ __forceinline__ uint32_t find_job(const LiftJob* jobs, uint32_t njobs, uint64_t item) {
  uint32_t lo = 0, hi = njobs;
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) / 2;
    if (jobs[mid].item_begin <= item) lo = mid;
    else hi = mid;
  }
  return lo;
}

// One wave (64 lanes) per work item: a tile of 64 columns of one group
// segment of one outer slab of one chunk-step.
template <class T, bool kLds>
__global__ __launch_bounds__(64) void lift_step_kernel(const LiftJob* __restrict__ jobs, uint32_t njobs, bool fwd,
                                                       uint32_t lds_n) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  uint64_t item = blockIdx.x + (uint64_t)blockIdx.y * gridDim.x;
  uint32_t ji = find_job(jobs, njobs, item);
  const LiftJob job = jobs[ji];
  uint64_t local = item - job.item_begin;
  if (local >= job.items) return;
  uint64_t t = local % job.tiles, rest = local / job.tiles;
  uint64_t s = rest % job.segs, o = rest / job.segs;
  uint64_t j = t * 64 + threadIdx.x;
  if (j >= job.inner) return;
  uint64_t g0 = s * job.group;
  uint64_t n = job.len - g0 < job.group ? job.len - g0 : job.group;
  uint64_t elem = o * job.len * job.inner + g0 * job.inner + j;
  T* plane = (T*)job.plane;
  Col<T> g{plane + elem, job.inner};

  if (kLds) {
    T* lx = (T*)lds_raw;
    Col<T> x{lx + threadIdx.x, 64};
    Col<T> tmp{lx + (uint64_t)lds_n * 64 + threadIdx.x, 64};
    if (fwd && job.src) {
      for (uint64_t k = 0; k < n; ++k) x[k] = (T)load_dtype(job.src + (elem + k * job.inner) * job.dsize, job.dtype);
    } else {
      for (uint64_t k = 0; k < n; ++k) x[k] = g[k];
    }
    if (job.kind != 3) transform_column<T>(job, fwd, x, tmp, n);
    if (!fwd && job.dst) {
      bool ok = true;
      for (uint64_t k = 0; k < n; ++k) ok &= store_dtype<T>(job.dst + (elem + k * job.inner) * job.dsize, job.dtype, x[k]);
      if (!ok) atomicOr(job.flag, 1);
    } else {
      for (uint64_t k = 0; k < n; ++k) g[k] = x[k];
    }
  } else {
    Col<T> tmp{(T*)job.tmp + elem, job.inner};
    if (fwd && job.src)
      for (uint64_t k = 0; k < n; ++k) g[k] = (T)load_dtype(job.src + (elem + k * job.inner) * job.dsize, job.dtype);
    if (job.kind != 3) transform_column<T>(job, fwd, g, tmp, n);
    if (!fwd && job.dst) {
      bool ok = true;
      for (uint64_t k = 0; k < n; ++k) ok &= store_dtype<T>(job.dst + (elem + k * job.inner) * job.dsize, job.dtype, g[k]);
      if (!ok) atomicOr(job.flag, 1);
    }
  }
}

// LDS tile kernel: one 256-thread workgroup per 64-column tile of one group
// segment (n <= the LDS limit).  The segment is loaded once with coalesced
// 256-byte rows, every level runs with the four waves spread over the
// segment's sample pairs (predict / update phases separated by barriers,
// ping-pong between two LDS buffers), and the result is stored once (narrowed
// to the dtype on the last inverse step).  Same arithmetic as the per-column
// reference loops above.
template <class T>
__global__ __launch_bounds__(256) void lift_tile_kernel(const LiftJob* __restrict__ jobs, uint32_t njobs, bool fwd,
                                                        uint32_t nmax) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T* X = (T*)lds_raw;
  T* Y = X + (size_t)nmax * 64;
  uint64_t item = blockIdx.x + (uint64_t)blockIdx.y * gridDim.x;
  const LiftJob& job = jobs[find_job(jobs, njobs, item)];
  uint64_t local = item - job.item_begin;
  if (local >= job.items) return;
  const uint32_t tid = threadIdx.x, col = tid & 63, grp = tid >> 6;
  uint64_t t = local % job.tiles, rest = local / job.tiles;
  uint64_t sg = rest % job.segs, o = rest / job.segs;
  const uint64_t j = t * 64 + col;
  const bool valid = j < job.inner;
  const uint64_t g0 = sg * job.group;
  const uint32_t n = (uint32_t)(job.len - g0 < job.group ? job.len - g0 : job.group);
  const uint64_t elem0 = o * job.len * job.inner + g0 * job.inner + j;
  T* plane = (T*)job.plane;
  const uint64_t inner = job.inner;
  // load (unrolled: eight rows' loads in flight per thread)
#pragma unroll 8
  for (uint32_t k = grp; k < n; k += 4) {
    T v = 0;
    if (valid) {
      if (fwd && job.src) v = (T)load_dtype(job.src + (elem0 + k * inner) * job.dsize, job.dtype);
      else v = plane[elem0 + k * inner];
    }
    X[k * 64 + col] = v;
  }
  __syncthreads();
  if (job.kind != 3 && n >= 2) {
    if (job.kind == 0) {
      if (fwd) {
        for (uint32_t k = grp; k < n; k += 4) Y[k * 64 + col] = k == 0 ? X[col] : X[k * 64 + col] - X[(k - 1) * 64 + col];
        __syncthreads();
        for (uint32_t k = grp; k < n; k += 4) X[k * 64 + col] = Y[k * 64 + col];
      } else if (grp == 0) {
        for (uint32_t k = 1; k < n; ++k) X[k * 64 + col] = wadd(X[k * 64 + col], X[(k - 1) * 64 + col]);
      }
      __syncthreads();
    } else {
      const uint32_t bc = band_count(n, job.levels);
      for (uint32_t step = 0; step < bc; ++step) {
        const uint32_t lv = fwd ? step : bc - 1 - step;
        const uint32_t len = (uint32_t)((n + (1u << lv) - 1) >> lv);
        const uint32_t ns = (len + 1) / 2, nd = len / 2;
        auto x = [&](uint32_t k) -> T& { return X[k * 64 + col]; };
        auto y = [&](uint32_t k) -> T& { return Y[k * 64 + col]; };
        if (job.kind == 1) {  // haar (kernel.rs:167-200)
          for (uint32_t i = grp; i < nd; i += 4) {
            if (fwd) {
              T d = x(2 * i + 1) - x(2 * i);
              y(ns + i) = d;
              y(i) = x(2 * i) + (d >> 1);
            } else {
              T d = x(ns + i);
              T x0 = wsub(x(i), (T)(d >> 1));
              y(2 * i) = x0;
              y(2 * i + 1) = wadd(x0, d);
            }
          }
          if ((len & 1) && grp == 0) {
            if (fwd) y(ns - 1) = x(len - 1);
            else y(len - 1) = x(ns - 1);
          }
        } else if (fwd) {  // 5/3 analysis (kernel.rs:214-240)
          for (uint32_t i = grp; i < nd; i += 4) {
            uint32_t r = 2 * i + 2 < len ? 2 * i + 2 : 2 * len - 2 - (2 * i + 2);
            y(ns + i) = x(2 * i + 1) - ((x(2 * i) + x(r) + 1) >> 1);
          }
          __syncthreads();
          for (uint32_t i = grp; i < ns; i += 4) {
            uint32_t il = i == 0 ? 0 : i - 1, ir = i < nd - 1 ? i : nd - 1;
            y(i) = x(2 * i) + ((y(ns + il) + y(ns + ir) + 2) >> 2);
          }
        } else {  // 5/3 synthesis (kernel.rs:243-272), wrapping
          for (uint32_t i = grp; i < ns; i += 4) {
            uint32_t il = i == 0 ? 0 : i - 1, ir = i < nd - 1 ? i : nd - 1;
            y(2 * i) = wsub(x(i), (T)(wadd(wadd(x(ns + il), x(ns + ir)), (T)2) >> 2));
          }
          __syncthreads();
          for (uint32_t i = grp; i < nd; i += 4) {
            uint32_t r = 2 * i + 2 < len ? 2 * i + 2 : 2 * len - 2 - (2 * i + 2);
            y(2 * i + 1) = wadd(x(ns + i), (T)(wadd(wadd(y(2 * i), y(r)), (T)1) >> 1));
          }
        }
        __syncthreads();
        for (uint32_t k = grp; k < len; k += 4) x(k) = y(k);
        __syncthreads();
      }
    }
  }
  // store
  bool ok = true;
#pragma unroll 8
  for (uint32_t k = grp; k < n; k += 4) {
    if (!valid) break;
    T v = X[k * 64 + col];
    if (!fwd && job.dst) ok &= store_dtype<T>(job.dst + (elem0 + k * inner) * job.dsize, job.dtype, v);
    else plane[elem0 + k * inner] = v;
  }
  if (!ok) atomicOr(job.flag, 1);
}

// One inverse level on registers: s[0..ns) low band, d[0..nd) high band,
// x[0..len) the synthesised samples (same arithmetic as level_inv).  Every
// index is a compile-time constant; entries past ns / nd / len hold values
// that are never stored.
template <uint32_t H>
__device__ __forceinline__ void synth_regs(uint32_t kind, const int32_t (&s)[H], const int32_t (&d)[H], uint32_t nd,
                                           uint32_t len, int32_t (&x)[2 * H]) {
  if (kind == 1) {  // haar: x0 = s - (d >> 1), x1 = x0 + d; odd length keeps the last low sample
#pragma unroll
    for (uint32_t i = 0; i < H; ++i) {
      int32_t x0 = wsub(s[i], (int32_t)(d[i] >> 1));
      x[2 * i] = i < nd ? x0 : s[i];
      x[2 * i + 1] = wadd(x0, d[i]);
    }
  } else {  // 5/3: even samples from the two neighbouring details (mirrored), then odd ones
#pragma unroll
    for (uint32_t i = 0; i < H; ++i) {
      int32_t dl = d[i == 0 ? 0 : i - 1];
      int32_t dr = i < nd ? d[i] : dl;
      x[2 * i] = wsub(s[i], (int32_t)(wadd(wadd(dl, dr), 2) >> 2));
    }
#pragma unroll
    for (uint32_t i = 0; i < H; ++i) {
      int32_t r = 2 * i + 2 < len ? x[2 * i + 2 < 2 * H ? 2 * i + 2 : 2 * i] : x[2 * i];
      x[2 * i + 1] = wadd(d[i], (int32_t)(wadd(wadd(x[2 * i], r), 1) >> 1));
    }
  }
}

template <class D> __device__ __forceinline__ bool fits(int32_t v);
template <> __device__ __forceinline__ bool fits<uint8_t>(int32_t v) { return v >= 0 && v <= 255; }
template <> __device__ __forceinline__ bool fits<int8_t>(int32_t v) { return v >= -128 && v <= 127; }
template <> __device__ __forceinline__ bool fits<uint16_t>(int32_t v) { return v >= 0 && v <= 65535; }
template <> __device__ __forceinline__ bool fits<int16_t>(int32_t v) { return v >= -32768 && v <= 32767; }
template <> __device__ __forceinline__ bool fits<uint32_t>(int32_t v) { return v >= 0; }
template <> __device__ __forceinline__ bool fits<int32_t>(int32_t) { return true; }

template <class D, uint32_t N>
__device__ __forceinline__ bool store_rows(D* p, uint64_t inner, const int32_t (&v)[N], uint32_t n) {
  bool ok = true;
#pragma unroll
  for (uint32_t k = 0; k < N; ++k) {
    if (k < n) {
      p[k * inner] = (D)v[k];
      ok &= fits<D>(v[k]);
    }
  }
  return ok;
}

// Register-resident inverse step of int32 planes: one lane per column of one
// segment of at most N rows and at most two synthesis levels, four
// independent waves per workgroup (one work item each).  Every row of the
// segment is loaded at once (the bands from their run-time row offsets into
// compile-time register slots), both levels run on registers, and each row is
// stored once, narrowed to the dtype on the last step.  No LDS, no barriers:
// occupancy is set by registers alone and each lane has up to N loads in
// flight.
// Register limits (amdgpu_waves_per_eu): the 64-row form at four waves per
// SIMD (128 registers; three at the unconstrained 133) and the 32-row form at
// eight (64; seven at 68) keep more rows in flight than they spill (32 bytes).
template <uint32_t N>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(N >= 64 ? 4 : N >= 32 ? 8 : 1))) void
lift_col_kernel(const LiftJob* __restrict__ jobs, uint32_t njobs) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint64_t item = (blockIdx.x + (uint64_t)blockIdx.y * gridDim.x) * 4 + wave;
  const LiftJob& job = jobs[find_job(jobs, njobs, item)];
  const uint64_t local = item - job.item_begin;
  if (local >= job.items) return;
  const uint64_t t = local % job.tiles, rest = local / job.tiles;
  const uint64_t sg = rest % job.segs, o = rest / job.segs;
  const uint64_t inner = job.inner;
  if (t * 64 + lane >= inner) return;
  const uint64_t g0 = sg * job.group;
  const uint32_t n = (uint32_t)(job.len - g0 < job.group ? job.len - g0 : job.group);
  const uint64_t row0 = (o * job.len + g0) * inner + t * 64 + lane;
  const int32_t* in = (const int32_t*)job.plane + row0;
  int32_t out[N];
  const uint32_t bc = job.kind == 1 || job.kind == 2 ? band_count(n, job.levels) : 0;
  if (bc == 0) {
#pragma unroll
    for (uint32_t k = 0; k < N; ++k) out[k] = k < n ? in[k * inner] : 0;
    if (job.kind == 0) {  // delta: wrapping prefix sum down the column
#pragma unroll
      for (uint32_t k = 1; k < N; ++k) out[k] = wadd(out[k], out[k - 1]);
    }
  } else {
    // level 0 bands: low rows [0, len1), high rows [len1, n)
    const uint32_t len1 = (n + 1) / 2, nd1 = n / 2;
    int32_t x1[N / 2], d1[N / 2];
#pragma unroll
    for (uint32_t i = 0; i < N / 2; ++i) d1[i] = i < nd1 ? in[(len1 + i) * inner] : 0;
    if (bc == 2) {  // level 1 bands inside the low rows: [0, ns2), [ns2, len1)
      const uint32_t ns2 = (len1 + 1) / 2, nd2 = len1 / 2;
      int32_t s2[N / 4], d2[N / 4];
#pragma unroll
      for (uint32_t i = 0; i < N / 4; ++i) {
        s2[i] = i < ns2 ? in[i * inner] : 0;
        d2[i] = i < nd2 ? in[(ns2 + i) * inner] : 0;
      }
      synth_regs<N / 4>(job.kind, s2, d2, nd2, len1, x1);
    } else {
#pragma unroll
      for (uint32_t i = 0; i < N / 2; ++i) x1[i] = i < len1 ? in[i * inner] : 0;
    }
    synth_regs<N / 2>(job.kind, x1, d1, nd1, n, out);
  }
  bool ok = true;
  if (job.dst) {
    uint8_t* p = job.dst + row0 * job.dsize;
    switch (job.dtype) {
      case NDIC_U8: ok = store_rows<uint8_t, N>(p, inner, out, n); break;
      case NDIC_I8: ok = store_rows<int8_t, N>((int8_t*)p, inner, out, n); break;
      case NDIC_U16: ok = store_rows<uint16_t, N>((uint16_t*)p, inner, out, n); break;
      case NDIC_I16: ok = store_rows<int16_t, N>((int16_t*)p, inner, out, n); break;
      case NDIC_U32: ok = store_rows<uint32_t, N>((uint32_t*)p, inner, out, n); break;
      default: ok = store_rows<int32_t, N>((int32_t*)p, inner, out, n); break;
    }
  } else {
    store_rows<int32_t, N>((int32_t*)job.plane + row0, inner, out, n);
  }
  if (!ok) atomicOr(job.flag, 1);
}

// Grid-stride widen (dtype -> plane) with per-chunk min/max for the budget
// check, or narrow (plane -> dtype) with a range flag.
template <class T>
__global__ void convert_kernel(const uint8_t* __restrict__ src, uint32_t src_dtype, T* __restrict__ plane,
                               uint8_t* __restrict__ dst, uint32_t dst_dtype, uint64_t n, int64_t* minmax,
                               int32_t* flag) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  int64_t lo = INT64_MAX, hi = INT64_MIN;
  bool ok = true;
  size_t ss = src ? (src_dtype == 0 || src_dtype == 1 ? 1 : src_dtype < 4 ? 2 : src_dtype < 6 ? 4 : 8) : 0;
  size_t ds = dst ? (dst_dtype == 0 || dst_dtype == 1 ? 1 : dst_dtype < 4 ? 2 : dst_dtype < 6 ? 4 : 8) : 0;
  for (; i < n; i += stride) {
    T v;
    if (src) {
      int64_t w = load_dtype(src + i * ss, src_dtype);
      if (src_dtype == NDIC_U64 && *(const uint64_t*)(src + i * 8) > (uint64_t)INT64_MAX) ok = false;
      if (src_dtype == NDIC_U32 && w > (int64_t)INT32_MAX) ok = false;
      v = (T)w;
      if (plane) plane[i] = v;
    } else {
      v = plane[i];
    }
    lo = (int64_t)v < lo ? (int64_t)v : lo;
    hi = (int64_t)v > hi ? (int64_t)v : hi;
    if (dst) ok &= store_dtype<T>(dst + i * ds, dst_dtype, v);
  }
  if (minmax) {
    // wave, then workgroup reduction: one atomic pair per workgroup (atomics
    // on one address serialise, so one per wave costs more than the pass)
    for (int off = 32; off > 0; off >>= 1) {
      int64_t a = __shfl_xor(lo, off), b = __shfl_xor(hi, off);
      lo = a < lo ? a : lo;
      hi = b > hi ? b : hi;
    }
    __shared__ int64_t wlo[4], whi[4];  // blockDim.x == 256: four waves
    const uint32_t wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
      wlo[wv] = lo;
      whi[wv] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (uint32_t k = 1; k < blockDim.x / 64; ++k) {
        lo = wlo[k] < lo ? wlo[k] : lo;
        hi = whi[k] > hi ? whi[k] : hi;
      }
      atomicMin((long long*)&minmax[0], (long long)lo);
      atomicMax((long long*)&minmax[1], (long long)hi);
    }
  }
  if (!ok && flag) atomicOr(flag, 1);
}

}  // namespace

size_t lift_lds_limit(bool wide) { return wide ? 64 : 128; }

uint32_t lift_col_limit() { return 64; }

void launch_lift_step(const LiftJob* d_jobs, uint32_t njobs, uint64_t items, bool fwd, bool wide, uint32_t max_n,
                      bool col, hipStream_t stream) {
  if (items == 0) return;
  if (col) {
    if (fwd || wide || max_n > lift_col_limit()) fail(NDIC_IO, "internal: register lift launch out of its range");
    uint64_t blocks = ceil_div(items, 4);
    uint32_t bx = (uint32_t)(blocks < 65535 ? blocks : 65535);
    uint32_t by = (uint32_t)ceil_div(blocks, bx);
    auto k = max_n <= 4 ? lift_col_kernel<4> : max_n <= 8 ? lift_col_kernel<8> : max_n <= 16 ? lift_col_kernel<16>
           : max_n <= 32 ? lift_col_kernel<32> : lift_col_kernel<64>;
    hipLaunchKernelGGL(k, dim3(bx, by), dim3(256), 0, stream, d_jobs, njobs);
    NDIC_HIP(hipGetLastError());
    return;
  }
  uint32_t gx = (uint32_t)(items < 65535 ? items : 65535);
  uint32_t gy = (uint32_t)ceil_div(items, gx);
  bool lds = max_n <= lift_lds_limit(wide);
  size_t shmem = lds ? (size_t)2 * max_n * 64 * (wide ? 8 : 4) : 0;
  if (lds) {
    if (wide)
      hipLaunchKernelGGL((lift_tile_kernel<int64_t>), dim3(gx, gy), dim3(256), shmem, stream, d_jobs, njobs, fwd, max_n);
    else
      hipLaunchKernelGGL((lift_tile_kernel<int32_t>), dim3(gx, gy), dim3(256), shmem, stream, d_jobs, njobs, fwd, max_n);
  } else if (wide) {
    if (lds) hipLaunchKernelGGL((lift_step_kernel<int64_t, true>), dim3(gx, gy), dim3(64), shmem, stream, d_jobs, njobs, fwd, max_n);
    else hipLaunchKernelGGL((lift_step_kernel<int64_t, false>), dim3(gx, gy), dim3(64), 0, stream, d_jobs, njobs, fwd, max_n);
  } else {
    if (lds) hipLaunchKernelGGL((lift_step_kernel<int32_t, true>), dim3(gx, gy), dim3(64), shmem, stream, d_jobs, njobs, fwd, max_n);
    else hipLaunchKernelGGL((lift_step_kernel<int32_t, false>), dim3(gx, gy), dim3(64), 0, stream, d_jobs, njobs, fwd, max_n);
  }
  NDIC_HIP(hipGetLastError());
}

void launch_convert(const uint8_t* src, uint32_t src_dtype, void* plane, bool wide, uint8_t* dst, uint32_t dst_dtype,
                    uint64_t n, int64_t* minmax, int32_t* flag, hipStream_t stream) {
  if (n == 0) return;
  uint32_t blocks = (uint32_t)(ceil_div(n, 256) < 1024 ? ceil_div(n, 256) : 1024);
  if (wide)
    hipLaunchKernelGGL(convert_kernel<int64_t>, dim3(blocks), dim3(256), 0, stream, src, src_dtype, (int64_t*)plane,
                       dst, dst_dtype, n, minmax, flag);
  else
    hipLaunchKernelGGL(convert_kernel<int32_t>, dim3(blocks), dim3(256), 0, stream, src, src_dtype, (int32_t*)plane,
                       dst, dst_dtype, n, minmax, flag);
  NDIC_HIP(hipGetLastError());
}

}  // namespace ndic
