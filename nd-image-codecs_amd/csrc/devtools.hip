// Device-memory helpers and the synthetic-volume generator used by the test
// suite and the benchmark (exported through ndic_hip_dev.h).  They run on the
// library's own stream and HIP runtime, so a caller needs no second GPU
// framework in the process.
#include <cmath>

#include "common.hpp"
#include "engine.hpp"
#include "../../include/ndic_hip_dev.h"

namespace ndic {
namespace {

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

// (0, 1] from 24 random bits
__device__ __forceinline__ float unit(uint64_t r) { return ((float)(r >> 40) + 1.0f) * (1.0f / 16777216.0f); }

__global__ __launch_bounds__(256) void synth_kernel(uint8_t* __restrict__ out, uint32_t Z, uint32_t Y, uint32_t X,
                                                    int32_t dtype, int32_t content, ndic_synth_blobs blobs,
                                                    uint64_t seed) {
  uint64_t n = (uint64_t)Z * Y * X;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    int32_t v = 0;
    if (content == NDIC_SYNTH_RANDOM) {
      uint64_t r = mix64(seed ^ (i * 0xd6e8feb86659fd93ull));
      if (dtype == NDIC_U8) v = (int32_t)(r & 0xff);
      else if (dtype == NDIC_U16) v = (int32_t)(r & 0xffff);
      else v = (int32_t)(r & 0xffff) - 32768;
    } else if (content == NDIC_SYNTH_BLOBS) {
      uint32_t x = (uint32_t)(i % X);
      uint64_t t = i / X;
      uint32_t y = (uint32_t)(t % Y), z = (uint32_t)(t / Y);
      float lam = 1000.0f;
      for (int b = 0; b < NDIC_SYNTH_MAX_BLOBS && b < blobs.count; ++b) {
        float dz = (float)z - blobs.cz[b], dy = (float)y - blobs.cy[b], dx = (float)x - blobs.cx[b];
        float s2 = 2.0f * blobs.sigma[b] * blobs.sigma[b];
        lam += blobs.amp[b] * __expf(-(dz * dz + dy * dy + dx * dx) / s2);
      }
      // Poisson(lam) by its normal approximation (lam >= 1000 here)
      uint64_t r1 = mix64(seed ^ (2 * i + 1) * 0xd6e8feb86659fd93ull), r2 = mix64(r1);
      float g = sqrtf(-2.0f * __logf(unit(r1))) * __cosf(6.2831853f * unit(r2));
      float f = rintf(lam + sqrtf(lam) * g);
      int32_t v12 = (int32_t)fminf(fmaxf(f, 0.0f), 4095.0f);
      if (dtype == NDIC_U8) v = v12 >> 4;
      else if (dtype == NDIC_U16) v = v12 << 4;
      else v = (v12 << 4) - 32768;
    }
    if (dtype == NDIC_U8) out[i] = (uint8_t)v;
    else ((uint16_t*)out)[i] = (uint16_t)v;
  }
}

__global__ __launch_bounds__(256) void mismatch_kernel(const uint8_t* __restrict__ a, const uint8_t* __restrict__ b,
                                                       uint64_t n, unsigned long long* count) {
  uint64_t local = 0;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint64_t nv = n / 16;
  const uint4* a4 = (const uint4*)a;
  const uint4* b4 = (const uint4*)b;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < nv; i += stride) {
    uint4 p = a4[i], q = b4[i];
    local += (p.x != q.x) + (p.y != q.y) + (p.z != q.z) + (p.w != q.w);
  }
  for (uint64_t i = nv * 16 + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += stride)
    local += a[i] != b[i];
  if (local) atomicAdd(count, (unsigned long long)local);
}

// Content hash of n bytes: the sum (mod 2^64) over the 64-bit little-endian
// words w_i (the last one zero-padded) of mix64(w_i ^ (i * K)), plus
// mix64(n ^ K); order-dependent through the word index.  Added into *slot.
constexpr uint64_t kHashK = 0x9e3779b97f4a7c15ull;
__global__ __launch_bounds__(256) void hash_kernel(const uint8_t* __restrict__ p, uint64_t n,
                                                   unsigned long long* slot) {
  const uint64_t nw = (n + 7) / 8, full = n / 8;
  uint64_t local = 0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < nw; i += stride) {
    uint64_t w = 0;
    if (i < full && ((uintptr_t)p & 7) == 0) {
      w = ((const uint64_t*)p)[i];
    } else if (i < full) {
      for (int b = 0; b < 8; ++b) w |= (uint64_t)p[8 * i + b] << (8 * b);
    } else {
      for (uint64_t b = 8 * i; b < n; ++b) w |= (uint64_t)p[b] << (8 * (b - 8 * i));
    }
    local += mix64(w ^ (i * kHashK));
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) local += mix64(n ^ kHashK);
  for (int o = 32; o >= 1; o >>= 1) local += __shfl_xor(local, o);
  if ((threadIdx.x & 63) == 0 && local) atomicAdd(slot, (unsigned long long)local);
}

__global__ __launch_bounds__(256) void copy_kernel(uint4* __restrict__ dst, const uint4* __restrict__ src, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

uint32_t grid_for(uint64_t n) {
  uint64_t g = (n + 255) / 256;
  return (uint32_t)(g < 8192 ? (g ? g : 1) : 8192);
}

}  // namespace
}  // namespace ndic

using namespace ndic;

extern "C" ndic_status ndic_dev_malloc(size_t bytes, void** out, ndic_error* err) {
  return guard_call(err, [&] {
    Context::get();
    *out = nullptr;
    if (bytes == 0) return;
    NDIC_HIP(hipMalloc(out, bytes));
  });
}

extern "C" void ndic_dev_free(void* p) {
  if (p) (void)hipFree(p);
}

extern "C" ndic_status ndic_dev_memcpy(void* dst, const void* src, size_t n, int32_t kind, ndic_error* err) {
  return guard_call(err, [&] {
    Context& ctx = Context::get();
    std::lock_guard<std::recursive_mutex> lock(ctx.mu);
    hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    if (n) NDIC_HIP(hipMemcpyAsync(dst, src, n, k, ctx.stream));
    NDIC_HIP(hipStreamSynchronize(ctx.stream));
  });
}

extern "C" ndic_status ndic_dev_memset(void* dst, int32_t value, size_t n, ndic_error* err) {
  return guard_call(err, [&] {
    Context& ctx = Context::get();
    std::lock_guard<std::recursive_mutex> lock(ctx.mu);
    if (n) NDIC_HIP(hipMemsetAsync(dst, value, n, ctx.stream));
    NDIC_HIP(hipStreamSynchronize(ctx.stream));
  });
}

extern "C" ndic_status ndic_dev_synchronize(ndic_error* err) {
  return guard_call(err, [&] {
    Context& ctx = Context::get();
    NDIC_HIP(hipStreamSynchronize(ctx.stream));
  });
}

extern "C" ndic_status ndic_dev_synth_chunk(void* d_out, uint32_t z, uint32_t y, uint32_t x, int32_t dtype,
                                            int32_t content, const ndic_synth_blobs* blobs, uint64_t seed,
                                            ndic_error* err) {
  return guard_call(err, [&] {
    if (dtype != NDIC_U8 && dtype != NDIC_U16 && dtype != NDIC_I16)
      fail(NDIC_INVALID_ARGUMENT, "synthetic volumes are u8, u16 or i16");
    Context& ctx = Context::get();
    std::lock_guard<std::recursive_mutex> lock(ctx.mu);
    uint64_t n = (uint64_t)z * y * x;
    if (n == 0) return;
    ndic_synth_blobs b{};
    if (blobs) b = *blobs;
    if (content == NDIC_SYNTH_ZERO) {
      NDIC_HIP(hipMemsetAsync(d_out, 0, n * (dtype == NDIC_U8 ? 1 : 2), ctx.stream));
    } else {
      hipLaunchKernelGGL(synth_kernel, dim3(grid_for(n)), dim3(256), 0, ctx.stream, (uint8_t*)d_out, z, y, x, dtype,
                         content, b, seed);
      NDIC_HIP(hipGetLastError());
    }
  });
}

extern "C" ndic_status ndic_dev_count_mismatch(const void* a, const void* b, size_t n, uint64_t* count,
                                               ndic_error* err) {
  return guard_call(err, [&] {
    Context& ctx = Context::get();
    std::lock_guard<std::recursive_mutex> lock(ctx.mu);
    unsigned long long* d = (unsigned long long*)ctx.small.ensure(256);
    NDIC_HIP(hipMemsetAsync(d, 0, 8, ctx.stream));
    if (n) {
      bool aligned = ((uintptr_t)a % 16 == 0) && ((uintptr_t)b % 16 == 0);
      if (!aligned) fail(NDIC_INVALID_ARGUMENT, "ndic_dev_count_mismatch needs 16-byte aligned buffers");
      hipLaunchKernelGGL(mismatch_kernel, dim3(grid_for(n / 16 + 1)), dim3(256), 0, ctx.stream, (const uint8_t*)a,
                         (const uint8_t*)b, (uint64_t)n, d);
      NDIC_HIP(hipGetLastError());
    }
    unsigned long long h = 0;
    NDIC_HIP(hipMemcpyAsync(&h, d, 8, hipMemcpyDeviceToHost, ctx.stream));
    NDIC_HIP(hipStreamSynchronize(ctx.stream));
    *count = h;
  });
}

extern "C" ndic_status ndic_dev_count_mismatch_async(const void* a, const void* b, size_t n, uint64_t* d_count,
                                                     ndic_error* err) {
  return guard_call(err, [&] {
    bool aligned = ((uintptr_t)a % 16 == 0) && ((uintptr_t)b % 16 == 0) && ((uintptr_t)d_count % 8 == 0);
    if (!aligned) fail(NDIC_INVALID_ARGUMENT, "ndic_dev_count_mismatch_async needs 16-byte aligned buffers");
    if (!n) return;
    Context& ctx = Context::get();
    std::lock_guard<std::recursive_mutex> lock(ctx.mu);
    hipLaunchKernelGGL(mismatch_kernel, dim3(grid_for(n / 16 + 1)), dim3(256), 0, ctx.stream, (const uint8_t*)a,
                       (const uint8_t*)b, (uint64_t)n, (unsigned long long*)d_count);
    NDIC_HIP(hipGetLastError());
  });
}

extern "C" ndic_status ndic_dev_hash_async(const void* p, size_t n, uint64_t* d_slot, ndic_error* err) {
  return guard_call(err, [&] {
    if ((uintptr_t)d_slot % 8) fail(NDIC_INVALID_ARGUMENT, "ndic_dev_hash_async needs an 8-byte aligned slot");
    Context& ctx = Context::get();
    std::lock_guard<std::recursive_mutex> lock(ctx.mu);
    hipLaunchKernelGGL(hash_kernel, dim3(grid_for((n + 7) / 8 + 1)), dim3(256), 0, ctx.stream, (const uint8_t*)p,
                       (uint64_t)n, (unsigned long long*)d_slot);
    NDIC_HIP(hipGetLastError());
  });
}

extern "C" ndic_status ndic_dev_copy(void* dst, const void* src, size_t n, ndic_error* err) {
  return guard_call(err, [&] {
    if (n % 16 || (uintptr_t)dst % 16 || (uintptr_t)src % 16)
      fail(NDIC_INVALID_ARGUMENT, "ndic_dev_copy moves 16-byte aligned multiples of 16 bytes");
    Context& ctx = Context::get();
    std::lock_guard<std::recursive_mutex> lock(ctx.mu);
    if (n) {
      hipLaunchKernelGGL(copy_kernel, dim3(grid_for(n / 16)), dim3(256), 0, ctx.stream, (uint4*)dst,
                         (const uint4*)src, (uint64_t)(n / 16));
      NDIC_HIP(hipGetLastError());
    }
    NDIC_HIP(hipStreamSynchronize(ctx.stream));
  });
}
