// common.hpp -- shared device/host helpers for the gfx950 layout engine.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>

namespace gvh {

constexpr int kWave = 64;

// Error plumbing: every C-ABI entry point returns 0 / negative and records a
// message that gvizhip_last_error() hands back.
void set_error(const char *fmt, ...);

#define GVH_CHECK(expr)                                                       \
  do {                                                                        \
    hipError_t _e = (expr);                                                   \
    if (_e != hipSuccess) {                                                   \
      ::gvh::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), \
                       __FILE__, __LINE__);                                   \
      return -1;                                                              \
    }                                                                         \
  } while (0)

#define GVH_CHECK_LAUNCH()                                                    \
  do {                                                                        \
    hipError_t _e = hipGetLastError();                                        \
    if (_e != hipSuccess) {                                                   \
      ::gvh::set_error("kernel launch failed: %s (%s:%d)",                    \
                       hipGetErrorString(_e), __FILE__, __LINE__);            \
      return -1;                                                              \
    }                                                                         \
  } while (0)

// Owned device allocation.
template <typename T> struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  int alloc(size_t count) {
    if (count <= n && p)
      return 0;
    release();
    size_t bytes = (count ? count : 1) * sizeof(T);
    if (hipMalloc((void **)&p, bytes) != hipSuccess) {
      p = nullptr;
      n = 0;
      set_error("hipMalloc of %zu bytes failed", bytes);
      return -1;
    }
    n = count;
    return 0;
  }
  void release() {
    if (p)
      (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  ~DevBuf() { release(); }
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
};

inline unsigned grid_for(int64_t n, int block) {
  int64_t g = (n + block - 1) / block;
  return (unsigned)(g > 0 ? g : 1);
}

// ------------------------------------------------------------- device math
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
    v += __shfl_xor(v, off, kWave);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
    v = fmax(v, __shfl_xor(v, off, kWave));
  return v;
}
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
    v = fmin(v, __shfl_xor(v, off, kWave));
  return v;
}

// Block reduction of up to 4 doubles (sum or max per slot). blockDim <= 1024.
template <int K>
__device__ __forceinline__ void block_reduce(double (&v)[K], const bool (&isMax)[K],
                                             double *smem /* K*16 */) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const int nwaves = (blockDim.x + kWave - 1) / kWave;
#pragma unroll
  for (int k = 0; k < K; k++)
    v[k] = isMax[k] ? wave_max(v[k]) : wave_sum(v[k]);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < K; k++)
      smem[k * 16 + wave] = v[k];
  __syncthreads();
  if (wave == 0) {
#pragma unroll
    for (int k = 0; k < K; k++) {
      double x = lane < nwaves ? smem[k * 16 + lane] : (isMax[k] ? -INFINITY : 0.0);
      v[k] = isMax[k] ? wave_max(x) : wave_sum(x);
    }
  }
  __syncthreads();
}

} // namespace gvh
