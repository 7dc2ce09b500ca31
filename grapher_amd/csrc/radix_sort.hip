// radix_sort.hip -- stable LSD radix sort of (uint64 key, uint32 value).
#include "radix_sort.hpp"

namespace gvh {
namespace {

constexpr int T = RadixSorter::kThreads;
constexpr int I = RadixSorter::kItems;
constexpr int TILE = RadixSorter::kTile;
constexpr int WAVES = T / kWave;

__global__ __launch_bounds__(T) void k_hist(const uint64_t *__restrict__ keys,
                                            int64_t n, int shift, int tiles,
                                            uint32_t *__restrict__ table) {
  __shared__ uint32_t hist[256];
  hist[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * TILE;
#pragma unroll 4
  for (int r = 0; r < I; r++) {
    int64_t idx = base + r * T + threadIdx.x;
    if (idx < n) {
      uint32_t d = (uint32_t)(keys[idx] >> shift) & 255u;
      atomicAdd(&hist[d], 1u);
    }
  }
  __syncthreads();
  table[(int64_t)threadIdx.x * tiles + blockIdx.x] = hist[threadIdx.x];
}

// Exclusive scan of table[0..m) by one block of 1024 threads.
__global__ __launch_bounds__(1024) void k_scan(uint32_t *__restrict__ table,
                                               int64_t m) {
  __shared__ uint32_t sums[1024];
  const int64_t chunk = (m + 1023) / 1024;
  const int64_t b = threadIdx.x * chunk;
  const int64_t e = b + chunk < m ? b + chunk : m;
  uint32_t s = 0;
  for (int64_t i = b; i < e; i++)
    s += table[i];
  sums[threadIdx.x] = s;
  __syncthreads();
  // Hillis-Steele inclusive scan over 1024 partial sums.
  for (int off = 1; off < 1024; off <<= 1) {
    uint32_t v = threadIdx.x >= (unsigned)off ? sums[threadIdx.x - off] : 0u;
    __syncthreads();
    sums[threadIdx.x] += v;
    __syncthreads();
  }
  uint32_t run = threadIdx.x ? sums[threadIdx.x - 1] : 0u;
  for (int64_t i = b; i < e; i++) {
    uint32_t v = table[i];
    table[i] = run;
    run += v;
  }
}

__global__ __launch_bounds__(T) void k_scatter(
    const uint64_t *__restrict__ keysIn, const uint32_t *__restrict__ valsIn,
    uint64_t *__restrict__ keysOut, uint32_t *__restrict__ valsOut, int64_t n,
    int shift, int tiles, const uint32_t *__restrict__ table) {
  __shared__ uint64_t sKey[TILE];
  __shared__ uint32_t sVal[TILE];
  __shared__ uint32_t gOff[256];
  __shared__ uint32_t tStart[256];
  __shared__ uint32_t running[256];
  __shared__ uint32_t waveCnt[WAVES][256];

  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int wave = tid / kWave;
  const int64_t base = (int64_t)blockIdx.x * TILE;
  const int tileLen = (int)((n - base) < TILE ? (n - base) : TILE);

  gOff[tid] = table[(int64_t)tid * tiles + blockIdx.x];
  running[tid] = 0;
  tStart[tid] = 0;
#pragma unroll
  for (int w = 0; w < WAVES; w++)
    waveCnt[w][tid] = 0;
  __syncthreads();

  // Tile histogram -> exclusive digit starts inside the tile.
  for (int r = 0; r < I; r++) {
    int li = r * T + tid;
    if (li < tileLen) {
      uint32_t d = (uint32_t)(keysIn[base + li] >> shift) & 255u;
      atomicAdd(&running[d], 1u);
    }
  }
  __syncthreads();
  if (tid < kWave) {
    // wave 0 scans 256 counts, 4 per lane
    uint32_t c0 = running[4 * tid], c1 = running[4 * tid + 1];
    uint32_t c2 = running[4 * tid + 2], c3 = running[4 * tid + 3];
    uint32_t s = c0 + c1 + c2 + c3;
    uint32_t incl = s;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      uint32_t v = __shfl_up(incl, off, kWave);
      if (lane >= off)
        incl += v;
    }
    uint32_t ex = incl - s;
    tStart[4 * tid] = ex;
    tStart[4 * tid + 1] = ex + c0;
    tStart[4 * tid + 2] = ex + c0 + c1;
    tStart[4 * tid + 3] = ex + c0 + c1 + c2;
  }
  __syncthreads();
  running[tid] = 0;
  __syncthreads();

  const uint64_t ltMask = (lane == 0) ? 0ull : (~0ull >> (kWave - lane));
  for (int r = 0; r < I; r++) {
    const int li = r * T + tid;
    const bool valid = li < tileLen;
    uint64_t key = valid ? keysIn[base + li] : 0ull;
    uint32_t val = valid ? valsIn[base + li] : 0u;
    uint32_t d = (uint32_t)(key >> shift) & 255u;
    uint64_t match = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      bool bit = (d >> b) & 1u;
      uint64_t bal = __ballot(bit);
      match &= bit ? bal : ~bal;
    }
    const uint32_t rank = (uint32_t)__popcll(match & ltMask);
    const bool leader = valid && (match & ltMask) == 0ull;
    if (leader)
      waveCnt[wave][d] = (uint32_t)__popcll(match);
    __syncthreads();
    if (valid) {
      uint32_t pos = running[d] + rank;
      for (int w = 0; w < wave; w++)
        pos += waveCnt[w][d];
      uint32_t slot = tStart[d] + pos;
      sKey[slot] = key;
      sVal[slot] = val;
    }
    __syncthreads();
    {
      uint32_t add = 0;
#pragma unroll
      for (int w = 0; w < WAVES; w++) {
        add += waveCnt[w][tid];
        waveCnt[w][tid] = 0;
      }
      running[tid] += add;
    }
    __syncthreads();
  }

  for (int li = tid; li < tileLen; li += T) {
    uint64_t key = sKey[li];
    uint32_t d = (uint32_t)(key >> shift) & 255u;
    uint32_t dst = gOff[d] + (uint32_t)(li - (int)tStart[d]);
    keysOut[dst] = key;
    valsOut[dst] = sVal[li];
  }
}

} // namespace

int RadixSorter::reserve(int64_t n) {
  if (n <= capacity)
    return 0;
  int64_t tiles = (n + kTile - 1) / kTile;
  if (keysAlt.alloc(n) || valsAlt.alloc(n) || table.alloc(256 * tiles))
    return -1;
  capacity = n;
  return 0;
}

int RadixSorter::sort(uint64_t *keys, uint32_t *vals, int64_t n, int beginBit,
                      int endBit, hipStream_t stream) {
  if (n <= 1)
    return 0;
  if (n > 0xFFFFFFFFll) {
    set_error("radix sort: %lld keys exceed the 32-bit index range",
              (long long)n);
    return -1;
  }
  if (reserve(n))
    return -1;
  const int tiles = (int)((n + kTile - 1) / kTile);
  uint64_t *kin = keys, *kout = keysAlt.p;
  uint32_t *vin = vals, *vout = valsAlt.p;
  int passes = 0;
  for (int shift = beginBit; shift < endBit; shift += 8, passes++) {
    hipLaunchKernelGGL(k_hist, dim3(tiles), dim3(kThreads), 0, stream, kin, n,
                       shift, tiles, table.p);
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, stream, table.p,
                       (int64_t)256 * tiles);
    hipLaunchKernelGGL(k_scatter, dim3(tiles), dim3(kThreads), 0, stream, kin,
                       vin, kout, vout, n, shift, tiles, table.p);
    GVH_CHECK_LAUNCH();
    uint64_t *tk = kin;
    kin = kout;
    kout = tk;
    uint32_t *tv = vin;
    vin = vout;
    vout = tv;
  }
  if (passes & 1) {
    GVH_CHECK(hipMemcpyAsync(keys, kin, sizeof(uint64_t) * n,
                             hipMemcpyDeviceToDevice, stream));
    GVH_CHECK(hipMemcpyAsync(vals, vin, sizeof(uint32_t) * n,
                             hipMemcpyDeviceToDevice, stream));
  }
  return 0;
}

} // namespace gvh
