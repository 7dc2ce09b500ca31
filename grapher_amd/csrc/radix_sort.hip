// radix_sort.hip -- stable LSD radix sort of (uint64 key, uint32 value).
#include "radix_sort.hpp"

namespace gvh {
namespace {

constexpr int T = RadixSorter::kThreads;
constexpr int I = RadixSorter::kItems;
constexpr int TILE = RadixSorter::kTile;
constexpr int WAVES = T / kWave;
static_assert(T % 256 == 0, "one or more threads per digit");

__global__ __launch_bounds__(T) void k_hist(const uint64_t *__restrict__ keys,
                                            int64_t n, int shift, int tiles,
                                            uint32_t *__restrict__ table) {
  __shared__ uint32_t hist[256];
  if (threadIdx.x < 256)
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
  if (threadIdx.x < 256)
    table[(int64_t)threadIdx.x * tiles + blockIdx.x] = hist[threadIdx.x];
}

// Exclusive scan of one digit's row of per-tile counts (one block per
// digit), in place; the row total goes to totals[digit]. Each thread owns a
// contiguous chunk of the row, so the row length is not bounded by LDS.
__global__ __launch_bounds__(T) void k_rowscan(uint32_t *__restrict__ table,
                                               int tiles,
                                               uint32_t *__restrict__ totals) {
  __shared__ uint32_t part[T];
  uint32_t *r = table + (int64_t)blockIdx.x * tiles;
  const int chunk = (tiles + T - 1) / T;
  const int64_t b = (int64_t)threadIdx.x * chunk;
  const int64_t e = b + chunk < tiles ? b + chunk : tiles;
  uint32_t sum = 0;
  for (int64_t i = b; i < e; i++)
    sum += r[i];
  part[threadIdx.x] = sum;
  __syncthreads();
  for (int off = 1; off < T; off <<= 1) {
    uint32_t v = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0u;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  uint32_t run = threadIdx.x ? part[threadIdx.x - 1] : 0u;
  for (int64_t i = b; i < e; i++) {
    uint32_t v = r[i];
    r[i] = run;
    run += v;
  }
  if (threadIdx.x == T - 1)
    totals[blockIdx.x] = part[T - 1];
}

__global__ __launch_bounds__(T) void k_scatter(
    const uint64_t *__restrict__ keysIn, const uint32_t *__restrict__ valsIn,
    uint64_t *__restrict__ keysOut, uint32_t *__restrict__ valsOut, int64_t n,
    int shift, int tiles, const uint32_t *__restrict__ table,
    const uint32_t *__restrict__ totals) {
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

  if (tid < 256) {
    gOff[tid] = table[(int64_t)tid * tiles + blockIdx.x];
    running[tid] = totals[tid];
    tStart[tid] = 0;
  }
  for (int i = tid; i < WAVES * 256; i += T)
    waveCnt[i / 256][i % 256] = 0;
  __syncthreads();

  // Digit bases: exclusive scan of the 256 digit totals (wave 0).
  if (tid < kWave) {
    uint32_t c0 = running[4 * tid], c1 = running[4 * tid + 1];
    uint32_t c2 = running[4 * tid + 2], c3 = running[4 * tid + 3];
    uint32_t s4 = c0 + c1 + c2 + c3;
    uint32_t incl = s4;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      uint32_t v = __shfl_up(incl, off, kWave);
      if (lane >= off)
        incl += v;
    }
    uint32_t ex = incl - s4;
    gOff[4 * tid] += ex;
    gOff[4 * tid + 1] += ex + c0;
    gOff[4 * tid + 2] += ex + c0 + c1;
    gOff[4 * tid + 3] += ex + c0 + c1 + c2;
  }
  __syncthreads();
  // This tile's digit counts, read back from the scanned table (the next
  // tile's offset minus this one's) -> exclusive digit starts in the tile.
  if (tid < 256) {
    const int64_t at = (int64_t)tid * tiles + blockIdx.x;
    const uint32_t next = (int)blockIdx.x + 1 < tiles ? table[at + 1] : totals[tid];
    running[tid] = next - table[at];
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
  if (tid < 256)
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
    if (tid < 256) {
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
  if (keysAlt.alloc(n) || valsAlt.alloc(n) || table.alloc(256 * tiles) ||
      totals.alloc(256))
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
    hipLaunchKernelGGL(k_rowscan, dim3(256), dim3(kThreads), 0, stream, table.p,
                       tiles, totals.p);
    hipLaunchKernelGGL(k_scatter, dim3(tiles), dim3(kThreads), 0, stream, kin,
                       vin, kout, vout, n, shift, tiles, table.p, totals.p);
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
