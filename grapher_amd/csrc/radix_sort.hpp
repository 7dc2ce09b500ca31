// radix_sort.hpp -- stable LSD radix sort of (uint64 key, uint32 value)
// pairs for gfx950, 8-bit digits.
//
// Per pass: (1) per-tile 256-bin digit histograms, (2) one block per digit
// scans that digit's row of the [256 x tiles] table, (3) a stable scatter: each
// tile ranks its keys with wave ballots (stable: tile order = key order),
// stages them digit-sorted in LDS and writes each digit's run contiguously.
#pragma once

#include "common.hpp"

namespace gvh {

#ifndef GVH_SORT_THREADS
#define GVH_SORT_THREADS 512
#endif
#ifndef GVH_SORT_ITEMS
#define GVH_SORT_ITEMS 4
#endif

struct RadixSorter {
  static constexpr int kThreads = GVH_SORT_THREADS;  // a multiple of 256
  static constexpr int kItems = GVH_SORT_ITEMS;      // keys per thread
  static constexpr int kTile = kThreads * kItems;    // keys per tile
  DevBuf<uint64_t> keysAlt;
  DevBuf<uint32_t> valsAlt;
  DevBuf<uint32_t> table; // 256 * tiles
  DevBuf<uint32_t> totals; // 256
  int64_t capacity = 0;

  int reserve(int64_t n);
  // Sorts keys[0..n) / vals[0..n) by key bits [beginBit, endBit). On return the
  // sorted data is in keys/vals (the alternate buffers are scratch).
  int sort(uint64_t *keys, uint32_t *vals, int64_t n, int beginBit, int endBit,
           hipStream_t stream);
};

} // namespace gvh
