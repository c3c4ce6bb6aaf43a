// Sort of (group, key, index) items: the large-batch orders of the instance
// path -- AP records grouped by class in score order (instance.py:207-213 and
// the per-class selection of :311-313) and per-image score order beyond the
// LDS sort (instance.py:140).
//
// The index is part of the order, so the order is total and the result is the
// stable sort of (group, key) without any stability argument.  Tiles of
// kTileItems items are sorted in LDS by a bitonic network (small tiles: the
// network is LDS-bandwidth bound, and many small workgroups spread it over
// the CUs); pairs of sorted runs are then merged, doubling the run width each
// pass.  A merge workgroup finds where its kSortTile outputs start and end in
// the two runs by a 64-way wave search over global memory (a few round
// trips), stages that slice in LDS and merges it, kSortItems outputs per lane.
#pragma once

#include <climits>

#include "common.h"

namespace pe {

struct SortItem {
  u64 key;
  uint32_t group;
  int32_t index;
};

__device__ __forceinline__ bool item_less(const SortItem& x, const SortItem& y) {
  if (x.group != y.group) return x.group < y.group;
  if (x.key != y.key) return x.key < y.key;
  return x.index < y.index;
}

constexpr int kSortThreads = 1024;
constexpr int kSortItems = 4;
constexpr int kSortTile = kSortThreads * kSortItems;  // outputs per merge workgroup
constexpr size_t kSortLds = sizeof(SortItem) * kSortTile;
constexpr int kTileThreads = 256;
constexpr int kTileItems = 1024;  // items per sort_tiles workgroup (the first run width)

// Each tile sorted in place in LDS; padding sorts last and is not written.
__global__ __launch_bounds__(kTileThreads) void sort_tiles(const SortItem* in, SortItem* out, i64 n) {
  __shared__ SortItem tile[kTileItems];
  const i64 base = static_cast<i64>(blockIdx.x) * kTileItems;
  const int count = static_cast<int>(n - base < kTileItems ? n - base : kTileItems);
  for (int e = threadIdx.x; e < kTileItems; e += kTileThreads)
    tile[e] = e < count ? in[base + e] : SortItem{~0ULL, ~0u, INT_MAX};
  __syncthreads();
  for (int k = 2; k <= kTileItems; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int p = threadIdx.x; p < kTileItems / 2; p += kTileThreads) {
        const int e = ((p & ~(j - 1)) << 1) | (p & (j - 1)), f = e + j;  // j is a power of two
        const SortItem x = tile[e], y = tile[f];
        if (item_less(y, x) == ((e & k) == 0)) {
          tile[e] = y;
          tile[f] = x;
        }
      }
      __syncthreads();
    }
  for (int e = threadIdx.x; e < count; e += kTileThreads) out[base + e] = tile[e];
}

// Merge sorted runs of `width` items pairwise into runs of 2 * width.  Each
// workgroup writes `span` = min(kSortTile, 2 * width) outputs (both powers
// of two, so a workgroup's outputs lie in one pair of runs).
__global__ __launch_bounds__(kSortThreads) void merge_runs(const SortItem* in, SortItem* out, i64 n, i64 width,
                                                           int span) {
  extern __shared__ uint4 lds4[];
  SortItem* buf = reinterpret_cast<SortItem*>(lds4);
  __shared__ i64 split[2];
  const i64 out0 = static_cast<i64>(blockIdx.x) * span;
  const i64 pair0 = out0 / (2 * width) * (2 * width);
  const i64 a0 = pair0, la = n - a0 < width ? n - a0 : width;
  const i64 b0 = a0 + la, lb = n - b0 < width ? (n - b0 > 0 ? n - b0 : 0) : width;
  const i64 end = out0 + span < pair0 + la + lb ? out0 + span : pair0 + la + lb;
  const i64 k0 = out0 - pair0, k1 = end - pair0;
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  if (wave < 2) {
    // Co-rank of output k: the number of run-A items among the first k
    // outputs, the first i with !(A[i] < B[k - i - 1]).  Each round probes 64
    // points of the remaining range at once (every probe is inside both runs).
    const i64 k = wave == 0 ? k0 : k1;
    long long lo = k - lb > 0 ? k - lb : 0, hi = k < la ? k : la;
    while (hi > lo) {
      const long long range = hi - lo;
      const long long p = lo + (static_cast<long long>(lane + 1) * range) / 65;
      const bool below = item_less(in[a0 + p], in[b0 + k - p - 1]);
      const int c = __popcll(__ballot(below));
      const long long p_last = __shfl(p, c > 0 ? c - 1 : 0, kWave);
      const long long p_first = __shfl(p, c < 64 ? c : 63, kWave);
      if (c > 0) lo = p_last + 1;
      if (c < 64) hi = p_first;
    }
    if (lane == 0) split[wave] = lo;
  }
  __syncthreads();
  const i64 ia0 = split[0], ia1 = split[1];
  const i64 jb0 = k0 - ia0, jb1 = k1 - ia1;
  const int na = static_cast<int>(ia1 - ia0), nb = static_cast<int>(jb1 - jb0);
  for (int e = threadIdx.x; e < na + nb; e += kSortThreads)
    buf[e] = e < na ? in[a0 + ia0 + e] : in[b0 + jb0 + (e - na)];
  __syncthreads();
  const int q0 = threadIdx.x * kSortItems, total = na + nb;
  if (q0 < total) {
    int lo = q0 - nb > 0 ? q0 - nb : 0, hi = q0 < na ? q0 : na;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (item_less(buf[mid], buf[na + q0 - mid - 1]))
        lo = mid + 1;
      else
        hi = mid;
    }
    int i = lo, j = q0 - lo;
    for (int u = 0; u < kSortItems && q0 + u < total; ++u) {
      const bool take_a = j >= nb || (i < na && item_less(buf[i], buf[na + j]));
      out[out0 + q0 + u] = take_a ? buf[i++] : buf[na + j++];
    }
  }
}

// Workspace of sort_items: two buffers of n items (the first holds the input).
inline size_t sort_items_bytes(i64 n) { return 2 * sizeof(SortItem) * static_cast<size_t>(n); }

// Sorts items[0 .. n) with `spare` as the other buffer; *result is the
// buffer that holds the sorted items.  Returns kOk or the error code.
inline int sort_items(SortItem* items, SortItem* spare, i64 n, hipStream_t stream, SortItem** result) {
  *result = items;
  if (n <= 1) return kOk;
  PE_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(merge_runs), hipFuncAttributeMaxDynamicSharedMemorySize,
                             static_cast<int>(kSortLds)));
  sort_tiles<<<static_cast<unsigned>((n + kTileItems - 1) / kTileItems), kTileThreads, 0, stream>>>(items, spare, n);
  PE_LAUNCH_CHECK();
  SortItem* src = spare;
  SortItem* dst = items;
  for (i64 width = kTileItems; width < n; width *= 2) {
    const int span = static_cast<int>(2 * width < kSortTile ? 2 * width : kSortTile);
    merge_runs<<<static_cast<unsigned>((n + span - 1) / span), kSortThreads, kSortLds, stream>>>(src, dst, n, width,
                                                                                                  span);
    PE_LAUNCH_CHECK();
    SortItem* t = src;
    src = dst;
    dst = t;
  }
  *result = src;
  return kOk;
}

}  // namespace pe
