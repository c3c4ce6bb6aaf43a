// Panoptic path on gfx950: joint slot histogram, areas/validation, PQ
// matching with semantic confusion, and the FP64 finaliser.
//
// Behaviour follows the reference's CPU path (evaluator.py:228-243 histogram,
// :381-417 matching, :448-468 compute, :471-544 per-image PQ) and its native
// boundary (histogram_kernel.cu:257-341 update_cuda, compute_kernel.cu:9-95).
//
// Design (MI355X-first, not a translation of the CUDA kernels):
//  * histogram: one workgroup owns a <= 65532-pixel chunk of one image and
//    privatises the G x P histogram in LDS as packed 16-bit counters (two per
//    32-bit word; a chunk can never carry a count past 0xFFFF), so G x P up to
//    64K bins fits in 128 KiB of the 160 KiB LDS.  Maps are streamed with
//    16-byte loads; each lane folds repeated keys in a two-entry register
//    cache (Voronoi-like maps repeat a key along rows and down columns), so
//    LDS atomics are issued per key change rather than per pixel.  Only
//    nonzero bins are flushed with 64-bit global atomics.
//  * score: one workgroup per image computes areas, matches every prediction
//    slot against the GT slots (integer strict-half test, IEEE FP64 IoU),
//    adds confusion directly from the pair counts and accumulates the matched
//    IoUs.  Dataset iou_sum: every matched IoU is an exact multiple of 2^-53,
//    so the class sums are kept exactly as 2^-53 fixed point (two u64 words
//    per class) and iou_sum is the correctly rounded FP64 of that exact total:
//    deterministic, independent of batching and launch order, and within
//    1e-12 relative of the reference's FP64 fold.  Per-image (reward) sums
//    fold in prediction-slot order as the reference does.
#include <climits>
#include <mutex>
#include <type_traits>

#include "common.h"
#include "panoptic_kernels.h"
#include "finalize.h"
#include "../../include/panoptic_hip.h"

namespace pe {

typedef int v4i __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------- map pixels

// A decoded COCO panoptic PNG pixel: bytes R, G, B whose segment ID is
// R + 256 * G + 65536 * B (README.md:67-68), i.e. the little-endian 24-bit
// integer of the three bytes.  Maps are B x H x W x 3 uint8.
struct Rgb24 {
  uint8_t c[3];
};
static_assert(sizeof(Rgb24) == 3, "an RGB pixel is three bytes");

// The integer a map pixel holds (an ID or a compact slot).
template <typename T>
struct MapValue {
  using type = T;
};
template <>
struct MapValue<Rgb24> {
  using type = int32_t;
};

template <typename T>
__device__ __forceinline__ typename MapValue<T>::type read_pixel(const T* map, i64 i) {
  return as_global(map)[i];
}
template <>
__device__ __forceinline__ int32_t read_pixel<Rgb24>(const Rgb24* map, i64 i) {
  const auto* p = as_global(reinterpret_cast<const uint8_t*>(map) + 3 * i);
  return static_cast<int32_t>(p[0] | (p[1] << 8) | (p[2] << 16));
}

// Pixels 4q .. 4q + 3 with vector loads that bypass the caches (maps are read
// once): int32 one 16-byte load, int64 two, RGB three dwords (12 bytes).  The
// map base must be 16-byte aligned (4-byte for RGB).
__device__ __forceinline__ void load_quad(const int32_t* map, i64 q, int32_t (&v)[4]) {
  const v4i x = __builtin_nontemporal_load(as_global(reinterpret_cast<const v4i*>(map)) + q);
  v[0] = x.x;
  v[1] = x.y;
  v[2] = x.z;
  v[3] = x.w;
}
__device__ __forceinline__ void load_quad(const i64* map, i64 q, i64 (&v)[4]) {
  typedef long long v2l __attribute__((ext_vector_type(2)));
  const auto* p = as_global(reinterpret_cast<const v2l*>(map)) + 2 * q;
  const v2l a = __builtin_nontemporal_load(p), b = __builtin_nontemporal_load(p + 1);
  v[0] = a.x;
  v[1] = a.y;
  v[2] = b.x;
  v[3] = b.y;
}
__device__ __forceinline__ void load_quad(const Rgb24* map, i64 q, int32_t (&v)[4]) {
  const auto* p = as_global(reinterpret_cast<const uint32_t*>(reinterpret_cast<const uint8_t*>(map) + 12 * q));
  const uint32_t w0 = __builtin_nontemporal_load(p), w1 = __builtin_nontemporal_load(p + 1),
                 w2 = __builtin_nontemporal_load(p + 2);
  v[0] = static_cast<int32_t>(w0 & 0xFFFFFFu);
  v[1] = static_cast<int32_t>((w0 >> 24) | ((w1 & 0xFFFFu) << 8));
  v[2] = static_cast<int32_t>((w1 >> 16) | ((w2 & 0xFFu) << 16));
  v[3] = static_cast<int32_t>(w2 >> 8);
}

// Base alignment load_quad needs for a map of pixel width `bytes`.
inline bool quad_aligned(const void* map, int bytes) {
  return reinterpret_cast<uintptr_t>(map) % (bytes == 3 ? 4 : 16) == 0;
}

// ------------------------------------------------------------ slot lookup

template <typename T>
__device__ __forceinline__ int compact_slot(T v, int S) {
  return (v >= 0 && v < static_cast<T>(S)) ? static_cast<int>(v) : -1;
}

// Original ID -> metadata slot: binary search of the sorted per-image ID
// table (evaluator.py:147-157 semantics: ID 0 is void; padding IDs are
// INT64_MAX with slot 0 and never resolve).
template <typename IdPtr, typename SlotPtr>
__device__ __forceinline__ int lookup_slot(i64 key, IdPtr ids, SlotPtr slots, int S) {
  if (key == 0) return 0;
  if (key < 0 || key == LLONG_MAX) return -1;
  int lo = 1, hi = S;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (ids[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  if (lo == S || ids[lo] != key) return -1;
  int s = slots[lo];
  return (s > 0 && s < S) ? s : -1;
}

// Sorted per-image (ID, slot) table in global memory: binary search.
struct SortedIds {
  const i64* ids;
  const int* slots;
  int S;
  __device__ __forceinline__ int find(i64 key) const { return lookup_slot(key, ids, slots, S); }
};

typedef const __attribute__((address_space(3))) i64* LdsIds;
typedef const __attribute__((address_space(3))) int* LdsInts;
constexpr i64 kEmptyId = -1;

// Multiplicative hash of an ID folded to 32 bits; `shift` = 32 - log2(size).
__device__ __forceinline__ uint32_t hash_id(i64 key, int shift) {
  const uint32_t x = static_cast<uint32_t>(key) ^ static_cast<uint32_t>(static_cast<u64>(key) >> 32);
  return (x * 0x9E3779B1u) >> shift;
}

// The same table as an open-addressing hash in LDS (linear probing, load
// factor <= 1/2): keys[h] is an ID (kEmptyId when free) and slot[h] what the
// binary search returns for it (the slot at the lowest sorted index holding
// the ID, or -1 when that slot is out of range).  A lookup is one round of
// two independent LDS reads instead of log2(S) dependent ones; that matters
// because a wave waits for its slowest lane whenever any lane's key changes.
struct LdsHashIds {
  LdsIds keys;
  LdsInts slot;
  uint32_t mask;
  int shift;
  __device__ __forceinline__ int find(i64 key) const {
    if (key == 0) return 0;
    if (key < 0 || key == LLONG_MAX) return -1;
    for (uint32_t h = hash_id(key, shift);; h = (h + 1) & mask) {
      const i64 k = keys[h];
      const int s = slot[h];
      if (k == key) return s;
      if (k == kEmptyId) return -1;
    }
  }
};

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(3))) u32x4* LdsBuckets;
constexpr uint32_t kEmptyKey32 = 0xFFFFFFFFu;
constexpr int kProbeOn = INT_MIN;  // a full bucket without the key: probe the next

// The table for maps whose pixels are 32-bit (int32 IDs, RGB): buckets of
// four keys, so one lookup is one pair of 16-byte LDS reads and almost never
// probes further.  Bucket h is two uint4 words: keys (kEmptyKey32 when free),
// then what the binary search returns for each key.  Inserts fill a bucket's
// entries in order and move to the next bucket only when all four are taken,
// so a bucket whose last entry is free ends every probe.  IDs above INT32_MAX
// are left out: no 32-bit pixel holds them.
struct LdsBucketIds {
  LdsBuckets bucket;
  uint32_t mask;
  int shift;
  __device__ __forceinline__ uint32_t home(uint32_t key) const { return (key * 0x9E3779B1u) >> shift; }
  // The slot of `key` in one bucket (selects, no branches), -1 when the
  // bucket is not full and lacks it, kProbeOn when it is full and lacks it.
  __device__ __forceinline__ static int match(u32x4 k, u32x4 s, uint32_t key) {
    int r = k.w == kEmptyKey32 ? -1 : kProbeOn;
    r = k.w == key ? static_cast<int>(s.w) : r;
    r = k.z == key ? static_cast<int>(s.z) : r;
    r = k.y == key ? static_cast<int>(s.y) : r;
    return k.x == key ? static_cast<int>(s.x) : r;
  }
  __device__ int probe(uint32_t key, uint32_t h) const {
    for (;; h = (h + 1) & mask) {
      const int r = match(bucket[2 * h], bucket[2 * h + 1], key);
      if (r != kProbeOn) return r;
    }
  }
  __device__ __forceinline__ int find(i64 key) const {
    if (key == 0) return 0;
    if (key < 0 || key > INT_MAX) return -1;
    return probe(static_cast<uint32_t>(key), home(static_cast<uint32_t>(key)));
  }
};

template <typename T, bool LOOKUP, typename Table = SortedIds>
struct SlotResolver {
  using V = typename MapValue<T>::type;
  Table table;
  int S;
  bool mapped;  // the map holds original IDs (else compact slots)
  V last_key = -1;  // never a key that resolves: IDs are >= 0
  int last_slot = -1;
  __device__ __forceinline__ int operator()(V v) {
    if (!LOOKUP || !mapped) return compact_slot(v, S);
    if (v != last_key) {
      last_key = v;
      last_slot = table.find(static_cast<i64>(v));
    }
    return last_slot;
  }
  // Four consecutive pixels.  Only a positive key that differs from the
  // positive key before it is looked up (void or invalid specks inside a
  // segment cost nothing), and the lookups run in rounds of one per lane:
  // the wave repeats the lookup code once per round while any lane has one
  // left, usually once per quad, rather than once per pixel position.
  __device__ __forceinline__ void quad(const V (&v)[4], int (&out)[4]) {
    if (!LOOKUP || !mapped) {
#pragma unroll
      for (int j = 0; j < 4; ++j) out[j] = compact_slot(v[j], S);
      return;
    }
    V prev = last_key;
    uint32_t pending = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      pending |= static_cast<uint32_t>(v[j] > 0 && v[j] != prev) << j;
      prev = v[j] > 0 ? v[j] : prev;
    }
    const uint32_t looked = pending;
    int r[4] = {0, 0, 0, 0};
    while (pending) {
      const int j = __builtin_ctz(pending);
      const V key = j == 0 ? v[0] : (j == 1 ? v[1] : (j == 2 ? v[2] : v[3]));
      const int slot = table.find(static_cast<i64>(key));
#pragma unroll
      for (int i = 0; i < 4; ++i) r[i] = j == i ? slot : r[i];
      pending &= pending - 1;
    }
    int slot = last_slot;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int here = v[j] == 0 ? 0 : (v[j] < 0 ? -1 : (((looked >> j) & 1u) ? r[j] : slot));
      out[j] = here;
      slot = v[j] > 0 ? here : slot;
    }
    last_key = prev;
    last_slot = slot;
  }
};

// ------------------------------------------------------- joint histogram

constexpr uint32_t kNoKey = 0xFFFFFFFFu;

__host__ __device__ inline size_t stage_bytes(int G, int P);

// Two-entry per-lane key cache in front of the LDS (or global) histogram.
template <bool GLOBAL>
struct KeyCache {
  uint32_t k0 = kNoKey, c0 = 0, k1 = kNoKey, c1 = 0;
  __device__ __forceinline__ static void flush(uint32_t* bins, u64* pairs, uint32_t key, uint32_t n) {
    if (GLOBAL) atomicAdd(pairs + key, static_cast<u64>(n));
    else atomicAdd(bins + (key >> 1), n << ((key & 1u) * 16u));
  }
  // One pixel's pair (ignored unless `valid`): a hit counts in registers; a
  // miss flushes the older entry and takes its place.  Selects only, apart
  // from the flush's atomic.
  __device__ __forceinline__ void add(uint32_t key, bool valid, uint32_t* bins, u64* pairs) {
    const bool hit0 = key == k0, hit1 = key == k1 && !hit0;
    const bool miss = valid && !hit0 && !hit1;
    const uint32_t out_key = k1, out_count = miss ? c1 : 0u;
    const uint32_t n0 = miss ? 1u : c0 + static_cast<uint32_t>(valid && hit0);
    const uint32_t n1 = miss ? c0 : c1 + static_cast<uint32_t>(valid && hit1);
    k1 = miss ? k0 : k1;
    k0 = miss ? key : k0;
    c0 = n0;
    c1 = n1;
    if (out_count) flush(bins, pairs, out_key, out_count);
  }
  __device__ __forceinline__ void drain(uint32_t* bins, u64* pairs) {
    if (c0) flush(bins, pairs, k0, c0);
    if (c1) flush(bins, pairs, k1, c1);
  }
};

// Whether both maps' pixels are 32-bit (the bucket tables apply).
template <typename GT, typename PR>
constexpr bool kKey32 = sizeof(typename MapValue<GT>::type) == 4 && sizeof(typename MapValue<PR>::type) == 4;

// The image's ID tables: hashes in LDS (STAGE; buckets for 32-bit pixels) or
// sorted tables in global memory.
template <bool STAGE, bool KEY32>
struct ImageTables {
  using Table = typename std::conditional<STAGE, typename std::conditional<KEY32, LdsBucketIds, LdsHashIds>::type,
                                          SortedIds>::type;
  Table gt, pred;
};

template <typename GT, typename PR, bool LOOKUP, bool VEC, bool GLOBAL, bool STAGE>
__device__ __forceinline__ void histogram_chunk(const HistArgs& a, int b, i64 begin, i64 end, uint32_t* bins,
                                                const ImageTables<STAGE, kKey32<GT, PR>>& t) {
  const GT* gt = static_cast<const GT*>(a.gt);
  const PR* pred = static_cast<const PR*>(a.pred);
  const int tid = threadIdx.x, nt = blockDim.x;  // 512, or 256 beside another stream's kernel
  u64* pairs = a.pairs + static_cast<i64>(b) * a.G * a.P;
  using Table = typename ImageTables<STAGE, kKey32<GT, PR>>::Table;
  SlotResolver<GT, LOOKUP, Table> gslot{t.gt, a.G, LOOKUP && a.gt_ids != nullptr};
  SlotResolver<PR, LOOKUP, Table> pslot{t.pred, a.P, LOOKUP && a.pred_ids != nullptr};
  KeyCache<GLOBAL> cache;
  bool bad = false;
  const uint32_t P = static_cast<uint32_t>(a.P);
  using GV = typename MapValue<GT>::type;
  using PV = typename MapValue<PR>::type;
  auto count = [&](int g, int p) {
    const bool valid = (g | p) >= 0;
    bad |= !valid;
    // Slots are below 2^24 wherever the bins are in LDS (G * P < 2^17).
    const uint32_t key = GLOBAL ? static_cast<uint32_t>(g) * P + static_cast<uint32_t>(p)
                                : __umul24(static_cast<uint32_t>(g), P) + static_cast<uint32_t>(p);
    cache.add(key, valid, bins, pairs);
  };
  auto visit = [&](GV gv, PV pv) {
    const int g = gslot(gv);
    count(g, pslot(pv));
  };
  if (VEC) {
    // Quads of four pixels (load_quad); scalar head/tail of <= 3 pixels.
    const i64 va = (begin + 3) & ~static_cast<i64>(3);
    const i64 vb = end & ~static_cast<i64>(3);
    const i64 head_end = va < end ? va : end;
    for (i64 i = begin + tid; i < head_end; i += nt) visit(read_pixel(gt, i), read_pixel(pred, i));
    if (va < vb) {
      const i64 q_end = vb >> 2;
      constexpr int U = 4;
      for (i64 base = (va >> 2) + tid; base < q_end; base += static_cast<i64>(nt) * U) {
        GV gv[U][4];
        PV pv[U][4];
        // All U quads are loaded before any is used (indices past the end
        // re-read the last quad and are not counted): a load under its own
        // bounds test would be waited for inside that test whenever its
        // pixels need unpacking (RGB), leaving one quad in flight per lane.
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const i64 q = min(base + static_cast<i64>(u) * nt, q_end - 1);
          load_quad(gt, q, gv[u]);
          load_quad(pred, q, pv[u]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const i64 q = base + static_cast<i64>(u) * nt;
          if (q < q_end) {
            int g[4], p[4];
            gslot.quad(gv[u], g);
            pslot.quad(pv[u], p);
#pragma unroll
            for (int j = 0; j < 4; ++j) count(g[j], p[j]);
          }
        }
      }
      for (i64 i = (vb > head_end ? vb : head_end) + tid; i < end; i += nt)
        visit(read_pixel(gt, i), read_pixel(pred, i));
    } else {
      for (i64 i = head_end + tid; i < end; i += nt) visit(read_pixel(gt, i), read_pixel(pred, i));
    }
  } else {
    for (i64 i = begin + tid; i < end; i += nt) visit(read_pixel(gt, i), read_pixel(pred, i));
  }
  cache.drain(bins, pairs);
  if (bad && a.error) atomicOr(a.error, 1);
}

// One chunk of one image: zero the LDS bins, count, flush to the image's
// pair table.  The pair table must be zero on entry (the scorer re-zeroes what
// it read).  Completion is published by the kernel boundary: a per-workgroup
// device-scope fence would write back the XCD's L2 for every chunk.
template <typename GT, typename PR, bool LOOKUP, bool VEC, bool GLOBAL, bool STAGE>
__device__ __forceinline__ void histogram_item(const HistArgs& a, int item, uint4* lds4) {
  uint32_t* bins = reinterpret_cast<uint32_t*>(lds4);
  const int tid = threadIdx.x, nt = blockDim.x;
  const int b = __builtin_amdgcn_readfirstlane(a.tiles.image_of(item));
  const int chunk = item - a.tiles.first_tile(b);
  const i64 image = a.tiles.begin(b);
  const i64 pixels = a.tiles.pixels(b);
  const i64 begin = image + static_cast<i64>(chunk) * a.tiles.tile;
  const i64 end = image + min(static_cast<i64>(chunk + 1) * a.tiles.tile, pixels);
  if (!GLOBAL) {
    for (int w = tid; w < a.words4; w += nt) lds4[w] = make_uint4(0, 0, 0, 0);
  }
  ImageTables<STAGE, kKey32<GT, PR>> t{};
  const i64 gt_row = static_cast<i64>(b) * a.G, pred_row = static_cast<i64>(b) * a.P;
  if constexpr (STAGE && kKey32<GT, PR>) {
    // Bucket tables after the bins: 2^bits buckets per map, each four keys
    // then four words (lowest sorted index while building, then the slot).
    const int Tg = 1 << a.gt_hash_bits, Tp = 1 << a.pred_hash_bits;
    auto* table = reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(lds4) + a.table_offset);
    for (int i = tid; i < (Tg + Tp) * 4; i += nt) {
      const int w = (i >> 2) * 8 + (i & 3);
      table[w] = kEmptyKey32;
      table[w + 4] = INT_MAX;
    }
    __syncthreads();
    for (int i = tid; i < a.G + a.P; i += nt) {
      const bool g = i < a.G;
      const i64* src = g ? a.gt_ids : a.pred_ids;
      const int index = g ? i : i - a.G;
      if (!src || index == 0) continue;  // index 0 is the void entry (ID 0)
      const i64 key = src[g ? gt_row + index : pred_row + index];
      if (key <= 0 || key > INT_MAX) continue;  // padding, or no 32-bit pixel holds it
      uint32_t* own = g ? table : table + 8 * Tg;
      const uint32_t mask = static_cast<uint32_t>((g ? Tg : Tp) - 1);
      const uint32_t k32 = static_cast<uint32_t>(key);
      bool placed = false;
      for (uint32_t h = hash_id(key, 32 - (g ? a.gt_hash_bits : a.pred_hash_bits)); !placed; h = (h + 1) & mask) {
        for (int e = 0; e < 4 && !placed; ++e) {
          const uint32_t prev = atomicCAS(own + 8 * h + e, kEmptyKey32, k32);
          if (prev == kEmptyKey32 || prev == k32) {
            atomicMin(reinterpret_cast<int*>(own) + 8 * h + 4 + e, index);
            placed = true;
          }
        }
      }
    }
    __syncthreads();
    for (int i = tid; i < (Tg + Tp) * 4; i += nt) {
      const int w = (i >> 2) * 8 + (i & 3);
      if (table[w] == kEmptyKey32) continue;
      const bool g = (i >> 2) < Tg;
      const int S = g ? a.G : a.P;
      const int lowest = static_cast<int>(table[w + 4]);
      const int s = g ? a.gt_slots[gt_row + lowest] : a.pred_slots[pred_row + lowest];
      table[w + 4] = static_cast<uint32_t>((s > 0 && s < S) ? s : -1);
    }
    t.gt = {(LdsBuckets)(table), static_cast<uint32_t>(Tg - 1), 32 - a.gt_hash_bits};
    t.pred = {(LdsBuckets)(table + 8 * Tg), static_cast<uint32_t>(Tp - 1), 32 - a.pred_hash_bits};
  } else if constexpr (STAGE) {
    // Hash the image's ID tables into LDS after the bins: Tg + Tp int64 keys,
    // then Tg + Tp int32 words (lowest sorted index while building, then the
    // resolved slot).
    const int Tg = 1 << a.gt_hash_bits, Tp = 1 << a.pred_hash_bits;
    auto* keys = reinterpret_cast<i64*>(reinterpret_cast<char*>(lds4) + a.table_offset);
    auto* word = reinterpret_cast<int*>(keys + Tg + Tp);
    for (int i = tid; i < Tg + Tp; i += nt) {
      keys[i] = kEmptyId;
      word[i] = INT_MAX;
    }
    __syncthreads();
    for (int i = tid; i < a.G + a.P; i += nt) {
      const bool g = i < a.G;
      const i64* src = g ? a.gt_ids : a.pred_ids;
      const int index = g ? i : i - a.G;
      if (!src || index == 0) continue;  // index 0 is the void entry (ID 0)
      const i64 key = src[g ? gt_row + index : pred_row + index];
      if (key <= 0 || key == LLONG_MAX) continue;  // padding never resolves
      i64* table = g ? keys : keys + Tg;
      int* lowest = g ? word : word + Tg;
      const uint32_t mask = static_cast<uint32_t>((g ? Tg : Tp) - 1);
      for (uint32_t h = hash_id(key, 32 - (g ? a.gt_hash_bits : a.pred_hash_bits));; h = (h + 1) & mask) {
        const u64 prev = atomicCAS(reinterpret_cast<unsigned long long*>(table + h),
                                   static_cast<unsigned long long>(kEmptyId), static_cast<unsigned long long>(key));
        if (static_cast<i64>(prev) == kEmptyId || static_cast<i64>(prev) == key) {
          atomicMin(lowest + h, index);
          break;
        }
      }
    }
    __syncthreads();
    for (int i = tid; i < Tg + Tp; i += nt) {
      if (keys[i] == kEmptyId) continue;
      const bool g = i < Tg;
      const int S = g ? a.G : a.P;
      const int s = g ? a.gt_slots[gt_row + word[i]] : a.pred_slots[pred_row + word[i]];
      word[i] = (s > 0 && s < S) ? s : -1;
    }
    t.gt = LdsHashIds{(LdsIds)(keys), (LdsInts)(word), static_cast<uint32_t>(Tg - 1), 32 - a.gt_hash_bits};
    t.pred = LdsHashIds{(LdsIds)(keys + Tg), (LdsInts)(word + Tg), static_cast<uint32_t>(Tp - 1),
                        32 - a.pred_hash_bits};
  } else {
    if (a.gt_ids) t.gt = SortedIds{a.gt_ids + gt_row, a.gt_slots + gt_row, a.G};
    if (a.pred_ids) t.pred = SortedIds{a.pred_ids + pred_row, a.pred_slots + pred_row, a.P};
  }
  if (!GLOBAL || STAGE) __syncthreads();
  histogram_chunk<GT, PR, LOOKUP, VEC, GLOBAL, STAGE>(a, b, begin, end, bins, t);
  if (!GLOBAL) {
    __syncthreads();
    u64* pairs = a.pairs + static_cast<i64>(b) * a.G * a.P;
    for (int w = tid; w < a.words4; w += nt) {
      uint4 v = lds4[w];
      uint32_t words[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (!words[j]) continue;
        uint32_t key = (static_cast<uint32_t>(w) * 4u + j) * 2u;
        uint32_t lo = words[j] & 0xFFFFu, hi = words[j] >> 16;
        if (lo) atomicAdd(pairs + key, static_cast<u64>(lo));
        if (hi) atomicAdd(pairs + key + 1, static_cast<u64>(hi));
      }
    }
  }
}

// One workgroup per chunk (surplus workgroups of a packed batch, whose chunk
// count is only known on the device, exit at once), or a persistent grid
// that walks the chunk list with a grid stride.
template <typename GT, typename PR, bool LOOKUP, bool VEC, bool GLOBAL, bool PERSIST, bool STAGE>
__global__ __launch_bounds__(kHistThreads) void joint_histogram(HistArgs a) {
  extern __shared__ uint4 lds4[];
  const int total = a.tiles.count(a.B);
  if (!PERSIST) {
    if (static_cast<int>(blockIdx.x) < total)
      histogram_item<GT, PR, LOOKUP, VEC, GLOBAL, STAGE>(a, blockIdx.x, lds4);
    return;
  }
  for (int item = blockIdx.x; item < total; item += gridDim.x) {
    histogram_item<GT, PR, LOOKUP, VEC, GLOBAL, STAGE>(a, item, lds4);
    __syncthreads();  // the next chunk re-zeroes the bins and re-stages the tables
  }
}

template <typename K>
static bool allow_lds(K kernel, size_t bytes) {
  if (bytes <= 64 * 1024) return true;
  return hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                             static_cast<int>(bytes)) == hipSuccess;
}

template <typename GT, typename PR, bool LOOKUP, bool VEC, bool GLOBAL>
static int launch_hist_kernel(const HistArgs& a, int blocks, size_t lds, hipStream_t stream) {
  // Original-ID batches are uniform (never packed), so they never take the
  // persistent grid; their ID tables go to LDS when they fit.
  void (*kernel)(HistArgs);
  if constexpr (LOOKUP)
    kernel = a.stage_tables ? joint_histogram<GT, PR, LOOKUP, VEC, GLOBAL, false, true>
                            : joint_histogram<GT, PR, LOOKUP, VEC, GLOBAL, false, false>;
  else
    kernel = a.persist ? joint_histogram<GT, PR, LOOKUP, VEC, GLOBAL, true, false>
                       : joint_histogram<GT, PR, LOOKUP, VEC, GLOBAL, false, false>;
  if (!allow_lds(kernel, lds)) return fail(kUnsupportedShape, "LDS request of %zu bytes refused", lds);
  ProfileScope timing("joint_histogram", stream);
  // Beside another stream's kernel, a narrower workgroup fits where one of
  // that kernel's workgroups retired.
  const int threads = stream_concurrent(stream) ? kConcurrentThreads : kHistThreads;
  kernel<<<blocks, threads, lds, stream>>>(a);
  PE_LAUNCH_CHECK();
  return kOk;
}

template <typename GT, typename PR, bool LOOKUP>
static int launch_hist_typed(const HistArgs& a, int blocks, size_t lds, bool vec, bool global,
                             hipStream_t stream) {
  if (global) return launch_hist_kernel<GT, PR, LOOKUP, false, true>(a, blocks, lds, stream);
  if (vec) return launch_hist_kernel<GT, PR, LOOKUP, true, false>(a, blocks, lds, stream);
  return launch_hist_kernel<GT, PR, LOOKUP, false, false>(a, blocks, lds, stream);
}

int hist_default_chunk(i64 total_pixels) {
  // Aim for >= 8 workgroups per compute unit; chunks are multiples of 4096
  // pixels and never exceed kMaxChunk (packed 16-bit LDS counters).
  i64 target = total_pixels / (8LL * compute_units());
  i64 chunk = (target + 4095) / 4096 * 4096;
  if (chunk < 4096) chunk = 4096;
  if (chunk > kMaxChunk) chunk = kMaxChunk;
  return static_cast<int>(chunk);
}

// Packed batches need tile_ints() ints of workspace for the chunk list.
int launch_joint_histogram(HistArgs a, int gt_bytes, int pred_bytes, int chunk, int* tile_ws, hipStream_t stream) {
  if (a.B == 0) return kOk;
  const i64 total_pixels = a.pix_off ? a.HW : a.HW * a.B;  // a packed batch passes its pixel total
  if (total_pixels == 0) return kOk;
  if (chunk <= 0) {
    // Tuned pixel chunk per workgroup (variant 0 sizes it from the batch).
    static const int kChunks[] = {0, 8192, 16384, 32768, kMaxChunk};
    chunk = kChunks[tuned_variant(kTuneJointHistogram)];
    if (chunk <= 0) chunk = hist_default_chunk(total_pixels);
  }
  chunk = std::min<int>(std::max<int>(chunk & ~3, 4096), kMaxChunk);
  TileList& t = a.tiles;
  t = TileList{};
  t.tile = chunk;
  i64 blocks;
  if (a.pix_off) {
    if (!tile_ws) return fail(kBadArgument, "a packed batch needs the chunk-list workspace");
    const i64 bound = tile_list_bound(a.B, total_pixels, 4096);
    if (bound >= INT_MAX) return fail(kUnsupportedShape, "too many histogram chunks");
    t.pix_off = a.pix_off;
    t.tile_off = tile_ws;
    t.tile_image = tile_ws + a.B + 1;
    int rc = launch_tile_list(a.pix_off, a.B, chunk, tile_ws, tile_ws + a.B + 1, stream);
    if (rc) return rc;
    // The exact chunk count lives on the device: launch the bound, or a
    // persistent grid of a few workgroups per compute unit.
    blocks = tile_list_bound(a.B, total_pixels, chunk);
    const int per_cu = packed_grid_per_cu();
    if (per_cu > 0) {
      blocks = std::min<i64>(blocks, static_cast<i64>(per_cu) * compute_units());
      a.persist = true;
    }
  } else {
    t.HW = a.HW;
    t.tiles_per_image = static_cast<int>((a.HW + chunk - 1) / chunk);
    blocks = static_cast<i64>(a.B) * t.tiles_per_image;
  }
  if (blocks > INT_MAX) return fail(kUnsupportedShape, "too many histogram workgroups");
  i64 bins = static_cast<i64>(a.G) * a.P;
  i64 words = (bins + 1) / 2;
  a.words4 = static_cast<int>((words + 3) / 4);
  size_t lds = static_cast<size_t>(a.words4) * 16;
  bool global = lds > kMaxHistLds;
  if (global) lds = 0;
  bool lookup = a.gt_ids || a.pred_ids;
  if (lookup && a.persist) return fail(kBadArgument, "original-ID maps cannot be packed");
  // ID hashes (12 bytes per entry, at least two entries per ID) after the
  // bins, if the total fits.
  // Maps of 32-bit pixels use buckets of four 32-bit keys (32 bytes per
  // bucket, at least two entries per ID); int64 maps use 12-byte entries.
  const bool key32 = gt_bytes <= 4 && pred_bytes <= 4;
  auto hash_bits = [key32](int n) {
    int bits = key32 ? 2 : 4;
    const i64 per = key32 ? 4 : 1;
    while (bits < 30 && per * (static_cast<i64>(1) << bits) < 2 * static_cast<i64>(n)) ++bits;
    return bits;
  };
  a.gt_hash_bits = hash_bits(a.G);
  a.pred_hash_bits = hash_bits(a.P);
  const size_t entry_bytes = key32 ? 32 : 12;
  const size_t table_bytes = align_up(((static_cast<size_t>(1) << a.gt_hash_bits) +
                                       (static_cast<size_t>(1) << a.pred_hash_bits)) * entry_bytes, 16);
  a.stage_tables = lookup && lds + table_bytes <= kMaxHistLds + 16 * 1024;
  if (lookup && !a.stage_tables && !global && table_bytes <= kMaxHistLds) {
    // The bins and the tables do not fit together: bins go to global memory.
    global = true;
    lds = 0;
    a.stage_tables = true;
  }
  a.table_offset = static_cast<int>(lds);
  if (a.stage_tables) lds += table_bytes;
  a.lds_bytes = lds;
  bool vec = quad_aligned(a.gt, gt_bytes) && quad_aligned(a.pred, pred_bytes);
  int nb = static_cast<int>(blocks);
#define PE_HIST(GT, PR)                                                                   \
  do {                                                                                    \
    return lookup ? launch_hist_typed<GT, PR, true>(a, nb, lds, vec, global, stream)     \
                  : launch_hist_typed<GT, PR, false>(a, nb, lds, vec, global, stream);   \
  } while (0)
  // RGB maps always carry original IDs (the caller checked the lookup table).
  if (gt_bytes == 3 && pred_bytes == 3) return launch_hist_typed<Rgb24, Rgb24, true>(a, nb, lds, vec, global, stream);
  if (gt_bytes == 3 && pred_bytes == 4) return launch_hist_typed<Rgb24, int32_t, true>(a, nb, lds, vec, global, stream);
  if (gt_bytes == 4 && pred_bytes == 3) return launch_hist_typed<int32_t, Rgb24, true>(a, nb, lds, vec, global, stream);
  if (gt_bytes == 4 && pred_bytes == 4) PE_HIST(int32_t, int32_t);
  if (gt_bytes == 4 && pred_bytes == 8) PE_HIST(int32_t, int64_t);
  if (gt_bytes == 8 && pred_bytes == 4) PE_HIST(int64_t, int32_t);
  PE_HIST(int64_t, int64_t);
#undef PE_HIST
}

// ------------------------------------------------------ areas / validation

// Per-image metadata staged in LDS for the scoring loops (falls back to the
// global arrays when G and P are too large for the workgroup's LDS).
struct Stage {
  i64* ga;
  i64* pa;
  double* val;
  int* gc;
  int* matched;
  int* lab;
  uint8_t* crowd;
  uint32_t* cnt;      // G x P pair counts in LDS (32-bit: per-image counts < 2^32)
  bool staged;
  bool staged_pairs;
};

__host__ __device__ inline size_t stage_bytes(int G, int P) {
  // ga, pa, val (8 bytes each), gc, matched, lab (4 bytes), crowd (1 byte)
  return 8 * static_cast<size_t>(G) + 16 * static_cast<size_t>(P) + 8 * static_cast<size_t>(G) +
         4 * static_cast<size_t>(P) + static_cast<size_t>(G) + 64;
}

// The staged and global layouts are built in separate functions and used in
// separate inlined call paths, so every pointer of one path has a single
// address space: the compiler then emits LDS (ds_*) instructions for the
// staged path instead of generic flat accesses, which would also make each
// wait cover the in-flight global atomics.
__device__ __forceinline__ Stage lds_stage(int G, int P, char* lds, size_t lds_bytes) {
  Stage st;
  char* p = lds;
  st.ga = reinterpret_cast<i64*>(p);
  p += 8 * static_cast<size_t>(G);
  st.pa = reinterpret_cast<i64*>(p);
  p += 8 * static_cast<size_t>(P);
  st.val = reinterpret_cast<double*>(p);
  p += 8 * static_cast<size_t>(P);
  st.gc = reinterpret_cast<int*>(p);
  p += 4 * static_cast<size_t>(G);
  st.matched = reinterpret_cast<int*>(p);
  p += 4 * static_cast<size_t>(G);
  st.lab = reinterpret_cast<int*>(p);
  p += 4 * static_cast<size_t>(P);
  st.crowd = reinterpret_cast<uint8_t*>(p);
  p += static_cast<size_t>(G);
  const size_t used = (static_cast<size_t>(p - lds) + 15) & ~static_cast<size_t>(15);
  st.staged = true;
  st.staged_pairs = used + 4 * static_cast<size_t>(G) * P <= lds_bytes;
  st.cnt = reinterpret_cast<uint32_t*>(lds + used);  // valid only when staged_pairs
  return st;
}

__device__ __forceinline__ Stage global_stage(const ScoreArgs& a, int b) {
  const int G = a.G, P = a.P;
  Stage st;
  st.staged = false;
  st.staged_pairs = false;
  st.cnt = nullptr;
  i64* areas = a.areas + static_cast<i64>(b) * (G + P);
  st.ga = areas;
  st.pa = areas + G;
  st.val = a.rec_val + static_cast<i64>(b) * P;
  st.gc = const_cast<int*>(a.gc + static_cast<i64>(b) * G);
  st.matched = a.matched + static_cast<i64>(b) * G;
  st.lab = a.rec_lab + static_cast<i64>(b) * P;
  st.crowd = const_cast<uint8_t*>(a.crowd + static_cast<i64>(b) * G);
  return st;
}

// Row sums (GT areas) with one wave per row; column sums (prediction areas)
// with one lane per column and eight loads in flight.
template <typename T>
__device__ __forceinline__ void image_areas(const T* pairs, int G, int P, i64* areas, i64* ga, i64* pa) {
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const int waves = blockDim.x / kWave;
  for (int g = wave; g < G; g += waves) {
    i64 s = 0;
    for (int p = lane; p < P; p += kWave) s += static_cast<i64>(pairs[static_cast<i64>(g) * P + p]);
    s = wave_sum(s);
    if (lane == 0) {
      areas[g] = s;
      ga[g] = s;
    }
  }
  for (int p = tid; p < P; p += blockDim.x) {
    i64 s = 0;
    int g = 0;
    for (; g + 8 <= G; g += 8) {
      T v[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) v[j] = pairs[static_cast<i64>(g + j) * P + p];
#pragma unroll
      for (int j = 0; j < 8; ++j) s += static_cast<i64>(v[j]);
    }
    for (; g < G; ++g) s += static_cast<i64>(pairs[static_cast<i64>(g) * P + p]);
    areas[G + p] = s;
    pa[p] = s;
  }
}

// Error bits 2 and 4 of the reference's validation (evaluator.py:208-224).
__device__ void image_checks(const i64* areas, const int* gc, const int* pc, const uint8_t* crowd,
                             int G, int P, int C, int reward, int* error) {
  int bits = 0;
  for (int g = threadIdx.x; g < G; g += blockDim.x) {
    int c = gc[g];
    if (c < -1 || c >= C || (g == 0 && (c != -1 || crowd[0]))) bits |= 2;
    if (g > 0 && ((areas[g] > 0) != (c >= 0))) bits |= 4;
  }
  for (int p = threadIdx.x; p < P; p += blockDim.x) {
    int c = pc[p];
    i64 pa = areas[G + p];
    if (c < -1 || c >= C || (p == 0 && c != -1)) bits |= 2;
    if (p > 0 && (reward ? (pa > 0 && c < 0) : ((pa > 0) != (c >= 0)))) bits |= 4;
  }
  if (bits) atomicOr(error, bits);
}

__global__ __launch_bounds__(kScoreThreads) void prepare_images(ScoreArgs a, int* error) {
  const int b = blockIdx.x;
  const int G = a.G, P = a.P;
  const i64 GP = static_cast<i64>(G) * P;
  i64* areas = a.areas + static_cast<i64>(b) * (G + P);
  // Areas go to the global scratch only (the staging copies are discarded).
  const u64* pairs = a.pairs + b * GP;
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  for (int g = wave; g < G; g += blockDim.x / kWave) {
    i64 s = 0;
    for (int p = lane; p < P; p += kWave) s += static_cast<i64>(pairs[static_cast<i64>(g) * P + p]);
    s = wave_sum(s);
    if (lane == 0) areas[g] = s;
  }
  for (int p = tid; p < P; p += blockDim.x) {
    i64 s = 0;
    int g = 0;
    for (; g + 8 <= G; g += 8) {
      u64 v[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) v[j] = pairs[static_cast<i64>(g + j) * P + p];
#pragma unroll
      for (int j = 0; j < 8; ++j) s += static_cast<i64>(v[j]);
    }
    for (; g < G; ++g) s += static_cast<i64>(pairs[static_cast<i64>(g) * P + p]);
    areas[G + p] = s;
  }
  __syncthreads();
  image_checks(areas, a.gc + static_cast<i64>(b) * G, a.pc + static_cast<i64>(b) * P,
               a.crowd + static_cast<i64>(b) * G, G, P, a.C, a.reward, error);
}

// ------------------------------------------------------------ PQ scoring

// Score one image with the whole workgroup: areas, PQ matching, confusion,
// TP/FP/FN and the per-image matched-IoU records.
template <typename T>
__device__ __forceinline__ void score_core(const ScoreArgs& a, int b, const Stage& st, const T* pairs) {
  const int G = a.G, P = a.P, C = a.C;
  const int tid = threadIdx.x, nthreads = blockDim.x;
  const int* pc = a.pc + static_cast<i64>(b) * P;
  i64* areas = a.areas + static_cast<i64>(b) * (G + P);
  const i64 row = a.reward ? static_cast<i64>(b) * C : 0;
  u64* tp = a.tp + row;
  u64* fp = a.fp + row;
  u64* fn = a.fn + row;
  if (a.areas_ready) {
    for (int g = tid; g < G; g += nthreads) st.ga[g] = areas[g];
    for (int p = tid; p < P; p += nthreads) st.pa[p] = areas[G + p];
  } else {
    image_areas(pairs, G, P, areas, st.ga, st.pa);
  }
  for (int g = tid; g < G; g += nthreads) st.matched[g] = 0;
  __syncthreads();
  for (int p = tid; p < P; p += nthreads) {
    const int pcls = pc[p];
    const bool valid_pred = pcls >= 0 && pcls < C;
    const i64 pa = st.pa[p];
    const i64 void_overlap = static_cast<i64>(pairs[p]);
    i64 ignored = void_overlap;
    bool pm = false;
    double sum_iou = 0.0;
    for (int g0 = 0; g0 < G; g0 += 8) {
      T v[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) v[j] = g0 + j < G ? pairs[static_cast<i64>(g0 + j) * P + p] : 0;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int g = g0 + j;
        if (g >= G) break;
        const i64 inter = static_cast<i64>(v[j]);
        const int tcls = st.gc[g];
        const bool valid_target = tcls >= 0 && tcls < C;
        if (!valid_target) continue;
        const bool is_crowd = st.crowd[g] != 0;
        if (!a.reward && inter && !is_crowd)
          atomicAdd(a.confusion + static_cast<i64>(tcls) * (C + 1) + (valid_pred ? pcls : C),
                    static_cast<u64>(inter));
        if (!valid_pred || tcls != pcls) continue;
        if (is_crowd) {
          ignored = void_overlap + inter;  // the last same-class crowd slot wins
        } else if (inter) {
          const i64 uni = st.ga[g] + pa - inter - void_overlap;
          if (inter > uni / 2) {  // IoU > 0.5 without rounding
            pm = true;
            sum_iou += static_cast<double>(inter) / static_cast<double>(uni);
            st.matched[g] = 1;
          }
        }
      }
    }
    int label = -1;
    if (valid_pred && (!a.reward || pa > 0)) {
      if (pm) {
        atomicAdd(tp + pcls, 1ULL);
        label = pcls;
      } else if (static_cast<double>(ignored) <= 0.5 * static_cast<double>(pa)) {
        atomicAdd(fp + pcls, 1ULL);
      }
    }
    st.lab[p] = label;
    st.val[p] = sum_iou;
  }
  __syncthreads();
  for (int g = tid; g < G; g += nthreads) {
    const int c = st.gc[g];
    if (c >= 0 && c < C && !st.crowd[g] && !st.matched[g]) atomicAdd(fn + c, 1ULL);
  }
  if (a.reward) {
    // Per-image sums: in-order fold over prediction slots.
    double* sums = a.iou_sum + row;
    for (int c = tid; c < C; c += nthreads) {
      double acc = sums[c];
      for (int p = 0; p < P; ++p)
        if (st.lab[p] == c) acc += st.val[p];
      sums[c] = acc;
    }
    return;
  }
  // Dataset mode: every matched IoU lies in (0.5, 1], an exact multiple of
  // 2^-53, so the class sums accumulate exactly as 2^-53 fixed point in two
  // 64-bit words (low 32 bits / high bits of each value): order-independent.
  for (int p = tid; p < P; p += nthreads) {
    const int c = st.lab[p];
    if (c < 0) continue;
    const u64 m = static_cast<u64>(st.val[p] * 9007199254740992.0);  // * 2^53, exact
    atomicAdd(a.iou_exact + 2 * c, m & 0xFFFFFFFFULL);
    atomicAdd(a.iou_exact + 2 * c + 1, m >> 32);
  }
}

// ---- parallel scoring (pair table, metadata and a per-class crowd table in LDS)
//
// Every (g, p) pair is independent once the slot areas are known: IoU > 0.5
// matches are unique per GT and per prediction slot (segments are disjoint),
// so each prediction slot has at most one matched IoU, and the crowd rule
// needs only the last same-class crowd slot, a per-class maximum.  The work
// is a sweep over the G x P table by the whole workgroup instead of one
// G-long loop per prediction slot.

struct Par {
  i64* ga;
  i64* pa;
  double* val;
  int* gc;
  int* pc;
  int* matched;
  int* hit;
  int* last;      // C: last crowd GT slot of each class, -1 if none
  uint8_t* crowd;
  uint32_t* cnt;  // G x P
};

__host__ __device__ inline size_t par_bytes(int G, int P, int C) {
  const size_t g = static_cast<size_t>(G), p = static_cast<size_t>(P), c = static_cast<size_t>(C);
  const size_t meta = 8 * g + 16 * p + 4 * g + 4 * p + 4 * g + 4 * p + 4 * c + g;
  return ((meta + 15) & ~static_cast<size_t>(15)) + 4 * g * p;
}

__device__ __forceinline__ Par par_layout(int G, int P, int C, char* lds) {
  Par s;
  char* q = lds;
  s.ga = reinterpret_cast<i64*>(q);
  q += 8 * static_cast<size_t>(G);
  s.pa = reinterpret_cast<i64*>(q);
  q += 8 * static_cast<size_t>(P);
  s.val = reinterpret_cast<double*>(q);
  q += 8 * static_cast<size_t>(P);
  s.gc = reinterpret_cast<int*>(q);
  q += 4 * static_cast<size_t>(G);
  s.pc = reinterpret_cast<int*>(q);
  q += 4 * static_cast<size_t>(P);
  s.matched = reinterpret_cast<int*>(q);
  q += 4 * static_cast<size_t>(G);
  s.hit = reinterpret_cast<int*>(q);
  q += 4 * static_cast<size_t>(P);
  s.last = reinterpret_cast<int*>(q);
  q += 4 * static_cast<size_t>(C);
  s.crowd = reinterpret_cast<uint8_t*>(q);
  q += static_cast<size_t>(G);
  const size_t used = (static_cast<size_t>(q - lds) + 15) & ~static_cast<size_t>(15);
  s.cnt = reinterpret_cast<uint32_t*>(lds + used);
  return s;
}

__device__ void score_parallel(const ScoreArgs& a, int b, char* lds) {
  const int G = a.G, P = a.P, C = a.C;
  const int tid = threadIdx.x, nthreads = blockDim.x;
  const int GP = G * P;  // fits the LDS, so far below 2^31
  const Par s = par_layout(G, P, C, lds);
  const int* gc = a.gc + static_cast<i64>(b) * G;
  const int* pc = a.pc + static_cast<i64>(b) * P;
  const uint8_t* crowd = a.crowd + static_cast<i64>(b) * G;
  i64* areas = a.areas + static_cast<i64>(b) * (G + P);
  u64* table = const_cast<u64*>(a.pairs) + static_cast<i64>(b) * GP;
  const i64 row = a.reward ? static_cast<i64>(b) * C : 0;
  const bool sum_areas = !a.areas_ready;
  for (int g = tid; g < G; g += nthreads) {
    s.gc[g] = gc[g];
    s.crowd[g] = crowd[g];
    s.matched[g] = 0;
    s.ga[g] = sum_areas ? 0 : areas[g];
  }
  for (int p = tid; p < P; p += nthreads) {
    s.pc[p] = pc[p];
    s.hit[p] = 0;
    s.val[p] = 0.0;
    s.pa[p] = sum_areas ? 0 : areas[G + p];
  }
  for (int c = tid; c < C; c += nthreads) s.last[c] = -1;
  __syncthreads();
  for (int g = tid; g < G; g += nthreads) {
    const int c = s.gc[g];
    if (s.crowd[g] && c >= 0 && c < C) atomicMax(s.last + c, g);
  }
  // Pair table -> LDS (sixteen loads in flight per lane), zeroing the global
  // table for the next update; slot areas as row/column sums of the nonzeros.
  constexpr int U = 16;
  for (int e0 = tid; e0 < GP; e0 += nthreads * U) {
    u64 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int e = e0 + u * nthreads;
      v[u] = e < GP ? table[e] : 0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int e = e0 + u * nthreads;
      if (e >= GP) break;
      s.cnt[e] = static_cast<uint32_t>(v[u]);
      if (!v[u]) continue;
      table[e] = 0;
      if (sum_areas) {
        const int g = e / P, p = e - g * P;
        atomicAdd(reinterpret_cast<u64*>(s.ga + g), v[u]);
        atomicAdd(reinterpret_cast<u64*>(s.pa + p), v[u]);
      }
    }
  }
  __syncthreads();
  if (sum_areas) {
    for (int g = tid; g < G; g += nthreads) areas[g] = s.ga[g];
    for (int p = tid; p < P; p += nthreads) areas[G + p] = s.pa[p];
  }
  // Pair sweep: confusion, and the unique IoU > 0.5 match of each slot.
  for (int e = tid; e < GP; e += nthreads) {
    const uint32_t inter = s.cnt[e];
    if (!inter) continue;
    const int g = e / P, p = e - g * P;
    const int tcls = s.gc[g];
    if (tcls < 0 || tcls >= C) continue;
    const int pcls = s.pc[p];
    const bool valid_pred = pcls >= 0 && pcls < C;
    const bool is_crowd = s.crowd[g] != 0;
    if (!a.reward && !is_crowd)
      atomicAdd(a.confusion + static_cast<i64>(tcls) * (C + 1) + (valid_pred ? pcls : C),
                static_cast<u64>(inter));
    if (!valid_pred || tcls != pcls || is_crowd) continue;
    const i64 void_overlap = s.cnt[p];
    const i64 uni = s.ga[g] + s.pa[p] - static_cast<i64>(inter) - void_overlap;
    if (static_cast<i64>(inter) > uni / 2) {  // IoU > 0.5 without rounding
      s.matched[g] = 1;
      s.hit[p] = 1;
      s.val[p] = static_cast<double>(inter) / static_cast<double>(uni);
    }
  }
  __syncthreads();
  u64* tp = a.tp + row;
  u64* fp = a.fp + row;
  u64* fn = a.fn + row;
  for (int p = tid; p < P; p += nthreads) {
    const int pcls = s.pc[p];
    int label = -1;
    if (pcls >= 0 && pcls < C) {
      const i64 pa = s.pa[p];
      const int lg = s.last[pcls];
      const i64 void_overlap = G > 0 ? static_cast<i64>(s.cnt[p]) : 0;
      const i64 ignored = void_overlap + (lg >= 0 ? static_cast<i64>(s.cnt[lg * P + p]) : 0);
      if (!a.reward || pa > 0) {
        if (s.hit[p]) {
          atomicAdd(tp + pcls, 1ULL);
          label = pcls;
        } else if (static_cast<double>(ignored) <= 0.5 * static_cast<double>(pa)) {
          atomicAdd(fp + pcls, 1ULL);
        }
      }
    }
    s.hit[p] = label;  // reused as the slot's label
  }
  for (int g = tid; g < G; g += nthreads) {
    const int c = s.gc[g];
    if (c >= 0 && c < C && !s.crowd[g] && !s.matched[g]) atomicAdd(fn + c, 1ULL);
  }
  __syncthreads();
  if (a.reward) {
    // Per-image sums: in-order fold over prediction slots.
    double* sums = a.iou_sum + row;
    for (int c = tid; c < C; c += nthreads) {
      double acc = sums[c];
      for (int p = 0; p < P; ++p)
        if (s.hit[p] == c) acc += s.val[p];
      sums[c] = acc;
    }
    return;
  }
  for (int p = tid; p < P; p += nthreads) {
    const int c = s.hit[p];
    if (c < 0) continue;
    const u64 m = static_cast<u64>(s.val[p] * 9007199254740992.0);  // * 2^53, exact
    atomicAdd(a.iou_exact + 2 * c, m & 0xFFFFFFFFULL);
    atomicAdd(a.iou_exact + 2 * c + 1, m >> 32);
  }
}

// Returns true when the image's global pair table was already zeroed.
__device__ bool score_image(const ScoreArgs& a, int b, char* lds, size_t lds_bytes) {
  const int G = a.G, P = a.P;
  const int tid = threadIdx.x, nthreads = blockDim.x;
  const i64 GP = static_cast<i64>(G) * P;
  const u64* pairs = a.pairs + b * GP;
  if (lds != nullptr && par_bytes(G, P, a.C) <= lds_bytes) {
    score_parallel(a, b, lds);
    return true;
  }
  if (lds == nullptr || stage_bytes(G, P) > lds_bytes) {
    score_core<u64>(a, b, global_stage(a, b), pairs);
    return false;
  }
  const Stage st = lds_stage(G, P, lds, lds_bytes);
  const int* gc = a.gc + static_cast<i64>(b) * G;
  const uint8_t* crowd = a.crowd + static_cast<i64>(b) * G;
  for (int g = tid; g < G; g += nthreads) {
    st.gc[g] = gc[g];
    st.crowd[g] = crowd[g];
  }
  if (!st.staged_pairs) {
    score_core<u64>(a, b, st, pairs);
    return false;
  }
  // One coalesced copy of the pair table into LDS (eight loads in flight per
  // lane); every loop below then reads LDS.  The global table is zeroed in
  // the same pass, ready for the next update.
  u64* table = const_cast<u64*>(pairs);
  constexpr int U = 8;
  for (i64 e0 = tid; e0 < GP; e0 += static_cast<i64>(nthreads) * U) {
    u64 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const i64 e = e0 + static_cast<i64>(u) * nthreads;
      v[u] = e < GP ? table[e] : 0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const i64 e = e0 + static_cast<i64>(u) * nthreads;
      if (e < GP) {
        st.cnt[e] = static_cast<uint32_t>(v[u]);
        if (v[u]) table[e] = 0;
      }
    }
  }
  __syncthreads();
  score_core<uint32_t>(a, b, st, st.cnt);
  return true;
}

// iou_sum[c] = the exact accumulated sum rounded once to FP64.
__device__ __forceinline__ double exact_to_double(u64 lo, u64 hi) {
  return (static_cast<double>(hi) * 4294967296.0 + static_cast<double>(lo)) * 0x1p-53;
}

// Caller-kept exact words (iou_add == 0): iou_sum is rewritten from the
// running exact total.  Per-update words in the workspace (iou_add != 0, the
// caller keeps no exact words): this update's exact sum is rounded once and
// added to the caller's iou_sum, one FP64 add per class and update, so a
// caller that zeroes iou_sum itself (reset) starts from zero.
__device__ void convert_iou_sum(const ScoreArgs& a) {
  for (int c = threadIdx.x; c < a.C; c += blockDim.x) {
    const double update = exact_to_double(a.iou_exact[2 * c], a.iou_exact[2 * c + 1]);
    a.iou_sum[c] = a.iou_add ? a.iou_sum[c] + update : update;
  }
}

__global__ __launch_bounds__(kScoreThreads) void score_images(ScoreArgs a) {
  extern __shared__ uint4 lds4[];
  if (a.guard && *a.guard) return;  // a failed validation never mutates totals
  if (score_image(a, blockIdx.x, reinterpret_cast<char*>(lds4), a.lds_bytes)) return;
  // Leave this image's pair table zeroed for the next update.
  __syncthreads();
  const i64 GP = static_cast<i64>(a.G) * a.P;
  u64* pairs = const_cast<u64*>(a.pairs) + blockIdx.x * GP;
  for (i64 e = threadIdx.x; e < GP; e += blockDim.x) pairs[e] = 0;
}

__global__ __launch_bounds__(kScoreThreads) void convert_iou(ScoreArgs a) {
  if (a.guard && *a.guard) return;
  convert_iou_sum(a);
}

int launch_score(ScoreArgs a, hipStream_t stream) {
  if (a.B == 0) return kOk;
  // The parallel layout when it fits; else metadata plus the 32-bit pair
  // table; else metadata only.
  size_t lds = par_bytes(a.G, a.P, a.C);
  if (lds > kMaxStageLds) lds = stage_bytes(a.G, a.P) + 4 * static_cast<size_t>(a.G) * a.P;
  if (lds > kMaxStageLds) lds = stage_bytes(a.G, a.P);
  if (lds > kMaxStageLds) lds = 0;
  if (!allow_lds(score_images, lds)) lds = 0;
  a.lds_bytes = lds;
  // Beside another stream's kernel, 1024-thread workgroups wait until a
  // compute unit has 16 wave slots free; narrow ones fit beside it.
  const int threads = stream_concurrent(stream) ? kConcurrentThreads : kScoreThreads;
  {
    ProfileScope timing("score_images", stream);
    score_images<<<a.B, threads, lds, stream>>>(a);
    PE_LAUNCH_CHECK();
  }
  if (!a.reward) {
    convert_iou<<<1, threads, 0, stream>>>(a);
    PE_LAUNCH_CHECK();
  }
  return kOk;
}

int launch_prepare(const ScoreArgs& a, int* error, hipStream_t stream) {
  if (a.B == 0) return kOk;
  prepare_images<<<a.B, kScoreThreads, 0, stream>>>(a, error);
  PE_LAUNCH_CHECK();
  return kOk;
}

// --------------------------------------------------------------- finaliser

__global__ __launch_bounds__(kFinalThreads) void finalize(FinalArgs a) {
  extern __shared__ uint4 lds4[];
  finalize_row(a, blockIdx.x, lds4);
}

int launch_finalize(const FinalArgs& a, i64 K, hipStream_t stream) {
  if (K <= 0) return kOk;
  size_t lds = 16 * static_cast<size_t>(a.C);
  if (lds > kMaxStageLds) return fail(kUnsupportedShape, "too many categories for the finaliser");
  if (!allow_lds(finalize, lds)) return fail(kUnsupportedShape, "finaliser LDS request refused");
  finalize<<<static_cast<unsigned>(K), kFinalThreads, lds, stream>>>(a);
  PE_LAUNCH_CHECK();
  return kOk;
}

// Per-image mean-class PQ from (B, C) statistics (evaluator.py:539-544):
// one wave per image, classes folded in order then combined across lanes in
// a fixed tree, so the result does not depend on the launch.
__global__ __launch_bounds__(256) void image_pq(const i64* tp, const i64* fp, const i64* fn,
                                                const double* sums, int B, int C, double* out,
                                                double* stats) {
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (wave >= B) return;
  const i64 row = static_cast<i64>(wave) * C;
  double pq_sum = 0.0, active = 0.0;
  for (int c = lane; c < C; c += kWave) {
    const i64 t = tp[row + c], f = fp[row + c], n = fn[row + c];
    const double s = sums[row + c];
    const double denom = static_cast<double>(t) + 0.5 * static_cast<double>(f + n);
    pq_sum += s / (denom < 1.0 ? 1.0 : denom);
    active += denom > 0 ? 1.0 : 0.0;
    if (stats) {
      double* o = stats + (row + c) * 4;
      o[0] = s;
      o[1] = static_cast<double>(t);
      o[2] = static_cast<double>(f);
      o[3] = static_cast<double>(n);
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    pq_sum += __shfl_xor(pq_sum, off, kWave);
    active += __shfl_xor(active, off, kWave);
  }
  if (lane == 0 && out) out[wave] = pq_sum / (active < 1.0 ? 1.0 : active);
}

// ---------------------------------------------------------------- packing

template <typename T, bool VEC>
__global__ __launch_bounds__(256) void pack_slots(const T* maps, i64 HW, const i64* ids, const int* slots,
                                                  int S, int blocks_per_image, int32_t* out, int* error) {
  const int b = blockIdx.x / blocks_per_image;
  const int part = blockIdx.x - b * blocks_per_image;
  const i64 span = (HW + blocks_per_image - 1) / blocks_per_image;
  const i64 begin = static_cast<i64>(b) * HW + part * span;
  const i64 end = static_cast<i64>(b) * HW + min(HW, (part + 1) * span);
  SlotResolver<T, true> resolve{SortedIds{ids + static_cast<i64>(b) * S, slots + static_cast<i64>(b) * S, S}, S,
                                true};
  bool bad = false;
  auto put = [&](i64 i, typename MapValue<T>::type v) {
    const int s = resolve(v);
    out[i] = s >= 0 ? s : 0;
    bad |= s < 0;
  };
  if (VEC) {
    // Whole quads with vector loads; the scalar head/tail has <= 3 pixels.
    const i64 va = min((begin + 3) & ~static_cast<i64>(3), end), vb = max(end & ~static_cast<i64>(3), va);
    for (i64 i = begin + threadIdx.x; i < va; i += blockDim.x) put(i, read_pixel(maps, i));
    for (i64 q = (va >> 2) + threadIdx.x; q < (vb >> 2); q += blockDim.x) {
      typename MapValue<T>::type v[4];
      load_quad(maps, q, v);
#pragma unroll
      for (int j = 0; j < 4; ++j) put(4 * q + j, v[j]);
    }
    for (i64 i = vb + threadIdx.x; i < end; i += blockDim.x) put(i, read_pixel(maps, i));
  } else {
    for (i64 i = begin + threadIdx.x; i < end; i += blockDim.x) put(i, read_pixel(maps, i));
  }
  if (bad) atomicOr(error, 1);
}

}  // namespace pe

using namespace pe;

extern "C" int64_t pe_update_workspace_bytes(int64_t B, int64_t G, int64_t P, int64_t C, int64_t packed_pixels) {
  Carve carve(nullptr);
  carve.take<int>(B * P);        // rec_lab (prediction labels, large G/P only)
  carve.take<double>(B * P);     // rec_val
  if (packed_pixels > 0) carve.take<int>(tile_list_ints(B, packed_pixels, 4096));  // chunk list
  carve.take<u64>(2 * C);        // per-update exact IoU words (iou_exact == NULL)
  return static_cast<int64_t>(align_up(carve.used, 256));
}

namespace pe {
int update_impl(const void* gt, int gt_bytes, const void* pred, int pred_bytes, int64_t B, int64_t H, int64_t W,
                const int64_t* pix_off, const int32_t* gc, const int32_t* pc, const uint8_t* crowd, int64_t G,
                int64_t P, int64_t C, int64_t* pairs, int64_t* areas, int32_t* matched, int64_t* tp, int64_t* fp,
                int64_t* fn, double* iou_sum, int64_t* confusion, const int64_t* gt_ids, const int32_t* gt_slots,
                const int64_t* pred_ids, const int32_t* pred_slots, int32_t* error, uint64_t* iou_exact,
                void* workspace, int64_t workspace_bytes, int validate, int reward, int pairs_zeroed,
                void* stream_ptr) {
  PE_REQUIRE(B >= 0 && H >= 0 && W >= 0, "negative shape");
  PE_REQUIRE(G > 0 && G < INT_MAX && P > 0 && P < INT_MAX && C > 0 && C < INT_MAX, "invalid capacities");
  PE_REQUIRE(B < INT_MAX && G * P < (1LL << 32), "pair histogram too large");
  auto width_ok = [](int bytes) { return bytes == 3 || bytes == 4 || bytes == 8; };
  PE_REQUIRE(width_ok(gt_bytes) && width_ok(pred_bytes), "maps must be int32/int64 or RGB uint8 x 3");
  PE_REQUIRE((gt_ids == nullptr) == (gt_slots == nullptr) && (pred_ids == nullptr) == (pred_slots == nullptr),
             "lookup IDs and slots must be supplied together");
  PE_REQUIRE((gt_bytes != 3 || gt_ids) && (pred_bytes != 3 || pred_ids), "RGB maps need their ID lookup table");
  PE_REQUIRE(!(gt_bytes == 3 && pred_bytes == 8) && !(gt_bytes == 8 && pred_bytes == 3),
             "an RGB map pairs with an RGB or int32 map");
  const i64 packed_pixels = pix_off ? H * W : 0;
  PE_REQUIRE(workspace_bytes >= pe_update_workspace_bytes(B, G, P, C, packed_pixels), "workspace too small");
  PE_REQUIRE(!validate || error, "validation needs the error word");
  PE_REQUIRE(reward || confusion, "dataset mode needs the confusion accumulator");
  hipStream_t stream = static_cast<hipStream_t>(stream_ptr);
  if (B == 0) return kOk;
  Carve carve(workspace);
  ScoreArgs s{};
  s.pairs = reinterpret_cast<const u64*>(pairs);
  s.gc = gc;
  s.pc = pc;
  s.crowd = crowd;
  s.areas = reinterpret_cast<i64*>(areas);
  s.matched = matched;
  s.tp = reinterpret_cast<u64*>(tp);
  s.fp = reinterpret_cast<u64*>(fp);
  s.fn = reinterpret_cast<u64*>(fn);
  s.iou_sum = iou_sum;
  s.confusion = reinterpret_cast<u64*>(confusion);
  s.rec_lab = carve.take<int>(B * P);
  s.rec_val = carve.take<double>(B * P);
  s.iou_exact = reinterpret_cast<u64*>(iou_exact);
  s.B = static_cast<int>(B);
  s.G = static_cast<int>(G);
  s.P = static_cast<int>(P);
  s.C = static_cast<int>(C);
  s.reward = reward;
  const size_t pair_bytes = static_cast<size_t>(B * G * P) * 8;
  int rc = kOk;
  if (!pairs_zeroed && (rc = launch_zero(pairs, pair_bytes, stream))) return rc;
  if (validate && (rc = launch_zero(error, 4, stream))) return rc;
  HistArgs h{};
  h.gt = gt;
  h.pred = pred;
  h.B = static_cast<int>(B);
  h.HW = H * W;  // a packed batch: the pixel total
  h.pix_off = reinterpret_cast<const i64*>(pix_off);
  h.G = s.G;
  h.P = s.P;
  h.pairs = reinterpret_cast<u64*>(pairs);
  h.gt_ids = gt_ids;
  h.gt_slots = gt_slots;
  h.pred_ids = pred_ids;
  h.pred_slots = pred_slots;
  h.error = validate ? error : nullptr;
  int* tile_ws = pix_off ? carve.take<int>(tile_list_ints(B, packed_pixels, 4096)) : nullptr;
  u64* update_exact = carve.take<u64>(2 * C);
  if (!reward && !iou_exact) {
    // No caller-kept exact words: accumulate this update's exact sums in the
    // workspace and add their rounded value to iou_sum.
    s.iou_exact = update_exact;
    s.iou_add = 1;
    if ((rc = launch_zero(update_exact, 16 * static_cast<size_t>(C), stream))) return rc;
  }
  if ((rc = launch_joint_histogram(h, gt_bytes, pred_bytes, 0, tile_ws, stream))) return rc;
  if (validate) {
    if ((rc = launch_prepare(s, error, stream))) return rc;
    int32_t host_error = 0;
    PE_HIP(hipMemcpyAsync(&host_error, error, sizeof(int32_t), hipMemcpyDeviceToHost, stream));
    PE_HIP(hipStreamSynchronize(stream));
    if (host_error) {
      // Nothing accumulated; restore the zeroed pair table for the next call.
      if ((rc = launch_zero(pairs, pair_bytes, stream))) return rc;
      return host_error;
    }
    s.areas_ready = 1;
  }
  return launch_score(s, stream);
}
}  // namespace pe

extern "C" int pe_update(const void* gt, int gt_bytes, const void* pred, int pred_bytes, int64_t B,
                         int64_t H, int64_t W, const int64_t* pix_off, const int32_t* gc, const int32_t* pc,
                         const uint8_t* crowd,
                         int64_t G, int64_t P, int64_t C, int64_t* pairs, int64_t* areas, int32_t* matched,
                         int64_t* tp, int64_t* fp, int64_t* fn, double* iou_sum, int64_t* confusion,
                         const int64_t* gt_ids, const int32_t* gt_slots, const int64_t* pred_ids,
                         const int32_t* pred_slots, int32_t* error, uint64_t* iou_exact, void* workspace,
                         int64_t workspace_bytes, int validate, int reward, int pairs_zeroed,
                         void* stream_ptr) {
  return update_impl(gt, gt_bytes, pred, pred_bytes, B, H, W, pix_off, gc, pc, crowd, G, P, C, pairs, areas,
                     matched, tp, fp, fn, iou_sum, confusion, gt_ids, gt_slots, pred_ids, pred_slots, error,
                     iou_exact, workspace, workspace_bytes, validate, reward, pairs_zeroed, stream_ptr);
}

// ------------------------------------------------ update without allocation
//
// pe_update_noalloc's workspace: one array per GPU, held until the process
// ends.  Host code never passes the array's name to a kernel (on the host it
// is only a stand-in symbol); its device address is looked up once per
// device with hipGetSymbolAddress and used everywhere instead.
constexpr int64_t kStaticScratchBytes = 8LL << 20;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex scratch_mutex;
void* scratch_address[kMaxScratchDevices] = {};

// Device address of the current device's scratch array (looked up on the
// first call on that device, with that device current).
int current_scratch(void** out) {
  int device = -1;
  PE_HIP(hipGetDevice(&device));
  PE_REQUIRE(device >= 0 && device < kMaxScratchDevices, "device index %d outside the scratch table", device);
  std::lock_guard<std::mutex> lock(scratch_mutex);
  if (!scratch_address[device]) {
    void* address = nullptr;
    PE_HIP(hipGetSymbolAddress(&address, HIP_SYMBOL(scratch)));
    PE_REQUIRE(address != nullptr, "no device address for the scratch array");
    scratch_address[device] = address;
  }
  *out = scratch_address[device];
  return kOk;
}
}  // namespace

extern "C" int64_t pe_update_noalloc_max_workspace_bytes(void) { return kStaticScratchBytes; }

extern "C" int pe_update_noalloc(const void* gt, int gt_bytes, const void* pred, int pred_bytes, int64_t B,
                                 int64_t H, int64_t W, const int64_t* pix_off, const int32_t* gc, const int32_t* pc,
                                 const uint8_t* crowd, int64_t G, int64_t P, int64_t C, int64_t* pairs,
                                 int64_t* areas, int32_t* matched, int64_t* tp, int64_t* fp, int64_t* fn,
                                 double* iou_sum, int64_t* confusion, const int64_t* gt_ids,
                                 const int32_t* gt_slots, const int64_t* pred_ids, const int32_t* pred_slots,
                                 int32_t* error, uint64_t* iou_exact, int validate, int reward, int pairs_zeroed,
                                 void* stream_ptr) {
  PE_REQUIRE(B >= 0 && H >= 0 && W >= 0 && G > 0 && P > 0 && C > 0 && C < INT_MAX, "invalid shape");
  const int64_t needed = pe_update_workspace_bytes(B, G, P, C, pix_off ? H * W : 0);
  if (needed > kStaticScratchBytes)
    return fail(kWorkspaceTooSmall,
                "pe_update_noalloc: this batch needs %lld workspace bytes, above the static scratch maximum of "
                "%lld bytes; use pe_update with a caller workspace",
                static_cast<long long>(needed), static_cast<long long>(kStaticScratchBytes));
  void* workspace = nullptr;
  const int rc = current_scratch(&workspace);
  if (rc) return rc;
  // Nothing in the workspace needs to start at zero: the per-update exact IoU
  // words are cleared by update_impl, the other regions are written before
  // they are read.
  return update_impl(gt, gt_bytes, pred, pred_bytes, B, H, W, pix_off, gc, pc, crowd, G, P, C, pairs, areas,
                     matched, tp, fp, fn, iou_sum, confusion, gt_ids, gt_slots, pred_ids, pred_slots, error,
                     iou_exact, workspace, kStaticScratchBytes, validate, reward, pairs_zeroed, stream_ptr);
}

namespace pe {
int make_final_args(const int64_t* tp, const int64_t* fp, const int64_t* fn, const double* sums,
                    const int64_t* confusion, const uint8_t* things, int64_t K, int64_t C, double* per_class,
                    double* summaries, int64_t* counts, double* miou, int64_t* out_tp, int64_t* out_fp,
                    int64_t* out_fn, double* out_sums, int64_t* out_confusion, FinalArgs* out) {
  PE_REQUIRE(C > 0 && C < INT_MAX && K >= 0 && K < INT_MAX, "invalid compute shape");
  PE_REQUIRE((out_tp == nullptr) == (out_fp == nullptr) && (out_tp == nullptr) == (out_fn == nullptr) &&
                 (out_tp == nullptr) == (out_sums == nullptr),
             "state copies must be requested together");
  FinalArgs& a = *out;
  a = FinalArgs{};
  a.tp = reinterpret_cast<const i64*>(tp);
  a.fp = reinterpret_cast<const i64*>(fp);
  a.fn = reinterpret_cast<const i64*>(fn);
  a.sums = sums;
  a.confusion = reinterpret_cast<const i64*>(confusion);
  a.things = things;
  a.C = static_cast<int>(C);
  a.per_class = per_class;
  a.summaries = summaries;
  a.counts = reinterpret_cast<i64*>(counts);
  a.miou = miou;
  a.out_tp = reinterpret_cast<i64*>(out_tp);
  a.out_fp = reinterpret_cast<i64*>(out_fp);
  a.out_fn = reinterpret_cast<i64*>(out_fn);
  a.out_sums = out_sums;
  a.out_confusion = reinterpret_cast<i64*>(out_confusion);
  return kOk;
}
}  // namespace pe

extern "C" int pe_compute(const int64_t* tp, const int64_t* fp, const int64_t* fn, const double* sums,
                          const int64_t* confusion, const uint8_t* things, int64_t K, int64_t C,
                          double* per_class, double* summaries, int64_t* counts, double* miou,
                          int64_t* out_tp, int64_t* out_fp, int64_t* out_fn, double* out_sums,
                          int64_t* out_confusion, void* stream) {
  FinalArgs a;
  const int rc = make_final_args(tp, fp, fn, sums, confusion, things, K, C, per_class, summaries, counts, miou,
                                 out_tp, out_fp, out_fn, out_sums, out_confusion, &a);
  if (rc) return rc;
  return launch_finalize(a, K, static_cast<hipStream_t>(stream));
}

extern "C" int pe_pack(const void* maps, int map_bytes, int64_t B, int64_t HW, const int64_t* ids,
                       const int32_t* slots, int64_t S, int32_t* out, int32_t* error, void* stream_ptr) {
  PE_REQUIRE(map_bytes == 3 || map_bytes == 4 || map_bytes == 8, "maps must be int32/int64 or RGB uint8 x 3");
  PE_REQUIRE(S > 0 && S < INT_MAX && B >= 0 && HW >= 0, "invalid lookup shapes");
  hipStream_t stream = static_cast<hipStream_t>(stream_ptr);
  int rc = launch_zero(error, 4, stream);
  if (rc || B == 0 || HW == 0) return rc;
  i64 per_image = std::max<i64>(1, std::min<i64>((HW + 8191) / 8192, (4LL * compute_units() + B - 1) / B));
  PE_REQUIRE(B * per_image < INT_MAX, "too many images");
  int blocks = static_cast<int>(B * per_image);
  const bool vec = quad_aligned(maps, map_bytes);
#define PE_PACK(T)                                                                                          \
  (vec ? pack_slots<T, true> : pack_slots<T, false>)<<<blocks, 256, 0, stream>>>(                           \
      static_cast<const T*>(maps), HW, ids, slots, static_cast<int>(S), static_cast<int>(per_image), out, error)
  if (map_bytes == 3) PE_PACK(Rgb24);
  else if (map_bytes == 4) PE_PACK(int32_t);
  else PE_PACK(int64_t);
#undef PE_PACK
  PE_LAUNCH_CHECK();
  return kOk;
}

extern "C" int pe_image_pq(const int64_t* tp, const int64_t* fp, const int64_t* fn, const double* sums,
                           int64_t B, int64_t C, double* out, double* stats, void* stream_ptr) {
  PE_REQUIRE(B >= 0 && B < INT_MAX / 64 && C > 0 && C < INT_MAX, "invalid reward shape");
  if (B == 0) return kOk;
  int blocks = static_cast<int>((B * kWave + 255) / 256);
  image_pq<<<blocks, 256, 0, static_cast<hipStream_t>(stream_ptr)>>>(
      reinterpret_cast<const i64*>(tp), reinterpret_cast<const i64*>(fp), reinterpret_cast<const i64*>(fn),
      sums, static_cast<int>(B), static_cast<int>(C), out, stats);
  PE_LAUNCH_CHECK();
  return kOk;
}
