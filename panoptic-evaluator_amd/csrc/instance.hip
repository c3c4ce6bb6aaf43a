// Instance path on gfx950: mask intersections (int8 MFMA), detection x
// target-slot histograms, COCO greedy matching and AP/AR accumulation.
//
// Behaviour follows the reference's CPU path (instance.py:255-331 matching and
// accumulation, :230-248 summaries, combined.py:63-161 shared-target rules)
// and its native boundary (instance_kernel.cu:282-381).
//
// Design (MI355X-first):
//  * mask x mask intersections: bool masks are bytes 0/1, i.e. valid int8
//    MFMA operands as they lie in HBM.  Default (staged) form: a 512-thread
//    workgroup covers 128 detections x 64 GT masks over a pixel range; each
//    256-pixel block of every row is staged in LDS and wave w multiplies
//    detection tile w by every GT tile with 16x16x64 int8 MFMAs (exact int32
//    accumulation).  Register form (tuner alternative): one wave per
//    16-detection tile with operands loaded straight into VGPRs.  Partial
//    tiles are added atomically (counts never exceed the pixel count).  Mask
//    areas come from the same loads (byte sums), so every mask byte is read
//    from HBM once.
//  * detection x slot histogram: one workgroup owns a pixel tile of one image,
//    keeps the target slots of the tile in registers and streams every
//    detection mask over it with 16-byte loads (all-zero chunks skip at once),
//    counting (detection, slot) in packed 16-bit LDS bins.
//  * matching: one wave per (image, class); lane L < 40 owns the chain
//    (area range L / 10, IoU threshold L % 10); detections are processed in
//    score order with a per-chain GT bitmask in LDS.
//  * AP: up to 4096 records, one workgroup per class gathers its records,
//    sorts them by (score descending, input index) with a bitonic network in
//    LDS and scans them per (threshold, area, maxDets) job; larger record sets
//    are first grouped per class in global score order by stable sorts.  The
//    cumulative TP/FP scan runs forward and the precision envelope backward,
//    matching the reference's searchsorted/cummax result exactly.

#include <atomic>
#include <cfloat>
#include <climits>
#include <type_traits>
#include <cmath>

#include "common.h"
#include "finalize.h"
#include "merge_sort.h"
#include "../../include/panoptic_hip.h"

namespace pe {

typedef int v4i __attribute__((ext_vector_type(4)));

// --------------------------------------------------------- MFMA overlaps


__device__ __forceinline__ int byte_sum4(int w) {
  // four bytes, each 0 or 1
  return static_cast<int>((static_cast<unsigned>(w) * 0x01010101u) >> 24);
}

__device__ __forceinline__ v4i load16(const uint8_t* row, i64 k, i64 HW, bool valid_row, bool aligned) {
  v4i v = {0, 0, 0, 0};
  if (!valid_row || k >= HW) return v;
  if (aligned && k + 16 <= HW) return *as_global(reinterpret_cast<const v4i*>(row + k));
  const auto* g = as_global(row);
  uint8_t bytes[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) bytes[j] = k + j < HW ? g[k + j] : 0;
  __builtin_memcpy(&v, bytes, 16);
  return v;
}

// Images of a mask-overlap batch: detection masks pm[b] (D_b x pixels_b) and
// GT masks gm[b] (G_b x pixels_b); inter[b] is a D_b x G_b int32 block at
// inter_off[b]; areas at det_off[b] / gt_off[b].
// One image (pm_table == nullptr) takes its pointers and counts from the
// arguments themselves.
struct OverlapArgs {
  const uint8_t* const* pm_table;
  const uint8_t* const* gm_table;
  const i64* det_off;     // B + 1
  const i64* gt_off;      // B + 1
  const i64* inter_off;   // B + 1
  const i64* pixels;      // B pixel counts, or nullptr: every image has HW
  const uint8_t* pm;      // one image
  const uint8_t* gm;
  int D, G;
  i64 HW;
  int chunks;             // pixel chunks of the largest image
  int dgroups, ggroups;   // work items: dgroups x ggroups x (images * chunks)
  int items;
  int xcd_span;           // > 0: work items per XCD (grid padded to kXcds * xcd_span)
  i64 chunk_px;           // pixels per work item, for kernels whose chunk is set at launch
  // Optional (batch form, bit-packed masks): per image, 2 int32 per mask, a
  // word range [first, end) outside which the mask's row is zero.
  const int32_t* const* pm_ext;
  const int32_t* const* gm_ext;
  int32_t* inter;
  int32_t* pa;            // may be null
  int32_t* ga;            // may be null
  int aligned;
};

// XCDs of a gfx950 device; the dispatcher deals workgroup i to XCD i % kXcds.
constexpr int kXcds = 8;

// One work item of an overlap launch: detection group (fastest), GT group,
// then image * chunks + chunk, so the workgroups reading one chunk's GT masks
// are neighbours.  With xcd_span > 0 neighbouring items run on one XCD
// (workgroup i takes item (i % kXcds) * xcd_span + i / kXcds) and share its
// L2; otherwise the dispatcher spreads them over every XCD.
struct OverlapItem {
  int b;
  int dgroup, ggroup, chunk;
  int D, G;
  i64 HW;
  const uint8_t* pm;
  const uint8_t* gm;
  int32_t* inter;
  int32_t* pa;
  int32_t* ga;
};

__device__ __forceinline__ bool overlap_item(const OverlapArgs& o, int dets_per_group, int gts_per_group,
                                             i64 chunk_px, OverlapItem& it) {
  const int wg = static_cast<int>(blockIdx.x);
  const int item = o.xcd_span > 0 ? (wg % kXcds) * o.xcd_span + wg / kXcds : wg;
  if (item >= o.items) return false;
  it.dgroup = item % o.dgroups;
  const int rest = item / o.dgroups;
  it.ggroup = rest % o.ggroups;
  const int z = rest / o.ggroups;
  const int b = z / o.chunks;
  it.b = b;
  it.chunk = z - b * o.chunks;
  const bool table = o.pm_table != nullptr;
  const i64 d_first = table ? o.det_off[b] : 0, g_first = table ? o.gt_off[b] : 0;
  it.D = table ? static_cast<int>(o.det_off[b + 1] - d_first) : o.D;
  it.G = table ? static_cast<int>(o.gt_off[b + 1] - g_first) : o.G;
  it.HW = o.pixels ? o.pixels[b] : o.HW;
  // With D == 0 or G == 0 the first group of the other side still counts areas.
  if (it.chunk * chunk_px >= it.HW || it.dgroup * dets_per_group >= max(it.D, 1) ||
      it.ggroup * gts_per_group >= max(it.G, 1))
    return false;
  it.pm = table ? o.pm_table[b] : o.pm;
  it.gm = table ? o.gm_table[b] : o.gm;
  it.inter = o.inter + (table ? o.inter_off[b] : 0);
  it.pa = o.pa ? o.pa + d_first : nullptr;
  it.ga = o.ga ? o.ga + g_first : nullptr;
  return true;
}

// inter[d0 + i, g0 + j] += one 16 x 16 MFMA result tile; lane holds rows
// 4 * (lane / 16) + r, column lane % 16.
__device__ __forceinline__ void add_overlap_tile(v4i acc, int lane, int d0, int g0, int D, int G,
                                                 int32_t* inter) {
  const int g = g0 + (lane & 15);
  if (g >= G) return;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int d = d0 + 4 * (lane >> 4) + r;
    const int v = acc[r];
    if (d < D && v) atomicAdd(inter + static_cast<i64>(d) * G + g, v);
  }
}

// inter[d, g] += popcount(pm[d] & gm[g]) over one wave's pixel chunk, for a
// 16-detection tile and up to kOverlapGTiles 16-GT tiles (kOverlapChunk
// pixels per wave, a multiple of 64), loaded straight into MFMA operands.
template <int kOverlapGTiles, int kOverlapChunk, int kOverlapAhead>
__global__ __launch_bounds__(64) void mask_overlap_mfma(OverlapArgs o) {
  const int lane = threadIdx.x;
  OverlapItem it;
  if (!overlap_item(o, 16, 16 * kOverlapGTiles, kOverlapChunk, it)) return;
  const int dtile = it.dgroup, ggroup = it.ggroup;
  const int D = it.D, G = it.G;
  const i64 HW = it.HW;
  const i64 k_begin = static_cast<i64>(it.chunk) * kOverlapChunk;
  const uint8_t* pm = it.pm;
  const uint8_t* gm = it.gm;
  int32_t* inter = it.inter;
  int32_t* pa = it.pa;
  int32_t* ga = it.ga;
  const int row = lane & 15, kgroup = lane >> 4;
  const int d = dtile * 16 + row;
  const bool dvalid = d < D;
  const uint8_t* arow = pm + static_cast<i64>(dvalid ? d : 0) * HW;
  const uint8_t* brow[kOverlapGTiles];
  bool bvalid[kOverlapGTiles];
#pragma unroll
  for (int t = 0; t < kOverlapGTiles; ++t) {
    const int g = (ggroup * kOverlapGTiles + t) * 16 + row;
    bvalid[t] = g < G;
    brow[t] = gm + static_cast<i64>(bvalid[t] ? g : 0) * HW;
  }
  int tiles = G - ggroup * kOverlapGTiles * 16;
  tiles = tiles <= 0 ? 0 : min(kOverlapGTiles, (tiles + 15) / 16);
  v4i acc[kOverlapGTiles];
#pragma unroll
  for (int t = 0; t < kOverlapGTiles; ++t) acc[t] = v4i{0, 0, 0, 0};
  int asum = 0, bsum[kOverlapGTiles];
#pragma unroll
  for (int t = 0; t < kOverlapGTiles; ++t) bsum[t] = 0;
  const bool count_a = pa != nullptr && ggroup == 0;
  const bool count_b = ga != nullptr && dtile == 0;
  const i64 k_end = min(k_begin + kOverlapChunk, HW);
  const bool al = o.aligned != 0;
  // kOverlapAhead K steps of loads are issued before their MFMAs.
  for (i64 k0 = k_begin; k0 < k_end; k0 += 64 * kOverlapAhead) {
    v4i a[kOverlapAhead], bv[kOverlapAhead][kOverlapGTiles];
#pragma unroll
    for (int u = 0; u < kOverlapAhead; ++u) {
      const i64 k = k0 + u * 64 + kgroup * 16;
      a[u] = load16(arow, k, k_end, dvalid, al);
#pragma unroll
      for (int t = 0; t < kOverlapGTiles; ++t) bv[u][t] = t < tiles ? load16(brow[t], k, k_end, bvalid[t], al)
                                                                     : v4i{0, 0, 0, 0};
    }
#pragma unroll
    for (int u = 0; u < kOverlapAhead; ++u) {
      if (count_a) asum += byte_sum4(a[u].x) + byte_sum4(a[u].y) + byte_sum4(a[u].z) + byte_sum4(a[u].w);
#pragma unroll
      for (int t = 0; t < kOverlapGTiles; ++t) {
        if (t >= tiles) break;
        if (count_b)
          bsum[t] += byte_sum4(bv[u][t].x) + byte_sum4(bv[u][t].y) + byte_sum4(bv[u][t].z) + byte_sum4(bv[u][t].w);
        acc[t] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[u], bv[u][t], acc[t], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int t = 0; t < kOverlapGTiles; ++t) {
    if (t >= tiles) break;
    add_overlap_tile(acc[t], lane, dtile * 16, (ggroup * kOverlapGTiles + t) * 16, D, G, inter);
  }
  // Areas: the four lanes of a row (kgroups) hold disjoint pixels.
  if (count_a) {
    asum += __shfl_xor(asum, 16, kWave);
    asum += __shfl_xor(asum, 32, kWave);
    if (kgroup == 0 && dvalid && asum) atomicAdd(pa + d, asum);
  }
  if (count_b) {
#pragma unroll
    for (int t = 0; t < kOverlapGTiles; ++t) {
      if (t >= tiles) break;
      int sum = bsum[t];
      sum += __shfl_xor(sum, 16, kWave);
      sum += __shfl_xor(sum, 32, kWave);
      const int g = (ggroup * kOverlapGTiles + t) * 16 + row;
      if (kgroup == 0 && g < G && sum) atomicAdd(ga + g, sum);
    }
  }
}

// The same contraction with both operands staged in LDS: a workgroup of DT
// waves covers 16 * DT detections x 16 * GT GT masks over a Chunk-pixel
// range, in blocks of KB pixels.  Every row of a block is loaded as one
// contiguous KB-byte run (16 bytes per lane, neighbouring lanes on
// neighbouring bytes), so a load instruction touches few rows (pages, L2
// channels) and each mask byte is fetched once per workgroup; wave w then
// multiplies detection tile w by every GT tile out of LDS.  Rows are padded
// by 16 bytes so that the 16 rows of an MFMA operand fall in distinct banks.
// The next block's loads are in flight while the current one is multiplied.
// Expands 32 mask bits into 32 bytes of 0/1 (byte j = bit j): each nibble
// n becomes (n * 0x00204081) & 0x01010101, the four shifted copies of n
// (by 0, 7, 14, 21 bits) never overlapping.
__device__ __forceinline__ void expand_bits(uint32_t w, v4i& lo, v4i& hi) {
  uint32_t d[8];
#pragma unroll
  for (int n = 0; n < 8; ++n) d[n] = (((w >> (4 * n)) & 0xFu) * 0x00204081u) & 0x01010101u;
  lo = v4i{static_cast<int>(d[0]), static_cast<int>(d[1]), static_cast<int>(d[2]), static_cast<int>(d[3])};
  hi = v4i{static_cast<int>(d[4]), static_cast<int>(d[5]), static_cast<int>(d[6]), static_cast<int>(d[7])};
}

// BITS: the masks are bit rows (mask_words(HW) words each); every 32-pixel
// word is expanded to 0/1 bytes on its way into LDS, so the same int8 MFMA
// loop runs on 1/8 of the HBM bytes.  Ahead: blocks whose loads are in
// flight (bit rows are so short per block that one block of look-ahead does
// not cover the memory latency).
// DPW: detection tiles per wave (2 halves the LDS reads of the GT operands
// per MFMA, with half the waves).
template <int DT, int GT, int KB, int Chunk, bool BITS, int Ahead = 1, int DPW = 1>
__global__ __launch_bounds__(64 * DT / DPW) void mask_overlap_staged(OverlapArgs o) {
  static_assert(DT % DPW == 0, "whole detection tiles per wave");
  constexpr int kThreads = 64 * DT / DPW;
  constexpr int kRows = 16 * (DT + GT);
  constexpr int kStride = KB + 16;
  constexpr int kPiecePx = BITS ? 32 : 16;  // pixels per loaded piece (one word / 16 bytes)
  constexpr int kPerRow = KB / kPiecePx;    // pieces of one row's block
  constexpr int kSlots = kRows * kPerRow / kThreads;
  using Piece = typename std::conditional<BITS, uint32_t, v4i>::type;
  static_assert(KB % 64 == 0 && Chunk % (KB * Ahead) == 0, "blocks are whole MFMA steps");
  static_assert(kThreads % kPerRow == 0 && kPerRow <= kWave && (kPerRow & (kPerRow - 1)) == 0,
                "the pieces of one row sit in one wave");
  static_assert(kRows * kPerRow % kThreads == 0, "every thread loads the same number of pieces");
  __shared__ __align__(16) uint8_t tile[kRows * kStride];
  OverlapItem it;
  if (!overlap_item(o, 16 * DT, 16 * GT, Chunk, it)) return;
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const int D = it.D, G = it.G;
  const int d_base = it.dgroup * 16 * DT, g_base = it.ggroup * 16 * GT;
  const i64 HW = it.HW;
  const i64 row_len = BITS ? 4 * mask_words(HW) : HW;  // bytes per mask row
  const i64 k_begin = static_cast<i64>(it.chunk) * Chunk;
  const i64 k_end = min(k_begin + Chunk, HW);
  const bool al = o.aligned != 0;
  // Slot i of this thread: piece tid + i * kThreads of the block, i.e. row
  // (tid + i * kThreads) / kPerRow at pixel (tid % kPerRow) * kPiecePx.
  const int col = (tid % kPerRow) * kPiecePx;
  const uint8_t* rowp[kSlots];
  bool rvalid[kSlots];
  int lds[kSlots];
  int area[kSlots];
#pragma unroll
  for (int i = 0; i < kSlots; ++i) {
    const int r = (tid + i * kThreads) / kPerRow;
    const bool det = r < 16 * DT;
    const int idx = det ? d_base + r : g_base + r - 16 * DT;
    rvalid[i] = idx < (det ? D : G);
    rowp[i] = (det ? it.pm : it.gm) + static_cast<i64>(rvalid[i] ? idx : 0) * row_len;
    lds[i] = r * kStride + col;
    area[i] = 0;
  }
  auto load = [&](int i, i64 k) -> Piece {
    if constexpr (BITS) {
      // Bits past HW are zero in every row, so a partial last word is whole.
      // Ordinary (cached) loads: a block takes 32 bytes of a row, and the
      // next three blocks read the rest of the same 128-byte line.
      return rvalid[i] && k < k_end ? *(as_global(reinterpret_cast<const uint32_t*>(rowp[i])) + (k >> 5)) : 0u;
    } else {
      return load16(rowp[i], k, k_end, rvalid[i], al);
    }
  };
  int tiles = G - g_base;
  tiles = tiles <= 0 ? 0 : min(GT, (tiles + 15) / 16);
  // Wave w multiplies detection tiles w * DPW .. w * DPW + DPW - 1.
  const int dtiles = min(DPW, (D - d_base - wave * DPW * 16 + 15) / 16);  // uniform per wave; <= 0: idle
  const bool compute = dtiles > 0;
  const int row = lane & 15, kgroup = lane >> 4;
  const uint8_t* a_row = tile + (wave * DPW * 16 + row) * kStride + kgroup * 16;
  const uint8_t* b_row = tile + (16 * DT + row) * kStride + kgroup * 16;
  v4i acc[DPW][GT];
#pragma unroll
  for (int d = 0; d < DPW; ++d)
#pragma unroll
    for (int t = 0; t < GT; ++t) acc[d][t] = v4i{0, 0, 0, 0};
  Piece buf[Ahead][kSlots];
#pragma unroll
  for (int ahead = 0; ahead < Ahead; ++ahead)
#pragma unroll
    for (int i = 0; i < kSlots; ++i) buf[ahead][i] = load(i, k_begin + ahead * KB + col);
  for (i64 k0 = k_begin; k0 < k_end; k0 += KB * Ahead) {
#pragma unroll
    for (int ahead = 0; ahead < Ahead; ++ahead) {
      const i64 k = k0 + ahead * KB;
      if (k >= k_end) break;  // uniform
      __syncthreads();        // the previous block has been multiplied
#pragma unroll
      for (int i = 0; i < kSlots; ++i) {
        if constexpr (BITS) {
          v4i lo, hi;
          expand_bits(buf[ahead][i], lo, hi);
          *reinterpret_cast<v4i*>(tile + lds[i]) = lo;
          *reinterpret_cast<v4i*>(tile + lds[i] + 16) = hi;
          area[i] += __builtin_popcount(buf[ahead][i]);
        } else {
          const v4i v = buf[ahead][i];
          *reinterpret_cast<v4i*>(tile + lds[i]) = v;
          area[i] += byte_sum4(v.x) + byte_sum4(v.y) + byte_sum4(v.z) + byte_sum4(v.w);
        }
      }
      __syncthreads();
      if (k + Ahead * KB < k_end) {
#pragma unroll
        for (int i = 0; i < kSlots; ++i) buf[ahead][i] = load(i, k + Ahead * KB + col);
      }
      if (compute) {
#pragma unroll
        for (int step = 0; step < KB / 64; ++step) {
          v4i a[DPW];
#pragma unroll
          for (int d = 0; d < DPW; ++d)
            a[d] = d < dtiles ? *reinterpret_cast<const v4i*>(a_row + d * 16 * kStride + step * 64) : v4i{0, 0, 0, 0};
#pragma unroll
          for (int t = 0; t < GT; ++t) {
            if (t >= tiles) break;
            const v4i b = *reinterpret_cast<const v4i*>(b_row + t * 16 * kStride + step * 64);
#pragma unroll
            for (int d = 0; d < DPW; ++d)
              if (d < dtiles) acc[d][t] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[d], b, acc[d][t], 0, 0, 0);
          }
        }
      }
    }
  }
  if (compute) {
#pragma unroll
    for (int d = 0; d < DPW; ++d) {
      if (d >= dtiles) break;
#pragma unroll
      for (int t = 0; t < GT; ++t) {
        if (t >= tiles) break;
        add_overlap_tile(acc[d][t], lane, d_base + (wave * DPW + d) * 16, g_base + t * 16, D, G, it.inter);
      }
    }
  }
  // Areas: the kPerRow neighbouring lanes of a slot share its row.
  const bool count_a = it.pa != nullptr && it.ggroup == 0;
  const bool count_b = it.ga != nullptr && it.dgroup == 0;
#pragma unroll
  for (int i = 0; i < kSlots; ++i) {
    int sum = area[i];
#pragma unroll
    for (int m = 1; m < kPerRow; m <<= 1) sum += __shfl_xor(sum, m, kWave);
    const int r = (tid + i * kThreads) / kPerRow;
    const bool det = r < 16 * DT;
    if (col == 0 && rvalid[i] && sum && (det ? count_a : count_b))
      atomicAdd(det ? it.pa + d_base + r : it.ga + g_base + r - 16 * DT, sum);
  }
}

struct OverlapVariant {
  int dets;     // detections per work item
  int gts;      // GT masks per work item
  int chunk;    // pixels per work item
  int threads;
  bool xcd;
  void (*kernel)(OverlapArgs);
};
// Order matches the tuning registry's descriptions.
const OverlapVariant kOverlap[] = {
    {128, 64, 8192, 512, true, mask_overlap_staged<8, 4, 256, 8192, false>},
    {128, 64, 4096, 512, true, mask_overlap_staged<8, 4, 256, 4096, false>},
    {64, 64, 8192, 256, true, mask_overlap_staged<4, 4, 256, 8192, false>},
    {128, 32, 8192, 512, true, mask_overlap_staged<8, 2, 512, 8192, false>},
    {16, 64, 4096, 64, true, mask_overlap_mfma<4, 4096, 4>},
    {16, 64, 8192, 64, true, mask_overlap_mfma<4, 8192, 4>},
    {16, 128, 16384, 64, true, mask_overlap_mfma<8, 16384, 1>},
    {16, 64, 8192, 64, false, mask_overlap_mfma<4, 8192, 4>},
    {128, 64, 4096, 256, true, mask_overlap_staged<8, 4, 256, 4096, false, 1, 2>},
};

int launch_overlap(OverlapArgs o, int64_t B, int64_t max_d, int64_t max_g, int64_t max_hw, int64_t total_d,
                   int64_t total_g, int64_t total_inter, hipStream_t stream) {
  int rc;
  if ((rc = launch_zero3(reinterpret_cast<uint32_t*>(o.inter), total_inter, reinterpret_cast<uint32_t*>(o.pa),
                         total_d, reinterpret_cast<uint32_t*>(o.ga), total_g, stream)))
    return rc;
  if (B == 0 || max_hw == 0 || (max_d == 0 && max_g == 0)) return kOk;
  const OverlapVariant& v = kOverlap[tuned_variant(kTuneMaskOverlap)];
  const i64 chunks = (max_hw + v.chunk - 1) / v.chunk;
  const i64 dgroups = std::max<i64>(1, (max_d + v.dets - 1) / v.dets);
  const i64 ggroups = std::max<i64>(1, (max_g + v.gts - 1) / v.gts);
  const i64 items = dgroups * ggroups * B * chunks;
  if (items >= INT_MAX - kXcds) return fail(kUnsupportedShape, "too many masks or pixels for one overlap launch");
  o.chunks = static_cast<int>(chunks);
  o.dgroups = static_cast<int>(dgroups);
  o.ggroups = static_cast<int>(ggroups);
  o.items = static_cast<int>(items);
  o.xcd_span = v.xcd ? static_cast<int>((items + kXcds - 1) / kXcds) : 0;
  const i64 grid = v.xcd ? static_cast<i64>(o.xcd_span) * kXcds : items;
  ProfileScope timing("mask_overlap_mfma", stream);
  v.kernel<<<static_cast<unsigned>(grid), v.threads, 0, stream>>>(o);
  PE_LAUNCH_CHECK();
  return kOk;
}

// --------------------------------------------------------- bit-packed masks

// Four mask bytes (any nonzero byte is set) -> four bits, byte j -> bit j.
__device__ __forceinline__ uint32_t byte_bits4(uint32_t x) {
  const uint32_t set = (((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x) & 0x80808080u;  // high bit of nonzero bytes
  return ((set >> 7) * 0x01020408u) >> 24;
}

// Bool masks (N rows of HW bytes) -> bit rows of `stride` = mask_words(HW)
// words; bits past HW are zero.
__global__ __launch_bounds__(256) void pack_mask_bits(const uint8_t* masks, i64 N, i64 HW, i64 stride, int aligned,
                                                      uint32_t* out) {
  const i64 total = N * stride;
  for (i64 w = static_cast<i64>(blockIdx.x) * 256 + threadIdx.x; w < total; w += static_cast<i64>(gridDim.x) * 256) {
    const i64 n = w / stride, k = (w - n * stride) * 32;
    uint32_t word = 0;
    if (k < HW) {
      const uint8_t* row = masks + n * HW;
      if (aligned && k + 32 <= HW) {
        const auto* p = as_global(reinterpret_cast<const v4i*>(row + k));
        const v4i lo = __builtin_nontemporal_load(p), hi = __builtin_nontemporal_load(p + 1);
        const uint32_t d[8] = {static_cast<uint32_t>(lo.x), static_cast<uint32_t>(lo.y),
                               static_cast<uint32_t>(lo.z), static_cast<uint32_t>(lo.w),
                               static_cast<uint32_t>(hi.x), static_cast<uint32_t>(hi.y),
                               static_cast<uint32_t>(hi.z), static_cast<uint32_t>(hi.w)};
#pragma unroll
        for (int j = 0; j < 8; ++j) word |= byte_bits4(d[j]) << (4 * j);
      } else {
        const auto* g = as_global(row);
        for (int j = 0; j < 32 && k + j < HW; ++j) word |= static_cast<uint32_t>(g[k + j] != 0) << j;
      }
    }
    out[w] = word;
  }
}

// inter[d, g] += popcount(pm[d] & gm[g]) over bit rows: a workgroup of 256
// threads covers 16 * DR detections x 16 * GR GT masks over ChunkW words
// (32 pixels each), staging blocks of KW words of every row in LDS (each row
// block is a contiguous run, loaded 16 bytes per lane).  Thread (td, tg)
// owns detections td + 16 j and GT masks tg + 16 i, so the 16 lanes of one
// LDS read pass touch 16 consecutive rows (distinct banks with the 4-word
// row padding) or one row (a broadcast).  Areas are popcounts of the staged
// words.  The next block's loads are in flight during the popcounts.
template <int DR, int GR, int KW, int ChunkW>
__global__ __launch_bounds__(256) void mask_overlap_bits(OverlapArgs o) {
  constexpr int kThreads = 256;
  constexpr int kDets = 16 * DR, kGts = 16 * GR;
  constexpr int kRows = kDets + kGts;
  constexpr int kStride = KW + 4;  // words
  constexpr int kPerRow = KW / 4;  // 16-byte pieces of one row block
  constexpr int kSlots = kRows * kPerRow / kThreads;
  static_assert(KW % 4 == 0 && ChunkW % KW == 0, "blocks are whole 16-byte pieces");
  static_assert(kThreads % kPerRow == 0 && kPerRow <= kWave && (kPerRow & (kPerRow - 1)) == 0,
                "the pieces of one row sit in one wave");
  static_assert(kRows * kPerRow % kThreads == 0, "every thread loads the same number of pieces");
  __shared__ __align__(16) uint32_t tile[kRows * kStride];
  OverlapItem it;
  if (!overlap_item(o, kDets, kGts, static_cast<i64>(ChunkW) * 32, it)) return;
  const int tid = threadIdx.x;
  const int D = it.D, G = it.G;
  const int d_base = it.dgroup * kDets, g_base = it.ggroup * kGts;
  const i64 words = mask_words(it.HW);
  const i64 w_begin = static_cast<i64>(it.chunk) * ChunkW;
  const i64 w_end = min(w_begin + ChunkW, words);  // both multiples of 4
  const auto* pm = reinterpret_cast<const uint32_t*>(it.pm);
  const auto* gm = reinterpret_cast<const uint32_t*>(it.gm);
  // Slot i of this thread: row tid / kPerRow + i * kRowStep of the block, at
  // word (tid % kPerRow) * 4.  Rows and pointers are recomputed per load to
  // keep registers for the counts.
  constexpr int kRowStep = kThreads / kPerRow;
  const int col = (tid % kPerRow) * 4, row0 = tid / kPerRow;
  auto slot_row = [&](int i, int& idx, bool& det) {
    const int r = row0 + i * kRowStep;
    det = r < kDets;
    idx = det ? d_base + r : g_base + r - kDets;
    return r;
  };
  auto load = [&](int i, i64 w) {
    int idx;
    bool det;
    slot_row(i, idx, det);
    if (idx >= (det ? D : G) || w + col >= w_end) return v4i{0, 0, 0, 0};
    const uint32_t* row = (det ? pm : gm) + static_cast<i64>(idx) * words;
    return __builtin_nontemporal_load(as_global(reinterpret_cast<const v4i*>(row + w + col)));
  };
  int area[kSlots];
#pragma unroll
  for (int i = 0; i < kSlots; ++i) area[i] = 0;
  // Detection blocks j and GT blocks i that hold any mask (uniform).
  const int jd = min(DR, (D - d_base + 15) / 16), ig = min(GR, max(0, (G - g_base + 15) / 16));
  const int td = tid >> 4, tg = tid & 15;
  int acc[DR][GR];
#pragma unroll
  for (int j = 0; j < DR; ++j)
#pragma unroll
    for (int i = 0; i < GR; ++i) acc[j][i] = 0;
  v4i buf[kSlots];
#pragma unroll
  for (int i = 0; i < kSlots; ++i) buf[i] = load(i, w_begin);
  typedef unsigned v2u __attribute__((ext_vector_type(2)));
  for (i64 w = w_begin; w < w_end; w += KW) {
    __syncthreads();  // the previous block has been counted
#pragma unroll
    for (int i = 0; i < kSlots; ++i) {
      *reinterpret_cast<v4i*>(tile + (row0 + i * kRowStep) * kStride + col) = buf[i];
      area[i] += __builtin_popcount(static_cast<uint32_t>(buf[i].x)) +
                 __builtin_popcount(static_cast<uint32_t>(buf[i].y)) +
                 __builtin_popcount(static_cast<uint32_t>(buf[i].z)) +
                 __builtin_popcount(static_cast<uint32_t>(buf[i].w));
    }
    __syncthreads();
    if (w + KW < w_end) {
#pragma unroll
      for (int i = 0; i < kSlots; ++i) buf[i] = load(i, w + KW);
    }
#pragma unroll 2
    for (int c = 0; c < KW; c += 2) {
      v2u dv[DR], gv[GR];
#pragma unroll
      for (int j = 0; j < DR; ++j)
        if (j < jd) dv[j] = *reinterpret_cast<const v2u*>(tile + (td + 16 * j) * kStride + c);
#pragma unroll
      for (int i = 0; i < GR; ++i)
        if (i < ig) gv[i] = *reinterpret_cast<const v2u*>(tile + (kDets + tg + 16 * i) * kStride + c);
#pragma unroll
      for (int j = 0; j < DR; ++j) {
        if (j >= jd) break;
#pragma unroll
        for (int i = 0; i < GR; ++i) {
          if (i >= ig) break;
          acc[j][i] += __builtin_popcount(dv[j].x & gv[i].x) + __builtin_popcount(dv[j].y & gv[i].y);
        }
      }
    }
  }
#pragma unroll
  for (int j = 0; j < DR; ++j) {
    const int d = d_base + td + 16 * j;
#pragma unroll
    for (int i = 0; i < GR; ++i) {
      const int g = g_base + tg + 16 * i;
      if (d < D && g < G && acc[j][i]) atomicAdd(it.inter + static_cast<i64>(d) * G + g, acc[j][i]);
    }
  }
  // Areas: the kPerRow neighbouring lanes of a slot share its row.
  const bool count_a = it.pa != nullptr && it.ggroup == 0;
  const bool count_b = it.ga != nullptr && it.dgroup == 0;
#pragma unroll
  for (int i = 0; i < kSlots; ++i) {
    int sum = area[i];
#pragma unroll
    for (int m = 1; m < kPerRow; m <<= 1) sum += __shfl_xor(sum, m, kWave);
    int idx;
    bool det;
    slot_row(i, idx, det);
    if (col == 0 && idx < (det ? D : G) && sum && (det ? count_a : count_b))
      atomicAdd(det ? it.pa + idx : it.ga + idx, sum);
  }
}

// Bit-packed masks with the detection operands built in registers: each wave
// loads its detection rows' words straight from HBM and expands the 16 bits
// an MFMA lane needs into its 16 operand bytes, so LDS holds only the GT
// rows, which every wave shares.  Per 256-pixel block the LDS traffic is the
// GT rows' expanded bytes (written once, read by every wave) instead of all
// rows' (c1ip: ~80 KB instead of 144 KB per block with DPW = 2).
// Workgroup: DT / DPW waves cover 16 * DT detections x 16 * GT GT masks over
// a Chunk-pixel range; wave w owns detection tiles w * DPW .. w * DPW + DPW - 1.
template <int DT, int GT, int Chunk, int DPW>
__global__ __launch_bounds__(64 * DT / DPW) void mask_overlap_bits_reg(OverlapArgs o) {
  static_assert(DT % DPW == 0, "whole detection tiles per wave");
  constexpr int kThreads = 64 * DT / DPW;
  constexpr int KB = 256;                       // pixels per block: 4 MFMA steps, 8 words per row
  constexpr int kStride = KB + 16;              // expanded GT row in LDS, padded
  constexpr int kRows = 16 * GT;
  constexpr int kPerRow = KB / 32;              // words of one row's block
  constexpr int kSlots = kRows * kPerRow / kThreads;
  static_assert(Chunk % KB == 0, "blocks are whole MFMA steps");
  static_assert(kThreads % kPerRow == 0 && kRows * kPerRow % kThreads == 0, "every thread stages whole words");
  __shared__ __align__(16) uint8_t tile[kRows * kStride];
  OverlapItem it;
  if (!overlap_item(o, 16 * DT, 16 * GT, Chunk, it)) return;
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const int D = it.D, G = it.G;
  const int d_base = it.dgroup * 16 * DT, g_base = it.ggroup * 16 * GT;
  const i64 HW = it.HW;
  const i64 row_words = mask_words(HW);
  const i64 k_begin = static_cast<i64>(it.chunk) * Chunk;
  const i64 k_end = min(k_begin + Chunk, HW);
  // GT staging: slot i of this thread is word (tid + i * kThreads) % 8 of GT
  // row (tid + i * kThreads) / 8.
  const int col = tid % kPerRow;
  const uint32_t* grow[kSlots];
  bool gvalid[kSlots];
  int glds[kSlots];
  int garea[kSlots];
#pragma unroll
  for (int i = 0; i < kSlots; ++i) {
    const int r = (tid + i * kThreads) / kPerRow;
    gvalid[i] = g_base + r < G;
    grow[i] = reinterpret_cast<const uint32_t*>(it.gm) + static_cast<i64>(gvalid[i] ? g_base + r : 0) * row_words;
    glds[i] = r * kStride + col * 32;
    garea[i] = 0;
  }
  // Detection rows of this wave: row lane % 16 of tile d; the 16 operand
  // bytes of step s are bits 16 * (kgroup & 1) .. + 15 of word
  // 2 * s + (kgroup >> 1) of the block.
  const int row = lane & 15, kgroup = lane >> 4;
  const int dtiles = min(DPW, (D - d_base - wave * DPW * 16 + 15) / 16);  // uniform per wave; <= 0: idle
  const uint32_t* drow[DPW];
  bool dvalid[DPW];
  int darea[DPW];
#pragma unroll
  for (int d = 0; d < DPW; ++d) {
    const int idx = d_base + (wave * DPW + d) * 16 + row;
    dvalid[d] = idx < D;
    drow[d] = reinterpret_cast<const uint32_t*>(it.pm) + static_cast<i64>(dvalid[d] ? idx : 0) * row_words;
    darea[d] = 0;
  }
  // One block's words: GT words for the LDS stage, detection words (8 per
  // row, two 16-byte loads; the second only inside the row) for registers.
  auto load_gt = [&](int i, i64 k0) -> uint32_t {
    const i64 k = k0 + 32 * col;
    return gvalid[i] && k < k_end ? *(as_global(grow[i]) + (k >> 5)) : 0u;
  };
  auto load_det = [&](int d, i64 k0, v4i& lo, v4i& hi) {
    const bool on = dvalid[d] && k0 < k_end;
    const auto* p = reinterpret_cast<const __attribute__((address_space(1))) v4i*>(as_global(drow[d]) + (k0 >> 5));
    lo = on ? *p : v4i{0, 0, 0, 0};
    hi = on && k0 + 128 < HW ? *(p + 1) : v4i{0, 0, 0, 0};
  };
  int tiles = G - g_base;
  tiles = tiles <= 0 ? 0 : min(GT, (tiles + 15) / 16);
  const bool compute = dtiles > 0;
  const uint8_t* b_row = tile + row * kStride + kgroup * 16;
  v4i acc[DPW][GT];
#pragma unroll
  for (int d = 0; d < DPW; ++d)
#pragma unroll
    for (int t = 0; t < GT; ++t) acc[d][t] = v4i{0, 0, 0, 0};
  uint32_t gbuf[kSlots];
  v4i dlo[DPW], dhi[DPW];
#pragma unroll
  for (int i = 0; i < kSlots; ++i) gbuf[i] = load_gt(i, k_begin);
#pragma unroll
  for (int d = 0; d < DPW; ++d) load_det(d, k_begin, dlo[d], dhi[d]);
  for (i64 k0 = k_begin; k0 < k_end; k0 += KB) {
    __syncthreads();  // the previous block has been multiplied
#pragma unroll
    for (int i = 0; i < kSlots; ++i) {
      v4i lo, hi;
      expand_bits(gbuf[i], lo, hi);
      *reinterpret_cast<v4i*>(tile + glds[i]) = lo;
      *reinterpret_cast<v4i*>(tile + glds[i] + 16) = hi;
      garea[i] += __builtin_popcount(gbuf[i]);
    }
    // This block's detection words move to registers of their own, so the
    // next block's loads can be issued before the multiplications.
    uint32_t words[DPW][4];
#pragma unroll
    for (int d = 0; d < DPW; ++d) {
      const bool odd = kgroup >> 1;
      words[d][0] = static_cast<uint32_t>(odd ? dlo[d].y : dlo[d].x);
      words[d][1] = static_cast<uint32_t>(odd ? dlo[d].w : dlo[d].z);
      words[d][2] = static_cast<uint32_t>(odd ? dhi[d].y : dhi[d].x);
      words[d][3] = static_cast<uint32_t>(odd ? dhi[d].w : dhi[d].z);
    }
    __syncthreads();
    if (k0 + KB < k_end) {
#pragma unroll
      for (int i = 0; i < kSlots; ++i) gbuf[i] = load_gt(i, k0 + KB);
#pragma unroll
      for (int d = 0; d < DPW; ++d) load_det(d, k0 + KB, dlo[d], dhi[d]);
    }
    if (compute) {
#pragma unroll
      for (int step = 0; step < KB / 64; ++step) {
        v4i a[DPW];
#pragma unroll
        for (int d = 0; d < DPW; ++d) {
          const uint32_t half = (words[d][step] >> (16 * (kgroup & 1))) & 0xFFFFu;
          darea[d] += __builtin_popcount(half);
          uint32_t q[4];
#pragma unroll
          for (int n = 0; n < 4; ++n) q[n] = (((half >> (4 * n)) & 0xFu) * 0x00204081u) & 0x01010101u;
          a[d] = v4i{static_cast<int>(q[0]), static_cast<int>(q[1]), static_cast<int>(q[2]), static_cast<int>(q[3])};
        }
#pragma unroll
        for (int t = 0; t < GT; ++t) {
          if (t >= tiles) break;
          const v4i b = *reinterpret_cast<const v4i*>(b_row + t * 16 * kStride + step * 64);
#pragma unroll
          for (int d = 0; d < DPW; ++d)
            if (d < dtiles) acc[d][t] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[d], b, acc[d][t], 0, 0, 0);
        }
      }
    }
  }
  if (compute) {
#pragma unroll
    for (int d = 0; d < DPW; ++d) {
      if (d >= dtiles) break;
#pragma unroll
      for (int t = 0; t < GT; ++t) {
        if (t >= tiles) break;
        add_overlap_tile(acc[d][t], lane, d_base + (wave * DPW + d) * 16, g_base + t * 16, D, G, it.inter);
      }
    }
  }
  // Areas.  Detection rows: the four kgroup lanes of a row hold disjoint
  // quarters of its bits (counted only in waves that compute).  GT rows: the
  // eight neighbouring lanes of a slot share its row.
  if (it.pa != nullptr && it.ggroup == 0 && compute) {
#pragma unroll
    for (int d = 0; d < DPW; ++d) {
      int sum = darea[d];
      sum += __shfl_xor(sum, 16, kWave);
      sum += __shfl_xor(sum, 32, kWave);
      const int idx = d_base + (wave * DPW + d) * 16 + row;
      if (kgroup == 0 && d < dtiles && dvalid[d] && sum) atomicAdd(it.pa + idx, sum);
    }
  }
  const bool count_b = it.ga != nullptr && it.dgroup == 0;
#pragma unroll
  for (int i = 0; i < kSlots; ++i) {
    int sum = garea[i];
#pragma unroll
    for (int m = 1; m < kPerRow; m <<= 1) sum += __shfl_xor(sum, m, kWave);
    const int r = (tid + i * kThreads) / kPerRow;
    if (col == 0 && gvalid[i] && sum && count_b) atomicAdd(it.ga + g_base + r, sum);
  }
}

// Bit-packed masks on the FP4 matrix cores: gfx950's scaled f8f6f4 MFMAs
// (v_mfma_scale_f32_32x32x64_f8f6f4 for M = 32, ..._16x16x128_... for M = 16)
// with both operands in e2m1 and scales 1.0.  A mask bit becomes the e2m1
// value 1.0 (nibble 0b0010) or 0.  Lane l supplies 32 pixels of row l % M of
// each operand: nibble j of operand dword i holds bit 4j + i of the lane's
// mask word, so a word becomes its operand with one shift and one AND per
// dword.  Detection and GT operands use the same pixel-to-k map, so a product
// counts |det AND gt| over the lane group's pixels whatever order the hardware
// gives the k slots.  A work item's sums are integers of at most Chunk, which
// FP32 holds exactly.
// No LDS and no barriers: every wave is one work item (M * DT detections x
// M * GT GT masks over a pixel range of chunk_px, chosen at launch so that
// one round of waves covers the batch: each wave streams a long range, and the
// start-up latency of its first loads is paid once) and keeps DT x GT M x M
// tiles in accumulation registers.  A block is 4 * (64 / M) words of every row: lane
// group q = l / M loads words 4q .. 4q + 3 of its row as one 16-byte load, so
// a row's block is one contiguous run, and the k-step s of the block takes
// word s of every lane.  Ahead blocks are in flight.
// Result layout: M = 32, lane l, register r: row 8 (r / 4) + 4 (l / 32) + r % 4,
// column l % 32; M = 16: row 4 (l / 16) + r, column l % 16 (rows: detections,
// columns: GT masks).
typedef int v8i __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef float v16f __attribute__((ext_vector_type(16)));

__device__ __forceinline__ v8i fp4_operand(uint32_t w) {
  return v8i{static_cast<int>((w << 1) & 0x22222222u), static_cast<int>(w & 0x22222222u),
             static_cast<int>((w >> 1) & 0x22222222u), static_cast<int>((w >> 2) & 0x22222222u), 0, 0, 0, 0};
}

template <int M>
struct Fp4Tile;
template <>
struct Fp4Tile<32> {
  using Acc = v16f;
  static constexpr int kRegs = 16;
  __device__ static Acc mma(v8i a, v8i b, Acc c) {
    return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, b, c, 4, 4, 0, 127, 0, 127);
  }
  __device__ static int row(int lane, int r) { return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }
};
template <>
struct Fp4Tile<16> {
  using Acc = v4f;
  static constexpr int kRegs = 4;
  __device__ static Acc mma(v8i a, v8i b, Acc c) {
    return __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(a, b, c, 4, 4, 0, 127, 0, 127);
  }
  __device__ static int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

template <int M, int DT, int GT, int Ahead, int WG = 1>
__global__ __launch_bounds__(64 * WG) void mask_overlap_fp4(OverlapArgs o) {
  using Tile = Fp4Tile<M>;
  constexpr int kGroups = kWave / M;        // lane groups per row
  constexpr int kBlockWords = 4 * kGroups;  // words of a row per block
  constexpr int kRows = DT + GT;
  const i64 chunk_px = o.chunk_px;          // a multiple of 32 * kBlockWords (set at launch)
  OverlapItem it;
  if (!overlap_item(o, M * DT * WG, M * GT, chunk_px, it)) return;
  const int D = it.D, G = it.G;
  if (D == 0 && G == 0) return;  // nothing to count
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  const int row = lane % M, q = lane / M;
  // WG > 1: the waves of a workgroup take consecutive detection groups of one
  // pixel range and stay within Ahead blocks of each other (a barrier per
  // pass of the main loop), so the GT rows they share are read from HBM once
  // and found in the cache by the other waves.  A wave without detections
  // leaves at once (the barrier then waits only for the waves still running).
  const int d_base = (it.dgroup * WG + wave) * M * DT, g_base = it.ggroup * M * GT;
  if (wave > 0 && d_base >= D) return;
  const i64 row_words = mask_words(it.HW);  // a multiple of 4
  const i64 w_begin = static_cast<i64>(it.chunk) * (chunk_px / 32);
  const i64 w_end = min(w_begin + chunk_px / 32, row_words);
  const i64 w_full = w_begin + (w_end - w_begin) / kBlockWords * kBlockWords;  // end of the whole blocks
  // The rows of this wave's groups that can have pixels in [w_begin, w_end):
  // all of them, or with mask extents only those whose range meets it, listed
  // in LDS in row order.  Tiles are built from the lists, so rows that are
  // zero over the range are neither loaded nor multiplied.
  __shared__ int s_rows[WG][M * DT + M * GT];
  int* dlist = s_rows[wave];
  int* glist = dlist + M * DT;
  auto list_rows = [&](int* list, int first, int count, const int32_t* ext) {
    int n = 0;
    for (int i0 = 0; i0 < count; i0 += kWave) {
      const int i = i0 + lane;
      const bool keep =
          i < count && (!ext || (ext[2 * (first + i)] < w_end && ext[2 * (first + i) + 1] > w_begin));
      const u64 votes = __ballot(keep);
      if (keep) list[n + __popcll(votes & ((1ULL << lane) - 1))] = first + i;
      n += __popcll(votes);
    }
    return n;
  };
  const int nd = list_rows(dlist, d_base, max(0, min(M * DT, D - d_base)), o.pm_ext ? o.pm_ext[it.b] : nullptr);
  const int ng = list_rows(glist, g_base, max(0, min(M * GT, G - g_base)), o.gm_ext ? o.gm_ext[it.b] : nullptr);
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  const bool count_d = it.pa != nullptr && it.ggroup == 0;
  const bool count_g = it.ga != nullptr && it.dgroup == 0 && wave == 0;
  // Uniform: leave when the range holds no products and no areas to count.
  if (!((nd > 0 && ng > 0) || (nd > 0 && count_d) || (ng > 0 && count_g))) return;
  // Tiles holding at least one listed row (uniform per wave; 0 on a side
  // with none, whose item still counts the other side's areas).
  const int dtiles = (nd + M - 1) / M, gtiles = (ng + M - 1) / M;
  // Rows 0 .. DT-1: detection tiles, DT .. DT+GT-1: GT tiles.  Every load is
  // unconditional, so the compiler can keep Ahead blocks in flight: a tile
  // slot past its list reads the list's last row (or, for an empty list, the
  // other list's first row), and the results of such slots are never stored
  // (the flush and the areas check `valid`).
  const __attribute__((address_space(1))) uint32_t* src[kRows];
  bool valid[kRows];
#pragma unroll
  for (int j = 0; j < kRows; ++j) {
    const bool det = j < DT;
    const int li = det ? M * j + row : M * (j - DT) + row;
    const int n = det ? nd : ng;
    valid[j] = li < n;
    const int* own = det ? dlist : glist;
    const int* other = det ? glist : dlist;
    const int idx = n > 0 ? own[min(li, n - 1)] : other[0];
    const uint8_t* base = n > 0 ? (det ? it.pm : it.gm) : (det ? it.gm : it.pm);
    src[j] = as_global(reinterpret_cast<const uint32_t*>(base)) + static_cast<i64>(idx) * row_words;
  }
  // This lane's 4 words of the block at word wb: the piece index is clamped
  // into the row, so blocks fetched ahead past the end stay in bounds (they
  // are never used); `masked` zeroes pieces past w_end (the row's last,
  // partial block).
  auto load = [&](i64 wb, v4i* piece, bool masked) {
    const i64 w = wb + 4 * q;
    const bool inside = w < w_end;
    const i64 at = min(w, row_words - 4);
#pragma unroll
    for (int j = 0; j < kRows; ++j) {
      const v4i v = *reinterpret_cast<const __attribute__((address_space(1))) v4i*>(src[j] + at);
      piece[j] = !masked || inside ? v : v4i{0, 0, 0, 0};
    }
  };
  typename Tile::Acc acc[DT][GT];
#pragma unroll
  for (int d = 0; d < DT; ++d)
#pragma unroll
    for (int t = 0; t < GT; ++t) acc[d][t] = typename Tile::Acc{};
  int area[kRows];
#pragma unroll
  for (int j = 0; j < kRows; ++j) area[j] = 0;
  auto consume = [&](const v4i* piece) {
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      v8i op[kRows];
#pragma unroll
      for (int j = 0; j < kRows; ++j) {
        const uint32_t w = static_cast<uint32_t>(piece[j][s]);
        area[j] += __builtin_popcount(w);
        op[j] = fp4_operand(w);
      }
#pragma unroll
      for (int d = 0; d < DT; ++d) {
        if (d >= dtiles) break;
#pragma unroll
        for (int t = 0; t < GT; ++t) {
          if (t >= gtiles) break;
          acc[d][t] = Tile::mma(op[d], op[DT + t], acc[d][t]);
        }
      }
    }
  };
  // Every pass of the main loop consumes and refills exactly Ahead blocks (no
  // early exit inside it), so the compiler's count of loads in flight stays
  // exact and each block waits only for its own loads.  The blocks left after
  // it were fetched by its last pass.
  v4i buf[Ahead][kRows];
#pragma unroll
  for (int a = 0; a < Ahead; ++a) load(w_begin + a * kBlockWords, buf[a], false);
  i64 w0 = w_begin;
  for (; w0 + Ahead * kBlockWords <= w_full; w0 += Ahead * kBlockWords) {
    if constexpr (WG > 1) __builtin_amdgcn_s_barrier();  // progress only: no memory is shared
#pragma unroll
    for (int a = 0; a < Ahead; ++a) {
      consume(buf[a]);
      load(w0 + (a + Ahead) * kBlockWords, buf[a], false);
    }
  }
#pragma unroll
  for (int a = 0; a < Ahead - 1; ++a)
    if (w0 + a * kBlockWords < w_full) consume(buf[a]);
  if (w_full < w_end) {  // the row's last block, partial
    v4i piece[kRows];
    load(w_full, piece, true);
    consume(piece);
  }
#pragma unroll
  for (int d = 0; d < DT; ++d) {
    if (d >= dtiles) break;
#pragma unroll
    for (int t = 0; t < GT; ++t) {
      if (t >= gtiles) break;
      const int lg = M * t + row;
      if (lg >= ng) continue;
      const int g = glist[lg];
#pragma unroll
      for (int r = 0; r < Tile::kRegs; ++r) {
        const int ld = M * d + Tile::row(lane, r);
        const int v = static_cast<int>(acc[d][t][r]);
        if (ld < nd && v) atomicAdd(it.inter + static_cast<i64>(dlist[ld]) * G + g, v);
      }
    }
  }
  // Areas: the kGroups lanes of a row hold disjoint words of it (rows left
  // out of the lists are zero over the range).
#pragma unroll
  for (int j = 0; j < kRows; ++j) {
    const bool det = j < DT;
    int sum = area[j];
#pragma unroll
    for (int m = M; m < kWave; m <<= 1) sum += __shfl_xor(sum, m, kWave);
    if (q == 0 && valid[j] && sum && (det ? count_d : count_g)) {
      const int idx = det ? dlist[M * j + row] : glist[M * (j - DT) + row];
      atomicAdd(det ? it.pa + idx : it.ga + idx, sum);
    }
  }
}

struct BitsVariant {
  int dets, gts;
  i64 chunk_px;  // 0: sized at launch (one round of as many waves as fit on the device)
  int threads;
  void (*kernel)(OverlapArgs);
  int block_px = 0;  // chunk granularity of the launch-sized kernels
};
// Order matches the tuning registry's descriptions.
const BitsVariant kOverlapBits[] = {
    {128, 64, 4096, 512, mask_overlap_staged<8, 4, 256, 4096, true, 1>},
    {128, 64, 4096, 512, mask_overlap_staged<8, 4, 256, 4096, true, 4>},
    {128, 64, 8192, 512, mask_overlap_staged<8, 4, 256, 8192, true, 4>},
    {128, 64, 16384, 512, mask_overlap_staged<8, 4, 256, 16384, true, 8>},
    {128, 64, 4096, 256, mask_overlap_staged<8, 4, 256, 4096, true, 1, 2>},
    {128, 64, 16384, 256, mask_overlap_bits<8, 4, 32, 512>},
    {64, 64, 16384, 256, mask_overlap_bits<4, 4, 32, 512>},
    {128, 64, 4096, 256, mask_overlap_bits_reg<8, 4, 4096, 2>},
    {128, 64, 4096, 128, mask_overlap_bits_reg<8, 4, 4096, 4>},
    {128, 64, 8192, 256, mask_overlap_bits_reg<8, 4, 8192, 2>},
    {64, 64, 4096, 64, mask_overlap_fp4<16, 4, 4, 2>, 512},
    {64, 64, 0, 64, mask_overlap_fp4<16, 4, 4, 2>, 512},
    {64, 64, 0, 64, mask_overlap_fp4<32, 2, 2, 4>, 256},
    {64, 32, 0, 64, mask_overlap_fp4<16, 4, 2, 4>, 512},
    {64, 64, 0, 64, mask_overlap_fp4<32, 2, 2, 1>, 256},
    {128, 64, 0, 64, mask_overlap_fp4<16, 8, 4, 2>, 512},
    {128, 64, 0, 64, mask_overlap_fp4<16, 8, 4, 1>, 512},
    {128, 64, 0, 64, mask_overlap_fp4<32, 4, 2, 2>, 256},
    {128, 64, 0, 128, mask_overlap_fp4<16, 4, 4, 2, 2>, 512},
    {128, 64, 0, 128, mask_overlap_fp4<16, 4, 4, 3, 2>, 512},
};
constexpr int kOverlapBitsCount = sizeof(kOverlapBits) / sizeof(kOverlapBits[0]);

// Workgroups of variant v resident per compute unit (HIP's occupancy query,
// cached per device and variant).
int resident_per_cu(int v) {
  static std::atomic<int> cached[64][kOverlapBitsCount] = {};
  int device = 0;
  if (hipGetDevice(&device) != hipSuccess || device < 0 || device >= 64) device = 0;
  int n = cached[device][v].load(std::memory_order_relaxed);
  if (n <= 0) {
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kOverlapBits[v].kernel, kOverlapBits[v].threads, 0) !=
            hipSuccess ||
        n <= 0)
      n = 8;
    cached[device][v].store(n, std::memory_order_relaxed);
  }
  return n;
}

int launch_overlap_bits(OverlapArgs o, int64_t B, int64_t max_d, int64_t max_g, int64_t max_hw, int64_t total_d,
                        int64_t total_g, int64_t total_inter, hipStream_t stream) {
  int rc;
  if ((rc = launch_zero3(reinterpret_cast<uint32_t*>(o.inter), total_inter, reinterpret_cast<uint32_t*>(o.pa),
                         total_d, reinterpret_cast<uint32_t*>(o.ga), total_g, stream)))
    return rc;
  if (B == 0 || max_hw == 0 || (max_d == 0 && max_g == 0)) return kOk;
  const int variant = tuned_variant(kTuneMaskOverlapBits);
  const BitsVariant& v = kOverlapBits[variant];
  const i64 dgroups = std::max<i64>(1, (max_d + v.dets - 1) / v.dets);
  const i64 ggroups = std::max<i64>(1, (max_g + v.gts - 1) / v.gts);
  i64 chunk_px = v.chunk_px;
  if (chunk_px == 0) {
    // One round of workgroups: as many as the device holds at once, each with
    // an equal share of the batch's (image, group, pixel) work, at least four
    // blocks long.
    const i64 waves = static_cast<i64>(resident_per_cu(variant)) * compute_units();
    const i64 share = (B * dgroups * ggroups * max_hw + waves - 1) / waves;
    chunk_px = std::max<i64>(4 * v.block_px, (share + v.block_px - 1) / v.block_px * v.block_px);
    // FP32 accumulation is exact while every sum stays below 2^24: a work
    // item covers at most 2^22 pixels.
    chunk_px = std::min<i64>(chunk_px, i64{1} << 22);
  }
  const i64 chunks = (max_hw + chunk_px - 1) / chunk_px;
  const i64 items = dgroups * ggroups * B * chunks;
  if (items >= INT_MAX - kXcds) return fail(kUnsupportedShape, "too many masks or pixels for one overlap launch");
  o.chunks = static_cast<int>(chunks);
  o.dgroups = static_cast<int>(dgroups);
  o.ggroups = static_cast<int>(ggroups);
  o.items = static_cast<int>(items);
  o.xcd_span = static_cast<int>((items + kXcds - 1) / kXcds);
  o.chunk_px = chunk_px;
  ProfileScope timing("mask_overlap_bits", stream);
  v.kernel<<<static_cast<unsigned>(static_cast<i64>(o.xcd_span) * kXcds), v.threads, 0, stream>>>(o);
  PE_LAUNCH_CHECK();
  return kOk;
}

// ------------------------------------------------- detection x slot counts

constexpr int kDetThreads = 256;
constexpr size_t kDetMaxLds = 96 * 1024;

struct DetHistArgs {
  const uint8_t* const* masks;   // B device pointers (D_b x pixels_b bools, or bit rows)
  const i64* det_off;            // B + 1
  const i64* order;              // per detection: mask index within its image
  const int32_t* target;         // slots of every image, one after another
  int B;
  int S;
  TileList tiles;                // pixel tiles, each inside one image
  int dgroups;                   // detection groups per image
  int group;                     // detections per group
  u64* pairs;                    // (sum D) x S
  int aligned;
  // Optional, bit-packed masks only: per image, 2 int32 per mask (in mask
  // order): a word range [first, end) outside which the mask row is zero.
  const int32_t* const* extents;
};

// Pixel-by-pixel count of a chunk given as bits (bit j = pixel base + j),
// for the rare chunks with more distinct slots than SlotGroups keeps: one
// LDS add per set pixel, each slot read from the (cached) map when needed,
// so no per-pixel arrays stay live in registers.
template <typename TargetPtr>
__device__ __forceinline__ void count_pixels(uint32_t on, TargetPtr target, i64 base, i64 HW, int S, int rowbase,
                                             uint32_t* bins) {
  // Eight slots are fetched at once (indices clamped into the image), so a
  // chunk costs a few load latencies rather than one per set pixel.
  for (int j0 = 0; j0 < 32 && (on >> j0); j0 += 8) {
    int s[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) s[j] = target[min(base + j0 + j, HW - 1)];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int slot = s[j];
      if (!((on >> (j0 + j)) & 1u) || base + j0 + j >= HW || slot < 0 || slot >= S) continue;
      const int key = rowbase + slot;
      atomicAdd(bins + (key >> 1), 1u << ((key & 1) * 16));
    }
  }
}

// The distinct valid target slots of an N-pixel chunk (N <= 32) with the
// pixels each covers (bit j = pixel j), found once per tile.  Voronoi-like
// maps give one segment plus scattered void pixels, sometimes a second
// segment or two at a boundary: up to kGroups slots are kept; a chunk with more is
// `complex` and is counted pixel by pixel.
constexpr int kGroups = 5;
struct SlotGroups {
  int slot[kGroups];
  uint32_t mask[kGroups];
  bool complex;
};

// Adds slot s at pixel j of a chunk to its groups.
__device__ __forceinline__ void group_insert(SlotGroups& g, int s, int j, int S) {
  if (s < 0 || s >= S) return;
  bool placed = false;
#pragma unroll
  for (int k = 0; k < kGroups; ++k) {
    if (placed) continue;
    if (g.mask[k] == 0) {
      g.slot[k] = s;
      g.mask[k] = 1u << j;
      placed = true;
    } else if (g.slot[k] == s) {
      g.mask[k] |= 1u << j;
      placed = true;
    }
  }
  g.complex |= !placed;
}

// Slot groups of the N-pixel chunk starting at pixel `base` of the map.  A
// chunk inside a 16-byte aligned map is read as N / 4 vectors from one
// address, all in flight together; a chunk at the image's end (at most one
// lane per image) is read pixel by pixel.
template <int N, typename TargetPtr>
__device__ __forceinline__ SlotGroups chunk_groups(TargetPtr target, i64 base, i64 HW, bool vec, int S) {
  SlotGroups g;
#pragma unroll
  for (int k = 0; k < kGroups; ++k) {
    g.slot[k] = 0;
    g.mask[k] = 0;
  }
  g.complex = false;
  if (vec && base + N <= HW) {
    const auto* v = reinterpret_cast<const __attribute__((address_space(1))) v4i*>(target + base);
    v4i raw[N / 4];
#pragma unroll
    for (int q = 0; q < N / 4; ++q) raw[q] = v[q];
#pragma unroll
    for (int q = 0; q < N / 4; ++q) {
      group_insert(g, raw[q].x, 4 * q, S);
      group_insert(g, raw[q].y, 4 * q + 1, S);
      group_insert(g, raw[q].z, 4 * q + 2, S);
      group_insert(g, raw[q].w, 4 * q + 3, S);
    }
  } else {
#pragma unroll 1
    for (int j = 0; j < N; ++j)
      if (base + j < HW) group_insert(g, target[base + j], j, S);
  }
  return g;
}

// One lane's count of a detection's N-pixel chunk (bit j = pixel j set): one
// popcount and at most one LDS add per distinct slot of the chunk.  A complex
// chunk re-reads its target slots (cached: the tile's map was just read).
template <typename TargetPtr>
__device__ __forceinline__ void count_lane(uint32_t on, const SlotGroups& g, TargetPtr target, i64 base, i64 HW,
                                           int S, int rowbase, uint32_t* bins) {
  if (g.complex) {
    count_pixels(on, target, base, HW, S, rowbase, bins);
    return;
  }
#pragma unroll
  for (int k = 0; k < kGroups; ++k) {
    const uint32_t c = __popc(on & g.mask[k]);
    if (c) {
      const int key = rowbase + g.slot[k];
      atomicAdd(bins + (key >> 1), c << ((key & 1) * 16));
    }
  }
}

__device__ __forceinline__ bool any_count(const uint32_t* c) {
  uint32_t any = 0;
#pragma unroll
  for (int k = 0; k < kGroups; ++k) any |= c[k];
  return any != 0;
}

// The same count for every lane of the wave together (`on` = 0 for lanes
// with nothing to add).  The lanes of a wave cover neighbouring pixels, so
// most add to the same (detection, slot) bin, and lane-by-lane adds would
// serialise on that LDS address.  Instead each distinct key is summed across
// the wave (DPP reduction) and added once; a wave holds few distinct keys
// (the segments under the mask plus void).
template <typename TargetPtr>
__device__ __forceinline__ void count_wave(uint32_t on, const SlotGroups& g, TargetPtr target, i64 base, i64 HW,
                                           int S, int rowbase, uint32_t* bins) {
  uint32_t c[kGroups];
  int key[kGroups];
#pragma unroll
  for (int k = 0; k < kGroups; ++k) {
    c[k] = g.complex ? 0u : static_cast<uint32_t>(__popc(on & g.mask[k]));
    key[k] = rowbase + g.slot[k];
  }
  if (g.complex && on) count_pixels(on, target, base, HW, S, rowbase, bins);
  u64 pending = __ballot(any_count(c));
  while (pending) {
    const int leader = __builtin_ctzll(pending);
    int mine = key[kGroups - 1];
#pragma unroll
    for (int k = kGroups - 2; k >= 0; --k) mine = c[k] ? key[k] : mine;
    const int k_u = __builtin_amdgcn_readlane(mine, leader);
    uint32_t contrib = 0;
#pragma unroll
    for (int k = 0; k < kGroups; ++k)
      if (key[k] == k_u) {
        contrib += c[k];
        c[k] = 0;
      }
    const uint32_t total = __reduce_add_sync(~0ULL, contrib);
    if (lane_id() == leader) atomicAdd(bins + (k_u >> 1), total << ((k_u & 1) * 16));
    pending = __ballot(any_count(c));
  }
}

// 16 mask bytes (bool: each 0 or 1) as 16 bits: in each 32-bit word the product
// with 0x01020408 moves byte i's bit to bit 24 + i and leaves no carry into
// bits 24..27.
__device__ __forceinline__ uint32_t byte_bits(v4i m) {
  const uint32_t w0 = (static_cast<uint32_t>(m.x) * 0x01020408u) >> 24;
  const uint32_t w1 = (static_cast<uint32_t>(m.y) * 0x01020408u) >> 24;
  const uint32_t w2 = (static_cast<uint32_t>(m.z) * 0x01020408u) >> 24;
  const uint32_t w3 = (static_cast<uint32_t>(m.w) * 0x01020408u) >> 24;
  return (w0 & 0xFu) | ((w1 & 0xFu) << 4) | ((w2 & 0xFu) << 8) | ((w3 & 0xFu) << 12);
}

// Masks are read exactly once: stream them past the caches.
__device__ __forceinline__ v4i load16_stream(const uint8_t* row, i64 k, i64 HW, bool aligned) {
  const auto* g = as_global(row);
  if (aligned && k + 16 <= HW) return __builtin_nontemporal_load(as_global(reinterpret_cast<const v4i*>(row + k)));
  v4i v = {0, 0, 0, 0};
  if (k >= HW) return v;
  uint8_t bytes[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) bytes[j] = k + j < HW ? g[k + j] : 0;
  __builtin_memcpy(&v, bytes, 16);
  return v;
}

// uint4 units of the packed 16-bit (detection, slot) bins of a group.
__host__ __device__ constexpr i64 det_bins4(i64 group, i64 S) { return ((group * S + 1) / 2 + 3) / 4; }

// Shared set-up of one (tile, detection group) item: zeroed bins and the
// group's mask row offsets (element units) staged in LDS behind them, so the
// row loop reads them from LDS instead of waiting on order[] before every
// batch of mask loads.  Returns false when the group is empty.
struct DetItem {
  int b, tile, nd;
  i64 first, d0, HW;
  uint32_t* bins;
  i64* row_off;
  int words;
};

__device__ __forceinline__ bool begin_item(const DetHistArgs& a, int item, uint4* lds4, bool bits, DetItem& it) {
  const int tile_item = item / a.dgroups, dgroup = item - tile_item * a.dgroups;
  it.b = __builtin_amdgcn_readfirstlane(a.tiles.image_of(tile_item));
  it.tile = tile_item - a.tiles.first_tile(it.b);
  it.first = a.det_off[it.b];
  const int D = static_cast<int>(a.det_off[it.b + 1] - it.first);
  it.d0 = static_cast<i64>(dgroup) * a.group;
  it.nd = min(a.group, D - static_cast<int>(it.d0));
  if (it.nd <= 0) return false;  // uniform across the workgroup
  it.HW = a.tiles.pixels(it.b);
  it.bins = reinterpret_cast<uint32_t*>(lds4);
  it.row_off = reinterpret_cast<i64*>(lds4 + det_bins4(a.group, a.S));
  it.words = (it.nd * a.S + 1) / 2;
  const int words4 = (it.words + 3) / 4;
  const i64 stride = bits ? mask_words(it.HW) : it.HW;
  for (int w = threadIdx.x; w < words4; w += kDetThreads) lds4[w] = make_uint4(0, 0, 0, 0);
  for (int d = threadIdx.x; d < it.nd; d += kDetThreads)
    it.row_off[d] = (a.order ? a.order[it.first + it.d0 + d] : it.d0 + d) * stride;
  __syncthreads();
  return true;
}

// Flush the item's nonzero 16-bit bins into the (detection, slot) table.
__device__ __forceinline__ void end_item(const DetHistArgs& a, const DetItem& it) {
  __syncthreads();
  u64* out = a.pairs + (it.first + it.d0) * a.S;
  for (int w = threadIdx.x; w < it.words; w += kDetThreads) {
    const uint32_t v = it.bins[w];
    if (!v) continue;
    const uint32_t lo = v & 0xFFFFu, hi = v >> 16;
    if (lo) atomicAdd(out + 2 * w, static_cast<u64>(lo));
    if (hi) atomicAdd(out + 2 * w + 1, static_cast<u64>(hi));
  }
}

// Bool masks.  CH = 16-pixel chunks per lane (tile = 16 * 256 * CH pixels);
// U = masks whose loads are in flight together.  Every load is unconditional,
// with its index clamped into the image or group and the value masked after:
// a load behind a branch gets its own wait, which would serialise the loads.
template <int CH, int U>
__device__ __forceinline__ void detection_item_bytes(const DetHistArgs& a, int item, uint4* lds4) {
  constexpr int kTile = 16 * kDetThreads * CH;
  DetItem it;
  if (!begin_item(a, item, lds4, false, it)) return;
  const int nd = it.nd;
  const i64 HW = it.HW;
  const i64 p0 = static_cast<i64>(it.tile) * kTile;
  const auto* target = as_global(a.target + a.tiles.begin(it.b));
  const bool vec = (reinterpret_cast<uintptr_t>(target) & 15) == 0;  // base is a multiple of 16 pixels
  const uint8_t* masks = a.masks[it.b];
  i64 base[CH];
  bool inside[CH];
  SlotGroups groups[CH];
#pragma unroll
  for (int h = 0; h < CH; ++h) {
    base[h] = p0 + (static_cast<i64>(h) * kDetThreads + threadIdx.x) * 16;
    inside[h] = base[h] < HW;
    groups[h] = chunk_groups<16>(target, base[h], HW, vec, a.S);
  }
  if (a.aligned) {
    // 16-byte aligned rows whose pixel counts are multiples of 16: a lane's
    // chunk inside the image is one whole aligned vector.
    i64 offset[CH];
#pragma unroll
    for (int h = 0; h < CH; ++h) offset[h] = inside[h] ? base[h] : 0;
    for (int r0 = 0; r0 < nd; r0 += U) {
      v4i raw[U][CH];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const uint8_t* mask = masks + it.row_off[min(r0 + u, nd - 1)];
#pragma unroll
        for (int h = 0; h < CH; ++h)
          raw[u][h] = __builtin_nontemporal_load(as_global(reinterpret_cast<const v4i*>(mask + offset[h])));
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int h = 0; h < CH; ++h) {
          const v4i& v = raw[u][h];
          if (r0 + u < nd && inside[h] && (v.x | v.y | v.z | v.w) != 0)
            count_lane(byte_bits(v), groups[h], target, base[h], HW, a.S, (r0 + u) * a.S, it.bins);
        }
    }
  } else {
    for (int r0 = 0; r0 < nd; r0 += U) {
      v4i m[U][CH];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (r0 + u < nd) {
          const uint8_t* mask = masks + it.row_off[r0 + u];
#pragma unroll
          for (int h = 0; h < CH; ++h) m[u][h] = load16_stream(mask, base[h], HW, false);
        } else {
#pragma unroll
          for (int h = 0; h < CH; ++h) m[u][h] = v4i{0, 0, 0, 0};
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int h = 0; h < CH; ++h)
          if ((m[u][h].x | m[u][h].y | m[u][h].z | m[u][h].w) != 0)
            count_lane(byte_bits(m[u][h]), groups[h], target, base[h], HW, a.S, (r0 + u) * a.S, it.bins);
    }
  }
  end_item(a, it);
}

// Bit-packed masks (rows of mask_words(pixels) 32-bit words, bit p % 32 of
// word p / 32 = pixel p).  A lane owns W consecutive words (32 * W pixels,
// tile = 32 * W * 256 pixels) and reads them with one vector load per mask,
// so a workgroup keeps U * W * 1 KB of mask bits in flight.  AGG: sum each
// (detection, slot) key across the wave before the LDS add.
template <int W>
struct WordsT;
template <>
struct WordsT<1> {
  using T = uint32_t;
};
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
template <>
struct WordsT<2> {
  using T = u32x2;
};
template <>
struct WordsT<4> {
  using T = u32x4;
};
template <int W>
__device__ __forceinline__ uint32_t word_of(const typename WordsT<W>::T& v, int w) {
  if constexpr (W == 1) {
    return v;
  } else if constexpr (W == 2) {
    return w == 0 ? v.x : v.y;
  } else {
    return w == 0 ? v.x : w == 1 ? v.y : w == 2 ? v.z : v.w;
  }
}

template <int W, int U, bool AGG>
__device__ __forceinline__ void detection_item_bits(const DetHistArgs& a, int item, uint4* lds4) {
  constexpr int kTile = 32 * W * kDetThreads;
  using Vec = typename WordsT<W>::T;
  DetItem it;
  if (!begin_item(a, item, lds4, true, it)) return;
  const int nd = it.nd;
  const i64 HW = it.HW;
  const i64 lane_px = static_cast<i64>(it.tile) * kTile + static_cast<i64>(threadIdx.x) * 32 * W;
  const auto* target = as_global(a.target + a.tiles.begin(it.b));
  const bool vec = (reinterpret_cast<uintptr_t>(target) & 15) == 0;  // lane_px is a multiple of 32
  bool inside[W];
  SlotGroups groups[W];
#pragma unroll
  for (int w = 0; w < W; ++w) {
    inside[w] = lane_px + 32 * w < HW;
    groups[w] = chunk_groups<32>(target, lane_px + 32 * w, HW, vec, a.S);
  }
  // Rows are padded to whole 16-byte units, so the W words of a lane whose
  // first pixel is inside the image are inside the row.
  const i64 word0 = inside[0] ? lane_px >> 5 : 0;
  const auto* rows = as_global(reinterpret_cast<const uint32_t*>(a.masks[it.b]));
  // With mask extents, only the detections whose nonzero words can meet this
  // tile are read: their indices are listed in LDS (in any order: the counts
  // are integer adds) after the row offsets; list[group] is the count.
  int* list = reinterpret_cast<int*>(it.row_off + a.group);
  int m = nd;
  if (a.extents) {
    const int32_t* ext = a.extents[it.b];
    const i64 tw0 = static_cast<i64>(it.tile) * (W * kDetThreads), tw1 = tw0 + W * kDetThreads;
    if (threadIdx.x == 0) list[a.group] = 0;
    __syncthreads();
    for (int d = threadIdx.x; d < nd; d += kDetThreads) {
      const i64 mask = a.order ? a.order[it.first + it.d0 + d] : it.d0 + d;
      if (ext[2 * mask] < tw1 && ext[2 * mask + 1] > tw0) list[atomicAdd(list + a.group, 1)] = d;
    }
    __syncthreads();
    m = list[a.group];
  }
  for (int r0 = 0; r0 < m; r0 += U) {
    Vec raw[U];
    int det[U];
#pragma unroll
    for (int u = 0; u < U; ++u) det[u] = a.extents ? list[min(r0 + u, m - 1)] : min(r0 + u, nd - 1);
#pragma unroll
    for (int u = 0; u < U; ++u)
      raw[u] = __builtin_nontemporal_load(reinterpret_cast<const __attribute__((address_space(1))) Vec*>(
          rows + it.row_off[det[u]] + word0));
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int w = 0; w < W; ++w) {
        const uint32_t on = r0 + u < m && inside[w] ? word_of<W>(raw[u], w) : 0u;
        if constexpr (AGG) {
          count_wave(on, groups[w], target, lane_px + 32 * w, HW, a.S, det[u] * a.S, it.bins);
        } else {
          if (on) count_lane(on, groups[w], target, lane_px + 32 * w, HW, a.S, det[u] * a.S, it.bins);
        }
      }
  }
  end_item(a, it);
}

// One workgroup per item (grid >= item count; surplus workgroups of a packed
// batch, whose count is only known on the device, exit at once), or a
// persistent grid that walks the items with a grid stride.
template <int CH, int U, bool PERSIST>
__global__ __launch_bounds__(kDetThreads) void detection_histogram(DetHistArgs a) {
  extern __shared__ uint4 lds4[];
  const int total = a.tiles.count(a.B) * a.dgroups;
  if (!PERSIST) {
    if (static_cast<int>(blockIdx.x) < total) detection_item_bytes<CH, U>(a, blockIdx.x, lds4);
    return;
  }
  for (int item = blockIdx.x; item < total; item += gridDim.x) {
    detection_item_bytes<CH, U>(a, item, lds4);
    __syncthreads();  // the next item re-zeroes the bins
  }
}

template <int W, int U, bool AGG, bool PERSIST>
__global__ __launch_bounds__(kDetThreads) void detection_histogram_bits(DetHistArgs a) {
  extern __shared__ uint4 lds4[];
  const int total = a.tiles.count(a.B) * a.dgroups;
  if (!PERSIST) {
    if (static_cast<int>(blockIdx.x) < total) detection_item_bits<W, U, AGG>(a, blockIdx.x, lds4);
    return;
  }
  for (int item = blockIdx.x; item < total; item += gridDim.x) {
    detection_item_bits<W, U, AGG>(a, item, lds4);
    __syncthreads();
  }
}

struct DetVariant {
  int tile;                            // pixels per tile
  void (*kernel)(DetHistArgs);
  void (*persistent)(DetHistArgs);
};
#define PE_DET_VARIANT(CH, U) \
  { 16 * kDetThreads * CH, detection_histogram<CH, U, false>, detection_histogram<CH, U, true> }
#define PE_DET_BITS(W, U, AGG) \
  { 32 * W * kDetThreads, detection_histogram_bits<W, U, AGG, false>, detection_histogram_bits<W, U, AGG, true> }
// Order matches the tuning registry's descriptions.
const DetVariant kDetHist[] = {
    PE_DET_VARIANT(1, 4), PE_DET_VARIANT(1, 8), PE_DET_VARIANT(2, 2), PE_DET_VARIANT(2, 4), PE_DET_VARIANT(4, 2),
};
const DetVariant kDetHistBits[] = {
    PE_DET_BITS(2, 8, false), PE_DET_BITS(2, 16, false), PE_DET_BITS(1, 16, false), PE_DET_BITS(4, 4, false),
    PE_DET_BITS(4, 8, false), PE_DET_BITS(1, 8, false), PE_DET_BITS(2, 8, true),
};
#undef PE_DET_VARIANT
#undef PE_DET_BITS

// ------------------------------------------------------------ COCO matching

struct MatchArgs {
  int B, C;
  const i64* det_off;
  const i64* gt_off;
  const i64* inter_off;
  const int32_t* det_cls;     // score order
  const double* det_area;     // pixel areas, score order
  const i64* det_row;         // row of the intersection block for a detection (NULL: identity)
  const int32_t* gt_cls;      // -1: not a candidate
  const double* gt_area;      // pixel areas (IoU)
  const uint8_t* gt_crowd;
  const double* gt_ann;       // annotation areas (area ranges)
  const void* inter;          // per image D_b x G_b intersections
  int inter_bytes;            // 4 or 8
  int8_t* status;             // (sum D) x 40
  int32_t* ranks;             // sum D
  i64* counts;                // C x 4 (accumulated)
  int max_gt;
  int words;                  // used-bitmask words per chain
  const int* guard;           // skip everything when *guard != 0
  const uint8_t* gt_things;   // optional: only GT of thing classes are candidates
  const i64* gt_area_i64;     // optional: int64 pixel areas at b * gt_area_stride + j
  i64 gt_area_stride;
  int zero_inter;             // zero every detection's intersection row after matching
  // Grouped form (coco_match_grouped): per-image class lists made beforehand
  // by match_groups; GT are the S target slots of each image and the rows
  // are int64 rows of S columns (image b's block starts at row det_off[b]).
  const int4* group_hdr;      // B x (C + 1): det start, det count, gt start, gt count
  const int2* group_det;      // sum D: (score-order index, row) grouped by class
  const int32_t* group_gt;    // B x S: slot indices grouped by class
  int S;
};

// GT candidate class (or -1), pixel area and annotation area of GT j of image b.
__device__ __forceinline__ int gt_class(const MatchArgs& a, i64 g) {
  const int c = a.gt_cls[g];
  if (a.gt_things && (c < 0 || c >= a.C || !a.gt_things[c])) return -1;
  return c;
}

__device__ __forceinline__ double gt_pixel_area(const MatchArgs& a, int b, i64 g0, int j) {
  return a.gt_area_i64 ? static_cast<double>(a.gt_area_i64[b * a.gt_area_stride + j]) : a.gt_area[g0 + j];
}

__device__ __forceinline__ double gt_ann_area(const MatchArgs& a, int b, i64 g0, int j) {
  return a.gt_ann ? a.gt_ann[g0 + j] : gt_pixel_area(a, b, g0, j);
}

__device__ __forceinline__ void area_range(int a, double& lo, double& hi) {
  lo = a == 2 ? 1024.0 : a == 3 ? 9216.0 : 0.0;
  hi = a == 1 ? 1024.0 : a == 2 ? 9216.0 : 1e10;
}

__device__ __forceinline__ double iou_threshold(int t) {
  // COCO's np.linspace(.5, .95, 10) values (instance.py:252), incl. .8999999999999999
  const double table[10] = {0.5, 0.55, 0.6, 0.65, 0.7, 0.75, 0.8, 0.85, 0.8999999999999999, 0.95};
  return table[t];
}

// Matching state staged in LDS before the chains run (when it fits): the
// intersections of every (ranked detection, candidate GT) pair and the GT
// candidates' crowd flags and areas.  The chains are serial over detections;
// staged, each step reads LDS instead of waiting on a global load.
constexpr int kMatchStagePairs = 512;
constexpr int kMatchStageGts = 128;
constexpr size_t kMatchStageBytes = 8 * kMatchStagePairs + 17 * kMatchStageGts;

__device__ __forceinline__ void match_chains(const MatchArgs& a, int b, i64 d0, i64 g0, int G, i64 ib, int nd,
                                             int ng, bool staged, const int* dets, const i64* det_rw,
                                             const double* det_pa, const double* pair_inter, const double* gt_px,
                                             const double* gt_an, const uint8_t* gt_cr, const int* gts,
                                             uint32_t* used);

// Zero the int64 intersection rows of image b's detections of class c (the
// class-0 wave also takes detections of no valid class), so a caller that
// keeps the table gets it back all-zero.
__device__ __forceinline__ void zero_rows(const MatchArgs& a, int b, int c, i64 d0, int D, int G, i64 ib) {
  const int lane = threadIdx.x;
  i64* rows = static_cast<i64*>(const_cast<void*>(a.inter)) + ib;
  for (int base = 0; base < D; base += kWave) {
    const int i = min(base + lane, D - 1);
    const int cls = a.det_cls[d0 + i];
    const i64 row = a.det_row ? a.det_row[d0 + i] : i;
    const bool mine = base + lane < D && (cls == c || (c == 0 && (cls < 0 || cls >= a.C)));
    u64 pending = __ballot(mine);
    while (pending) {
      const int src = __builtin_ctzll(pending);
      pending &= pending - 1;
      const i64 r = __shfl(row, src, kWave);
      for (int g = lane; g < G; g += kWave) rows[r * G + g] = 0;
    }
  }
}

__global__ __launch_bounds__(64) void coco_match(MatchArgs a) {
  extern __shared__ uint4 lds4[];
  if (a.guard && *a.guard) return;
  const int b = blockIdx.x / a.C, c = blockIdx.x - (blockIdx.x / a.C) * a.C;
  const int lane = threadIdx.x;
  double* det_pa = reinterpret_cast<double*>(lds4);   // 100: pixel area of ranked detection r
  i64* det_rw = reinterpret_cast<i64*>(det_pa + 100);  // 100: its intersection row
  int* dets = reinterpret_cast<int*>(det_rw + 100);    // 100
  double* pair_inter = reinterpret_cast<double*>(dets + 100);  // kMatchStagePairs
  double* gt_px = pair_inter + kMatchStagePairs;               // kMatchStageGts: pixel areas
  double* gt_an = gt_px + kMatchStageGts;                      // kMatchStageGts: annotation areas
  uint8_t* gt_cr = reinterpret_cast<uint8_t*>(gt_an + kMatchStageGts);  // kMatchStageGts
  int* gts = reinterpret_cast<int*>(gt_cr + kMatchStageGts);   // max_gt
  uint32_t* used = reinterpret_cast<uint32_t*>(gts + a.max_gt);  // 40 x words
  const i64 d0 = a.det_off[b], g0 = a.gt_off[b];
  const int D = static_cast<int>(a.det_off[b + 1] - d0);
  const int G = static_cast<int>(a.gt_off[b + 1] - g0);
  // Detections of class c in score order; the first 100 get ranks 1..100.
  // Every other detection gets rank 0 and match flags -1 here as well (those
  // of class c beyond the first 100 from this wave, those of no valid class
  // from the class-0 wave); the top 100 get their flags from the chains below.
  int nd = 0;
  auto unranked = [&](i64 d) {
    a.ranks[d] = 0;
    int8_t* row = a.status + d * 40;
#pragma unroll
    for (int k = 0; k < 40; ++k) row[k] = -1;
  };
  // Four wave-wide chunks per step, their class loads issued together
  // (indices clamped into the image; lanes past D are then ignored).
  constexpr int kScan = 4;
  for (int base = 0; base < D; base += kWave * kScan) {
    int cls[kScan];
    double area[kScan];
    i64 row[kScan];
#pragma unroll
    for (int k = 0; k < kScan; ++k) {
      const int i = min(base + k * kWave + lane, D - 1);
      cls[k] = a.det_cls[d0 + i];
      area[k] = a.det_area ? a.det_area[d0 + i] : 0.0;
      row[k] = a.det_row ? a.det_row[d0 + i] : i;
    }
#pragma unroll
    for (int k = 0; k < kScan; ++k) {
      const int i = base + k * kWave + lane;
      const bool sel = i < D && cls[k] == c;
      const u64 mask = __ballot(sel);
      if (sel) {
        const int rank = nd + __popcll(mask & ((1ULL << lane) - 1));
        if (rank < 100) {
          dets[rank] = i;
          det_pa[rank] = area[k];
          det_rw[rank] = row[k];
          a.ranks[d0 + i] = rank + 1;
        } else {
          unranked(d0 + i);
        }
      } else if (i < D && c == 0 && (cls[k] < 0 || cls[k] >= a.C)) {
        unranked(d0 + i);
      }
      nd += __popcll(mask);
    }
  }
  nd = min(nd, 100);
  // Ground truth of class c (metadata order) and positives per area range.
  // As above: class, crowd flag and annotation area of four chunks loaded
  // together before any is used.  A GT is a candidate only if its class is
  // c, so the thing-class filter needs only c's own flag.
  int ng = 0;
  int positives[4] = {0, 0, 0, 0};
  const bool c_candidate = !a.gt_things || a.gt_things[c];
  for (int base = 0; base < G && c_candidate; base += kWave * kScan) {
    int cls[kScan];
    bool crowded[kScan];
    double ann[kScan];
#pragma unroll
    for (int k = 0; k < kScan; ++k) {
      const int j = min(base + k * kWave + lane, G - 1);
      cls[k] = a.gt_cls[g0 + j];
      crowded[k] = a.gt_crowd && a.gt_crowd[g0 + j];
      ann[k] = gt_ann_area(a, b, g0, j);
    }
#pragma unroll
    for (int k = 0; k < kScan; ++k) {
      const int j = base + k * kWave + lane;
      const bool sel = j < G && cls[k] == c;
      const u64 mask = __ballot(sel);
      if (sel) gts[ng + __popcll(mask & ((1ULL << lane) - 1))] = j;
      ng += __popcll(mask);
      const bool eligible = sel && !crowded[k];
      const double area = eligible ? ann[k] : 0.0;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        double lo, hi;
        area_range(r, lo, hi);
        positives[r] += __popcll(__ballot(eligible && area >= lo && area <= hi));
      }
    }
  }
  if (lane == 0 && ng)
    for (int r = 0; r < 4; ++r)
      if (positives[r]) atomicAdd(reinterpret_cast<u64*>(a.counts) + c * 4 + r, static_cast<u64>(positives[r]));
  const i64 ib = a.inter_off[b];
  if (!a.det_area && nd > 0) {
    // Detection areas are the row sums of the int64 intersection rows (the
    // shared-target table spans every target slot): four rows at a time, each
    // lane's columns loaded together, then summed across the wave.
    __syncthreads();  // det_rw[] complete
    const i64* rows = static_cast<const i64*>(a.inter) + ib;
    const int ranked = min(nd, 100);
    for (int r0 = 0; r0 < ranked; r0 += 4) {
      i64 sum[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const i64* row = rows + det_rw[min(r0 + u, ranked - 1)] * G;
        sum[u] = 0;
        for (int g = lane; g < G; g += kWave) sum[u] += row[g];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const i64 total = wave_sum(sum[u]);
        if (lane == 0 && r0 + u < ranked) det_pa[r0 + u] = static_cast<double>(total);
      }
    }
  }
  if (nd == 0) {
    if (a.zero_inter) zero_rows(a, b, c, d0, D, G, ib);
    return;
  }
  for (int w = lane; w < 40 * a.words; w += kWave) used[w] = 0;
  const bool staged = ng <= kMatchStageGts && nd * ng <= kMatchStagePairs;
  if (staged && ng > 0) {
    __syncthreads();  // gts[] complete
    for (int q = lane; q < ng; q += kWave) {
      const int j = gts[q];
      gt_cr[q] = a.gt_crowd && a.gt_crowd[g0 + j] != 0;
      gt_px[q] = gt_pixel_area(a, b, g0, j);
      gt_an[q] = gt_ann_area(a, b, g0, j);
    }
    // Four pairs per lane per step, loads issued together (clamped).
    const int pairs = nd * ng;
    for (int e0 = lane; e0 < pairs; e0 += 4 * kWave) {
      double v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int e = min(e0 + u * kWave, pairs - 1), r = e / ng, q = e - r * ng;
        const i64 idx = ib + det_rw[r] * G + gts[q];
        v[u] = a.inter_bytes == 8 ? static_cast<double>(static_cast<const i64*>(a.inter)[idx])
                                  : static_cast<double>(static_cast<const int32_t*>(a.inter)[idx]);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (e0 + u * kWave < pairs) pair_inter[e0 + u * kWave] = v[u];
    }
  }
  __syncthreads();
  if (lane < 40) match_chains(a, b, d0, g0, G, ib, nd, ng, staged, dets, det_rw, det_pa, pair_inter, gt_px, gt_an,
                              gt_cr, gts, used);
  if (a.zero_inter) {
    __syncthreads();  // every intersection read is done
    zero_rows(a, b, c, d0, D, G, ib);
  }
}

// The 40 greedy chains of one (image, class) wave: lane = area range * 10 +
// IoU threshold; detections in score order.
__device__ __forceinline__ void match_chains(const MatchArgs& a, int b, i64 d0, i64 g0, int G, i64 ib, int nd,
                                             int ng, bool staged, const int* dets, const i64* det_rw,
                                             const double* det_pa, const double* pair_inter, const double* gt_px,
                                             const double* gt_an, const uint8_t* gt_cr, const int* gts,
                                             uint32_t* used) {
  const int lane = threadIdx.x & (kWave - 1);
  const int range = lane / 10, t = lane - (lane / 10) * 10;
  double lo, hi;
  area_range(range, lo, hi);
  const double thr = iou_threshold(t);
  uint32_t* taken = used + lane * a.words;
  for (int r = 0; r < nd; ++r) {
    const int i = dets[r];
    const i64 row = det_rw[r];
    const double pa = det_pa[r];
    int best = -1, best_ignore = 2;
    double best_iou = thr;
    // Candidates four at a time: their loads (crowd flag, intersection, pixel
    // and annotation areas) issue together; the greedy choice then visits
    // them in metadata order, as one at a time.
    constexpr int Q = 4;
    for (int q0 = 0; q0 < ng; q0 += Q) {
      bool crowd[Q];
      double inter[Q], garea[Q], ann[Q];
#pragma unroll
      for (int u = 0; u < Q; ++u) {
        const int q = min(q0 + u, ng - 1);
        if (staged) {
          crowd[u] = gt_cr[q] != 0;
          inter[u] = pair_inter[r * ng + q];
          garea[u] = gt_px[q];
          ann[u] = gt_an[q];
        } else {
          const int j = gts[q];
          crowd[u] = a.gt_crowd && a.gt_crowd[g0 + j] != 0;
          const i64 idx = ib + row * G + j;
          inter[u] = a.inter_bytes == 8 ? static_cast<double>(static_cast<const i64*>(a.inter)[idx])
                                        : static_cast<double>(static_cast<const int32_t*>(a.inter)[idx]);
          garea[u] = gt_pixel_area(a, b, g0, j);
          ann[u] = gt_ann_area(a, b, g0, j);
        }
      }
#pragma unroll
      for (int u = 0; u < Q; ++u) {
        const int q = q0 + u;
        if (q >= ng) break;
        if (!crowd[u] && ((taken[q >> 5] >> (q & 31)) & 1u)) continue;
        double uni = crowd[u] ? pa : pa + garea[u] - inter[u];
        if (uni < 1.0) uni = 1.0;
        const double iou = inter[u] / uni;
        const int ignore = crowd[u] || ann[u] < lo || ann[u] > hi;
        // Non-ignored GT first, then larger IoU (>= running best), later
        // metadata entry on equal IoU: COCO's scan order.
        if (iou >= thr && (ignore < best_ignore || (ignore == best_ignore && iou >= best_iou))) {
          best = q;
          best_ignore = ignore;
          best_iou = iou;
        }
      }
    }
    int8_t flag = -1;
    if (best >= 0) {
      taken[best >> 5] |= 1u << (best & 31);
      flag = best_ignore ? -1 : 1;
    } else if (lo <= pa && pa <= hi) {
      flag = 0;
    }
    a.status[(d0 + i) * 40 + lane] = flag;
  }
}

// ------------------------------------------------ grouped COCO matching
//
// The shared-target (combined) path splits matching in two.  match_groups
// needs only the score order and the target slot classes, so it runs beside
// the detection histogram: per image it groups the detections (score order
// kept within a class) and the candidate GT slots by class, writes every
// detection's rank and the flags of the unranked ones.  coco_match_grouped,
// after the histogram, then reads its (image, class) lists directly: three
// dependent rounds of loads (header, lists, intersections and GT metadata)
// instead of scanning the image's detections and slots.

struct GroupArgs {
  int B, C, S;
  const i64* det_off;
  const int32_t* det_cls;    // score order
  const i64* det_row;        // score order -> row (mask index within the image)
  const int32_t* gt_cls;     // B x S
  const uint8_t* gt_things;  // C or NULL: only GT of thing classes are candidates
  int8_t* status;            // (sum D) x 40
  int32_t* ranks;            // sum D
  int4* hdr;
  int2* det_list;
  int32_t* gt_list;
  int det_slots, gt_slots;   // LDS sort sizes (powers of two >= the largest D, >= S)
};

// Ascending bitonic sort of m (a power of two) 32-bit keys in LDS.
__device__ __forceinline__ void bitonic_u32(uint32_t* keys, int m) {
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int k = 2; k <= m; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int e = tid; e < m; e += nt) {
        const int f = e ^ j;
        if (f <= e) continue;
        const uint32_t ke = keys[e], kf = keys[f];
        if ((ke > kf) == ((e & k) == 0)) {
          keys[e] = kf;
          keys[f] = ke;
        }
      }
      __syncthreads();
    }
}

// Keys are (class bucket << 12) | index: a stable grouping by class of at
// most 4096 entries; bucket C holds the entries of no (candidate) class.
constexpr int kGroupIndexBits = 12;
constexpr int kGroupMaxEntries = 1 << kGroupIndexBits;

__global__ __launch_bounds__(256) void match_groups(GroupArgs g) {
  extern __shared__ uint4 lds4[];
  uint32_t* dkey = reinterpret_cast<uint32_t*>(lds4);
  uint32_t* gkey = dkey + g.det_slots;
  int* dstart = reinterpret_cast<int*>(gkey + g.gt_slots);
  int* dend = dstart + (g.C + 1);
  int* gstart = dend + (g.C + 1);
  int* gend = gstart + (g.C + 1);
  const int b = blockIdx.x, tid = threadIdx.x, nt = blockDim.x, C = g.C, S = g.S;
  const i64 d0 = g.det_off[b];
  const int D = static_cast<int>(min<i64>(g.det_off[b + 1] - d0, g.det_slots));  // the host sizes det_slots
  for (int c = tid; c <= C; c += nt) dstart[c] = dend[c] = gstart[c] = gend[c] = 0;
  for (int e = tid; e < g.det_slots; e += nt) {
    uint32_t key = 0xFFFFFFFFu;
    if (e < D) {
      const int c = g.det_cls[d0 + e];
      key = (static_cast<uint32_t>(c >= 0 && c < C ? c : C) << kGroupIndexBits) | static_cast<uint32_t>(e);
    }
    dkey[e] = key;
  }
  for (int j = tid; j < g.gt_slots; j += nt) {
    uint32_t key = 0xFFFFFFFFu;
    if (j < S) {
      const int c = g.gt_cls[static_cast<i64>(b) * S + j];
      const bool candidate = c >= 0 && c < C && (!g.gt_things || g.gt_things[c]);
      key = (static_cast<uint32_t>(candidate ? c : C) << kGroupIndexBits) | static_cast<uint32_t>(j);
    }
    gkey[j] = key;
  }
  __syncthreads();
  bitonic_u32(dkey, g.det_slots);
  bitonic_u32(gkey, g.gt_slots);
  for (int p = tid; p < D; p += nt) {
    const uint32_t k = dkey[p] >> kGroupIndexBits;
    if (p == 0 || (dkey[p - 1] >> kGroupIndexBits) != k) dstart[k] = p;
    if (p == D - 1 || (dkey[p + 1] >> kGroupIndexBits) != k) dend[k] = p + 1;
  }
  for (int p = tid; p < S; p += nt) {
    const uint32_t k = gkey[p] >> kGroupIndexBits;
    if (p == 0 || (gkey[p - 1] >> kGroupIndexBits) != k) gstart[k] = p;
    if (p == S - 1 || (gkey[p + 1] >> kGroupIndexBits) != k) gend[k] = p + 1;
  }
  __syncthreads();
  for (int p = tid; p < D; p += nt) {
    const uint32_t key = dkey[p];
    const int k = static_cast<int>(key >> kGroupIndexBits), e = static_cast<int>(key & (kGroupMaxEntries - 1));
    const int rank = p - dstart[k];
    g.det_list[d0 + p] = int2{e, static_cast<int>(g.det_row[d0 + e])};
    const bool ranked = k < C && rank < 100;
    g.ranks[d0 + e] = ranked ? rank + 1 : 0;
    if (!ranked) {
      // Rank 0 and flags -1 (five 8-byte words of -1; rows are 8-byte aligned).
      uint64_t* row = reinterpret_cast<uint64_t*>(g.status + (d0 + e) * 40);
#pragma unroll
      for (int w = 0; w < 5; ++w) row[w] = ~0ULL;
    }
  }
  for (int p = tid; p < S; p += nt)
    g.gt_list[static_cast<i64>(b) * S + p] = static_cast<int32_t>(gkey[p] & (kGroupMaxEntries - 1));
  int4* hdr = g.hdr + static_cast<i64>(b) * (C + 1);
  for (int c = tid; c <= C; c += nt) hdr[c] = int4{dstart[c], dend[c] - dstart[c], gstart[c], gend[c] - gstart[c]};
}

// Zero the int64 rows (S columns) of `count` listed detections of image b.
__device__ __forceinline__ void zero_listed_rows(const MatchArgs& a, const int2* list, int count, i64 block) {
  i64* rows = static_cast<i64*>(const_cast<void*>(a.inter)) + block;
  const int lane = threadIdx.x & (kWave - 1);
  for (int base = 0; base < count; base += kWave) {
    const int row = list[min(base + lane, count - 1)].y;
    const int here = min(kWave, count - base);
    for (int k = 0; k < here; ++k) {
      const i64 r = __shfl(row, k, kWave);
      for (int s = lane; s < a.S; s += kWave) rows[r * a.S + s] = 0;
    }
  }
}

// LDS of one matching wave (bytes, 16-byte multiple) for S target slots.
__host__ __device__ inline size_t match_wave_lds(int S) {
  const size_t words = (static_cast<size_t>(S) + 31) / 32;
  return align_up(20 * 100 + 8 * static_cast<size_t>(kMatchStagePairs) + 20 * static_cast<size_t>(S) +
                      align_up(static_cast<size_t>(S), 16) + 4 * 40 * words,
                  16);
}

// Orders a wave's LDS writes before its later LDS reads by other lanes (the
// waves of a workgroup may be matching different images).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// One wave matches class c of image b from the lists of match_groups, in
// three dependent load trips:
// trip 1 the class's header, trip 2 its lists, trip 3 GT metadata,
// intersections and detection row sums, then the chains; last it zeroes the
// rows of the class's detections.  `lds` is this wave's match_wave_lds(S).
__device__ void match_grouped_wave(const MatchArgs& a, int b, int c, char* lds) {
  const int C = a.C, S = a.S;
  const int lane = threadIdx.x & (kWave - 1);
  double* det_pa = reinterpret_cast<double*>(lds);     // 100
  i64* det_rw = reinterpret_cast<i64*>(det_pa + 100);  // 100
  int* dets = reinterpret_cast<int*>(det_rw + 100);    // 100
  double* pair_inter = reinterpret_cast<double*>(dets + 100);  // kMatchStagePairs
  double* gt_px = pair_inter + kMatchStagePairs;               // S
  double* gt_an = gt_px + S;                                   // S
  int* gts = reinterpret_cast<int*>(gt_an + S);                // S
  uint8_t* gt_cr = reinterpret_cast<uint8_t*>(gts + S);        // S
  uint32_t* used = reinterpret_cast<uint32_t*>(gt_cr + ((S + 15) & ~15));  // 40 x words
  // Load trip 1: the class's header (and the no-class bucket's for class 0).
  const int4* hdr = a.group_hdr + static_cast<i64>(b) * (C + 1);
  const int4 h = hdr[c];
  const int4 hx = hdr[c == 0 ? C : c];
  const i64 d0 = a.det_off[b], g0 = static_cast<i64>(b) * S, ib = d0 * S;
  const int nd_all = h.y, ng = h.w, extra = c == 0 ? hx.y : 0;
  if (nd_all == 0 && ng == 0 && extra == 0) return;  // uniform
  const int nd = min(nd_all, 100);
  // Load trip 2: the lists.
  for (int r = lane; r < nd; r += kWave) {
    const int2 v = a.group_det[d0 + h.x + r];
    dets[r] = v.x;
    det_rw[r] = v.y;
  }
  for (int q = lane; q < ng; q += kWave) gts[q] = a.group_gt[g0 + h.z + q];
  wave_sync();
  // Load trip 3: GT metadata (areas, crowd flags) and positives per area range.
  int positives[4] = {0, 0, 0, 0};
  for (int q0 = 0; q0 < ng; q0 += kWave) {
    const int q = q0 + lane;
    const int j = gts[min(q, ng - 1)];
    const bool crowded = a.gt_crowd && a.gt_crowd[g0 + j];
    const double px = static_cast<double>(a.gt_area_i64[b * a.gt_area_stride + j]);
    const double ann = a.gt_ann ? a.gt_ann[g0 + j] : px;
    if (q < ng) {
      gt_cr[q] = crowded;
      gt_px[q] = px;
      gt_an[q] = ann;
    }
    const bool eligible = q < ng && !crowded;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      double lo, hi;
      area_range(r, lo, hi);
      positives[r] += __popcll(__ballot(eligible && ann >= lo && ann <= hi));
    }
  }
  if (lane == 0 && ng)
    for (int r = 0; r < 4; ++r)
      if (positives[r]) atomicAdd(reinterpret_cast<u64*>(a.counts) + c * 4 + r, static_cast<u64>(positives[r]));
  const bool staged = nd * ng <= kMatchStagePairs;
  if (nd > 0) {
    const i64* rows = static_cast<const i64*>(a.inter) + ib;
    // Intersections of every (ranked detection, candidate) pair, four per
    // lane per step, loads issued together.
    if (staged && ng > 0) {
      const int pairs = nd * ng;
      for (int e0 = lane; e0 < pairs; e0 += 4 * kWave) {
        double v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int e = min(e0 + u * kWave, pairs - 1), r = e / ng, q = e - r * ng;
          v[u] = static_cast<double>(rows[det_rw[r] * S + gts[q]]);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
          if (e0 + u * kWave < pairs) pair_inter[e0 + u * kWave] = v[u];
      }
    }
    // Detection areas: the sums of their rows, four rows at a time.
    for (int r0 = 0; r0 < nd; r0 += 4) {
      i64 sum[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const i64* row = rows + det_rw[min(r0 + u, nd - 1)] * S;
        sum[u] = 0;
        for (int s = lane; s < S; s += kWave) sum[u] += row[s];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const i64 total = wave_sum(sum[u]);
        if (lane == 0 && r0 + u < nd) det_pa[r0 + u] = static_cast<double>(total);
      }
    }
    for (int w = lane; w < 40 * a.words; w += kWave) used[w] = 0;
    wave_sync();
    if (lane < 40)
      match_chains(a, b, d0, g0, S, ib, nd, ng, staged, dets, det_rw, det_pa, pair_inter, gt_px, gt_an, gt_cr, gts,
                   used);
  }
  // Every intersection read is done: zero the rows of the class's detections
  // (and, for class 0, of the detections of no valid class).
  wave_sync();
  if (nd_all) zero_listed_rows(a, a.group_det + d0 + h.x, nd_all, ib);
  if (extra) zero_listed_rows(a, a.group_det + d0 + hx.x, extra, ib);
}

__global__ __launch_bounds__(64) void coco_match_grouped(MatchArgs a) {
  extern __shared__ uint4 lds4[];
  const int b = blockIdx.x / a.C, c = blockIdx.x - (blockIdx.x / a.C) * a.C;
  match_grouped_wave(a, b, c, reinterpret_cast<char*>(lds4));
}

// ------------------------------------------------------------ AP / AR

__device__ __forceinline__ u64 score_key_desc(double s) {
  if (s == 0.0) s = 0.0;  // -0.0 and +0.0 compare equal in the reference's sort
  u64 bits = __double_as_longlong(s);
  u64 asc = (bits >> 63) ? ~bits : (bits | 0x8000000000000000ULL);
  return ~asc;  // ascending radix order == descending score
}

// Sort items of the AP records: class (records with ranks outside 1..100,
// which no maxDets setting counts, take class C and sort last), descending
// score, record index.
__global__ void ap_items(const double* scores, const int32_t* classes, const int32_t* ranks, i64 N, int C,
                         SortItem* items) {
  for (i64 i = static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x; i < N; i += static_cast<i64>(gridDim.x) * blockDim.x) {
    const int c = classes[i], r = ranks[i];
    const bool valid = r > 0 && r <= 100 && c >= 0 && c < C;
    items[i] = SortItem{score_key_desc(scores[i]), static_cast<uint32_t>(valid ? c : C), static_cast<int32_t>(i)};
  }
}

// Record order and each class's [start, end) from the sorted items.
__global__ void ap_order(const SortItem* sorted, i64 N, int C, int32_t* order, int32_t* start, int32_t* end) {
  for (i64 i = static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x; i < N; i += static_cast<i64>(gridDim.x) * blockDim.x) {
    const uint32_t k = sorted[i].group;
    order[i] = sorted[i].index;
    if (k >= static_cast<uint32_t>(C)) continue;
    if (i == 0 || sorted[i - 1].group != k) start[k] = static_cast<int32_t>(i);
    if (i == N - 1 || sorted[i + 1].group != k) end[k] = static_cast<int32_t>(i + 1);
  }
}

struct ApArgs {
  int C;
  const int32_t* order;      // record index per sorted position (class-grouped, score order)
  const int32_t* start;      // C
  const int32_t* end;        // C
  const int32_t* ranks;
  const int8_t* status;      // N x 40
  const i64* counts;         // C x 4
  double* precision;         // 10 x 101 x C x 4 x 3
  double* recall;            // 10 x C x 4 x 3
  double* partial;           // C x 24 summary accumulators
  // Optional: a device word that is zero between launches.  The last class
  // workgroup to finish (counted on it) folds the summaries and zeroes it.
  int* done;
  double* summary;           // 12 summaries (with done)
};

// One workgroup per class (instance.py:300-331).  Its first four waves scan
// the class's records, wave ar for area range ar, lane l < 30 for
// (t, maxDets) = (l / 3, l % 3): job j = t * 12 + ar * 3 + maxDets indexes
// the tables and the precision layout.  The records (score order) are staged
// through LDS in chunks: rank and the 40 match flags per record.  Forward
// pass: totals; backward pass: the precision envelope at the 101 recall
// thresholds.
constexpr int kApChunk = 1024;
constexpr int kApThreads = 1024;      // ap_accumulate (large batches)
constexpr int kApSmallThreads = 1024; // ap_small: more lanes for the gathers, the sort and the writer
constexpr int kApJobs = 120;

// LDS table of one class: for each recall target k and job j, the counts
// (tp, fp) of the record whose precision is the envelope there, packed in
// one Code (16 + 16 bits when counts stay below 65536, else 32 + 32), and per
// job the highest threshold its final recall reaches (-1: every threshold is
// 0; -2: no positives, every threshold is -1).
template <typename Code>
constexpr size_t ap_table_bytes() { return sizeof(Code) * 101 * kApJobs + sizeof(int) * kApJobs; }
constexpr size_t kApTableBytes = ap_table_bytes<uint32_t>();

// The least true-positive count t >= 1 with t / positives >= r * 0.01, both
// as the FP64 expressions the reference compares (t = positives qualifies:
// 100 * 0.01 rounds to 1.0).  Threshold r takes the envelope at the record
// where the running count reaches it (the first counted record for r = 0,
// where the envelope already equals the one at the first true positive:
// records before it have precision 0).
__device__ __forceinline__ int32_t recall_target(int r, i64 positives) {
  const double thr = static_cast<double>(r) * 0.01;
  const double npos = static_cast<double>(positives);
  i64 t = static_cast<i64>(ceil(thr * npos));
  t = t < 1 ? 1 : t > positives ? positives : t;
  while (t > 1 && static_cast<double>(t - 1) / npos >= thr) --t;
  while (t < positives && static_cast<double>(t) / npos < thr) ++t;
  return static_cast<int32_t>(t > INT_MAX ? INT_MAX : t);
}

// Whether precision tp / (tp + fp + eps) of (t, f) exceeds that of (et, ef)
// as FP64 values, compared on the integers.  For n = tp + fp >= 2, n + eps
// rounds to n, so the values are the correctly rounded fractions t / n;
// distinct fractions keep their order (or round equal, and then either
// record gives the same value).  n = 1 gives 1 / (1 + eps), just below the
// exact 1 of t = n >= 2 and above every other fraction; n = 0 gives 0.
// Count is int (counts below 4096: 24-bit products) or i64.
template <typename Count>
__device__ __forceinline__ bool precision_above(Count t, Count f, Count et, Count ef) {
  const Count n = t + f, en = et + ef;
  bool above, equal;
  if constexpr (sizeof(Count) == 4) {
    const uint32_t lhs = __umul24(static_cast<uint32_t>(t), static_cast<uint32_t>(en > 1 ? en : 1));
    const uint32_t rhs = __umul24(static_cast<uint32_t>(et), static_cast<uint32_t>(n > 1 ? n : 1));
    above = lhs > rhs;
    equal = lhs == rhs;
  } else {
    const u64 lhs = static_cast<u64>(t) * static_cast<u64>(en > 1 ? en : 1);
    const u64 rhs = static_cast<u64>(et) * static_cast<u64>(n > 1 ? n : 1);
    above = lhs > rhs;
    equal = lhs == rhs;
  }
  return above || (equal && n >= 2 && t == n && en == 1 && et == 1);
}

template <typename Code>
__device__ __forceinline__ double code_precision(Code code) {
  constexpr int kShift = 4 * sizeof(Code);
  const i64 tp = static_cast<i64>(code >> kShift), fp = static_cast<i64>(code & ((Code(1) << kShift) - 1));
  return static_cast<double>(tp) / (static_cast<double>(tp + fp) + DBL_EPSILON);
}

// The accumulation of class c over its records order[s0 .. s1) (global score
// order, ties by record index).  ORDER is the address space of ``order``
// (a global array or an LDS list), kept apart so accesses stay typed.
// The envelope (the running maximum precision, scanning backward) only
// changes at true positives, and a threshold only takes it at the true
// positive where the count reaches the threshold's target, so the backward
// pass does integer work at true positives and one LDS write per distinct
// target.  The workgroup then writes the class's 10 x 101 x 12 precision
// values (12 contiguous doubles per (t, r)) and each job folds its own
// thresholds, highest first (the order of the reference's loop).  STAGED:
// the caller has already put the records (at most one chunk) in LDS.
template <int THREADS, typename Code, typename OrderPtr>
__device__ __forceinline__ void ap_class(const ApArgs& a, int c, OrderPtr order, int s0, int s1, uint8_t* s_rank,
                                         int8_t* s_flag, Code* codes, int* last, bool staged = false) {
  static_assert(THREADS >= 4 * kWave, "four scan waves");
  using Count = std::conditional_t<sizeof(Code) == 4, int, i64>;
  constexpr int kShift = 4 * sizeof(Code);
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  // Recall targets of each area range: s_target[ar][k] the distinct targets
  // in descending order (0 ends the list), s_kr[ar][r] threshold r's k.
  __shared__ int32_t s_target[4][102];
  __shared__ uint8_t s_kr[4][101];
  if (wave < 4) {
    const i64 pos = a.counts[c * 4 + wave];
    const int x_lo = pos > 0 ? recall_target(lane, pos) : 1;                               // r = lane
    const int x_hi = pos > 0 && lane <= 36 ? recall_target(64 + lane, pos) : INT_MAX;      // r = 64 + lane
    // a threshold starts a new target where its target differs from r + 1's
    // (every lane takes part in both shuffles)
    const int down_lo = __shfl_down(x_lo, 1, kWave);
    const int first_hi = __shfl(x_hi, 0, kWave);
    const int next_hi = __shfl_down(x_hi, 1, kWave);
    const int next_lo = lane < 63 ? down_lo : first_hi;
    const bool new_lo = x_lo != next_lo;
    const bool new_hi = lane <= 36 && (lane == 36 || x_hi != next_hi);
    const u64 hi = __ballot(new_hi), lo = __ballot(new_lo);
    const int k_hi = __popcll(hi >> lane) - 1;
    const int k_lo = __popcll(lo >> lane) + __popcll(hi) - 1;
    if (lane <= 36) {
      s_kr[wave][64 + lane] = static_cast<uint8_t>(k_hi);
      if (new_hi) s_target[wave][k_hi] = x_hi;
    }
    s_kr[wave][lane] = static_cast<uint8_t>(k_lo);
    if (new_lo) s_target[wave][k_lo] = x_lo;
    if (lane == 0) s_target[wave][__popcll(hi) + __popcll(lo)] = 0;
  }
  // Scan lanes (the tables above are published by the forward pass's first
  // barrier, which runs whenever a scan does).
  const int ar = wave, t = lane / 3, m = lane % 3;
  const bool active = wave < 4 && lane < 30;
  const int job = t * 12 + ar * 3 + m;
  const int limit = m == 0 ? 1 : m == 1 ? 10 : 100;
  const i64 positives = active ? a.counts[c * 4 + ar] : 0;
  const i64 pstride = static_cast<i64>(a.C) * 12;
  const int lane_flag = ar * 10 + t;
  // Records of one chunk -> LDS: the rank and the 40 flags (five 8-byte words,
  // rows are 8-byte aligned) of each record, four gathers in flight per lane.
  // Out-of-range lanes load the last word (indices are clamped rather than
  // branched around, so every load of a step is issued before the first wait).
  auto stage = [&](int first, int count) {
    uint64_t* flag_words = reinterpret_cast<uint64_t*>(s_flag);
    const int words = count * 5;
    constexpr int U = 4;
    for (int e0 = tid; e0 < words; e0 += THREADS * U) {
      uint64_t v[U];
      int rank[U], idx[U];
#pragma unroll
      for (int u = 0; u < U; ++u) idx[u] = order[first + min(e0 + u * THREADS, words - 1) / 5];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int e = min(e0 + u * THREADS, words - 1);
        v[u] = reinterpret_cast<const uint64_t*>(a.status + static_cast<i64>(idx[u]) * 40)[e % 5];
        rank[u] = a.ranks[idx[u]];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int e = e0 + u * THREADS;
        if (e >= words) break;
        const int i = e / 5, w = e - i * 5;
        flag_words[e] = v[u];
        if (w == 0) s_rank[i] = static_cast<uint8_t>(rank[u]);
      }
    }
  };
  // forward: totals.  Records are read V at a time with every LDS load in
  // flight before any is used; the counts take no branches.
  constexpr int V = 4;
  Count tp = 0, fp = 0, n = 0;
  for (int first = s0; first < s1; first += kApChunk) {
    const int count = min(kApChunk, s1 - first);
    __syncthreads();
    if (!staged) stage(first, count);
    __syncthreads();
    if (active && positives)
      for (int i0 = 0; i0 < count; i0 += 2 * V) {
        int rank[2 * V], flag[2 * V];
#pragma unroll
        for (int v = 0; v < 2 * V; ++v) {
          const int iv = min(i0 + v, count - 1);
          rank[v] = s_rank[iv];
          flag[v] = s_flag[iv * 40 + lane_flag];
        }
#pragma unroll
        for (int v = 0; v < 2 * V; ++v)
          if (i0 + v >= count) rank[v] = 255;
#pragma unroll
        for (int v = 0; v < 2 * V; ++v) {
          const bool counted = rank[v] <= limit;
          tp += counted && flag[v] == 1;
          fp += counted && flag[v] == 0;
          n += counted;
        }
      }
  }
  const bool scan = active && positives && tp > 0;
  const double npos = static_cast<double>(positives);
  // Thresholds above the final recall get precision 0: the highest threshold
  // taken is the largest r with r * 0.01 <= final recall.
  int top = -1;
  if (scan) {
    const double final_recall = static_cast<double>(tp) / npos;
    top = min(max(static_cast<int>(final_recall * 100.0), 0), 100);
    while (top < 100 && static_cast<double>(top + 1) * 0.01 <= final_recall) ++top;
    while (top >= 0 && static_cast<double>(top) * 0.01 > final_recall) --top;
  }
  if (active) last[job] = !positives ? -2 : n == 0 ? -1 : top;
  // backward.  The next target (and the one after it, prefetched) stay in
  // registers; the scan of a lane ends with its last target.
  int k = scan ? s_kr[ar][top] : 0;
  int target = scan ? s_target[ar][k] : 0;
  int after = scan ? s_target[ar][k + 1] : 0;
  Count et = 0, ef = 0;  // the envelope's record (0 / (0 + eps) = 0.0)
  Code env_code = 0;
  Count ctp = tp, cfp = fp;
  const int last_first = s1 > s0 ? s0 + ((s1 - s0 - 1) / kApChunk) * kApChunk : s0;
  for (int first = last_first; first >= s0 && s1 > s0; first -= kApChunk) {
    const int count = min(kApChunk, s1 - first);
    if (s1 - s0 > kApChunk) {
      __syncthreads();
      stage(first, count);
      __syncthreads();
    }
    if (scan)
      for (int i = count - 1; i >= 0 && target > 0; i -= V) {
        int rank[V], flag[V];
#pragma unroll
        for (int v = 0; v < V; ++v) {
          const int iv = max(i - v, 0);
          rank[v] = s_rank[iv];
          flag[v] = s_flag[iv * 40 + lane_flag];
        }
#pragma unroll
        for (int v = 0; v < V; ++v)
          if (i - v < 0) rank[v] = 255;
#pragma unroll
        for (int v = 0; v < V; ++v) {
          const bool counted = rank[v] <= limit;
          const bool hit = counted && flag[v] == 1;
          // (ctp, cfp): the counts up to and including this record
          const bool up = hit && precision_above<Count>(ctp, cfp, et, ef);
          et = up ? ctp : et;
          ef = up ? cfp : ef;
          env_code = up ? static_cast<Code>(static_cast<Code>(ctp) << kShift | static_cast<Code>(cfp)) : env_code;
          if (hit && ctp == target) {
            codes[k * kApJobs + job] = env_code;
            ++k;
            target = after;
            after = s_target[ar][min(k + 1, 101)];
          }
          ctp -= hit;
          cfp -= counted && flag[v] == 0;
        }
      }
  }
  __syncthreads();
  // The class's precision values, and each job's sum over the thresholds it
  // reaches.  Thread (t, part, q) -- 960 threads, q = ar * 3 + m the job within
  // t -- takes thresholds r = part, part + 8, ... of job t * 12 + q: twelve
  // neighbouring threads write one (t, r) row's twelve values, and each thread
  // sums its own values, so the sums need no second pass over memory.  A job's
  // eight partial sums are then added in part order: a fixed order.
  static_assert(THREADS >= 960, "one thread per (threshold, part, job) triple");
  constexpr int kParts = 8, kPerPart = (101 + kParts - 1) / kParts;
  double* base = a.precision + c * 12;
  double* part_sums = reinterpret_cast<double*>(s_flag + 4096);  // kApJobs x kParts, beside sh below
  {
    const bool writer = tid < 10 * kParts * 12;
    const int q = tid % 12, part = (tid / 12) % kParts, tt = min(tid / (12 * kParts), 9);
    const int j = tt * 12 + q;
    const int reached = last[j];
    Code cw[kPerPart];
#pragma unroll
    for (int i = 0; i < kPerPart; ++i) cw[i] = codes[s_kr[q / 3][min(part + kParts * i, 100)] * kApJobs + j];
    // Envelopes with no true positive are exactly 0 (0 / (fp + eps)); a wave
    // whose values are all such skips the divisions.
    bool divide = false;
#pragma unroll
    for (int i = 0; i < kPerPart; ++i) {
      const int rr = part + kParts * i;
      divide |= writer && rr <= 100 && reached >= -1 && rr <= reached && (cw[i] >> kShift) != 0;
    }
    const bool any_divide = __ballot(divide) != 0;
    double sum = 0.0;
    if (writer) {
#pragma unroll
      for (int i = 0; i < kPerPart; ++i) {
        const int rr = part + kParts * i;
        if (rr > 100) break;
        const Code code = rr <= reached ? cw[i] : Code(0);
        double value = reached < -1 ? -1.0 : 0.0;
        if (any_divide && reached >= -1) value = code_precision(code);
        base[(static_cast<i64>(tt) * 101 + rr) * pstride + q] = value;
        if (rr <= reached) sum += value;
      }
      part_sums[j * kParts + part] = sum;
    }
  }
  __syncthreads();
  // Recall and the class's share of the 12 summaries: 24 (sum, count)
  // accumulators, each folded over the 120 jobs in job order.
  double part_sum = 0.0, part_n = 0.0, rv = -1.0;
  if (active) {
    if (positives) {
      double sum = 0.0;
#pragma unroll
      for (int p = 0; p < kParts; ++p) sum += part_sums[job * kParts + p];
      rv = static_cast<double>(tp) / npos;
      part_sum = sum;
      part_n = 101.0;
    }
    a.recall[static_cast<i64>(t) * pstride + c * 12 + ar * 3 + m] = rv;
  }
  double* sh = reinterpret_cast<double*>(s_flag);
  if (active) {
    sh[3 * job] = part_sum;
    sh[3 * job + 1] = part_n;
    sh[3 * job + 2] = rv;
  }
  __syncthreads();
  // Accumulator q: even q sums, odd q the matching counts, of summary q / 2:
  // 0 ap, 1 ap50, 2 ap75 (all areas, maxDets 100), 3-5 ap small/medium/large,
  // 6-8 ar1/ar10/ar100 (all areas), 9-11 ar small/medium/large (maxDets 100).
  // Each folds its jobs over thresholds t in increasing order.
  if (tid < 24) {
    const int q = tid, kk = q >> 1;
    const bool count = q & 1;
    const bool precision_side = kk < 6;
    int sar = 0, sm = 2, t0 = 0, t1 = 10;
    if (kk == 1) t1 = 1;
    else if (kk == 2) t0 = 5, t1 = 6;
    else if (kk >= 3 && kk < 6) sar = kk - 2;
    else if (kk >= 6 && kk < 9) sm = kk - 6;
    else if (kk >= 9) sar = kk - 8;
    double acc = 0.0;
    for (int st = t0; st < t1; ++st) {
      const int j = st * 12 + sar * 3 + sm;
      if (precision_side) {
        acc += count ? sh[3 * j + 1] : sh[3 * j];
      } else {
        const double srv = sh[3 * j + 2];
        if (srv >= 0) acc += count ? 1.0 : srv;
      }
    }
    if (a.done) {
      // Read by the last class workgroup, possibly on another XCD: the
      // exchange completes at the device's coherence point before this
      // workgroup counts itself done.
      atomicExch(reinterpret_cast<unsigned long long*>(a.partial + c * 24 + q),
                 static_cast<unsigned long long>(__double_as_longlong(acc)));
    } else {
      a.partial[c * 24 + q] = acc;
    }
  }
}

// Classes whose records were grouped by the global sort.  Dynamic LDS: the
// chunk's ranks and flags, then the 64-bit table.
constexpr size_t kApAccumulateLds = kApChunk + 40 * kApChunk + ap_table_bytes<u64>();
__global__ __launch_bounds__(kApThreads) void ap_accumulate(ApArgs a) {
  extern __shared__ uint4 lds4[];
  int8_t* s_flag = reinterpret_cast<int8_t*>(lds4);
  uint8_t* s_rank = reinterpret_cast<uint8_t*>(s_flag + 40 * kApChunk);
  u64* codes = reinterpret_cast<u64*>(s_rank + kApChunk);
  int* last = reinterpret_cast<int*>(codes + 101 * kApJobs);
  const int c = blockIdx.x;
  const int s0 = a.start[c], s1 = a.end[c];
  // Counts below 4096 take the 32-bit codes and 24-bit products.
  if (s1 - s0 < 4096)
    ap_class<kApThreads>(a, c, as_global(a.order), s0, s1, s_rank, s_flag, reinterpret_cast<uint32_t*>(codes), last);
  else
    ap_class<kApThreads>(a, c, as_global(a.order), s0, s1, s_rank, s_flag, codes, last);
}

// Batches of at most kApSmall records: each class workgroup gathers its own
// records (valid rank, this class) in index order, sorts them by (score
// descending, index) in LDS (a counted rank per record when they fit one per
// lane, else a bitonic network), and accumulates -- the same order the
// global sort of larger batches gives, in one launch.
constexpr int kApSmall = 4096;

struct ApSmallArgs {
  ApArgs ap;
  const double* scores;
  const int32_t* classes;
  int N;
  int sort_slots;        // power of two >= N
};

__device__ __forceinline__ void ap_small_class(const ApSmallArgs& s, int c, uint4* lds4) {
  const ApArgs& a = s.ap;
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  constexpr int kWaves = kApSmallThreads / kWave;
  u64* keys = reinterpret_cast<u64*>(lds4);
  int32_t* idx = reinterpret_cast<int32_t*>(keys + s.sort_slots);
  uint8_t* s_rank = reinterpret_cast<uint8_t*>(idx + s.sort_slots);
  int8_t* s_flag = reinterpret_cast<int8_t*>(s_rank + kApChunk);
  // Gather this class's records in index order: every thread issues its
  // loads for all of its records (i = it * kApSmallThreads + tid) at once, then a
  // single barrier publishes the per-(round, wave) counts for the offsets.
  constexpr int kRounds = kApSmall / kApSmallThreads;
  __shared__ int round_count[kRounds][kWaves];
  // Scores are loaded with the ranks and classes (one round trip).
  // Indices past N are clamped (not branched around) so that all 3 x kRounds
  // loads are in flight together; their values are then discarded.
  uint32_t mine = 0;
  int rank[kRounds], klass[kRounds];
  double score[kRounds];
#pragma unroll
  for (int it = 0; it < kRounds; ++it) {
    rank[it] = 0;
    klass[it] = -1;
    score[it] = 0.0;
  }
  if (s.N > 0) {
#pragma unroll
    for (int it = 0; it < kRounds; ++it) {
      const int i = min(it * kApSmallThreads + tid, s.N - 1);
      rank[it] = a.ranks[i];
      klass[it] = s.classes[i];
      score[it] = s.scores[i];
    }
#pragma unroll
    for (int it = 0; it < kRounds; ++it)
      if (it * kApSmallThreads + tid >= s.N) klass[it] = -1;
  }
#pragma unroll
  for (int it = 0; it < kRounds; ++it)
    mine |= static_cast<uint32_t>(klass[it] == c && rank[it] > 0 && rank[it] <= 100) << it;
  u64 ballot[kRounds];
#pragma unroll
  for (int it = 0; it < kRounds; ++it) {
    ballot[it] = __ballot((mine >> it) & 1u);
    if (lane == 0) round_count[it][wave] = __popcll(ballot[it]);
  }
  __syncthreads();
  int n = 0;
#pragma unroll
  for (int it = 0; it < kRounds; ++it) {
    int before = n;
    for (int w = 0; w < kWaves; ++w) {
      const int count = round_count[it][w];
      if (w < wave) before += count;
      n += count;
    }
    if ((mine >> it) & 1u) {
      const int slot = before + __popcll(ballot[it] & ((1ULL << lane) - 1));
      idx[slot] = it * kApSmallThreads + tid;
      keys[slot] = score_key_desc(score[it]);
    }
  }
  __syncthreads();
  if (n <= kApSmallThreads) {
    // Up to one record per lane: its sorted position is the number of
    // records before it in (key, index) order, counted over broadcast LDS
    // reads -- one barrier instead of a network of them.
    // The record's rank and 40 flags load meanwhile and go straight to its
    // sorted place in the staging area.
    int pos = 0, mine_idx = 0, mine_rank = 0;
    uint64_t words[5] = {0, 0, 0, 0, 0};
    if (tid < n) {
      const u64 k = keys[tid];
      mine_idx = idx[tid];
      const uint64_t* row = reinterpret_cast<const uint64_t*>(a.status + static_cast<i64>(mine_idx) * 40);
#pragma unroll
      for (int w = 0; w < 5; ++w) words[w] = row[w];
      mine_rank = a.ranks[mine_idx];
      for (int j = 0; j < n; ++j) {
        const u64 kj = keys[j];
        pos += kj < k || (kj == k && idx[j] < mine_idx);
      }
    }
    __syncthreads();
    if (tid < n) {
      idx[pos] = mine_idx;
      uint64_t* flag_words = reinterpret_cast<uint64_t*>(s_flag) + pos * 5;
#pragma unroll
      for (int w = 0; w < 5; ++w) flag_words[w] = words[w];
      s_rank[pos] = static_cast<uint8_t>(mine_rank);
    }
    __syncthreads();
  } else {
    int m = 1;
    while (m < n) m <<= 1;
    for (int e = n + tid; e < m; e += kApSmallThreads) {
      keys[e] = ~0ULL;
      idx[e] = INT_MAX;
    }
    __syncthreads();
    // Bitonic sort of (key, index) ascending: descending score, then index.
    for (int k = 2; k <= m; k <<= 1)
      for (int j = k >> 1; j > 0; j >>= 1) {
        for (int e = tid; e < m; e += kApSmallThreads) {
          const int f = e ^ j;
          if (f <= e) continue;
          const u64 ke = keys[e], kf = keys[f];
          const int ie = idx[e], jf = idx[f];
          const bool greater = ke > kf || (ke == kf && ie > jf);
          if (greater == ((e & k) == 0)) {
            keys[e] = kf;
            keys[f] = ke;
            idx[e] = jf;
            idx[f] = ie;
          }
        }
        __syncthreads();
      }
  }
  uint32_t* codes = reinterpret_cast<uint32_t*>(s_flag + 40 * kApChunk);
  ap_class<kApSmallThreads>(a, c, static_cast<const int32_t*>(idx), 0, n, s_rank, s_flag, codes,
                            reinterpret_cast<int*>(codes + 101 * kApJobs), n <= kApSmallThreads);
}

__device__ __forceinline__ void summaries_if_last(const ApArgs& a);

__global__ __launch_bounds__(kApSmallThreads) void ap_small(ApSmallArgs s) {
  extern __shared__ uint4 lds4[];
  ap_small_class(s, blockIdx.x, lds4);
  if (s.ap.done) summaries_if_last(s.ap);
}

// The same, with the panoptic finaliser's rows as extra workgroups of the
// launch (both use 1024 threads), so the two run side by side.
static_assert(kApSmallThreads == kFinalThreads, "the AP classes and the finaliser share one launch shape");
__global__ __launch_bounds__(kApSmallThreads) void ap_small_final(ApSmallArgs s, FinalArgs f) {
  extern __shared__ uint4 lds4[];
  const int block = blockIdx.x;
  if (block < s.ap.C)
    ap_small_class(s, block, lds4);
  else
    finalize_row(f, block - s.ap.C, lds4);
}

size_t ap_small_lds(int slots) { return 12 * static_cast<size_t>(slots) + kApChunk + 40 * kApChunk + kApTableBytes; }

// instance.py:236-248: the 12 summaries as means over valid entries (>= 0),
// from the per-class accumulators; classes are folded in a fixed order.
// COHERENT: the accumulators were written by other workgroups of the same
// launch (read at the device's coherence point).  Threads 0..255.
template <bool COHERENT>
__device__ __forceinline__ void fold_summaries(int C, const double* class_part, double* out) {
  __shared__ double red[24][10];
  const int tid = threadIdx.x;
  if (tid < 240) {
    const int q = tid / 10, k = tid - q * 10;
    double acc = 0.0;
    constexpr int U = 8;
    for (int c0 = k; c0 < C; c0 += 10 * U) {
      double v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int c = min(c0 + 10 * u, C - 1);
        const double* p = class_part + static_cast<i64>(c) * 24 + q;
        v[u] = COHERENT ? __longlong_as_double(static_cast<long long>(__hip_atomic_load(
                              reinterpret_cast<const unsigned long long*>(p), __ATOMIC_RELAXED,
                              __HIP_MEMORY_SCOPE_AGENT)))
                        : *p;
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (c0 + 10 * u < C) acc += v[u];
    }
    red[q][k] = acc;
  }
  __syncthreads();
  if (tid < 12) {
    double s = 0.0, n = 0.0;
    for (int k = 0; k < 10; ++k) {
      s += red[2 * tid][k];
      n += red[2 * tid + 1][k];
    }
    out[tid] = n > 0 ? s / n : -1.0;
  }
}

__global__ __launch_bounds__(256) void ap_summaries(int C, const double* class_part, double* out) {
  fold_summaries<false>(C, class_part, out);
}

// With a completion counter: the class workgroup that finishes last folds the
// summaries (no separate launch) and zeroes the counter for the next launch.
__device__ __forceinline__ void summaries_if_last(const ApArgs& a) {
  __shared__ int last;
  __syncthreads();  // this workgroup's accumulators are written (exchanged)
  if (threadIdx.x == 0) last = atomicAdd(a.done, 1) == a.C - 1;
  __syncthreads();
  if (!last) return;
  fold_summaries<true>(a.C, a.partial, a.summary);
  if (threadIdx.x == 0) atomicExch(a.done, 0);
}

// Matching deferred to the AP launch: class workgroup c first matches class c
// in every image of an update (wave w takes images w, w + waves, ..., each
// wave in its own LDS region), then accumulates its records; the finaliser
// rows run as extra workgroups, and (with a completion counter) the last
// class workgroup folds the summaries.  One launch replaces the matcher, the
// AP classes, the summaries and the finaliser.
struct MatchPhase {
  MatchArgs m;
  int waves;       // matching waves per workgroup
  int wave_lds;    // bytes of LDS per matching wave
};

__global__ __launch_bounds__(kApSmallThreads) void ap_small_match_final(ApSmallArgs s, FinalArgs f, MatchPhase p) {
  extern __shared__ uint4 lds4[];
  const int block = blockIdx.x;
  if (block < s.ap.C) {
    const int wave = threadIdx.x / kWave;
    if (wave < p.waves)
      for (int b = wave; b < p.m.B; b += p.waves)
        match_grouped_wave(p.m, b, block, reinterpret_cast<char*>(lds4) + static_cast<size_t>(wave) * p.wave_lds);
    // Flags and positives written above are read by other waves of this
    // workgroup below (one compute unit: the barrier's workgroup-scope fence
    // suffices).
    __syncthreads();
    ap_small_class(s, block, lds4);
    if (s.ap.done) summaries_if_last(s.ap);
  } else {
    finalize_row(f, block - s.ap.C, lds4);
  }
}

}  // namespace pe

using namespace pe;

extern "C" int pe_mask_overlap(const uint8_t* pm, int64_t D, const uint8_t* gm, int64_t G, int64_t HW,
                               int32_t* inter, int32_t* pa, int32_t* ga, void* stream_ptr) {
  PE_REQUIRE(D >= 0 && G >= 0 && HW >= 0 && D < INT_MAX && G < INT_MAX && HW < INT_MAX, "invalid mask shapes");
  OverlapArgs o{};
  o.pm = pm;
  o.gm = gm;
  o.D = static_cast<int>(D);
  o.G = static_cast<int>(G);
  o.HW = HW;
  o.inter = inter;
  o.pa = pa;
  o.ga = ga;
  o.aligned = HW % 16 == 0 && reinterpret_cast<uintptr_t>(pm) % 16 == 0 && reinterpret_cast<uintptr_t>(gm) % 16 == 0;
  return launch_overlap(o, 1, D, G, HW, D, G, D * G, static_cast<hipStream_t>(stream_ptr));
}

extern "C" int pe_mask_overlap_batch(int64_t B, const uint8_t* const* pm, const uint8_t* const* gm,
                                     const int64_t* det_off, const int64_t* gt_off, const int64_t* inter_off,
                                     const int64_t* pixels, int64_t HW, int64_t max_d, int64_t max_g,
                                     int64_t max_hw, int64_t total_d, int64_t total_g, int64_t total_inter,
                                     int32_t* inter, int32_t* pa, int32_t* ga, int aligned, void* stream_ptr) {
  PE_REQUIRE(B >= 0 && B < INT_MAX && max_d >= 0 && max_d < INT_MAX && max_g >= 0 && max_g < INT_MAX &&
                 max_hw >= 0 && total_d >= 0 && total_g >= 0 && total_inter >= 0,
             "invalid mask batch shapes");
  PE_REQUIRE(pixels || HW == max_hw, "uniform batches pass HW == max_hw");
  OverlapArgs o{};
  o.pm_table = pm;
  o.gm_table = gm;
  o.det_off = reinterpret_cast<const i64*>(det_off);
  o.gt_off = reinterpret_cast<const i64*>(gt_off);
  o.inter_off = reinterpret_cast<const i64*>(inter_off);
  o.pixels = reinterpret_cast<const i64*>(pixels);
  o.HW = HW;
  o.inter = inter;
  o.pa = pa;
  o.ga = ga;
  o.aligned = aligned;
  return launch_overlap(o, B, max_d, max_g, max_hw, total_d, total_g, total_inter,
                        static_cast<hipStream_t>(stream_ptr));
}

extern "C" int64_t pe_mask_words(int64_t pixels) { return pixels < 0 ? -1 : mask_words(pixels); }

extern "C" int pe_pack_masks(const uint8_t* masks, int64_t N, int64_t HW, uint32_t* bits, void* stream_ptr) {
  PE_REQUIRE(N >= 0 && HW >= 0, "invalid mask shapes");
  const i64 stride = mask_words(HW), total = N * stride;
  if (total == 0) return kOk;
  const int aligned = HW % 16 == 0 && reinterpret_cast<uintptr_t>(masks) % 16 == 0;
  const i64 blocks = std::min<i64>((total + 255) / 256, 64LL * compute_units());
  pack_mask_bits<<<static_cast<unsigned>(blocks), 256, 0, static_cast<hipStream_t>(stream_ptr)>>>(
      masks, N, HW, stride, aligned, bits);
  PE_LAUNCH_CHECK();
  return kOk;
}

extern "C" int pe_mask_overlap_bits_batch(int64_t B, const uint32_t* const* pm, const uint32_t* const* gm,
                                          const int64_t* det_off, const int64_t* gt_off, const int64_t* inter_off,
                                          const int64_t* pixels, int64_t HW, int64_t max_d, int64_t max_g,
                                          int64_t max_hw, int64_t total_d, int64_t total_g, int64_t total_inter,
                                          int32_t* inter, int32_t* pa, int32_t* ga, const int32_t* const* pm_ext,
                                          const int32_t* const* gm_ext, void* stream_ptr) {
  PE_REQUIRE(B >= 0 && B < INT_MAX && max_d >= 0 && max_d < INT_MAX && max_g >= 0 && max_g < INT_MAX &&
                 max_hw >= 0 && total_d >= 0 && total_g >= 0 && total_inter >= 0,
             "invalid mask batch shapes");
  PE_REQUIRE(pixels || HW == max_hw, "uniform batches pass HW == max_hw");
  PE_REQUIRE(B == 0 || (pm && gm && det_off && gt_off && inter_off), "missing batch tables");
  OverlapArgs o{};
  o.pm_table = reinterpret_cast<const uint8_t* const*>(pm);
  o.gm_table = reinterpret_cast<const uint8_t* const*>(gm);
  o.det_off = reinterpret_cast<const i64*>(det_off);
  o.gt_off = reinterpret_cast<const i64*>(gt_off);
  o.inter_off = reinterpret_cast<const i64*>(inter_off);
  o.pixels = reinterpret_cast<const i64*>(pixels);
  o.HW = HW;
  o.inter = inter;
  o.pa = pa;
  o.ga = ga;
  o.pm_ext = pm_ext;
  o.gm_ext = gm_ext;
  return launch_overlap_bits(o, B, max_d, max_g, max_hw, total_d, total_g, total_inter,
                             static_cast<hipStream_t>(stream_ptr));
}

extern "C" int64_t pe_detection_histogram_workspace_bytes(int64_t B, int64_t packed_pixels) {
  if (packed_pixels <= 0) return 0;
  return static_cast<int64_t>(align_up(4 * static_cast<size_t>(tile_list_ints(B, packed_pixels, 16 * kDetThreads)),
                                       256));
}

static int detection_histogram_launch(int64_t B, const uint8_t* const* masks, const int64_t* det_off,
                                      const int64_t* order, const int32_t* target, const int64_t* pix_off,
                                      int64_t HW, int64_t S, int64_t max_dets, int64_t total_dets,
                                      int64_t* pairs, int pairs_zeroed, int aligned, bool bits, void* workspace,
                                      int64_t workspace_bytes, void* stream_ptr,
                                      const int32_t* const* extents = nullptr) {
  PE_REQUIRE(B >= 0 && B < INT_MAX && HW > 0 && S > 0 && S < INT_MAX && max_dets >= 0 && total_dets >= 0,
             "invalid shapes");
  PE_REQUIRE(!pix_off || workspace_bytes >= pe_detection_histogram_workspace_bytes(B, HW),
             "workspace too small for a packed batch");
  hipStream_t stream = static_cast<hipStream_t>(stream_ptr);
  int rc;
  if (total_dets && !pairs_zeroed && (rc = launch_zero(pairs, 8 * static_cast<size_t>(total_dets * S), stream)))
    return rc;
  if (B == 0 || max_dets == 0) return kOk;
  DetHistArgs a{};
  a.masks = masks;
  a.extents = bits ? extents : nullptr;
  a.det_off = reinterpret_cast<const i64*>(det_off);
  a.order = reinterpret_cast<const i64*>(order);
  a.target = target;
  a.B = static_cast<int>(B);
  a.S = static_cast<int>(S);
  const DetVariant& v = bits ? kDetHistBits[tuned_variant(kTuneDetectionHistogramBits)]
                             : kDetHist[tuned_variant(kTuneDetectionHistogram)];
  const i64 tile = v.tile;
  // Detections per group: packed 16-bit counters of group x S plus one 8-byte
  // row offset per detection within the LDS budget (a tile holds at most
  // 16384 pixels, so no counter exceeds 0xFFFF).
  i64 group = static_cast<i64>(kDetMaxLds) / (2 * S + 8);
  PE_REQUIRE(group >= 1, "too many target slots for the detection histogram");
  if (detection_group_limit() > 0) group = std::min<i64>(group, detection_group_limit());
  group = std::min<i64>(group, max_dets);
  a.group = static_cast<int>(group);
  a.dgroups = static_cast<int>((max_dets + group - 1) / group);
  a.pairs = reinterpret_cast<u64*>(pairs);
  a.aligned = aligned;
  TileList& t = a.tiles;
  t.tile = static_cast<int>(tile);
  i64 blocks;
  bool persist = false;
  if (pix_off) {
    // HW is the pixel total of the packed batch.
    const i64 bound = tile_list_bound(B, HW, tile);
    PE_REQUIRE(bound * a.dgroups < INT_MAX, "too many detection tiles");
    int* ws = static_cast<int*>(workspace);
    t.pix_off = reinterpret_cast<const i64*>(pix_off);
    t.tile_off = ws;
    t.tile_image = ws + B + 1;
    if ((rc = launch_tile_list(t.pix_off, a.B, t.tile, ws, ws + B + 1, stream))) return rc;
    // The exact item count lives on the device: launch the bound, or a
    // persistent grid of a few workgroups per compute unit.
    blocks = bound * a.dgroups;
    const int per_cu = packed_grid_per_cu();
    if (per_cu > 0) {
      blocks = std::min<i64>(blocks, static_cast<i64>(per_cu) * compute_units());
      persist = true;
    }
  } else {
    t.HW = HW;
    t.tiles_per_image = static_cast<int>((HW + tile - 1) / tile);
    blocks = B * t.tiles_per_image * a.dgroups;
  }
  PE_REQUIRE(blocks < INT_MAX, "too many workgroups");
  // Bins, row offsets, and (bit-packed masks) the list of detections that
  // meet the tile plus its count.
  const size_t lds = static_cast<size_t>(det_bins4(group, S)) * 16 + 8 * static_cast<size_t>(group) +
                     (bits ? 4 * static_cast<size_t>(group + 1) : 0);
  auto kernel = persist ? v.persistent : v.kernel;
  if (lds > 64 * 1024)
    PE_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                               static_cast<int>(lds)));
  ProfileScope timing("detection_histogram", stream);
  kernel<<<static_cast<unsigned>(blocks), kDetThreads, lds, stream>>>(a);
  PE_LAUNCH_CHECK();
  return kOk;
}

extern "C" int pe_detection_histogram(int64_t B, const uint8_t* const* masks, const int64_t* det_off,
                                      const int64_t* order, const int32_t* target, const int64_t* pix_off,
                                      int64_t HW, int64_t S, int64_t max_dets, int64_t total_dets,
                                      int64_t* pairs, int pairs_zeroed, int aligned, void* workspace,
                                      int64_t workspace_bytes, void* stream_ptr) {
  return detection_histogram_launch(B, masks, det_off, order, target, pix_off, HW, S, max_dets, total_dets, pairs,
                                    pairs_zeroed, aligned, false, workspace, workspace_bytes, stream_ptr);
}

extern "C" int pe_detection_histogram_bits(int64_t B, const uint32_t* const* masks, const int64_t* det_off,
                                           const int64_t* order, const int32_t* target, const int64_t* pix_off,
                                           int64_t HW, int64_t S, int64_t max_dets, int64_t total_dets,
                                           int64_t* pairs, int pairs_zeroed, const int32_t* const* extents,
                                           void* workspace, int64_t workspace_bytes, void* stream_ptr) {
  return detection_histogram_launch(B, reinterpret_cast<const uint8_t* const*>(masks), det_off, order, target,
                                    pix_off, HW, S, max_dets, total_dets, pairs, pairs_zeroed, 1, true, workspace,
                                    workspace_bytes, stream_ptr, extents);
}

// [first, end) of the nonzero words of every bit row (first = end = 0 for an
// empty row): one workgroup per row, 16-byte loads, a workgroup min / max.
__global__ __launch_bounds__(256) void mask_extents(const uint32_t* bits, i64 row_words, int32_t* extents) {
  const i64 n = blockIdx.x;
  const auto* row = reinterpret_cast<const __attribute__((address_space(1))) v4i*>(as_global(bits + n * row_words));
  int lo = INT_MAX, hi = -1;
  for (i64 q = threadIdx.x; q < row_words / 4; q += 256) {
    const v4i v = row[q];
    const int w = static_cast<int>(4 * q);
    if (v.x) { lo = min(lo, w); hi = max(hi, w); }
    if (v.y) { lo = min(lo, w + 1); hi = max(hi, w + 1); }
    if (v.z) { lo = min(lo, w + 2); hi = max(hi, w + 2); }
    if (v.w) { lo = min(lo, w + 3); hi = max(hi, w + 3); }
  }
  for (int m = 1; m < kWave; m <<= 1) {
    lo = min(lo, __shfl_xor(lo, m, kWave));
    hi = max(hi, __shfl_xor(hi, m, kWave));
  }
  __shared__ int s_lo[4], s_hi[4];
  const int wave = threadIdx.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) {
    s_lo[wave] = lo;
    s_hi[wave] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < 4; ++k) {
      lo = min(lo, s_lo[k]);
      hi = max(hi, s_hi[k]);
    }
    extents[2 * n] = hi < 0 ? 0 : lo;
    extents[2 * n + 1] = hi < 0 ? 0 : hi + 1;
  }
}

extern "C" int pe_mask_extents(const uint32_t* bits, int64_t N, int64_t row_words, int32_t* extents,
                               void* stream_ptr) {
  PE_REQUIRE(N >= 0 && N < INT_MAX && row_words >= 0 && row_words % 4 == 0 && row_words < INT_MAX,
             "invalid bit-row shapes");
  PE_REQUIRE(N == 0 || (bits && extents), "null bit rows or extents");
  PE_REQUIRE(reinterpret_cast<uintptr_t>(bits) % 16 == 0, "bit rows must be 16-byte aligned");
  if (N == 0) return kOk;
  mask_extents<<<static_cast<unsigned>(N), 256, 0, static_cast<hipStream_t>(stream_ptr)>>>(bits, row_words, extents);
  PE_LAUNCH_CHECK();
  return kOk;
}

extern "C" int pe_coco_match(int64_t B, int64_t C, const int64_t* det_off, const int64_t* gt_off,
                             const int64_t* inter_off, const int32_t* det_cls, const double* det_area,
                             const int64_t* det_row, const int32_t* gt_cls, const double* gt_area,
                             const uint8_t* gt_crowd, const double* gt_ann, const void* inter, int inter_bytes,
                             int64_t total_dets, int64_t max_gt, int8_t* status, int32_t* ranks, int64_t* counts,
                             const int32_t* guard, const uint8_t* gt_things, const int64_t* gt_area_i64,
                             int64_t gt_area_stride, int zero_inter, void* stream_ptr) {
  PE_REQUIRE(B >= 0 && C > 0 && C < INT_MAX && B * C < INT_MAX && max_gt >= 0, "invalid matching shapes");
  PE_REQUIRE(inter_bytes == 4 || inter_bytes == 8, "intersections must be int32 or int64");
  PE_REQUIRE(det_area || inter_bytes == 8 || total_dets == 0,
             "detection areas may be omitted only for int64 intersection rows");
  PE_REQUIRE(!zero_inter || inter_bytes == 8, "only int64 intersection rows are zeroed");
  PE_REQUIRE(max_gt == 0 || gt_area || gt_area_i64, "GT pixel areas are required");
  hipStream_t stream = static_cast<hipStream_t>(stream_ptr);
  // The kernel writes every detection's rank and flags (rank 0 and flags -1
  // past the first 100 of a class, as on the CPU path).
  if (B == 0) return kOk;
  MatchArgs a{};
  a.B = static_cast<int>(B);
  a.C = static_cast<int>(C);
  a.det_off = reinterpret_cast<const i64*>(det_off);
  a.gt_off = reinterpret_cast<const i64*>(gt_off);
  a.inter_off = reinterpret_cast<const i64*>(inter_off);
  a.det_cls = det_cls;
  a.det_area = det_area;
  a.det_row = reinterpret_cast<const i64*>(det_row);
  a.gt_cls = gt_cls;
  a.gt_area = gt_area;
  a.gt_crowd = gt_crowd;
  a.gt_ann = gt_ann;
  a.inter = inter;
  a.inter_bytes = inter_bytes;
  a.status = status;
  a.ranks = ranks;
  a.counts = reinterpret_cast<i64*>(counts);
  a.max_gt = static_cast<int>(max_gt);
  a.words = static_cast<int>((max_gt + 31) / 32);
  a.guard = guard;
  a.gt_things = gt_things;
  a.gt_area_i64 = reinterpret_cast<const i64*>(gt_area_i64);
  a.gt_area_stride = gt_area_stride;
  a.zero_inter = zero_inter;
  const size_t lds = 16 * 100 + 4 * (100 + static_cast<size_t>(max_gt)) + kMatchStageBytes +
                     4 * 40 * static_cast<size_t>(a.words) + 16;
  PE_REQUIRE(lds <= 150 * 1024, "too many ground-truth instances per image for the matcher");
  if (lds > 64 * 1024)
    PE_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(coco_match), hipFuncAttributeMaxDynamicSharedMemorySize,
                               static_cast<int>(lds)));
  ProfileScope timing("coco_match", stream);
  coco_match<<<static_cast<unsigned>(B * C), 64, lds, stream>>>(a);
  PE_LAUNCH_CHECK();
  return kOk;
}

namespace {
struct GroupBuffers {
  int4* hdr;
  int2* det_list;
  int32_t* gt_list;
};

GroupBuffers carve_groups(void* base, i64 B, i64 C, i64 total_dets, i64 S, size_t* used = nullptr) {
  Carve carve(base);
  GroupBuffers g{};
  g.hdr = carve.take<int4>(B * (C + 1));
  g.det_list = carve.take<int2>(total_dets);
  g.gt_list = carve.take<int32_t>(B * S);
  if (used) *used = carve.used;
  return g;
}

int pow2_at_least(i64 n, int floor) {
  int m = floor;
  while (m < n) m <<= 1;
  return m;
}
}  // namespace

extern "C" int64_t pe_match_groups_bytes(int64_t B, int64_t C, int64_t total_dets, int64_t S) {
  if (B < 0 || C < 0 || total_dets < 0 || S < 0) return -1;
  size_t used = 0;
  carve_groups(nullptr, B, C, total_dets, S, &used);
  return static_cast<int64_t>(align_up(used, 256));
}

extern "C" int pe_match_groups(int64_t B, int64_t C, int64_t S, const int64_t* det_off, int64_t total_dets,
                               int64_t max_dets, const int32_t* det_cls, const int64_t* det_row,
                               const int32_t* gt_cls, const uint8_t* gt_things, int8_t* status, int32_t* ranks,
                               void* groups, int64_t groups_bytes, void* stream_ptr) {
  PE_REQUIRE(B >= 0 && B < INT_MAX && C > 0 && S >= 0 && total_dets >= 0 && max_dets >= 0,
             "invalid grouping shapes");
  PE_REQUIRE(max_dets <= kGroupMaxEntries && S <= kGroupMaxEntries && C < (1 << (32 - kGroupIndexBits)) - 1,
             "grouping holds at most 4096 detections and target slots per image");
  PE_REQUIRE(reinterpret_cast<uintptr_t>(status) % 8 == 0, "match flags must be 8-byte aligned");
  PE_REQUIRE(groups_bytes >= pe_match_groups_bytes(B, C, total_dets, S), "group buffer too small");
  const int det_slots = pow2_at_least(max_dets, 64), gt_slots = pow2_at_least(S, 64);
  const size_t lds = 4 * static_cast<size_t>(det_slots + gt_slots) + 16 * static_cast<size_t>(C + 1);
  PE_REQUIRE(lds <= 160 * 1024, "too many classes for the grouped matcher");
  if (B == 0) return kOk;
  const GroupBuffers w = carve_groups(groups, B, C, total_dets, S);
  GroupArgs g{};
  g.B = static_cast<int>(B);
  g.C = static_cast<int>(C);
  g.S = static_cast<int>(S);
  g.det_off = reinterpret_cast<const i64*>(det_off);
  g.det_cls = det_cls;
  g.det_row = reinterpret_cast<const i64*>(det_row);
  g.gt_cls = gt_cls;
  g.gt_things = gt_things;
  g.status = status;
  g.ranks = ranks;
  g.hdr = w.hdr;
  g.det_list = w.det_list;
  g.gt_list = w.gt_list;
  g.det_slots = det_slots;
  g.gt_slots = gt_slots;
  hipStream_t stream = static_cast<hipStream_t>(stream_ptr);
  if (lds > 64 * 1024)
    PE_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(match_groups),
                               hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
  match_groups<<<static_cast<unsigned>(B), 256, lds, stream>>>(g);
  PE_LAUNCH_CHECK();
  return kOk;
}

namespace {
int make_match_grouped(int64_t B, int64_t C, int64_t S, const int64_t* det_off, int64_t total_dets,
                       const void* groups, int64_t groups_bytes, const uint8_t* gt_crowd, const double* gt_ann,
                       const int64_t* gt_area_i64, int64_t gt_area_stride, int64_t* pairs, int8_t* status,
                       int64_t* counts, MatchArgs* out) {
  PE_REQUIRE(B >= 0 && C > 0 && C < INT_MAX && B * C < INT_MAX && S >= 0 && S <= kGroupMaxEntries &&
                 total_dets >= 0,
             "invalid matching shapes");
  PE_REQUIRE(groups_bytes >= pe_match_groups_bytes(B, C, total_dets, S), "group buffer too small");
  PE_REQUIRE(S == 0 || gt_area_i64, "GT pixel areas are required");
  PE_REQUIRE(match_wave_lds(static_cast<int>(S)) <= 150 * 1024, "too many target slots per image for the matcher");
  const GroupBuffers w = carve_groups(const_cast<void*>(groups), B, C, total_dets, S);
  MatchArgs a{};
  a.B = static_cast<int>(B);
  a.C = static_cast<int>(C);
  a.S = static_cast<int>(S);
  a.det_off = reinterpret_cast<const i64*>(det_off);
  a.gt_crowd = gt_crowd;
  a.gt_ann = gt_ann;
  a.gt_area_i64 = reinterpret_cast<const i64*>(gt_area_i64);
  a.gt_area_stride = gt_area_stride;
  a.inter = pairs;
  a.inter_bytes = 8;
  a.status = status;
  a.counts = reinterpret_cast<i64*>(counts);
  a.max_gt = static_cast<int>(S);
  a.words = static_cast<int>((S + 31) / 32);
  a.zero_inter = 1;
  a.group_hdr = w.hdr;
  a.group_det = w.det_list;
  a.group_gt = w.gt_list;
  *out = a;
  return kOk;
}
}  // namespace

int launch_match_grouped(const MatchArgs& a, hipStream_t stream) {
  if (a.B == 0) return kOk;
  const size_t lds = match_wave_lds(a.S);
  if (lds > 64 * 1024)
    PE_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(coco_match_grouped),
                               hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
  ProfileScope timing("coco_match", stream);
  coco_match_grouped<<<static_cast<unsigned>(static_cast<i64>(a.B) * a.C), 64, lds, stream>>>(a);
  PE_LAUNCH_CHECK();
  return kOk;
}

extern "C" int pe_coco_match_grouped(int64_t B, int64_t C, int64_t S, const int64_t* det_off, int64_t total_dets,
                                     const void* groups, int64_t groups_bytes, const uint8_t* gt_crowd,
                                     const double* gt_ann, const int64_t* gt_area_i64, int64_t gt_area_stride,
                                     int64_t* pairs, int8_t* status, int64_t* counts, void* stream_ptr) {
  MatchArgs a;
  const int rc = make_match_grouped(B, C, S, det_off, total_dets, groups, groups_bytes, gt_crowd, gt_ann,
                                    gt_area_i64, gt_area_stride, pairs, status, counts, &a);
  if (rc) return rc;
  return launch_match_grouped(a, static_cast<hipStream_t>(stream_ptr));
}

// Images with at most kOrderLds detections: one workgroup per image sorts
// (score key, index) pairs with a bitonic network in LDS and writes the
// outputs directly (one launch instead of keys + segmented sort + gather).
constexpr int kOrderLds = 4096;

// Per-image input arrays passed by value in the kernel arguments (no
// concatenated copy and no device table): image b's scores (f64) and classes
// (int32 or int64) start at scores[b] and classes[b].
constexpr int kOrderArgImages = 64;
struct ImageArrays {
  const double* scores[kOrderArgImages];
  const void* classes[kOrderArgImages];
  int class_bytes;
};

// Scores and classes come either concatenated (scores/classes, image b at
// det_off[b]) or per image (arrays, when ARRAYS).
template <bool ARRAYS>
__global__ __launch_bounds__(1024) void score_order_lds(const i64* det_off, const double* scores, const i64* classes,
                                                        ImageArrays arrays, int slots, i64* order,
                                                        double* sorted_scores, i64* sorted_classes,
                                                        int32_t* sorted_classes32) {
  extern __shared__ uint4 lds4[];
  u64* keys = reinterpret_cast<u64*>(lds4);
  int32_t* idx = reinterpret_cast<int32_t*>(keys + slots);
  const int b = blockIdx.x, tid = threadIdx.x, nthreads = blockDim.x;
  const i64 first = det_off[b];
  const int D = static_cast<int>(det_off[b + 1] - first);
  if (D <= 0) return;  // uniform
  if (D > slots) {
    // The caller's max_dets understated this image's count: the LDS holds
    // only `slots` entries, so nothing is sorted and every order entry of the
    // image is -1 (an index no caller accepts).
    for (int e = tid; e < D; e += nthreads) order[first + e] = -1;
    return;
  }
  const double* s = ARRAYS ? arrays.scores[b] : scores + first;
  int m = 1;
  while (m < D) m <<= 1;
  for (int e = tid; e < m; e += nthreads) {
    keys[e] = e < D ? score_key_desc(s[e]) : ~0ULL;
    idx[e] = e < D ? e : INT_MAX;
  }
  __syncthreads();
  // Ascending (key, index): descending score, ties in input order.
  for (int k = 2; k <= m; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int e = tid; e < m; e += nthreads) {
        const int f = e ^ j;
        if (f <= e) continue;
        const u64 ke = keys[e], kf = keys[f];
        const int ie = idx[e], jf = idx[f];
        const bool greater = ke > kf || (ke == kf && ie > jf);
        if (greater == ((e & k) == 0)) {
          keys[e] = kf;
          keys[f] = ke;
          idx[e] = jf;
          idx[f] = ie;
        }
      }
      __syncthreads();
    }
  for (int e = tid; e < D; e += nthreads) {
    const int src = idx[e];
    order[first + e] = src;
    sorted_scores[first + e] = s[src];
    i64 c;
    if (!ARRAYS)
      c = classes[first + src];
    else if (arrays.class_bytes == 8)
      c = static_cast<const i64*>(arrays.classes[b])[src];
    else
      c = static_cast<const int32_t*>(arrays.classes[b])[src];
    sorted_classes[first + e] = c;
    sorted_classes32[first + e] = static_cast<int32_t>(c);
  }
}

static int launch_score_order_lds(bool arrays, int64_t B, int64_t max_dets, const int64_t* det_off,
                                  const double* scores, const int64_t* classes, const ImageArrays& images,
                                  int64_t* order, double* sorted_scores, int64_t* sorted_classes,
                                  int32_t* sorted_classes32, hipStream_t stream) {
  int slots = 64;
  while (slots < max_dets) slots <<= 1;
  // One element per lane and stage; narrower beside another stream's kernel.
  const int threads = std::min(stream_concurrent(stream) ? kConcurrentThreads : 1024, slots);
  const size_t lds = 12 * static_cast<size_t>(slots);
  auto kernel = arrays ? score_order_lds<true> : score_order_lds<false>;
  if (lds > 64 * 1024)
    PE_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                               static_cast<int>(lds)));
  kernel<<<static_cast<unsigned>(B), threads, lds, stream>>>(
      reinterpret_cast<const i64*>(det_off), scores, reinterpret_cast<const i64*>(classes), images, slots,
      reinterpret_cast<i64*>(order), sorted_scores, reinterpret_cast<i64*>(sorted_classes), sorted_classes32);
  PE_LAUNCH_CHECK();
  return kOk;
}

namespace {
struct ApBuffers {
  SortItem* items;
  SortItem* spare;
  int32_t* order;
  int32_t* start;
  int32_t* end;
  double* partial;
};

ApBuffers carve_ap(void* workspace, i64 N, i64 C, size_t* used = nullptr) {
  Carve carve(workspace);
  ApBuffers w{};
  // The per-class gather of small batches needs no sort buffers.
  const i64 sorted = N > kApSmall ? N : 0;
  w.items = carve.take<SortItem>(sorted);
  w.spare = carve.take<SortItem>(sorted);
  w.order = carve.take<int32_t>(sorted);
  w.start = carve.take<int32_t>(C);
  w.end = carve.take<int32_t>(C);
  w.partial = carve.take<double>(C * 24);
  if (used) *used = carve.used;
  return w;
}
}  // namespace

extern "C" int64_t pe_ap_workspace_bytes(int64_t N, int64_t C) {
  size_t used = 0;
  carve_ap(nullptr, N, C, &used);
  return static_cast<int64_t>(align_up(used, 256));
}

// AP accumulation; with `fin`, also the K panoptic finaliser rows (in the
// same launch as the AP classes for small batches, else just before).
static int ap_compute(int64_t N, const double* scores, const int32_t* classes, const int32_t* ranks,
                      const int8_t* status, const int64_t* counts, int64_t C, double* precision, double* recall,
                      double* summary, void* workspace, int64_t workspace_bytes, void* stream_ptr,
                      const FinalArgs* fin, int64_t K, const MatchArgs* match = nullptr, int32_t* done = nullptr) {
  PE_REQUIRE(N >= 0 && N < INT_MAX && C > 0 && C <= INT_MAX / 240, "invalid AP shapes");
  PE_REQUIRE(workspace_bytes >= pe_ap_workspace_bytes(N, C), "workspace too small");
  PE_REQUIRE(reinterpret_cast<uintptr_t>(status) % 8 == 0, "match flags must be 8-byte aligned");
  hipStream_t stream = static_cast<hipStream_t>(stream_ptr);
  ApBuffers w = carve_ap(workspace, N, C);
  int rc;
  ApArgs a{};
  a.C = static_cast<int>(C);
  a.order = w.order;
  a.start = w.start;
  a.end = w.end;
  a.ranks = ranks;
  a.status = status;
  a.counts = reinterpret_cast<const i64*>(counts);
  a.precision = precision;
  a.recall = recall;
  a.partial = w.partial;
  if (N <= kApSmall) {
    // One launch: every class gathers and sorts its own records in LDS.
    ApSmallArgs sa{};
    sa.ap = a;
    sa.scores = scores;
    sa.classes = classes;
    sa.N = static_cast<int>(N);
    sa.sort_slots = 64;  // keeps the flag staging area 16-byte aligned
    while (sa.sort_slots < N) sa.sort_slots <<= 1;
    size_t lds = ap_small_lds(sa.sort_slots);
    MatchPhase phase{};
    if (match && fin && K > 0) {
      // As many matching waves as fit beside nothing else in LDS (at most
      // the workgroup's 16); the AP phase reuses the same LDS afterwards.
      phase.m = *match;
      phase.wave_lds = static_cast<int>(match_wave_lds(match->S));
      phase.waves = static_cast<int>(std::min<size_t>(kApSmallThreads / kWave, (150 * 1024) / phase.wave_lds));
    }
    if (phase.waves > 0) {
      lds = std::max({lds, 16 * static_cast<size_t>(fin->C), static_cast<size_t>(phase.waves) * phase.wave_lds});
      PE_REQUIRE(lds <= 160 * 1024, "too many categories for the finaliser");
      PE_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(ap_small_match_final),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
      // With a completion counter the last class workgroup also folds the
      // summaries.
      sa.ap.done = done;
      sa.ap.summary = summary;
      ProfileScope timing("ap_accumulate", stream);
      ap_small_match_final<<<static_cast<unsigned>(C + K), kApSmallThreads, lds, stream>>>(sa, *fin, phase);
      PE_LAUNCH_CHECK();
      if (done) return kOk;
    } else if (fin && K > 0) {
      if (match && (rc = launch_match_grouped(*match, stream))) return rc;
      lds = std::max(lds, 16 * static_cast<size_t>(fin->C));
      PE_REQUIRE(lds <= 160 * 1024, "too many categories for the finaliser");
      PE_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(ap_small_final),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
      ProfileScope timing("ap_accumulate", stream);
      ap_small_final<<<static_cast<unsigned>(C + K), kApSmallThreads, lds, stream>>>(sa, *fin);
      PE_LAUNCH_CHECK();
    } else {
      if (match && (rc = launch_match_grouped(*match, stream))) return rc;
      PE_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(ap_small), hipFuncAttributeMaxDynamicSharedMemorySize,
                                 static_cast<int>(lds)));
      // With a completion counter the last class workgroup folds the
      // summaries.
      sa.ap.done = done;
      sa.ap.summary = summary;
      ProfileScope timing("ap_accumulate", stream);
      ap_small<<<static_cast<unsigned>(C), kApSmallThreads, lds, stream>>>(sa);
      PE_LAUNCH_CHECK();
      if (done) return kOk;
    }
    ap_summaries<<<1, 256, 0, stream>>>(static_cast<int>(C), w.partial, summary);
    PE_LAUNCH_CHECK();
    return kOk;
  }
  if (match && (rc = launch_match_grouped(*match, stream))) return rc;
  if (fin && (rc = launch_finalize(*fin, K, stream))) return rc;
  if ((rc = launch_zero(w.start, 4 * static_cast<size_t>(C), stream))) return rc;
  if ((rc = launch_zero(w.end, 4 * static_cast<size_t>(C), stream))) return rc;
  {
    // One sort by (class, descending score, record index).
    const int blocks = static_cast<int>(std::min<i64>((N + 255) / 256, 4096));
    ap_items<<<blocks, 256, 0, stream>>>(scores, classes, ranks, N, static_cast<int>(C), w.items);
    PE_LAUNCH_CHECK();
    SortItem* sorted = nullptr;
    if ((rc = sort_items(w.items, w.spare, N, stream, &sorted))) return rc;
    ap_order<<<blocks, 256, 0, stream>>>(sorted, N, static_cast<int>(C), w.order, w.start, w.end);
    PE_LAUNCH_CHECK();
  }
  {
    PE_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(ap_accumulate),
                               hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(kApAccumulateLds)));
    ProfileScope timing("ap_accumulate", stream);
    ap_accumulate<<<static_cast<unsigned>(C), kApThreads, kApAccumulateLds, stream>>>(a);
    PE_LAUNCH_CHECK();
  }
  ap_summaries<<<1, 256, 0, stream>>>(static_cast<int>(C), w.partial, summary);
  PE_LAUNCH_CHECK();
  return kOk;
}

extern "C" int pe_ap_compute_done(int64_t N, const double* scores, const int32_t* classes, const int32_t* ranks,
                                  const int8_t* status, const int64_t* counts, int64_t C, double* precision,
                                  double* recall, double* summary, void* workspace, int64_t workspace_bytes,
                                  int32_t* done, void* stream_ptr) {
  PE_REQUIRE(done != nullptr, "null completion counter");
  return ap_compute(N, scores, classes, ranks, status, counts, C, precision, recall, summary, workspace,
                    workspace_bytes, stream_ptr, nullptr, 0, nullptr, done);
}

extern "C" int pe_ap_compute(int64_t N, const double* scores, const int32_t* classes, const int32_t* ranks,
                             const int8_t* status, const int64_t* counts, int64_t C, double* precision,
                             double* recall, double* summary, void* workspace, int64_t workspace_bytes,
                             void* stream_ptr) {
  return ap_compute(N, scores, classes, ranks, status, counts, C, precision, recall, summary, workspace,
                    workspace_bytes, stream_ptr, nullptr, 0);
}

extern "C" int pe_compute_with_ap(const int64_t* tp, const int64_t* fp, const int64_t* fn, const double* sums,
                                  const int64_t* confusion, const uint8_t* things, int64_t K, int64_t C,
                                  double* per_class, double* summaries, int64_t* counts, double* miou,
                                  int64_t* out_tp, int64_t* out_fp, int64_t* out_fn, double* out_sums,
                                  int64_t* out_confusion, int64_t N, const double* scores, const int32_t* classes,
                                  const int32_t* ranks, const int8_t* status, const int64_t* ap_counts,
                                  int64_t ap_C, double* precision, double* recall, double* summary,
                                  void* workspace, int64_t workspace_bytes, void* stream) {
  FinalArgs fin;
  const int rc = make_final_args(tp, fp, fn, sums, confusion, things, K, C, per_class, summaries, counts, miou,
                                 out_tp, out_fp, out_fn, out_sums, out_confusion, &fin);
  if (rc) return rc;
  return ap_compute(N, scores, classes, ranks, status, ap_counts, ap_C, precision, recall, summary, workspace,
                    workspace_bytes, stream, &fin, K);
}

extern "C" int pe_compute_with_ap_match(
    const int64_t* tp, const int64_t* fp, const int64_t* fn, const double* sums, const int64_t* confusion,
    const uint8_t* things, int64_t K, int64_t C, double* per_class, double* summaries, int64_t* counts, double* miou,
    int64_t* out_tp, int64_t* out_fp, int64_t* out_fn, double* out_sums, int64_t* out_confusion, int64_t N,
    const double* scores, const int32_t* classes, const int32_t* ranks, int8_t* status, int64_t* ap_counts,
    int64_t ap_C, double* precision, double* recall, double* summary, void* workspace, int64_t workspace_bytes,
    int64_t B, int64_t S, const int64_t* det_off, const void* groups, int64_t groups_bytes,
    const uint8_t* gt_crowd, const double* gt_ann, const int64_t* gt_area_i64, int64_t gt_area_stride,
    int64_t* pairs, int32_t* done, void* stream) {
  FinalArgs fin;
  int rc = make_final_args(tp, fp, fn, sums, confusion, things, K, C, per_class, summaries, counts, miou, out_tp,
                           out_fp, out_fn, out_sums, out_confusion, &fin);
  if (rc) return rc;
  MatchArgs match;
  if ((rc = make_match_grouped(B, ap_C, S, det_off, N, groups, groups_bytes, gt_crowd, gt_ann, gt_area_i64,
                               gt_area_stride, pairs, status, ap_counts, &match)))
    return rc;
  return ap_compute(N, scores, classes, ranks, status, ap_counts, ap_C, precision, recall, summary, workspace,
                    workspace_bytes, stream, &fin, K, &match, done);
}

// ------------------------------------------------------- small helpers

namespace pe {

// out[r] = sum of row r (int64 counts) as FP64: detection areas of the shared
// target histogram (combined.py:107-108).
__global__ __launch_bounds__(256) void row_sums(const i64* pairs, i64 rows, int cols, double* out) {
  const i64 r = (static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (r >= rows) return;
  i64 s = 0;
  for (int c = lane; c < cols; c += kWave) s += pairs[r * cols + c];
  s = wave_sum(s);
  if (lane == 0) out[r] = static_cast<double>(s);
}

// Instance input validation as an error bitmask (instance.py:104-126,
// combined.py:47-61, :92-93, :109-110):
//   1 non-finite score or annotation area, or negative annotation area
//   2 class outside 0..C-1        4 mask without pixels
//   8 detection class not a thing 16 instance_areas not finite / negative
__global__ __launch_bounds__(256) void instance_checks(i64 n_det, const i64* det_cls, const double* scores,
                                                       const double* det_area, i64 n_gt, const i64* gt_cls,
                                                       const double* gt_area, const double* ann, int C,
                                                       const uint8_t* things, i64 n_inst, const double* inst_areas,
                                                       int* error) {
  int bits = 0;
  const i64 stride = static_cast<i64>(gridDim.x) * blockDim.x;
  for (i64 i = static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x; i < n_det; i += stride) {
    const i64 c = det_cls[i];
    if (c < 0 || c >= C) bits |= 2;
    else if (things && !things[c]) bits |= 8;
    if (scores && !isfinite(scores[i])) bits |= 1;
    if (det_area && det_area[i] == 0.0) bits |= 4;
  }
  for (i64 j = static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x; j < n_gt; j += stride) {
    const i64 c = gt_cls[j];
    if (c < 0 || c >= C) bits |= 2;
    if (gt_area && gt_area[j] == 0.0) bits |= 4;
    if (ann && (!isfinite(ann[j]) || ann[j] < 0.0)) bits |= 1;
  }
  for (i64 j = static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x; j < n_inst; j += stride)
    if (!isfinite(inst_areas[j]) || inst_areas[j] < 0.0) bits |= 16;
  if (bits) atomicOr(error, bits);
}

}  // namespace pe

extern "C" int pe_row_sums(const int64_t* pairs, int64_t rows, int64_t cols, double* out, void* stream_ptr) {
  PE_REQUIRE(rows >= 0 && cols > 0 && cols < INT_MAX && rows < (1LL << 40), "invalid row-sum shape");
  if (rows == 0) return kOk;
  const i64 blocks = (rows * kWave + 255) / 256;
  PE_REQUIRE(blocks < INT_MAX, "too many rows");
  row_sums<<<static_cast<unsigned>(blocks), 256, 0, static_cast<hipStream_t>(stream_ptr)>>>(
      reinterpret_cast<const i64*>(pairs), rows, static_cast<int>(cols), out);
  PE_LAUNCH_CHECK();
  return kOk;
}

extern "C" int pe_instance_checks(int64_t n_det, const int64_t* det_cls, const double* scores, const double* det_area,
                                  int64_t n_gt, const int64_t* gt_cls, const double* gt_area, const double* ann,
                                  int64_t C, const uint8_t* things, int64_t n_inst, const double* inst_areas,
                                  int32_t* error, void* stream_ptr) {
  PE_REQUIRE(n_det >= 0 && n_gt >= 0 && n_inst >= 0 && C > 0 && C < INT_MAX, "invalid check shapes");
  hipStream_t stream = static_cast<hipStream_t>(stream_ptr);
  int rc = launch_zero(error, 4, stream);
  if (rc) return rc;
  const i64 n = std::max(n_det, std::max(n_gt, n_inst));
  if (n == 0) return kOk;
  const int blocks = static_cast<int>(std::min<i64>((n + 255) / 256, 1024));
  instance_checks<<<blocks, 256, 0, stream>>>(n_det, reinterpret_cast<const i64*>(det_cls), scores, det_area, n_gt,
                                              reinterpret_cast<const i64*>(gt_cls), gt_area, ann, static_cast<int>(C),
                                              things, n_inst, inst_areas, error);
  PE_LAUNCH_CHECK();
  return kOk;
}

// ------------------------------------------------------- per-image score order

namespace pe {

// Sort items of one image per workgroup: image, descending score, index.
__global__ __launch_bounds__(256) void score_items(const i64* det_off, const double* scores, SortItem* items) {
  const int b = blockIdx.x;
  const i64 first = det_off[b], last = det_off[b + 1];
  for (i64 i = first + threadIdx.x; i < last; i += blockDim.x)
    items[i] = SortItem{score_key_desc(scores[i]), static_cast<uint32_t>(b), static_cast<int32_t>(i)};
}

// One workgroup per image: gather classes/scores into score order and record
// the within-image mask index of every sorted position.
__global__ __launch_bounds__(256) void gather_score_order(const i64* det_off, const SortItem* sorted,
                                                          const double* scores, const i64* classes, i64* order,
                                                          double* sorted_scores, i64* sorted_classes,
                                                          int32_t* sorted_classes32) {
  const int b = blockIdx.x;
  const i64 first = det_off[b], last = det_off[b + 1];
  for (i64 i = first + threadIdx.x; i < last; i += blockDim.x) {
    const int32_t src = sorted[i].index;
    order[i] = src - first;
    sorted_scores[i] = scores[src];
    const i64 c = classes[src];
    sorted_classes[i] = c;
    sorted_classes32[i] = static_cast<int32_t>(c);
  }
}


}  // namespace pe

extern "C" int64_t pe_score_order_workspace_bytes(int64_t N, int64_t B) {
  Carve carve(nullptr);
  carve.take<SortItem>(N > 0 && B > 0 ? N : 0);
  carve.take<SortItem>(N > 0 && B > 0 ? N : 0);
  return static_cast<int64_t>(align_up(carve.used, 256));
}

extern "C" int pe_score_order(int64_t B, int64_t N, int64_t max_dets, const int64_t* det_off, const double* scores,
                              const int64_t* classes, int64_t* order, double* sorted_scores,
                              int64_t* sorted_classes, int32_t* sorted_classes32, void* workspace,
                              int64_t workspace_bytes, void* stream_ptr) {
  PE_REQUIRE(B >= 0 && B < INT_MAX && N >= 0 && N < INT_MAX, "invalid score-order shapes");
  PE_REQUIRE(workspace_bytes >= pe_score_order_workspace_bytes(N, B), "workspace too small");
  if (N == 0 || B == 0) return kOk;
  hipStream_t stream = static_cast<hipStream_t>(stream_ptr);
  if (max_dets > 0 && max_dets <= kOrderLds)
    return launch_score_order_lds(false, B, max_dets, det_off, scores, classes, ImageArrays{}, order,
                                  sorted_scores, sorted_classes, sorted_classes32, stream);
  Carve carve(workspace);
  SortItem* items = carve.take<SortItem>(N);
  SortItem* spare = carve.take<SortItem>(N);
  // Per image: descending score, ties in input order
  // (torch.argsort(descending=True, stable=True), instance.py:140).
  const i64* off = reinterpret_cast<const i64*>(det_off);
  score_items<<<static_cast<unsigned>(B), 256, 0, stream>>>(off, scores, items);
  PE_LAUNCH_CHECK();
  SortItem* sorted = nullptr;
  if (const int rc = sort_items(items, spare, N, stream, &sorted)) return rc;
  gather_score_order<<<static_cast<unsigned>(B), 256, 0, stream>>>(
      off, sorted, scores, reinterpret_cast<const i64*>(classes), reinterpret_cast<i64*>(order), sorted_scores,
      reinterpret_cast<i64*>(sorted_classes), sorted_classes32);
  PE_LAUNCH_CHECK();
  return kOk;
}

// The per-batch tables the matcher reads, in one launch: detection pixel
// areas in score order (f64), and the GT pixel areas (f64), classes (int64 and
// int32) and crowd flags of every image concatenated.  Image b's GT classes and
// crowd flags are read in place through the addresses in the arguments.
struct GtArrays {
  const void* classes[kOrderArgImages];
  const uint8_t* crowd[kOrderArgImages];
  int class_bytes;
};

// Image of row i of a batch laid out by offsets off[0..B] (off[B] > i).
__device__ __forceinline__ int image_of(const i64* off, int B, i64 i) {
  int lo = 0, hi = B - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= i) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__global__ __launch_bounds__(256) void instance_tables(int B, const i64* det_off, const i64* order, const int32_t* pa,
                                                       i64 N, const i64* gt_off, const int32_t* ga, i64 total_g,
                                                       GtArrays g, double* det_area, double* gt_area, i64* gt_cls,
                                                       int32_t* gt_cls32, uint8_t* crowd) {
  const i64 stride = static_cast<i64>(gridDim.x) * blockDim.x;
  for (i64 i = static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x; i < N + total_g; i += stride) {
    if (i < N) {
      const int b = image_of(det_off, B, i);
      det_area[i] = static_cast<double>(pa[det_off[b] + order[i]]);
    } else {
      const i64 j = i - N;
      const int b = image_of(gt_off, B, j);
      const i64 k = j - gt_off[b];
      const i64 c = g.class_bytes == 8 ? static_cast<const i64*>(g.classes[b])[k]
                                       : static_cast<i64>(static_cast<const int32_t*>(g.classes[b])[k]);
      gt_area[j] = static_cast<double>(ga[j]);
      gt_cls[j] = c;
      gt_cls32[j] = static_cast<int32_t>(c);
      crowd[j] = g.crowd[b][k];
    }
  }
}

extern "C" int pe_instance_tables(int64_t B, const int64_t* det_off, const int64_t* order, const int32_t* pa,
                                  int64_t N, const int64_t* gt_off, const int32_t* ga, int64_t total_g,
                                  const void* const* gt_classes, int class_bytes, const uint8_t* const* gt_crowd,
                                  double* det_area, double* gt_area, int64_t* gt_cls, int32_t* gt_cls32,
                                  uint8_t* crowd, void* stream_ptr) {
  PE_REQUIRE(B >= 0 && B <= kOrderArgImages, "at most 64 images per call");
  PE_REQUIRE(N >= 0 && total_g >= 0, "invalid table sizes");
  PE_REQUIRE(class_bytes == 4 || class_bytes == 8, "classes must be int32 or int64");
  PE_REQUIRE(B == 0 || (det_off && gt_off && gt_classes && gt_crowd), "null batch tables");
  PE_REQUIRE(N == 0 || (order && pa && det_area), "null detection tables");
  PE_REQUIRE(total_g == 0 || (ga && gt_area && gt_cls && gt_cls32 && crowd), "null GT tables");
  if (B == 0 || N + total_g == 0) return kOk;
  GtArrays g{};
  for (int64_t b = 0; b < B; ++b) {
    g.classes[b] = gt_classes[b];
    g.crowd[b] = gt_crowd[b];
  }
  g.class_bytes = class_bytes;
  const i64 blocks = std::min<i64>((N + total_g + 255) / 256, 4 * compute_units());
  hipStream_t stream = static_cast<hipStream_t>(stream_ptr);
  instance_tables<<<static_cast<unsigned>(blocks), 256, 0, stream>>>(
      static_cast<int>(B), reinterpret_cast<const i64*>(det_off), reinterpret_cast<const i64*>(order), pa, N,
      reinterpret_cast<const i64*>(gt_off), ga, total_g, g, det_area, gt_area, reinterpret_cast<i64*>(gt_cls),
      gt_cls32, crowd);
  PE_LAUNCH_CHECK();
  return kOk;
}

extern "C" int pe_score_order_images(int64_t B, int64_t max_dets, const int64_t* det_off,
                                     const double* const* scores, const void* const* classes, int class_bytes,
                                     int64_t* order, double* sorted_scores, int64_t* sorted_classes,
                                     int32_t* sorted_classes32, void* stream_ptr) {
  PE_REQUIRE(B >= 0 && B <= kOrderArgImages, "at most 64 images per call");
  PE_REQUIRE(max_dets >= 0 && max_dets <= kOrderLds, "at most 4096 detections per image");
  PE_REQUIRE(class_bytes == 4 || class_bytes == 8, "classes must be int32 or int64");
  PE_REQUIRE(B == 0 || (scores && classes), "null per-image arrays");
  if (B == 0 || max_dets == 0) return kOk;
  ImageArrays images{};
  for (int64_t b = 0; b < B; ++b) {
    images.scores[b] = scores[b];
    images.classes[b] = classes[b];
  }
  images.class_bytes = class_bytes;
  return launch_score_order_lds(true, B, max_dets, det_off, nullptr, nullptr, images, order, sorted_scores,
                                sorted_classes, sorted_classes32, static_cast<hipStream_t>(stream_ptr));
}
