// COCO run-length encoded (RLE) instance masks decoded into bit rows on
// gfx950.
//
// The reference hands its masks to pycocotools as RLE in its parity tests
// (tests/test_instance.py:110, examples/benchmark_instance.py:42) and takes
// bool bytes otherwise.  Here a caller holding COCO annotations keeps them
// compressed up to the GPU: the strings (or count lists) are uploaded as
// they are and decoded into the bit rows of bit-packed masks (bit p % 32 of
// word p / 32 is pixel p = y * W + x), which the instance and combined
// evaluators read directly.
//
// Format (pycocotools 2.0, common/maskApi.c): pixels in column-major order
// (pixel (y, x) at x * H + y); counts are alternating runs of 0s and 1s,
// starting with 0s, summing to H * W.  The compressed string stores count
// i > 2 as its difference from count i - 2, every value as little-endian
// 5-bit groups of its two's complement, one character 48 + group (+ 32 when
// a group follows), the last group's bit 4 being the sign.
//
// Kernels:
//  * rle_runs: one wave per mask.  The string is read 64 characters per step:
//    a ballot of the groups without continuation bit marks where counts end,
//    each ending lane sums its count's groups from the lanes before it
//    (shuffles), and the delta coding resolves as two inclusive scans, one
//    over even and one over odd count indices (count i > 2 is the sum of
//    the stored values of indices i, i - 2, ..., down to 1 or 2).  A third
//    scan gives every run's end position.  For every column whose first
//    pixel lies in a run, the wave writes the run's index (a per-column entry
//    point for the raster kernel); a run of 0s can span hundreds of columns,
//    so the entries of a step are spread over the lanes by a prefix sum of
//    the runs' column counts.  Malformed
//    input (a character outside the alphabet, a count of more than seven
//    groups, a negative count, a total other than H * W) marks the mask and
//    sets an error bit; the raster kernel leaves such masks all-zero.
//    The same launch clears every mask's bit row: one workgroup per mask
//    after the parsing workgroups, running beside them.
//  * rle_raster: four waves per (mask, 256 columns), lane = column.  Each
//    lane walks its column's runs down the image 32 rows at a time and forms
//    the column's 32-bit word of the block; a block with no set pixel in the
//    wave's 64 columns is skipped after one ballot, otherwise five butterfly
//    stages transpose each half-wave's 32 column words into 32 row words and
//    lane 32h + t writes row t's word of the half.  Only nonzero words are
//    written into the cleared rows: with W a multiple of 32 every word has
//    one writer (plain stores), else words straddle two and are merged with
//    atomicOr.
#include <climits>

#include "common.h"
#include "../../include/panoptic_hip.h"

namespace pe {

// Per-mask table row (int64 x 8), written by the host.
struct RleMask {
  i64 data_off;  // byte offset of the mask's characters or int32 counts in `data`
  i64 units;     // characters (format 0) or counts (format 1)
  i64 out_off;   // word offset of the mask's bit row in `bits`
  i64 col_off;   // offset of the mask's W column entries in the column table
  i64 height;
  i64 width;
  i64 format;    // 0: compressed string, 1: int32 counts
  i64 run_off;   // offset of the mask's run ends (units entries)
};
static_assert(sizeof(RleMask) == 64, "the host writes eight int64 per mask");

constexpr int kRleErrBadChar = 1;      // a character outside '0'..'o'
constexpr int kRleErrLongCount = 2;    // a count of more than seven groups, or an unterminated last count
constexpr int kRleErrNegative = 4;     // a negative count after delta decoding
constexpr int kRleErrTotal = 8;        // counts not summing to H * W

// Inclusive scans over the wave with DPP moves (a few cycles each, where a
// cross-lane shuffle through LDS takes about a hundred): shifts by 1, 2, 4
// and 8 lanes within each row of 16, then row 0's total broadcast into row 1
// and row 2's into row 3 (lane 15 / 47), then the total of rows 0-1 into
// rows 2 and 3 (lane 31).  Lanes whose source is outside their row, or whose
// row is masked off, add 0.
template <int Ctrl, int RowMask, bool Bound>
__device__ __forceinline__ int dpp_shift(int v, int old = 0) {
  return __builtin_amdgcn_update_dpp(old, v, Ctrl, RowMask, 0xF, Bound);
}
template <int Ctrl, int RowMask, bool Bound>
__device__ __forceinline__ i64 dpp_shift(i64 v) {
  const u64 u = static_cast<u64>(v);
  const uint32_t lo = static_cast<uint32_t>(dpp_shift<Ctrl, RowMask, Bound>(static_cast<int>(u)));
  const uint32_t hi = static_cast<uint32_t>(dpp_shift<Ctrl, RowMask, Bound>(static_cast<int>(u >> 32)));
  return static_cast<i64>((static_cast<u64>(hi) << 32) | lo);
}
template <typename T>
__device__ __forceinline__ T wave_inclusive_scan(T v) {
  v += dpp_shift<0x111, 0xF, true>(v);   // row_shr:1
  v += dpp_shift<0x112, 0xF, true>(v);   // row_shr:2
  v += dpp_shift<0x114, 0xF, true>(v);   // row_shr:4
  v += dpp_shift<0x118, 0xF, true>(v);   // row_shr:8
  v += dpp_shift<0x142, 0xA, false>(v);  // row_bcast:15 into rows 1 and 3
  v += dpp_shift<0x143, 0xC, false>(v);  // row_bcast:31 into rows 2 and 3
  return v;
}

// The last lane's value, read into every lane.
__device__ __forceinline__ int last_lane(int v) { return __builtin_amdgcn_readlane(v, kWave - 1); }

// One butterfly stage of a 32 x 32 bit transpose within each half-wave
// (lane = column, bit = row on entry): the lane K apart (ds_swizzle's xor
// mode) trades the off-diagonal K x K blocks, the lane with bit K clear
// taking its partner's rows below K into its rows from K up, the other lane
// the partner's rows from K up into its rows below K.
template <int K>
__device__ __forceinline__ uint32_t transpose_stage(uint32_t w, int lane) {
  constexpr uint32_t low = K == 16 ? 0x0000FFFFu : K == 8 ? 0x00FF00FFu : K == 4 ? 0x0F0F0F0Fu
                         : K == 2 ? 0x33333333u : 0x55555555u;
  const uint32_t partner = static_cast<uint32_t>(__builtin_amdgcn_ds_swizzle(static_cast<int>(w), 0x1F | (K << 10)));
  return (lane & K) ? (w & ~low) | ((partner >> K) & low) : (w & low) | ((partner & low) << K);
}

// Lane `src`'s value (src uniform), read into every lane.
__device__ __forceinline__ i64 read_lane(i64 v, int src) {
  const u64 u = static_cast<u64>(v);
  const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(u), src));
  const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(u >> 32), src));
  return static_cast<i64>((static_cast<u64>(hi) << 32) | lo);
}

// Inclusive running maximum over the wave (lanes without a source take -1).
__device__ __forceinline__ int wave_inclusive_max(int v) {
  v = max(v, dpp_shift<0x111, 0xF, false>(v, -1));
  v = max(v, dpp_shift<0x112, 0xF, false>(v, -1));
  v = max(v, dpp_shift<0x114, 0xF, false>(v, -1));
  v = max(v, dpp_shift<0x118, 0xF, false>(v, -1));
  v = max(v, dpp_shift<0x142, 0xA, false>(v, -1));
  v = max(v, dpp_shift<0x143, 0xC, false>(v, -1));
  return v;
}

// The wave's minimum: the same scan with min, lanes without a source (or
// masked rows) taking INT_MAX instead of 0.
__device__ __forceinline__ int wave_min(int v) {
  v = min(v, dpp_shift<0x111, 0xF, false>(v, INT_MAX));
  v = min(v, dpp_shift<0x112, 0xF, false>(v, INT_MAX));
  v = min(v, dpp_shift<0x114, 0xF, false>(v, INT_MAX));
  v = min(v, dpp_shift<0x118, 0xF, false>(v, INT_MAX));
  v = min(v, dpp_shift<0x142, 0xA, false>(v, INT_MAX));
  v = min(v, dpp_shift<0x143, 0xC, false>(v, INT_MAX));
  return last_lane(v);
}
__device__ __forceinline__ i64 last_lane(i64 v) {
  const u64 u = static_cast<u64>(v);
  const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(u), kWave - 1));
  const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(u >> 32), kWave - 1));
  return static_cast<i64>((static_cast<u64>(hi) << 32) | lo);
}

// Orders one wave's LDS writes and its later reads by other lanes.
__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__global__ __launch_bounds__(256) void rle_runs(i64 N, const RleMask* table, const uint8_t* data, int32_t* ends,
                                                int32_t* col_run, int32_t* nruns, int32_t* error, int32_t* extents,
                                                uint32_t* rows_out) {
  __shared__ int32_t staging[4][3][kWave];  // per wave: prefix offsets, first columns, run indices
  // Characters (or counts) of 16 steps are loaded together, indices clamped
  // into the mask's units, before the first is used: the steps are serial,
  // the loads need not be, and most strings fit one such block.
  constexpr int kSteps = 16;
  __shared__ int32_t block_units[4][kSteps * kWave];
  const unsigned parse_groups = static_cast<unsigned>((N + 3) / 4);
  if (blockIdx.x >= parse_groups) {
    // The workgroups after the parsing ones clear one bit row each; the
    // parse is latency-bound and leaves most of the bandwidth to them.
    const RleMask m = table[blockIdx.x - parse_groups];
    uint32_t* out = rows_out + m.out_off;
    const i64 words = m.height > 0 && m.width > 0 ? mask_words(m.height * m.width) : 0;
    if ((reinterpret_cast<uintptr_t>(out) & 15) == 0) {
      uint4* out4 = reinterpret_cast<uint4*>(out);
      for (i64 i = threadIdx.x; i < words / 4; i += blockDim.x) out4[i] = make_uint4(0, 0, 0, 0);
    } else {
      for (i64 i = threadIdx.x; i < words; i += blockDim.x) out[i] = 0;
    }
    return;
  }
  const int lane = threadIdx.x & (kWave - 1);
  int32_t* off_lds = staging[threadIdx.x >> 6][0];
  int32_t* x0_lds = staging[threadIdx.x >> 6][1];
  int32_t* idx_lds = staging[threadIdx.x >> 6][2];
  const i64 n = static_cast<i64>(blockIdx.x) * 4 + (threadIdx.x >> 6);
  if (n >= N) return;  // uniform per wave
  // The mask's extents start empty; the raster kernel widens them to its
  // nonzero words (a malformed mask stays empty).
  if (extents && lane == 0) {
    extents[2 * n] = INT_MAX;
    extents[2 * n + 1] = 0;
  }
  const RleMask m = table[n];
  const i64 H = m.height, W = m.width, total = H * W;
  const uint8_t* chars = data + m.data_off;
  const int32_t* given = reinterpret_cast<const int32_t*>(data + m.data_off);
  int32_t* run_end = ends + m.run_off;
  int32_t* cols = col_run + m.col_off;
  int bad = 0;
  i64 count_base = 0;                  // counts completed so far
  i64 pend_val = 0;                    // groups of a count still open at the end of a step
  int pend_k = 0;                      // how many groups it has
  i64 chain_even = 0, chain_odd = 0;   // running sums of the stored values (delta decoding)
  i64 position = 0;                    // end position of the last completed run
  // Each lane keeps its units of the block in its own LDS slots (no other
  // lane reads them), so the steps need not be unrolled.
  int32_t* unit = block_units[threadIdx.x >> 6] + lane;
  for (i64 block = 0; block < m.units; block += kSteps * kWave) {
    int loaded[kSteps];
#pragma unroll
    for (int s = 0; s < kSteps; ++s) {
      const i64 u = min<i64>(block + s * kWave + lane, m.units - 1);
      loaded[s] = m.format == 0 ? static_cast<int>(chars[u]) : given[u];
    }
#pragma unroll
    for (int s = 0; s < kSteps; ++s) unit[s * kWave] = loaded[s];
#pragma unroll 1
    for (int s = 0; s < kSteps; ++s) {
      const i64 base = block + s * kWave;
      if (base >= m.units) break;  // uniform
      const i64 u = base + lane;
      const bool in = u < m.units;
      i64 value = 0;    // this lane's count (valid where `ends_here`)
      bool ends_here;
      u64 emask;
      if (m.format == 0) {
        const int v = in ? unit[s * kWave] - 48 : 0;
        if (in && (v < 0 || v > 63)) bad |= kRleErrBadChar;
        const int group = v & 0x1F;
        ends_here = in && !(v & 0x20);
        emask = __ballot(ends_here);
        // The lanes of this step that belong to this lane's count: from the
        // lane after the previous ending lane (or lane 0, continuing the count
        // left open by the previous step).
        const u64 before = emask & ((1ULL << lane) - 1);
        const int start = before ? 64 - __clzll(before) : 0;
        const int k = lane - start + (start == 0 ? pend_k : 0);
        if (in && k > 6) bad |= kRleErrLongCount;
        const i64 term = in ? static_cast<i64>(static_cast<u64>(group) << (5 * min(k, 11))) : 0;
        i64 sum = term;
#pragma unroll
        for (int d = 1; d <= 6; ++d) {
          const i64 t = __shfl_up(term, d, kWave);
          if (lane - d >= start) sum += t;
        }
        if (start == 0) sum += pend_val;
        value = sum;
        if (ends_here && (group & 0x10)) value |= static_cast<i64>(-1) << (5 * min(k + 1, 62));
        // The count left open at the end of this step (if the last character
        // carries on): its groups so far.
        const int last = static_cast<int>(min<i64>(kWave, m.units - base)) - 1;
        const bool open = !((emask >> last) & 1ULL);
        const i64 open_sum = read_lane(sum, last);
        const int open_k = __builtin_amdgcn_readlane(k, last) + 1;
        pend_val = open ? open_sum : 0;
        pend_k = open ? open_k : 0;
      } else {
        ends_here = in;
        emask = __ballot(in);
        value = in ? static_cast<i64>(unit[s * kWave]) : 0;
      }
      // Count index of each ending lane, and the delta decoding of the string
      // form: count i > 2 adds count i - 2, i.e. the stored values of i, i - 2,
      // ... down to index 1 (odd) or 2 (even).
      const i64 idx = count_base + __popcll(emask & ((1ULL << lane) - 1));
      i64 count = value;
      if (m.format == 0) {
        const bool chained = ends_here && idx >= 1;
        const i64 even = wave_inclusive_scan<i64>(chained && !(idx & 1) ? value : 0);
        const i64 odd = wave_inclusive_scan<i64>(chained && (idx & 1) ? value : 0);
        if (chained) count = (idx & 1) ? chain_odd + odd : chain_even + even;
        chain_even += last_lane(even);
        chain_odd += last_lane(odd);
      }
      if (ends_here && count < 0) bad |= kRleErrNegative;
      const i64 clamped = ends_here && count > 0 ? count : 0;
      const i64 end = position + wave_inclusive_scan<i64>(clamped);
      position = last_lane(end);
      // Columns whose first pixel lies in this lane's run: [x0, x0 + ncols).
      int ncols = 0;
      i64 x0 = 0;
      if (ends_here) {
        const i64 lo = end - clamped;
        if (end > total) bad |= kRleErrTotal;
        if (idx < m.units) run_end[idx] = static_cast<int32_t>(min<i64>(end, INT_MAX));
        if (clamped > 0 && end <= total) {
          if (total <= INT_MAX) {  // uniform per mask: 32-bit division
            const uint32_t h = static_cast<uint32_t>(H);
            x0 = static_cast<uint32_t>(lo + H - 1) / h;
            ncols = static_cast<int>(max<i64>(0, static_cast<i64>(static_cast<uint32_t>(end - 1) / h) - x0 + 1));
          } else {
            x0 = (lo + H - 1) / H;
            ncols = static_cast<int>(max<i64>(0, (end - 1) / H - x0 + 1));
          }
        }
      }
      // The step's column entries are written by the whole wave (a run of 0s
      // can span hundreds of columns): entry q belongs to the last lane whose
      // exclusive prefix of column counts is <= q.
      const int incl = wave_inclusive_scan<int>(ncols);
      const int entries = last_lane(incl);
      if (entries) {
        // 64 entries at a time: the owner's lane is marked at its first
        // entry, and a running maximum over the marks (carried from the
        // previous 64) gives every entry's owner.
        const int off = incl - ncols;
        wave_sync_lds();
        x0_lds[lane] = static_cast<int32_t>(x0) - off;  // entry q is column x0_lds[owner] + q
        idx_lds[lane] = static_cast<int32_t>(idx);
        int owner = 0;
        for (int qb = 0; qb < entries; qb += kWave) {
          off_lds[lane] = -1;
          wave_sync_lds();
          if (ncols > 0 && off >= qb && off < qb + kWave) off_lds[off - qb] = lane;
          wave_sync_lds();
          const int o = max(wave_inclusive_max(off_lds[lane]), owner);
          owner = last_lane(o);
          if (qb + lane < entries) cols[x0_lds[o] + qb + lane] = idx_lds[o];
          wave_sync_lds();
        }
      }
      count_base += __popcll(emask);
    }
  }
  if (pend_k) bad |= kRleErrLongCount;
  if (position != total) bad |= kRleErrTotal;
  // Any lane's error marks the whole mask.
  const bool any_bad = __ballot(bad != 0) != 0;
  int bits = bad;
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) bits |= __shfl_xor(bits, d, kWave);
  if (lane == 0) {
    nruns[n] = any_bad ? -1 : static_cast<int32_t>(count_base);
    if (any_bad) atomicOr(error, bits);
  }
}

// Four waves per (mask, 256-column group); lane = column x.  Rows are
// produced 32 at a time: the column's word of the block from its runs, then
// a bit transpose gives 32-pixel row words.
__global__ __launch_bounds__(256) void rle_raster(i64 N, int groups, const RleMask* table, const int32_t* ends,
                                                  const int32_t* col_run, const int32_t* nruns, uint32_t* bits,
                                                  int32_t* extents) {
  const int lane = threadIdx.x & (kWave - 1);
  const i64 n = blockIdx.x / groups;
  const int band = (blockIdx.x - static_cast<int>(n) * groups) * 4 + (threadIdx.x >> 6);
  if (n >= N) return;
  const RleMask m = table[n];
  const int runs = nruns[n];
  const i64 H = m.height, W = m.width;
  const i64 x0 = static_cast<i64>(band) * kWave;
  if (runs <= 0 || x0 >= W) return;  // uniform per wave
  const i64 x = x0 + lane;
  const bool column = x < W;
  const int32_t* run_end = ends + m.run_off;
  uint32_t* out = bits + m.out_off;
  // The run holding the column's current row, with its [start, end), and the
  // next run's end loaded one advance ahead of its use.
  int r = column ? col_run[m.col_off + x] : 0;
  i64 r_end = column ? run_end[r] : 0;
  i64 r_start = column && r > 0 ? run_end[r - 1] : 0;
  i64 r_next = column && r + 1 < runs ? run_end[r + 1] : 0;
  const bool words_aligned = (W & 31) == 0;
  // The nonzero words this lane writes: their lowest and highest index.
  int w_lo = INT_MAX, w_hi = -1;
  auto note = [&](i64 w) {
    w_lo = min(w_lo, static_cast<int>(w));
    w_hi = max(w_hi, static_cast<int>(w));
  };
  const bool skippable = H <= INT_MAX;  // uniform per mask
  for (i64 y0 = 0; y0 < H; y0 += 32) {
    if (skippable) {
      // Jump to the first block in which some column of the wave can hold a
      // set pixel: a column inside a run of 0s has none before that run ends
      // (later, if the next runs are empty).  The run state stays valid: each
      // lane's run still ends at or after the new block's first pixel.
      i64 next = H;
      if (column) {
        if (r & 1) next = y0;
        else if (r + 1 < runs) next = min<i64>(H, r_end - x * H);
      }
      const i64 target = static_cast<i64>(wave_min(static_cast<int>(next))) & ~static_cast<i64>(31);
      if (target >= H) break;
      y0 = max(y0, target);
    }
    const int rows = static_cast<int>(min<i64>(32, H - y0));
    uint32_t word = 0;
    if (column) {
      const i64 lo = x * H + y0, hi = lo + rows;
      // Runs overlapping [lo, hi): set the 1-runs' bits, advance past the
      // runs that end inside the block.
      while (true) {
        if (r & 1) {
          const i64 a = max(r_start, lo), b = min(r_end, hi);
          if (b > a) {
            const int s = static_cast<int>(a - lo), len = static_cast<int>(b - a);
            word |= (len == 32 ? 0xFFFFFFFFu : ((1u << len) - 1u)) << s;
          }
        }
        if (r_end >= hi || r + 1 >= runs) break;
        ++r;
        r_start = r_end;
        r_end = r_next;
        r_next = r + 1 < runs ? run_end[r + 1] : 0;
      }
    }
    // Most blocks of a band hold no pixel of the mask: one ballot skips them.
    if (__ballot(word != 0) == 0) continue;
    // Transpose each half-wave's 32 x 32 bit block (lane = column, bit =
    // row): afterwards lane 32h + j holds row j's pixels of columns
    // x0 + 32h ... x0 + 32h + 31 (bit c = column x0 + 32h + c).
    uint32_t segment = word;
    segment = transpose_stage<16>(segment, lane);
    segment = transpose_stage<8>(segment, lane);
    segment = transpose_stage<4>(segment, lane);
    segment = transpose_stage<2>(segment, lane);
    segment = transpose_stage<1>(segment, lane);
    const int j = lane & 31;
    if (j < rows && segment) {
      const i64 p = (y0 + j) * W + x0 + (lane & 32);  // first pixel of the segment
      const i64 w0 = p >> 5;
      const int shift = static_cast<int>(p & 31);
      if (words_aligned) {
        out[w0] = segment;
        note(w0);
      } else {
        const uint32_t lo_part = segment << shift, hi_part = shift ? segment >> (32 - shift) : 0;
        if (lo_part) {
          atomicOr(out + w0, lo_part);
          note(w0);
        }
        if (hi_part) {
          atomicOr(out + w0 + 1, hi_part);
          note(w0 + 1);
        }
      }
    }
  }
  if (extents) {
    for (int sh = 1; sh < kWave; sh <<= 1) {
      w_lo = min(w_lo, __shfl_xor(w_lo, sh, kWave));
      w_hi = max(w_hi, __shfl_xor(w_hi, sh, kWave));
    }
    if (lane == 0 && w_hi >= 0) {
      atomicMin(extents + 2 * n, w_lo);
      atomicMax(extents + 2 * n + 1, w_hi + 1);
    }
  }
}

}  // namespace pe

using namespace pe;

extern "C" int64_t pe_rle_workspace_bytes(int64_t N, int64_t units, int64_t columns) {
  if (N < 0 || units < 0 || columns < 0) return -1;
  Carve carve(nullptr);
  carve.take<int32_t>(units);
  carve.take<int32_t>(columns);
  carve.take<int32_t>(N);
  carve.take<int32_t>(1);
  return static_cast<int64_t>(align_up(carve.used, 256));
}

extern "C" int pe_rle_decode(int64_t N, const int64_t* table, const uint8_t* data, int64_t units, int64_t columns,
                             int64_t max_width, uint32_t* bits, int32_t* error, int32_t* extents,
                             void* workspace, int64_t workspace_bytes, void* stream_ptr) {
  PE_REQUIRE(N >= 0 && N < INT_MAX / 4 && units >= 0 && columns >= 0 && max_width >= 0 && max_width < INT_MAX,
             "invalid RLE shapes");
  PE_REQUIRE(workspace_bytes >= pe_rle_workspace_bytes(N, units, columns), "workspace too small");
  PE_REQUIRE(error != nullptr, "the error word is required");
  if (N == 0) return kOk;
  hipStream_t stream = static_cast<hipStream_t>(stream_ptr);
  Carve carve(workspace);
  int32_t* ends = carve.take<int32_t>(units);
  int32_t* col_run = carve.take<int32_t>(columns);
  int32_t* nruns = carve.take<int32_t>(N);
  const auto* rows = reinterpret_cast<const RleMask*>(table);
  {
    ProfileScope timing("rle_runs", stream);
    rle_runs<<<static_cast<unsigned>((N + 3) / 4 + N), 256, 0, stream>>>(N, rows, data, ends, col_run, nruns,
                                                                          error, extents, bits);
    PE_LAUNCH_CHECK();
  }
  const int groups = static_cast<int>((max_width + 4 * kWave - 1) / (4 * kWave));
  if (groups == 0) return kOk;
  PE_REQUIRE(N * groups < INT_MAX, "too many RLE masks for one launch");
  ProfileScope timing("rle_raster", stream);
  rle_raster<<<static_cast<unsigned>(N * groups), 256, 0, stream>>>(N, groups, rows, ends, col_run, nruns, bits,
                                                                     extents);
  PE_LAUNCH_CHECK();
  return kOk;
}
