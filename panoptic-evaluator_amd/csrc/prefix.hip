// Generation-order prefix rewards on gfx950 (segmentation_prefixes).
//
// Behaviour follows the reference's CPU path (prefix.py:11-122: later masks
// overwrite earlier ones on the canvas; visible-segment PQ/semantic stats at
// checkpoints; independent-mask greedy instance matching per IoU threshold)
// and its native boundary (prefix_kernel.cu:142-199).
//
// Design (MI355X-first; replaces 2 launches per mask with 2 launches per
// batch of images):
//  * events: one pixel pass over every mask of every image.  Each lane owns
//    16 pixels, keeps their GT slots and current top mask in registers and
//    streams the masks in generation order with 16-byte loads.  Every covered
//    pixel is a transfer (mask i, GT slot g, previous top p), counted in
//    raw[i][g] (intersection of the original mask) and in the triangular
//    per-step table dec[i][g][p] (p <= i).  GT areas come from the same pass.
//  * scoring: the G x (N+1) pair table of the visible canvas at step k is
//    a sum: column j >= 1 is raw[j-1] minus the decrements of the steps
//    j..k-1, column 0 the GT areas minus those of steps 0..k-1.  So the
//    checkpoints of an image are split into slices, one workgroup each: it
//    builds the table at its first checkpoint in one pass over the earlier
//    steps' counts (LDS atomics, no per-step barrier), then moves from
//    checkpoint to checkpoint the same way and computes the full
//    PQ/confusion of each (iou_sum folded in slot order, as the reference's
//    scatter_add_).  The greedy instance matching (one wave per IoU
//    threshold) runs in workgroups of its own.  When the table does not fit
//    in LDS, one workgroup per image replays every step in global scratch.
#include <climits>

#include "common.h"
#include "../../include/panoptic_hip.h"

namespace pe {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int kPrefixThreads = 256;
// One 16-pixel vector per lane keeps the per-pixel state small (about 60
// VGPRs, 7 waves per SIMD): occupancy hides the mask loads' latency.
constexpr int kPrefixVecs = 1;  // 16-pixel vectors per lane
constexpr int kPrefixTile = 16 * kPrefixThreads * kPrefixVecs;
static_assert(kPrefixTile < 65536, "a step's count per pixel tile must fit the 16-bit table entries");

struct PrefixArgs {
  int B, G, C, T;
  TileList tiles;               // pixel tiles, each inside one image
  const int32_t* target;        // every image's slots, one after another
  const uint8_t* const* masks;  // B pointers (N_b x HW)
  const i64* n_off;             // B + 1: mask offsets
  const int32_t* cls;           // sum N: classes in generation order
  const int32_t* gc;            // B x G
  const uint8_t* crowd;         // B x G
  const uint8_t* things;        // C
  const double* thresholds;     // T
  const i64* k_off;             // B + 1: checkpoint offsets
  const int32_t* points;        // sum K: checkpoints (strictly increasing per image)
  // Per-image scratch; image b's sections start at (in elements)
  //   raw: n_off[b] * G, dec: G * sum_{b' < b} N_b'(N_b' + 1) / 2,
  //   pairs: (n_off[b] + b) * G, pa/lab/val: n_off[b] + b, used: b * T * G.
  int32_t* raw;                         // N_b x G
  int32_t* dec;                         // step-major triangles: step i holds G x (i + 1)
  int32_t* pairs;                       // G x (N_b + 1)
  i64* pa;                              // N_b + 1
  int32_t* ga;                          // B x G
  uint8_t* used;                        // T x G
  int32_t* matched;                     // B x G
  i64* itp;                             // B x T x C
  i64* ifp;                             // B x T x C
  i64* positives;                       // B x C
  int32_t* lab;                         // N_b + 1
  double* val;                          // N_b + 1
  // outputs per snapshot row
  i64* tp; i64* fp; i64* fn; double* sums; i64* confusion;
  i64* out_itp; i64* out_ifp; i64* out_ifn; double* precision; double* recall;
  int32_t* error;
  int aligned;
  // Staged scoring: B x match_groups matching workgroups, then one workgroup
  // per slice of at most slice_rows checkpoints (slices.pix_off = k_off).
  int match_groups;
  int slice_rows;
  TileList slices;
};

__device__ __forceinline__ v4i load16(const uint8_t* row, i64 k, i64 HW, bool aligned) {
  v4i v = {0, 0, 0, 0};
  if (k >= HW) return v;
  const auto* g = as_global(row);
  if (aligned && k + 16 <= HW) return __builtin_nontemporal_load(as_global(reinterpret_cast<const v4i*>(row + k)));
  uint8_t bytes[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) bytes[j] = k + j < HW ? g[k + j] : 0;
  __builtin_memcpy(&v, bytes, 16);
  return v;
}

__device__ __forceinline__ i64 tri(i64 i) { return i * (i + 1) / 2; }

__device__ __forceinline__ i64 dec_base(const PrefixArgs& a, int b) {
  i64 total = 0;
  for (int q = 0; q < b; ++q) total += tri(a.n_off[q + 1] - a.n_off[q]);
  return total * a.G;
}

// Masks whose loads are in flight together (processing stays in generation
// order: only the loads run ahead).
constexpr int kPrefixAhead = 6;

// Per-step event table in LDS (TABLE): the transitions of step i are counted
// in a G x (N + 1) table and marked in a bitmap of the same keys (no-return
// LDS atomics only), and after a barrier the workgroup adds each marked
// entry once to the global raw/dec tables.  Without it (large G x N) every
// pixel run goes to global atomics.
constexpr size_t kPrefixStageMax = 120 * 1024;    // LDS for the staged scoring state
constexpr int kPrefixTableMax = 32768;            // G x (N + 1) 16-bit entries in LDS

// A step adds at most one tile's pixels (kPrefixTile < 2^16) to an entry, so
// the table packs two 16-bit counts per 32-bit word; the bitmap word of key
// k is k / 32, so both keys of a table word share one bitmap word.
__host__ __device__ inline size_t prefix_lds_bytes(int G, int P1, bool table) {
  const size_t keys = static_cast<size_t>(G) * P1;
  size_t bytes = 4 * static_cast<size_t>(G);
  if (table) bytes += 4 * ((keys + 1) / 2 + (keys + 31) / 32);
  return (bytes + 15) & ~static_cast<size_t>(15);
}

// One workgroup per pixel tile of one image (grid >= tile count; surplus
// workgroups of a packed batch, whose count is only known on the device, exit
// at once).
template <bool TABLE>
__global__ __launch_bounds__(kPrefixThreads) void prefix_events(PrefixArgs a) {
  extern __shared__ uint4 lds4[];
  uint32_t* ga = reinterpret_cast<uint32_t*>(lds4);
  if (static_cast<int>(blockIdx.x) >= a.tiles.count(a.B)) return;
  const int b = __builtin_amdgcn_readfirstlane(a.tiles.image_of(blockIdx.x));
  const int tile = blockIdx.x - a.tiles.first_tile(b);
  const i64 HW = a.tiles.pixels(b);
  const int G = a.G;
  for (int g = threadIdx.x; g < G; g += kPrefixThreads) ga[g] = 0;
  __syncthreads();
  const i64 n0 = a.n_off[b];
  const int N = static_cast<int>(a.n_off[b + 1] - n0);
  const auto* target = as_global(a.target + a.tiles.begin(b));
  // Per pixel: its GT slot and the mask on top so far (0: none).  With the
  // table G and P1 are at most 2^15, so both share one register (state =
  // slot | top << 16, slot 0xFFFF for none); otherwise slots / top.
  uint32_t state[kPrefixVecs][16];
  int slots[kPrefixVecs][16], top[kPrefixVecs][16];
  i64 base[kPrefixVecs];
  bool bad = false;
#pragma unroll
  for (int h = 0; h < kPrefixVecs; ++h) {
    base[h] = static_cast<i64>(tile) * kPrefixTile + (static_cast<i64>(h) * kPrefixThreads + threadIdx.x) * 16;
    // All sixteen loads issue before the first is used (indices past the
    // image are clamped to its last pixel and their values dropped).
    int raw[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) raw[j] = target[min(base[h] + j, HW - 1)];
    int run_g = -1, run = 0;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      int s = -1;
      if (base[h] + j < HW) {
        s = raw[j];
        if (s < 0 || s >= G) {
          bad = true;
          s = -1;
        }
      }
      if (TABLE) {
        state[h][j] = s < 0 ? 0xFFFFu : static_cast<uint32_t>(s);
      } else {
        slots[h][j] = s;
        top[h][j] = 0;
      }
      if (s != run_g) {
        if (run_g >= 0) atomicAdd(ga + run_g, static_cast<uint32_t>(run));
        run_g = s;
        run = 0;
      }
      run += s >= 0;
    }
    if (run_g >= 0) atomicAdd(ga + run_g, static_cast<uint32_t>(run));
  }
  if (bad) atomicOr(a.error, 1);
  const uint8_t* masks = a.masks[b];
  int32_t* raw = a.raw + n0 * G;
  int32_t* dec = a.dec + dec_base(a, b);
  const int P1 = N + 1;
  const bool aligned = a.aligned != 0;
  // LDS step table (TABLE): G x P1 counts and the bitmap of the keys counted
  // in the current step (both all zero between steps).  Only steps whose
  // mask covers some pixel of the tile are processed.
  const int table_words = (G * P1 + 1) / 2, mark_words = (G * P1 + 31) / 32;
  uint32_t* table = ga + G;
  uint32_t* marks = table + table_words;
  __shared__ uint32_t wave_touch[2][kPrefixThreads / kWave];
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  if (TABLE) {
    for (int e = threadIdx.x; e < table_words + mark_words; e += kPrefixThreads) table[e] = 0;
    __syncthreads();
  }
  auto add_run = [&](int i, int32_t* dec_i, int key, int run) {
    if (TABLE) {
      atomicAdd(table + (key >> 1), static_cast<uint32_t>(run) << ((key & 1) * 16));
      atomicOr(marks + (key >> 5), 1u << (key & 31));
    } else {
      const int rg = key / P1, prev = key - rg * P1;
      atomicAdd(raw + static_cast<i64>(i) * G + rg, run);
      atomicAdd(dec_i + static_cast<i64>(rg) * (i + 1) + prev, run);
    }
  };
  for (int i0 = 0; i0 < N; i0 += kPrefixAhead) {
    v4i ahead[kPrefixAhead][kPrefixVecs];
#pragma unroll
    for (int u = 0; u < kPrefixAhead; ++u) {
      if (i0 + u < N) {
        const uint8_t* mask = masks + static_cast<i64>(i0 + u) * HW;
#pragma unroll
        for (int h = 0; h < kPrefixVecs; ++h) ahead[u][h] = load16(mask, base[h], HW, aligned);
      } else {
#pragma unroll
        for (int h = 0; h < kPrefixVecs; ++h) ahead[u][h] = v4i{0, 0, 0, 0};
      }
    }
    // Which of these steps touch the tile at all (one barrier per group; the
    // per-wave bits alternate between two buffers across groups).
    uint32_t touched = (1u << kPrefixAhead) - 1;
    if (TABLE) {
      uint32_t mine = 0;
#pragma unroll
      for (int u = 0; u < kPrefixAhead; ++u)
#pragma unroll
        for (int h = 0; h < kPrefixVecs; ++h)
          if ((ahead[u][h].x | ahead[u][h].y | ahead[u][h].z | ahead[u][h].w) != 0) mine |= 1u << u;
      uint32_t bits = 0;
#pragma unroll
      for (int u = 0; u < kPrefixAhead; ++u)
        if (__ballot((mine >> u) & 1u)) bits |= 1u << u;
      const int parity = (i0 / kPrefixAhead) & 1;
      if (lane == 0) wave_touch[parity][wave] = bits;
      __syncthreads();
      touched = 0;
#pragma unroll
      for (int w = 0; w < kPrefixThreads / kWave; ++w) touched |= wave_touch[parity][w];
    }
#pragma unroll
    for (int u = 0; u < kPrefixAhead; ++u) {
    const int i = i0 + u;
    if (i >= N) break;
    if (!((touched >> u) & 1u)) continue;  // uniform: no pixel of mask i in this tile
    const v4i* m = ahead[u];
    int32_t* dec_i = dec + static_cast<i64>(G) * tri(i);
#pragma unroll
    for (int h = 0; h < kPrefixVecs; ++h) {
      if ((m[h].x | m[h].y | m[h].z | m[h].w) == 0) continue;
      const uint32_t words[4] = {static_cast<uint32_t>(m[h].x), static_cast<uint32_t>(m[h].y),
                                 static_cast<uint32_t>(m[h].z), static_cast<uint32_t>(m[h].w)};
      int run_key = -1, run = 0;
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const bool on = (words[j >> 2] >> ((j & 3) * 8)) & 0xFFu;
        int key = -1;
        if (TABLE) {
          const uint32_t st = state[h][j], g = st & 0xFFFFu;
          if (on && g != 0xFFFFu) {
            key = static_cast<int>(g * P1 + (st >> 16));
            state[h][j] = g | (static_cast<uint32_t>(i + 1) << 16);
          }
        } else {
          const int g = slots[h][j];
          if (on && g >= 0) {
            key = g * P1 + top[h][j];
            top[h][j] = i + 1;
          }
        }
        if (key != run_key) {
          if (run_key >= 0) add_run(i, dec_i, run_key, run);
          run_key = key;
          run = 0;
        }
        run += key >= 0;
      }
      if (run_key >= 0) add_run(i, dec_i, run_key, run);
    }
    if (TABLE) {
      // Flush this step's marked entries once each: the owner of a bitmap
      // word reads and clears it and the table words it covers, then adds
      // their counts to the global tables.  Skipped entirely when the tile
      // saw no pixel of mask i (the decision is uniform after the barrier).
      __syncthreads();
      for (int w = threadIdx.x; w < mark_words; w += kPrefixThreads) {
        const uint32_t marked = marks[w];
        if (!marked) continue;
        marks[w] = 0;
        uint32_t pairs = (marked | (marked >> 1)) & 0x55555555u;  // even bit p: table word 16 w + p / 2
        while (pairs) {
          const int p = __builtin_ctz(pairs);
          pairs &= pairs - 1;
          const int word = 16 * w + p / 2;
          const uint32_t counts = table[word];
          table[word] = 0;
#pragma unroll
          for (int half = 0; half < 2; ++half) {
            const int32_t v = static_cast<int32_t>((counts >> (16 * half)) & 0xFFFFu);
            if (!v) continue;
            const int key = 2 * word + half, rg = key / P1, prev = key - rg * P1;
            atomicAdd(raw + static_cast<i64>(i) * G + rg, v);
            atomicAdd(dec_i + static_cast<i64>(rg) * (i + 1) + prev, v);
          }
        }
      }
      __syncthreads();
    }
    }
  }
  __syncthreads();
  for (int g = threadIdx.x; g < G; g += kPrefixThreads)
    if (ga[g]) atomicAdd(a.ga + static_cast<i64>(b) * G + g, static_cast<int32_t>(ga[g]));
}

// Per-image scoring state: the G x (N + 1) pair table of the visible canvas
// and its metadata.  Built either entirely in LDS (STAGED) or entirely in the
// global scratch, so that every access of one instantiation has a single
// address space (LDS accesses then compile to ds_* instructions).
struct PrefixView {
  int32_t* pairs;      // G x P1
  i64* pa;             // P1
  int32_t* lab;        // P1
  double* val;         // P1
  int32_t* ga;         // G
  int32_t* gc;         // G
  uint8_t* crowd;      // G
  int32_t* matched;    // G
  uint8_t* used;       // T x G
  i64* itp;            // T x C
  i64* ifp;            // T x C
  i64* positives;      // C
};

__host__ __device__ inline size_t prefix_view_bytes(int G, int P1, int C, int T) {
  const size_t g = static_cast<size_t>(G), p = static_cast<size_t>(P1), c = static_cast<size_t>(C);
  const size_t t = static_cast<size_t>(T);
  // 8-byte arrays first, then 4-byte, then bytes
  return 8 * (p + p + 2 * t * c + c) + 4 * (g * p + p + 3 * g) + (g + t * g) + 16;
}

__device__ __forceinline__ PrefixView lds_view(char* lds, int G, int P1, int C, int T) {
  PrefixView v;
  char* q = lds;
  auto take = [&](size_t bytes) {
    char* r = q;
    q += bytes;
    return r;
  };
  v.pa = reinterpret_cast<i64*>(take(8 * static_cast<size_t>(P1)));
  v.val = reinterpret_cast<double*>(take(8 * static_cast<size_t>(P1)));
  v.itp = reinterpret_cast<i64*>(take(8 * static_cast<size_t>(T) * C));
  v.ifp = reinterpret_cast<i64*>(take(8 * static_cast<size_t>(T) * C));
  v.positives = reinterpret_cast<i64*>(take(8 * static_cast<size_t>(C)));
  v.pairs = reinterpret_cast<int32_t*>(take(4 * static_cast<size_t>(G) * P1));
  v.lab = reinterpret_cast<int32_t*>(take(4 * static_cast<size_t>(P1)));
  v.ga = reinterpret_cast<int32_t*>(take(4 * static_cast<size_t>(G)));
  v.gc = reinterpret_cast<int32_t*>(take(4 * static_cast<size_t>(G)));
  v.matched = reinterpret_cast<int32_t*>(take(4 * static_cast<size_t>(G)));
  v.crowd = reinterpret_cast<uint8_t*>(take(static_cast<size_t>(G)));
  v.used = reinterpret_cast<uint8_t*>(take(static_cast<size_t>(T) * G));
  return v;
}

__device__ __forceinline__ PrefixView global_view(const PrefixArgs& a, int b, i64 n0) {
  const int G = a.G, C = a.C, T = a.T;
  PrefixView v;
  v.pairs = a.pairs + (n0 + b) * G;
  v.pa = a.pa + n0 + b;
  v.lab = a.lab + n0 + b;
  v.val = a.val + n0 + b;
  v.ga = a.ga + static_cast<i64>(b) * G;
  v.gc = const_cast<int32_t*>(a.gc + static_cast<i64>(b) * G);
  v.crowd = const_cast<uint8_t*>(a.crowd + static_cast<i64>(b) * G);
  v.matched = a.matched + static_cast<i64>(b) * G;
  v.used = a.used + static_cast<i64>(b) * T * G;
  v.itp = a.itp + static_cast<i64>(b) * T * C;
  v.ifp = a.ifp + static_cast<i64>(b) * T * C;
  v.positives = a.positives + static_cast<i64>(b) * C;
  return v;
}

// Instance outputs of checkpoint row r for thresholds t0, t0 + step, ...:
// counts, FN = positives - TP, precision and recall per (threshold, class),
// over the lanes `first`, `first + stride`, ...
__device__ __forceinline__ void prefix_instance_row(const PrefixArgs& a, const PrefixView& v, i64 r, int t0, int t_step,
                                                    int first, int stride) {
  const int C = a.C, TC = a.T * C;
  for (int t = t0; t < a.T; t += t_step)
    for (int c = first; c < C; c += stride) {
      const int e = t * C + c;
      const i64 t_ = v.itp[e], f_ = v.ifp[e], n_ = v.positives[c];
      a.out_itp[r * TC + e] = t_;
      a.out_ifp[r * TC + e] = f_;
      a.out_ifn[r * TC + e] = n_ - t_;
      a.precision[r * TC + e] = static_cast<double>(t_) / static_cast<double>(t_ + f_ < 1 ? 1 : t_ + f_);
      a.recall[r * TC + e] = static_cast<double>(t_) / static_cast<double>(n_ < 1 ? 1 : n_);
    }
}

// Panoptic and semantic statistics of the visible canvas (snapshot row r),
// plus the instance outputs when INSTANCE (otherwise written by
// prefix_instance_waves).
template <bool INSTANCE>
__device__ __forceinline__ void prefix_snapshot(const PrefixArgs& a, int k, i64 r, int P1, const PrefixView& v,
                                                const int32_t* cls) {
  const int G = a.G, C = a.C, tid = threadIdx.x;
  // kLanes neighbouring lanes share one canvas column and split the GT
  // slots; their partial results combine with shuffles inside the group.
  constexpr int kLanes = 8;
  const int sub = tid % kLanes, first_col = tid / kLanes, col_step = blockDim.x / kLanes;
  for (int j = first_col; j <= k; j += col_step) {
    i64 s = 0;
    for (int g = sub; g < G; g += kLanes) s += v.pairs[static_cast<i64>(g) * P1 + j];
#pragma unroll
    for (int m = 1; m < kLanes; m <<= 1) s += __shfl_xor(s, m, kWave);
    if (sub == 0) v.pa[j] = s;
  }
  for (int g = tid; g < G; g += blockDim.x) v.matched[g] = 0;
  __syncthreads();
  u64* tp = reinterpret_cast<u64*>(a.tp + r * C);
  u64* fp = reinterpret_cast<u64*>(a.fp + r * C);
  u64* fn = reinterpret_cast<u64*>(a.fn + r * C);
  u64* conf = reinterpret_cast<u64*>(a.confusion + r * C * (C + 1));
  for (int j = first_col; j <= k; j += col_step) {
    const i64 pa = v.pa[j];
    const int label = (j > 0 && pa > 0) ? cls[j - 1] : -1;
    const bool valid_pred = label >= 0 && label < C;
    const i64 void_overlap = v.pairs[j];
    // At most one GT slot can hold more than half of the union (GT segments
    // are disjoint), so the matched IoU sum has at most one term; for crowd
    // slots of the label's class the last one (highest slot) decides.
    bool pm = false;
    double sum = 0.0;
    int crowd_g = -1;
    i64 crowd_inter = 0;
    for (int g = sub; g < G; g += kLanes) {
      const int tcls = v.gc[g];
      if (tcls < 0 || tcls >= C) continue;
      const i64 inter = v.pairs[static_cast<i64>(g) * P1 + j];
      const bool is_crowd = v.crowd[g] != 0;
      if (inter && !is_crowd) atomicAdd(conf + static_cast<i64>(tcls) * (C + 1) + (valid_pred ? label : C),
                                        static_cast<u64>(inter));
      if (!valid_pred || tcls != label) continue;
      if (is_crowd) {
        crowd_g = g;
        crowd_inter = inter;
      } else if (inter) {
        const i64 uni = static_cast<i64>(v.ga[g]) + pa - inter - void_overlap;
        if (inter > uni / 2) {
          pm = true;
          sum += static_cast<double>(inter) / static_cast<double>(uni);
          v.matched[g] = 1;
        }
      }
    }
#pragma unroll
    for (int m = 1; m < kLanes; m <<= 1) {
      pm |= __shfl_xor(static_cast<int>(pm), m, kWave) != 0;
      sum += __shfl_xor(sum, m, kWave);
      const int o_g = __shfl_xor(crowd_g, m, kWave);
      const i64 o_inter = __shfl_xor(crowd_inter, m, kWave);
      if (o_g > crowd_g) {
        crowd_g = o_g;
        crowd_inter = o_inter;
      }
    }
    if (sub == 0) {
      const i64 ignored = crowd_g >= 0 ? void_overlap + crowd_inter : void_overlap;
      int out = -1;
      if (valid_pred) {
        if (pm) {
          atomicAdd(tp + label, 1ULL);
          out = label;
        } else if (static_cast<double>(ignored) <= 0.5 * static_cast<double>(pa)) {
          atomicAdd(fp + label, 1ULL);
        }
      }
      v.lab[j] = out;
      v.val[j] = sum;
    }
  }
  __syncthreads();
  for (int g = tid; g < G; g += blockDim.x) {
    const int c = v.gc[g];
    if (c >= 0 && c < C && !v.crowd[g] && !v.matched[g]) atomicAdd(fn + c, 1ULL);
  }
  // Matched columns in increasing order (wave 0, reusing pa's storage), so
  // each class folds its IoUs in column order over that short list.
  int* list = reinterpret_cast<int*>(v.pa);
  __shared__ int list_len;
  if (tid < kWave) {
    int count = 0;
    for (int j0 = 0; j0 <= k; j0 += kWave) {
      const int j = j0 + tid;
      const bool hit = j <= k && v.lab[j] >= 0;
      const u64 ballot = __ballot(hit);
      if (hit) list[count + __popcll(ballot & ((1ULL << tid) - 1))] = j;
      count += __popcll(ballot);
    }
    if (tid == 0) list_len = count;
  }
  __syncthreads();
  double* sums = a.sums + r * C;
  const int matched_cols = list_len;
  for (int c = tid; c < C; c += blockDim.x) {
    double acc = 0.0;
    for (int e = 0; e < matched_cols; ++e) {
      const int j = list[e];
      if (v.lab[j] == c) acc += v.val[j];
    }
    sums[c] = acc;
  }
  if (INSTANCE) prefix_instance_row(a, v, r, 0, 1, tid, blockDim.x);
  __syncthreads();
}

// Independent-mask greedy matching of every mask in generation order
// (prefix.py semantics: same-class non-crowd GT not yet taken, IoU >=
// threshold, the highest IoU and then the lowest slot wins; a crowd match
// only ignores), with wave w owning thresholds t_first + w, + t_step, ...:
// the lanes cover the GT slots, a wave argmax picks the match and lane 0
// counts it.  The matching never reads the canvas, so it runs apart from
// it, without block barriers, and writes its checkpoint rows itself.
// `used` entries are only ever touched by the lane owning that slot, and the
// counts only by lane 0, in LDS, so a wave sees its own writes in order.
__device__ __forceinline__ void prefix_instance_waves(const PrefixArgs& a, const PrefixView& v, int N,
                                                      const int32_t* raw, const int32_t* cls, i64 kfirst,
                                                      i64 klast, int t_first, int t_step) {
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const int G = a.G, C = a.C;
  for (int t = t_first + wave; t < a.T; t += t_step) {
    const double thr = a.thresholds[t];
    i64 q = kfirst;
    if (q < klast && a.points[q] == 0) prefix_instance_row(a, v, q++, t, a.T, lane, kWave);
    for (int i = 0; i < N; ++i) {
      const int c = cls[i];
      if (c >= 0 && c < C && a.things[c]) {
        const int32_t* row = raw + static_cast<i64>(i) * G;
        i64 s = 0;
        for (int g = lane; g < G; g += kWave) s += row[g];
        const double area = static_cast<double>(wave_sum(s));
        int best = INT_MAX;
        double best_iou = -1.0;
        bool ignored = false;
        for (int g = lane; g < G; g += kWave) {
          if (v.gc[g] != c) continue;
          const i64 inter = row[g];
          double uni = v.crowd[g] ? area : area + static_cast<double>(v.ga[g]) - static_cast<double>(inter);
          if (uni < 1.0) uni = 1.0;
          const double iou = static_cast<double>(inter) / uni;
          if (iou < thr) continue;
          if (v.crowd[g]) ignored = true;
          else if (!v.used[t * G + g] && iou > best_iou) {
            best = g;
            best_iou = iou;
          }
        }
#pragma unroll
        for (int m = kWave / 2; m > 0; m >>= 1) {
          const double o_iou = __shfl_xor(best_iou, m, kWave);
          const int o_best = __shfl_xor(best, m, kWave);
          if (o_iou > best_iou || (o_iou == best_iou && o_best < best)) {
            best_iou = o_iou;
            best = o_best;
          }
        }
        ignored = __ballot(ignored) != 0;
        if (best != INT_MAX) {
          if (lane == (best & (kWave - 1))) v.used[t * G + best] = 1;
          if (lane == 0) ++v.itp[t * C + c];
        } else if (!ignored && lane == 0) {
          ++v.ifp[t * C + c];
        }
      }
      if (q < klast && a.points[q] == i + 1) prefix_instance_row(a, v, q++, t, a.T, lane, kWave);
    }
  }
}

// GT metadata of image b into the LDS view.
__device__ __forceinline__ void prefix_stage_targets(const PrefixArgs& a, int b, const PrefixView& v) {
  const int G = a.G;
  const int32_t* gc = a.gc + static_cast<i64>(b) * G;
  const uint8_t* crowd = a.crowd + static_cast<i64>(b) * G;
  const int32_t* ga = a.ga + static_cast<i64>(b) * G;
  for (int g = threadIdx.x; g < G; g += blockDim.x) {
    v.gc[g] = gc[g];
    v.crowd[g] = crowd[g];
    v.ga[g] = ga[g];
  }
  __syncthreads();
}

// Validation of the target metadata and the generated classes (error bits).
__device__ __forceinline__ void prefix_validate(const PrefixArgs& a, const PrefixView& v, int N, const int32_t* raw,
                                                const int32_t* cls) {
  const int G = a.G, C = a.C, tid = threadIdx.x;
  int bits = 0;
  for (int g = tid; g < G; g += blockDim.x) {
    const int c = v.gc[g];
    if (c < -1 || c >= C || (g == 0 && (c != -1 || v.crowd[0]))) bits |= 2;
    if (g > 0 && ((v.ga[g] > 0) != (c >= 0))) bits |= 4;
  }
  for (int i = tid; i < N; i += blockDim.x) {
    if (cls[i] < 0 || cls[i] >= C) bits |= 8;
    i64 area = 0;
    for (int g = 0; g < G; ++g) area += raw[static_cast<i64>(i) * G + g];
    if (area == 0) bits |= 16;
  }
  if (bits) atomicOr(a.error, bits);
}

// Matching state: nothing taken, no counts, positives per thing class.
__device__ __forceinline__ void prefix_clear_matching(const PrefixArgs& a, const PrefixView& v) {
  const int G = a.G, C = a.C, T = a.T, tid = threadIdx.x;
  for (int e = tid; e < T * G; e += blockDim.x) v.used[e] = 0;
  for (int e = tid; e < T * C; e += blockDim.x) {
    v.itp[e] = 0;
    v.ifp[e] = 0;
  }
  for (int c = tid; c < C; c += blockDim.x) {
    i64 n = 0;
    if (a.things[c])
      for (int g = 0; g < G; ++g) n += v.gc[g] == c && !v.crowd[g];
    v.positives[c] = n;
  }
}

// Empty canvas: column 0 (not covered) holds the GT areas.
__device__ __forceinline__ void prefix_clear_canvas(const PrefixView& v, int G, int P1) {
  for (int e = threadIdx.x; e < G * P1; e += blockDim.x) {
    const int g = e / P1;
    v.pairs[e] = (e - g * P1) == 0 ? v.ga[g] : 0;
  }
}

// Step i whose triangle holds cell y (tri(i) <= y < tri(i + 1)).
__device__ __forceinline__ int prefix_step_of(i64 y) {
  int i = static_cast<int>((sqrt(8.0 * static_cast<double>(y) + 1.0) - 1.0) * 0.5);
  while (tri(i + 1) <= y) ++i;
  while (i > 0 && tri(i) > y) --i;
  return i;
}

constexpr int kPrefixLoads = 8;  // loads in flight per thread while moving the canvas

// Moves the LDS canvas from step `from` to step `to` (from <= to): the
// columns of masks from..to-1 gain their raw counts and every earlier column
// loses the decrements of steps from..to-1.  Each entry is a sum of these
// terms, so all of them are added at once with LDS atomics.  The steps'
// triangles are one contiguous range of dec, read as 16-byte vectors from the
// 16-byte boundary at or below its start (the scratch section after the
// triangles keeps the last vector inside the workspace).
__device__ __forceinline__ void prefix_advance(const PrefixArgs& a, const PrefixView& v, const int32_t* raw,
                                               const int32_t* dec, int from, int to, int P1) {
  const int G = a.G, tid = threadIdx.x;
  const i64 rlo = static_cast<i64>(from) * G, rcount = static_cast<i64>(to - from) * G;
  for (i64 e0 = tid; e0 < rcount; e0 += kPrefixThreads * kPrefixLoads) {
    int32_t r[kPrefixLoads];
#pragma unroll
    for (int u = 0; u < kPrefixLoads; ++u) {
      const i64 e = e0 + u * kPrefixThreads;
      r[u] = e < rcount ? raw[rlo + e] : 0;
    }
#pragma unroll
    for (int u = 0; u < kPrefixLoads; ++u) {
      if (!r[u]) continue;
      const i64 x = rlo + e0 + u * kPrefixThreads;
      const int i = static_cast<int>(x / G), g = static_cast<int>(x - static_cast<i64>(i) * G);
      atomicAdd(v.pairs + g * P1 + i + 1, r[u]);
    }
  }
  const i64 lo = G * tri(from), count = G * tri(to) - lo;
  if (count <= 0) return;
  const uintptr_t at = reinterpret_cast<uintptr_t>(dec + lo);
  const v4i* first = reinterpret_cast<const v4i*>(at & ~static_cast<uintptr_t>(15));
  const int lead = static_cast<int>((at & 15) / 4);
  const i64 vectors = (lead + count + 3) / 4;
  for (i64 v0 = tid; v0 < vectors; v0 += kPrefixThreads * kPrefixLoads) {
    v4i d[kPrefixLoads];
#pragma unroll
    for (int u = 0; u < kPrefixLoads; ++u) {
      const i64 vi = v0 + u * kPrefixThreads;
      d[u] = vi < vectors ? __builtin_nontemporal_load(as_global(first + vi)) : v4i{0, 0, 0, 0};
    }
#pragma unroll
    for (int u = 0; u < kPrefixLoads; ++u)
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int n = d[u][c];
        const i64 x = 4 * (v0 + u * kPrefixThreads) + c - lead;
        if (!n || x < 0 || x >= count) continue;  // outside: counts of other steps
        const i64 cell = lo + x;
        const int i = prefix_step_of(cell / G);
        const int r = static_cast<int>(cell - G * tri(i)), g = r / (i + 1);
        atomicSub(v.pairs + g * P1 + (r - g * (i + 1)), n);
      }
  }
}

// Matching workgroup `group` of image b: thresholds group * waves + w.
__device__ __forceinline__ void prefix_match_job(const PrefixArgs& a, int b, int group, const PrefixView& v) {
  const i64 n0 = a.n_off[b];
  const int N = static_cast<int>(a.n_off[b + 1] - n0);
  const int32_t* raw = a.raw + n0 * a.G;
  const int32_t* cls = a.cls + n0;
  const int waves = kPrefixThreads / kWave;
  prefix_stage_targets(a, b, v);
  if (group == 0) prefix_validate(a, v, N, raw, cls);
  prefix_clear_matching(a, v);
  __syncthreads();
  prefix_instance_waves(a, v, N, raw, cls, a.k_off[b], a.k_off[b + 1], group * waves, a.match_groups * waves);
}

// Canvas workgroup of image b for checkpoint rows q0..q1-1.
__device__ __forceinline__ void prefix_canvas_job(const PrefixArgs& a, int b, i64 q0, i64 q1, const PrefixView& v) {
  const i64 n0 = a.n_off[b];
  const int P1 = static_cast<int>(a.n_off[b + 1] - n0) + 1;
  const int32_t* raw = a.raw + n0 * a.G;
  const int32_t* dec = a.dec + dec_base(a, b);
  const int32_t* cls = a.cls + n0;
  prefix_stage_targets(a, b, v);
  prefix_clear_canvas(v, a.G, P1);
  __syncthreads();
  int at = 0;
  for (i64 q = q0; q < q1; ++q) {
    const int k = a.points[q];
    if (k < at || k >= P1) {  // not increasing or beyond the masks: rows left unscored
      if (threadIdx.x == 0) atomicOr(a.error, 32);
      return;
    }
    prefix_advance(a, v, raw, dec, at, k, P1);
    __syncthreads();
    prefix_snapshot<false>(a, k, q, P1, v, cls);  // ends with a barrier
    at = k;
  }
}

// One workgroup per image with the scoring state in the global scratch:
// column i + 1 becomes raw[i], dec[i] is subtracted, the greedy match of
// mask i runs with one lane per IoU threshold, and every checkpoint is
// scored on the way.
__device__ __forceinline__ void prefix_score_image(const PrefixArgs& a, int b, const PrefixView& v) {
  const int tid = threadIdx.x;
  const int G = a.G, C = a.C, T = a.T;
  const i64 n0 = a.n_off[b];
  const int N = static_cast<int>(a.n_off[b + 1] - n0);
  const int P1 = N + 1;
  const int32_t* raw = a.raw + n0 * G;
  const int32_t* dec = a.dec + dec_base(a, b);
  const int32_t* cls = a.cls + n0;
  const i64 kfirst = a.k_off[b], klast = a.k_off[b + 1];
  prefix_validate(a, v, N, raw, cls);
  prefix_clear_canvas(v, G, P1);
  prefix_clear_matching(a, v);
  __syncthreads();
  i64 q = kfirst;
  if (q < klast && a.points[q] == 0) {
    prefix_snapshot<true>(a, 0, q, P1, v, cls);
    ++q;
  }
  for (int i = 0; i < N; ++i) {
    const int32_t* dec_i = dec + static_cast<i64>(G) * tri(i);
    const int width = i + 1;
    // Earlier masks lose what mask i covers: G x (i + 1) decrements, eight
    // loads in flight per thread.
    constexpr int U = 8;
    const int total = G * width;
    for (int e0 = tid; e0 < total; e0 += kPrefixThreads * U) {
      int32_t d[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int e = e0 + u * kPrefixThreads;
        d[u] = e < total ? dec_i[e] : 0;
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (!d[u]) continue;
        const int e = e0 + u * kPrefixThreads;
        const int g = e / width;
        v.pairs[g * P1 + (e - g * width)] -= d[u];
      }
    }
    for (int g = tid; g < G; g += blockDim.x) v.pairs[g * P1 + i + 1] = raw[static_cast<i64>(i) * G + g];
    __syncthreads();
    const int c = cls[i];
    if (c >= 0 && c < C && a.things[c]) {
      double area = 0.0;
      {
        i64 s = 0;
        for (int g = 0; g < G; ++g) s += raw[static_cast<i64>(i) * G + g];
        area = static_cast<double>(s);
      }
      for (int t = tid; t < T; t += blockDim.x) {
        const double thr = a.thresholds[t];
        int best = -1;
        double best_iou = -1.0;
        bool ignored = false;
        for (int g = 0; g < G; ++g) {
          if (v.gc[g] != c) continue;
          const i64 inter = raw[static_cast<i64>(i) * G + g];
          double uni = v.crowd[g] ? area : area + static_cast<double>(v.ga[g]) - static_cast<double>(inter);
          if (uni < 1.0) uni = 1.0;
          const double iou = static_cast<double>(inter) / uni;
          if (iou < thr) continue;
          if (v.crowd[g]) ignored = true;
          else if (!v.used[t * G + g] && iou > best_iou) {
            best = g;
            best_iou = iou;
          }
        }
        if (best >= 0) {
          v.used[t * G + best] = 1;
          ++v.itp[t * C + c];
        } else if (!ignored) {
          ++v.ifp[t * C + c];
        }
      }
    }
    __syncthreads();
    if (q < klast && a.points[q] == i + 1) {
      prefix_snapshot<true>(a, i + 1, q, P1, v, cls);
      ++q;
    }
  }
}

// STAGED: the scoring state of every workgroup lives in LDS (sized for the
// largest image of the launch); matching workgroups come first, then the
// canvas slices (surplus workgroups, whose count is only known on the
// device, exit at once).  Otherwise one workgroup per image in the global
// scratch.
template <bool STAGED>
__global__ __launch_bounds__(kPrefixThreads) void prefix_score(PrefixArgs a) {
  extern __shared__ uint4 lds4[];
  int job = blockIdx.x;
  if (!STAGED) {
    prefix_score_image(a, job, global_view(a, job, a.n_off[job]));
    return;
  }
  const int match_jobs = a.B * a.match_groups;
  int b;
  if (job < match_jobs) {
    b = job / a.match_groups;
  } else {
    job -= match_jobs;
    if (job >= a.slices.count(a.B)) return;
    b = a.slices.image_of(job);
  }
  const int P1 = static_cast<int>(a.n_off[b + 1] - a.n_off[b]) + 1;
  const PrefixView v = lds_view(reinterpret_cast<char*>(lds4), a.G, P1, a.C, a.T);
  if (static_cast<int>(blockIdx.x) < match_jobs) {
    prefix_match_job(a, b, job - b * a.match_groups, v);
  } else {
    const i64 q0 = a.k_off[b] + static_cast<i64>(job - a.slices.first_tile(b)) * a.slice_rows;
    prefix_canvas_job(a, b, q0, min(q0 + a.slice_rows, a.k_off[b + 1]), v);
  }
}

}  // namespace pe

using namespace pe;

namespace {
struct PrefixScratch {
  int32_t* raw;
  int32_t* dec;
  int32_t* pairs;
  i64* pa;
  int32_t* ga;
  uint8_t* used;
  int32_t* matched;
  i64* itp;
  i64* ifp;
  i64* positives;
  int32_t* lab;
  double* val;
  int* tile_list;     // packed batches: tile_off then tile_image
  int* slice_list;    // staged scoring: slice_off then slice_image
  size_t zero_bytes;  // raw + dec + ga, contiguous at the start
  size_t used_bytes;
};

PrefixScratch carve_prefix(void* base, i64 B, i64 G, i64 C, i64 T, i64 sum_n, i64 sum_tri, i64 packed_pixels) {
  Carve carve(base);
  PrefixScratch s{};
  s.raw = carve.take<int32_t>(sum_n * G);
  s.dec = carve.take<int32_t>(sum_tri * G);  // followed by ga: 16-byte reads past its end stay inside
  s.ga = carve.take<int32_t>(B * G);
  s.zero_bytes = carve.used;
  s.pairs = carve.take<int32_t>((sum_n + B) * G);
  s.pa = carve.take<i64>(sum_n + B);
  s.lab = carve.take<int32_t>(sum_n + B);
  s.val = carve.take<double>(sum_n + B);
  s.used = carve.take<uint8_t>(B * T * G);
  s.matched = carve.take<int32_t>(B * G);
  s.itp = carve.take<i64>(B * T * C);
  s.ifp = carve.take<i64>(B * T * C);
  s.positives = carve.take<i64>(B * C);
  s.tile_list = packed_pixels > 0 ? carve.take<int>(tile_list_ints(B, packed_pixels, kPrefixTile)) : nullptr;
  // At most N_b + 1 checkpoints per image, slices of at least one row.
  s.slice_list = carve.take<int>(tile_list_ints(B, sum_n + B, 1));
  s.used_bytes = carve.used;
  return s;
}
}  // namespace

extern "C" int64_t pe_prefix_workspace_bytes(int64_t B, int64_t G, int64_t C, int64_t T, int64_t sum_n,
                                             int64_t sum_tri, int64_t packed_pixels) {
  return static_cast<int64_t>(
      align_up(carve_prefix(nullptr, B, G, C, T, sum_n, sum_tri, packed_pixels).used_bytes, 256));
}

extern "C" int pe_prefix(int64_t B, const int32_t* target, const int64_t* pix_off, int64_t HW, int64_t G, int64_t C,
                         const uint8_t* const* masks, const int64_t* n_off, int64_t sum_n, int64_t sum_tri,
                         int64_t max_n,
                         const int32_t* cls, const int32_t* gc, const uint8_t* crowd, const uint8_t* things,
                         const double* thresholds, int64_t T, const int64_t* k_off, const int32_t* points,
                         int64_t total_rows, int64_t* tp, int64_t* fp, int64_t* fn, double* sums,
                         int64_t* confusion, int64_t* out_itp, int64_t* out_ifp, int64_t* out_ifn,
                         double* precision, double* recall, int32_t* error, void* workspace,
                         int64_t workspace_bytes, int aligned, void* stream_ptr) {
  PE_REQUIRE(B >= 0 && HW > 0 && G > 0 && G < INT_MAX && C > 0 && C < INT_MAX && T > 0 && T * C < INT_MAX,
             "prefix dimensions outside supported range");
  PE_REQUIRE(sum_n >= 0 && sum_tri >= 0 && total_rows >= 0 && total_rows <= sum_n + B && max_n >= 0 &&
                 max_n <= sum_n,
             "invalid prefix sizes");
  PE_REQUIRE(G * (max_n + 1) < INT_MAX && G * (max_n * (max_n + 1) / 2) < INT_MAX, "prefix tables too large");
  const i64 packed_pixels = pix_off ? HW : 0;
  PE_REQUIRE(workspace_bytes >= pe_prefix_workspace_bytes(B, G, C, T, sum_n, sum_tri, packed_pixels),
             "workspace too small");
  hipStream_t stream = static_cast<hipStream_t>(stream_ptr);
  int rc;
  if ((rc = launch_zero(error, 4, stream))) return rc;
  if (total_rows) {
    const size_t rows = static_cast<size_t>(total_rows);
    if ((rc = launch_zero(tp, 8 * rows * C, stream))) return rc;
    if ((rc = launch_zero(fp, 8 * rows * C, stream))) return rc;
    if ((rc = launch_zero(fn, 8 * rows * C, stream))) return rc;
    if ((rc = launch_zero(confusion, 8 * rows * C * (C + 1), stream))) return rc;
  }
  if (B == 0) return kOk;
  PrefixScratch s = carve_prefix(workspace, B, G, C, T, sum_n, sum_tri, packed_pixels);
  if ((rc = launch_zero(workspace, s.zero_bytes, stream))) return rc;
  PrefixArgs a{};
  a.B = static_cast<int>(B);
  a.G = static_cast<int>(G);
  a.C = static_cast<int>(C);
  a.T = static_cast<int>(T);
  a.target = target;
  a.masks = masks;
  a.n_off = reinterpret_cast<const i64*>(n_off);
  a.cls = cls;
  a.gc = gc;
  a.crowd = crowd;
  a.things = things;
  a.thresholds = thresholds;
  a.k_off = reinterpret_cast<const i64*>(k_off);
  a.points = points;
  a.raw = s.raw;
  a.dec = s.dec;
  a.pairs = s.pairs;
  a.pa = s.pa;
  a.ga = s.ga;
  a.used = s.used;
  a.matched = s.matched;
  a.itp = s.itp;
  a.ifp = s.ifp;
  a.positives = s.positives;
  a.lab = s.lab;
  a.val = s.val;
  a.tp = reinterpret_cast<i64*>(tp);
  a.fp = reinterpret_cast<i64*>(fp);
  a.fn = reinterpret_cast<i64*>(fn);
  a.sums = sums;
  a.confusion = reinterpret_cast<i64*>(confusion);
  a.out_itp = reinterpret_cast<i64*>(out_itp);
  a.out_ifp = reinterpret_cast<i64*>(out_ifp);
  a.out_ifn = reinterpret_cast<i64*>(out_ifn);
  a.precision = precision;
  a.recall = recall;
  a.error = error;
  a.aligned = aligned;
  TileList& t = a.tiles;
  t.tile = kPrefixTile;
  i64 blocks;
  if (pix_off) {
    // HW is the pixel total of the packed batch.
    t.pix_off = reinterpret_cast<const i64*>(pix_off);
    t.tile_off = s.tile_list;
    t.tile_image = s.tile_list + B + 1;
    if ((rc = launch_tile_list(t.pix_off, a.B, kPrefixTile, s.tile_list, s.tile_list + B + 1, stream))) return rc;
    blocks = tile_list_bound(B, HW, kPrefixTile);
  } else {
    t.HW = HW;
    t.tiles_per_image = static_cast<int>((HW + kPrefixTile - 1) / kPrefixTile);
    blocks = B * t.tiles_per_image;
  }
  PE_REQUIRE(blocks < INT_MAX, "too many workgroups");
  // The LDS step table when G x (max_n + 1) fits.
  bool table = static_cast<i64>(G) * (max_n + 1) <= kPrefixTableMax;
  size_t lds = prefix_lds_bytes(static_cast<int>(G), static_cast<int>(table ? max_n + 1 : 1), table);
  if (table && lds > 64 * 1024 &&
      hipFuncSetAttribute(reinterpret_cast<const void*>(prefix_events<true>),
                          hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)) != hipSuccess) {
    table = false;
    lds = prefix_lds_bytes(static_cast<int>(G), 1, false);
  }
  PE_REQUIRE(lds <= 64 * 1024 || table, "too many target slots for the prefix kernel");
  {
    ProfileScope timing("prefix_events", stream);
    if (table)
      prefix_events<true><<<static_cast<unsigned>(blocks), kPrefixThreads, lds, stream>>>(a);
    else
      prefix_events<false><<<static_cast<unsigned>(blocks), kPrefixThreads, lds, stream>>>(a);
    PE_LAUNCH_CHECK();
  }
  {
    // Scoring state in LDS when the largest image's fits.
    const size_t view = prefix_view_bytes(static_cast<int>(G), static_cast<int>(max_n + 1), static_cast<int>(C),
                                          static_cast<int>(T));
    const bool staged = view <= kPrefixStageMax &&
                        hipFuncSetAttribute(reinterpret_cast<const void*>(prefix_score<true>),
                                            hipFuncAttributeMaxDynamicSharedMemorySize,
                                            static_cast<int>(kPrefixStageMax)) == hipSuccess;
    if (staged) {
      // Slices of about total_rows / (2 x CUs) checkpoints (two staged
      // workgroups fit a CU), matching workgroups of one wave per threshold.
      const i64 waves = kPrefixThreads / kWave;
      const i64 target = 2 * static_cast<i64>(compute_units());
      a.match_groups = static_cast<int>((T + waves - 1) / waves);
      a.slice_rows = static_cast<int>(std::max<i64>(1, (total_rows + target - 1) / target));
      TileList& sl = a.slices;
      sl.pix_off = a.k_off;
      sl.tile = a.slice_rows;
      sl.tile_off = s.slice_list;
      sl.tile_image = s.slice_list + B + 1;
      if ((rc = launch_tile_list(a.k_off, a.B, a.slice_rows, s.slice_list, s.slice_list + B + 1, stream)))
        return rc;
      const i64 jobs = B * a.match_groups + tile_list_bound(B, total_rows, a.slice_rows);
      PE_REQUIRE(jobs < INT_MAX, "too many workgroups");
      ProfileScope timing("prefix_score", stream);
      prefix_score<true><<<static_cast<unsigned>(jobs), kPrefixThreads, view, stream>>>(a);
    } else {
      ProfileScope timing("prefix_score", stream);
      prefix_score<false><<<static_cast<unsigned>(B), kPrefixThreads, 0, stream>>>(a);
    }
    PE_LAUNCH_CHECK();
  }
  return kOk;
}
