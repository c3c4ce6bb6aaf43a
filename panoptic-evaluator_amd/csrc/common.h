// Shared device/host helpers for the gfx950 evaluation kernels.
//
// Status convention of every extern "C" entry point:
//   0            success
//   1 | 2 | 4    input-validation bitmask (only from entry points that validate)
//   < 0          argument or HIP runtime error; pe_last_error() has the text.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdio>
#include <cstring>

namespace pe {

using u64 = unsigned long long;
using i64 = int64_t;

constexpr int kWave = 64;  // CDNA wavefront width

enum Status : int {
  kOk = 0,
  kBadArgument = -1,
  kWorkspaceTooSmall = -2,
  kUnsupportedShape = -3,
  kHipError = -4,
};

// Per-thread message buffer behind pe_last_error().
char* error_buffer();
int fail(int code, const char* fmt, ...);

#define PE_HIP(expr)                                                               \
  do {                                                                             \
    hipError_t pe_status_ = (expr);                                                \
    if (pe_status_ != hipSuccess)                                                  \
      return ::pe::fail(::pe::kHipError, "%s failed: %s", #expr,                   \
                        hipGetErrorString(pe_status_));                            \
  } while (0)

#define PE_LAUNCH_CHECK()                                                          \
  do {                                                                             \
    hipError_t pe_status_ = hipGetLastError();                                     \
    if (pe_status_ != hipSuccess)                                                  \
      return ::pe::fail(::pe::kHipError, "kernel launch failed: %s",               \
                        hipGetErrorString(pe_status_));                            \
  } while (0)

#define PE_REQUIRE(cond, ...)                                                      \
  do {                                                                             \
    if (!(cond)) return ::pe::fail(::pe::kBadArgument, __VA_ARGS__);              \
  } while (0)

// Compute units of the current device (cached per device ordinal).
int compute_units();

// Whether `stream` was marked (pe_stream_set_concurrent) as running beside
// another stream's long kernel: latency-bound kernels launched on it then
// use workgroups small enough to fit beside that kernel's workgroups.
bool stream_concurrent(hipStream_t stream);
// Workgroup size of the latency-bound kernels on a concurrent stream.
constexpr int kConcurrentThreads = 256;

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Bump allocator over a caller-provided workspace.  Every slice is 256-byte
// aligned so vector loads/stores on any slice stay aligned.
struct Carve {
  char* base;
  size_t used = 0;
  explicit Carve(void* p) : base(static_cast<char*>(p)) {}
  template <typename T>
  T* take(size_t count) {
    used = align_up(used, 256);
    T* out = reinterpret_cast<T*>(base ? base + used : nullptr);
    used += count * sizeof(T);
    return out;
  }
};

// ---------------------------------------------------------------- device side

__device__ inline i64 wave_sum(i64 v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Butterfly sum: every lane ends with the same value (a + b == b + a), and the
// pairing is fixed, so the result does not depend on scheduling.
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

__device__ inline int lane_id() { return __lane_id(); }

// Global-address-space view of a pointer.  Pointers read from memory inside
// a loop (per-image mask tables, offsets) are otherwise generic, and generic
// (flat) loads also count against the LDS wait counter.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(1))) T* as_global(const T* p) {
  return (const __attribute__((address_space(1))) T*)(p);
}

// Zero-fill with 16-byte stores (graph-capture safe replacement for memset).
__global__ void zero_words(uint4* data, i64 count16, uint32_t* tail, i64 tail_words);

int launch_zero(void* ptr, size_t bytes, hipStream_t stream);
// Clears words [0, w_i) of each non-null p_i in one launch.
int launch_zero3(uint32_t* p0, i64 w0, uint32_t* p1, i64 w1, uint32_t* p2, i64 w2, hipStream_t stream);

// ---- work lists over packed batches of uneven images
//
// A packed batch stores its images one after another; image b covers pixels
// pix_off[b] .. pix_off[b + 1] of the flat maps.  Kernels walk it as a list of
// tiles, each inside one image: image b has ceil(pixels_b / tile) tiles,
// tile_off[b] is their exclusive prefix and tile_image[t] the image of tile t.
// Uniform batches (pix_off == nullptr) use arithmetic instead of the tables.
struct TileList {
  const i64* pix_off;     // B + 1, or nullptr: every image has HW pixels
  i64 HW;
  int tile;               // pixels per tile
  int tiles_per_image;    // uniform batches
  const int* tile_off;    // B + 1 (packed)
  const int* tile_image;  // tile_off[B] entries (packed)

  __device__ __forceinline__ int count(int B) const { return pix_off ? tile_off[B] : B * tiles_per_image; }
  __device__ __forceinline__ int image_of(int t) const { return pix_off ? tile_image[t] : t / tiles_per_image; }
  __device__ __forceinline__ int first_tile(int b) const { return pix_off ? tile_off[b] : b * tiles_per_image; }
  __device__ __forceinline__ i64 begin(int b) const { return pix_off ? pix_off[b] : static_cast<i64>(b) * HW; }
  __device__ __forceinline__ i64 pixels(int b) const { return pix_off ? pix_off[b + 1] - pix_off[b] : HW; }
};

// Upper bound of the tile count of a packed batch (tiles of >= min_tile pixels).
inline i64 tile_list_bound(i64 B, i64 total_pixels, i64 min_tile) { return total_pixels / min_tile + B; }
// Ints of workspace a packed tile list needs (tile_off then tile_image).
inline i64 tile_list_ints(i64 B, i64 total_pixels, i64 min_tile) {
  return B + 1 + tile_list_bound(B, total_pixels, min_tile);
}
// Builds tile_off / tile_image on the device (one workgroup); graph-capture safe.
int launch_tile_list(const i64* pix_off, int B, int tile, int* tile_off, int* tile_image, hipStream_t stream);
// Byte fill kernel (graph-capture safe replacement for hipMemsetD8Async).
int launch_fill_u8(void* ptr, uint8_t value, size_t bytes, hipStream_t stream);

}  // namespace pe

namespace pe {

// Benchmark timing: when enabled (pe_profile_enable), a pair of HIP events is
// recorded on the launch stream around every launch of a named kernel.
// Intended for eager runs; leave disabled while capturing HIP graphs.
bool profiling_enabled();
struct ProfileScope {
  const char* name;
  hipStream_t stream;
  hipEvent_t start = nullptr;
  ProfileScope(const char* kernel, hipStream_t s);
  ~ProfileScope();
};

// Bit-packed masks: a mask of `pixels` pixels is a row of mask_words(pixels)
// 32-bit words, bit p % 32 of word p / 32 being pixel p; rows are padded with
// zero bits to a multiple of 128 pixels so that every row is 16-byte aligned.
__host__ __device__ inline i64 mask_words(i64 pixels) { return (pixels + 127) / 128 * 4; }

// Launch-shape tuning.  Each tunable kernel has a fixed list of variants
// (template instantiations and launch shapes, all computing identical
// results); the selected variant is used by every later launch in the
// process.  Selection is process-wide and meant to be set once, before
// launches are issued (the autotuner times each variant and keeps the best).
enum TuneKernel { kTuneDetectionHistogram = 0, kTuneJointHistogram, kTuneMaskOverlap, kTunePackedGrid,
                  kTuneDetectionGroup, kTuneMaskOverlapBits, kTuneDetectionHistogramBits, kTuneKernels };
struct TuneSpec {
  const char* name;
  int count;
  int fallback;                   // default variant
  const char* const* describe;    // count entries
};
const TuneSpec& tune_spec(TuneKernel kernel);
int tuned_variant(TuneKernel kernel);
// Workgroups per compute unit of the persistent grids over packed batches
// (0: one workgroup per work item).
int packed_grid_per_cu();
// Detections per workgroup of the detection histogram (0: as many as the LDS
// budget holds); more groups re-read the target tile but need less LDS.
int detection_group_limit();

}  // namespace pe
