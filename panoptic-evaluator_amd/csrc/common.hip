// Host helpers shared by every translation unit: error text, device queries
// and the zero-fill kernel used instead of hipMemsetAsync (so that callers may
// capture any entry point in a HIP graph).
#include <cstdarg>
#include <algorithm>
#include <atomic>
#include <mutex>
#include <vector>

#include "common.h"
#include "../../include/panoptic_hip.h"

namespace pe {

char* error_buffer() {
  static thread_local char buffer[512];
  return buffer;
}

int fail(int code, const char* fmt, ...) {
  va_list args;
  va_start(args, fmt);
  vsnprintf(error_buffer(), 512, fmt, args);
  va_end(args);
  return code;
}

int compute_units() {
  static int cached[64] = {0};
  int device = 0;
  if (hipGetDevice(&device) != hipSuccess || device < 0 || device >= 64) return 256;
  if (!cached[device]) {
    int count = 0;
    if (hipDeviceGetAttribute(&count, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
        count <= 0)
      count = 256;
    cached[device] = count;
  }
  return cached[device];
}

__global__ void zero_words(uint4* data, i64 count16, uint32_t* tail, i64 tail_words) {
  i64 stride = static_cast<i64>(gridDim.x) * blockDim.x;
  for (i64 i = static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x; i < count16; i += stride)
    data[i] = make_uint4(0, 0, 0, 0);
  if (blockIdx.x == 0)
    for (i64 i = threadIdx.x; i < tail_words; i += blockDim.x) tail[i] = 0;
}

int launch_zero(void* ptr, size_t bytes, hipStream_t stream) {
  if (!bytes) return kOk;
  // Callers pass 4-byte multiples; the body uses 16-byte stores when aligned.
  char* p = static_cast<char*>(ptr);
  size_t lead = (16 - (reinterpret_cast<uintptr_t>(p) & 15)) & 15;
  if (lead > bytes) lead = bytes;
  if (lead % 4) return fail(kBadArgument, "zero fill needs 4-byte alignment");
  if (lead) {
    zero_words<<<1, 64, 0, stream>>>(nullptr, 0, reinterpret_cast<uint32_t*>(p), lead / 4);
    p += lead;
    bytes -= lead;
  }
  i64 count16 = bytes / 16;
  i64 tail_words = (bytes % 16) / 4;
  int blocks = static_cast<int>(std::min<i64>((count16 + 255) / 256, 4 * compute_units()));
  if (blocks < 1) blocks = 1;
  zero_words<<<blocks, 256, 0, stream>>>(reinterpret_cast<uint4*>(p), count16,
                                         reinterpret_cast<uint32_t*>(p + count16 * 16), tail_words);
  PE_LAUNCH_CHECK();
  return kOk;
}

// Up to three word ranges cleared by one launch (coalesced 4-byte stores).
struct ZeroRanges {
  uint32_t* ptr[3];
  i64 words[3];
};

__global__ void zero_ranges(ZeroRanges r) {
  const i64 a = r.words[0], ab = a + r.words[1], total = ab + r.words[2];
  const i64 stride = static_cast<i64>(gridDim.x) * blockDim.x;
  for (i64 i = static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += stride) {
    if (i < a)
      r.ptr[0][i] = 0;
    else if (i < ab)
      r.ptr[1][i - a] = 0;
    else
      r.ptr[2][i - ab] = 0;
  }
}

int launch_zero3(uint32_t* p0, i64 w0, uint32_t* p1, i64 w1, uint32_t* p2, i64 w2, hipStream_t stream) {
  ZeroRanges r{{p0, p1, p2}, {p0 ? w0 : 0, p1 ? w1 : 0, p2 ? w2 : 0}};
  const i64 total = r.words[0] + r.words[1] + r.words[2];
  if (total == 0) return kOk;
  const int blocks = static_cast<int>(std::min<i64>((total + 255) / 256, 4 * compute_units()));
  zero_ranges<<<blocks, 256, 0, stream>>>(r);
  PE_LAUNCH_CHECK();
  return kOk;
}

// One workgroup: per-thread runs of images are summed, the run totals scanned
// in LDS, then every image writes its tiles' image index.
constexpr int kTileListThreads = 1024;

__global__ __launch_bounds__(kTileListThreads) void build_tile_list(const i64* pix_off, int B, int tile,
                                                                      int* tile_off, int* tile_image) {
  __shared__ int run_total[kTileListThreads];
  const int tid = threadIdx.x;
  const int per = (B + kTileListThreads - 1) / kTileListThreads;
  const int b0 = min(tid * per, B), b1 = min(b0 + per, B);
  int sum = 0;
  for (int b = b0; b < b1; ++b) sum += static_cast<int>((pix_off[b + 1] - pix_off[b] + tile - 1) / tile);
  run_total[tid] = sum;
  __syncthreads();
  // Inclusive Hillis-Steele scan of the run totals.
  for (int d = 1; d < kTileListThreads; d <<= 1) {
    const int v = tid >= d ? run_total[tid - d] : 0;
    __syncthreads();
    run_total[tid] += v;
    __syncthreads();
  }
  int off = run_total[tid] - sum;
  for (int b = b0; b < b1; ++b) {
    tile_off[b] = off;
    off += static_cast<int>((pix_off[b + 1] - pix_off[b] + tile - 1) / tile);
  }
  if (tid == kTileListThreads - 1) tile_off[B] = run_total[tid];
  __syncthreads();
  for (int b = b0; b < b1; ++b) {
    const int first = tile_off[b], last = tile_off[b + 1];
    for (int t = first; t < last; ++t) tile_image[t] = b;
  }
}

int launch_tile_list(const i64* pix_off, int B, int tile, int* tile_off, int* tile_image, hipStream_t stream) {
  if (B <= 0) return kOk;
  if (tile <= 0) return fail(kBadArgument, "tile size must be positive");
  build_tile_list<<<1, kTileListThreads, 0, stream>>>(pix_off, B, tile, tile_off, tile_image);
  PE_LAUNCH_CHECK();
  return kOk;
}

__global__ void fill_bytes(uint8_t* data, uint8_t value, i64 count) {
  const i64 stride = static_cast<i64>(gridDim.x) * blockDim.x;
  for (i64 i = static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride) data[i] = value;
}

int launch_fill_u8(void* ptr, uint8_t value, size_t bytes, hipStream_t stream) {
  if (!bytes) return kOk;
  int blocks = static_cast<int>(std::min<i64>((static_cast<i64>(bytes) + 255) / 256, 4 * compute_units()));
  fill_bytes<<<blocks, 256, 0, stream>>>(static_cast<uint8_t*>(ptr), value, static_cast<i64>(bytes));
  PE_LAUNCH_CHECK();
  return kOk;
}

}  // namespace pe

extern "C" const char* pe_last_error(void) { return pe::error_buffer(); }

extern "C" int pe_abi_version(void) { return PE_ABI_VERSION; }

// ------------------------------------------------- streaming-read ceiling

namespace pe {
typedef int v4i_read __attribute__((ext_vector_type(4)));

// Reads n16 16-byte vectors with four non-temporal loads in flight per lane
// (grid-stride) and folds them into a value that is stored only if it equals
// a sentinel the compiler cannot see (a kernel argument), so the loads stay.
__global__ __launch_bounds__(256) void stream_read(const v4i_read* data, i64 n16, int sentinel, int* sink) {
  constexpr int U = 4;
  int acc = 0;
  const i64 stride = static_cast<i64>(gridDim.x) * blockDim.x;
  for (i64 i = static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x; i < n16; i += stride * U) {
    v4i_read v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const i64 k = i + u * stride;
      v[u] = k < n16 ? __builtin_nontemporal_load(data + k) : v4i_read{0, 0, 0, 0};
    }
#pragma unroll
    for (int u = 0; u < U; ++u) acc += v[u].x ^ v[u].y ^ v[u].z ^ v[u].w;
  }
  if (acc == sentinel) sink[0] = acc;
}
}  // namespace pe

extern "C" int pe_stream_read(const void* data, int64_t bytes, int blocks, int* sink, void* stream) {
  PE_REQUIRE(data && sink, "null buffer");
  PE_REQUIRE(reinterpret_cast<uintptr_t>(data) % 16 == 0 && bytes > 0 && bytes % 16 == 0,
             "data must be 16-byte aligned and a positive multiple of 16 bytes");
  PE_REQUIRE(blocks > 0 && blocks <= 65536, "blocks must be in 1..65536");
  hipStream_t s = static_cast<hipStream_t>(stream);
  pe::ProfileScope timing("stream_read", s);
  pe::stream_read<<<blocks, 256, 0, s>>>(static_cast<const pe::v4i_read*>(data), bytes / 16, 0x5eed1234, sink);
  PE_LAUNCH_CHECK();
  return pe::kOk;
}

// ---------------------------------------------------- stream fork and join

extern "C" int pe_event_create(void** event) {
  PE_REQUIRE(event, "event output pointer is null");
  hipEvent_t e = nullptr;
  PE_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  *event = e;
  return pe::kOk;
}

extern "C" int pe_event_destroy(void* event) {
  if (event) PE_HIP(hipEventDestroy(static_cast<hipEvent_t>(event)));
  return pe::kOk;
}

extern "C" int pe_stream_wait(void* waiting, void* signalling, void* event) {
  PE_REQUIRE(event, "event is null");
  PE_HIP(hipEventRecord(static_cast<hipEvent_t>(event), static_cast<hipStream_t>(signalling)));
  PE_HIP(hipStreamWaitEvent(static_cast<hipStream_t>(waiting), static_cast<hipEvent_t>(event), 0));
  return pe::kOk;
}

// ------------------------------------------------- concurrent streams

namespace pe {
namespace {
std::mutex concurrent_lock;
std::vector<hipStream_t> concurrent_streams;
std::atomic<int> concurrent_count{0};
}  // namespace

bool stream_concurrent(hipStream_t stream) {
  if (concurrent_count.load(std::memory_order_relaxed) == 0) return false;
  std::lock_guard<std::mutex> guard(concurrent_lock);
  for (hipStream_t s : concurrent_streams)
    if (s == stream) return true;
  return false;
}
}  // namespace pe

extern "C" int pe_stream_set_concurrent(void* stream, int on) {
  using namespace pe;
  PE_REQUIRE(stream, "the default stream cannot be marked");
  std::lock_guard<std::mutex> guard(concurrent_lock);
  hipStream_t s = static_cast<hipStream_t>(stream);
  auto it = std::find(concurrent_streams.begin(), concurrent_streams.end(), s);
  if (on && it == concurrent_streams.end()) concurrent_streams.push_back(s);
  if (!on && it != concurrent_streams.end()) concurrent_streams.erase(it);
  concurrent_count.store(static_cast<int>(concurrent_streams.size()), std::memory_order_relaxed);
  return kOk;
}

// ------------------------------------------------------------ timing hooks

#include <atomic>
#include <map>
#include <string>
#include <utility>
#include <vector>

namespace pe {
namespace {
struct Profile {
  std::mutex lock;
  bool enabled = false;
  std::map<std::string, std::vector<std::pair<hipEvent_t, hipEvent_t>>> spans;
  std::vector<hipEvent_t> pool;
  hipEvent_t take() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
  }
};
Profile& profile() {
  static Profile p;
  return p;
}
}  // namespace

bool profiling_enabled() { return profile().enabled; }

ProfileScope::ProfileScope(const char* kernel, hipStream_t s) : name(kernel), stream(s) {
  Profile& p = profile();
  if (!p.enabled) return;
  std::lock_guard<std::mutex> guard(p.lock);
  start = p.take();
  if (start && hipEventRecord(start, stream) != hipSuccess) start = nullptr;
}

ProfileScope::~ProfileScope() {
  if (!start) return;
  Profile& p = profile();
  std::lock_guard<std::mutex> guard(p.lock);
  hipEvent_t stop = p.take();
  if (stop && hipEventRecord(stop, stream) == hipSuccess) p.spans[name].emplace_back(start, stop);
}

}  // namespace pe

extern "C" int pe_profile_enable(int on) {
  pe::profile().enabled = on != 0;
  return 0;
}

extern "C" int pe_profile_reset(void) {
  pe::Profile& p = pe::profile();
  std::lock_guard<std::mutex> guard(p.lock);
  for (auto& item : p.spans)
    for (auto& span : item.second) {
      p.pool.push_back(span.first);
      p.pool.push_back(span.second);
    }
  p.spans.clear();
  return 0;
}

extern "C" int pe_profile_read(const char* kernel, double* total_ms, int64_t* launches) {
  pe::Profile& p = pe::profile();
  std::lock_guard<std::mutex> guard(p.lock);
  double total = 0.0;
  int64_t count = 0;
  auto it = p.spans.find(kernel);
  if (it != p.spans.end()) {
    for (auto& span : it->second) {
      if (hipEventSynchronize(span.second) != hipSuccess)
        return pe::fail(pe::kHipError, "event synchronize failed");
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, span.first, span.second) != hipSuccess)
        return pe::fail(pe::kHipError, "event elapsed time failed");
      total += ms;
      ++count;
    }
  }
  if (total_ms) *total_ms = total;
  if (launches) *launches = count;
  return 0;
}

// ------------------------------------------------------------------ tuning

namespace pe {
namespace {
const char* const kDetVariants[] = {"1 chunk x 4 masks", "1 chunk x 8 masks", "2 chunks x 2 masks",
                                    "2 chunks x 4 masks", "4 chunks x 2 masks"};
const char* const kJointVariants[] = {"auto chunk", "8192 px", "16384 px", "32768 px", "65532 px"};
const char* const kOverlapVariants[] = {
    "LDS-staged 128 dets x 64 GT x 8192 px, 256 px blocks",
    "LDS-staged 128 dets x 64 GT x 4096 px, 256 px blocks",
    "LDS-staged 64 dets x 64 GT x 8192 px, 256 px blocks",
    "LDS-staged 128 dets x 32 GT x 8192 px, 512 px blocks",
    "registers 16 dets x 64 GT x 4096 px, 4 steps ahead, XCD-grouped",
    "registers 16 dets x 64 GT x 8192 px, 4 steps ahead, XCD-grouped",
    "registers 16 dets x 128 GT x 16384 px, XCD-grouped",
    "registers 16 dets x 64 GT x 8192 px, 4 steps ahead, spread",
    "LDS-staged 128 dets x 64 GT x 4096 px, 256 px blocks, 2 detection tiles per wave"};
const char* const kGridVariants[] = {"one workgroup per work item", "persistent: 8 workgroups per CU",
                                     "persistent: 4 workgroups per CU"};
const char* const kGroupVariants[] = {"LDS budget (96 KB)", "128 detections", "64 detections", "32 detections"};
const char* const kBitsVariants[] = {
    "bits expanded to int8 MFMA operands in LDS, 128 dets x 64 GT x 4096 px",
    "bits expanded to int8 MFMA operands in LDS, 128 dets x 64 GT x 4096 px, 4 blocks ahead",
    "bits expanded to int8 MFMA operands in LDS, 128 dets x 64 GT x 8192 px, 4 blocks ahead",
    "bits expanded to int8 MFMA operands in LDS, 128 dets x 64 GT x 16384 px, 8 blocks ahead",
    "bits expanded to int8 MFMA operands in LDS, 128 dets x 64 GT x 4096 px, 2 detection tiles per wave",
    "popcounts of ANDed words, 128 dets x 64 GT x 16384 px",
    "popcounts of ANDed words, 64 dets x 64 GT x 16384 px",
    "detection operands expanded in registers, GT in LDS, 128 dets x 64 GT x 4096 px, 2 detection tiles per wave",
    "detection operands expanded in registers, GT in LDS, 128 dets x 64 GT x 4096 px, 4 detection tiles per wave",
    "detection operands expanded in registers, GT in LDS, 128 dets x 64 GT x 8192 px, 2 detection tiles per wave",
    "FP4 MFMA 16x16x128, no LDS, 64 dets x 64 GT x 4096 px, 2 blocks ahead",
    "FP4 MFMA 16x16x128, no LDS, 64 dets x 64 GT, one round of waves, 2 blocks ahead",
    "FP4 MFMA 32x32x64, no LDS, 64 dets x 64 GT, one round of waves, 4 blocks ahead",
    "FP4 MFMA 16x16x128, no LDS, 64 dets x 32 GT, one round of waves, 4 blocks ahead",
    "FP4 MFMA 32x32x64, no LDS, 64 dets x 64 GT, one round of waves, 1 block ahead",
    "FP4 MFMA 16x16x128, no LDS, 128 dets x 64 GT, one round of waves, 2 blocks ahead",
    "FP4 MFMA 16x16x128, no LDS, 128 dets x 64 GT, one round of waves, 1 block ahead",
    "FP4 MFMA 32x32x64, no LDS, 128 dets x 64 GT, one round of waves, 2 blocks ahead",
    "FP4 MFMA 16x16x128, no LDS, 2 waves of 64 dets x 64 GT in step, one round, 2 blocks ahead",
    "FP4 MFMA 16x16x128, no LDS, 2 waves of 64 dets x 64 GT in step, one round, 3 blocks ahead"};
const char* const kDetBitsVariants[] = {
    "2 words x 8 masks", "2 words x 16 masks", "1 word x 16 masks", "4 words x 4 masks", "4 words x 8 masks",
    "1 word x 8 masks", "2 words x 8 masks, wave-summed bins"};
const TuneSpec kSpecs[kTuneKernels] = {
    {"detection_histogram", 5, 0, kDetVariants},
    {"joint_histogram", 5, 0, kJointVariants},
    {"mask_overlap_mfma", 9, 0, kOverlapVariants},
    {"packed_grid", 3, 0, kGridVariants},
    {"detection_group", 4, 0, kGroupVariants},
    {"mask_overlap_bits", 20, 18, kBitsVariants},
    {"detection_histogram_bits", 7, 0, kDetBitsVariants},
};
std::atomic<int> g_selected[kTuneKernels] = {{-1}, {-1}, {-1}, {-1}, {-1}, {-1}, {-1}};

int tune_index(const char* name) {
  if (!name) return -1;
  for (int k = 0; k < kTuneKernels; ++k)
    if (std::string(name) == kSpecs[k].name) return k;
  return -1;
}
}  // namespace

const TuneSpec& tune_spec(TuneKernel kernel) { return kSpecs[kernel]; }

int packed_grid_per_cu() {
  static const int kPerCu[] = {0, 8, 4};
  return kPerCu[tuned_variant(kTunePackedGrid)];
}

int detection_group_limit() {
  static const int kGroup[] = {0, 128, 64, 32};
  return kGroup[tuned_variant(kTuneDetectionGroup)];
}

int tuned_variant(TuneKernel kernel) {
  const int v = g_selected[kernel].load(std::memory_order_relaxed);
  return v < 0 ? kSpecs[kernel].fallback : v;
}

}  // namespace pe

extern "C" int pe_tune_count(const char* kernel) {
  const int k = pe::tune_index(kernel);
  return k < 0 ? 0 : pe::kSpecs[k].count;
}

extern "C" int pe_tune_set(const char* kernel, int variant) {
  const int k = pe::tune_index(kernel);
  PE_REQUIRE(k >= 0, "unknown tunable kernel");
  PE_REQUIRE(variant >= -1 && variant < pe::kSpecs[k].count, "variant out of range");
  pe::g_selected[k].store(variant, std::memory_order_relaxed);
  return 0;
}

extern "C" int pe_tune_get(const char* kernel) {
  const int k = pe::tune_index(kernel);
  return k < 0 ? -1 : pe::tuned_variant(static_cast<pe::TuneKernel>(k));
}

extern "C" const char* pe_tune_describe(const char* kernel, int variant) {
  const int k = pe::tune_index(kernel);
  if (k < 0 || variant < 0 || variant >= pe::kSpecs[k].count) return "";
  return pe::kSpecs[k].describe[variant];
}
