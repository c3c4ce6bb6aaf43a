// Reward server: continuous batching of per-image PQ requests (the
// reference's panoptic_quality, evaluator.py:471-544) and prefix-reward
// requests (segmentation_prefixes, prefix.py:11-122) on one GPU.  No
// counterpart in the reference; it is the serving form of the same kernels.
//
// A host worker thread owns the whole loop:
//  * admission by HBM budget -- the budget is a device arena allocated once;
//    a batch is admitted when its merged layout (inputs packed one image after
//    another, padded slot metadata, scratch, statistics and outputs) fits in
//    the arena.  Requests wait first-in first-out while it does not; a request
//    larger than the whole arena runs alone in memory of its own once nothing
//    else is in flight;
//  * continuous batching -- every pass of the loop merges all waiting requests
//    of a kind that fit into one launch (images of any size: pe_update or
//    pe_prefix over a packed batch) on the next of a few streams of default
//    priority, so batches admitted while earlier ones run overlap with them;
//  * packing without host round trips -- one pinned upload carries the batch's
//    tables and a copy list, and one gather kernel copies every image's maps
//    and metadata into the batch layout;
//  * completion -- the worker polls each launch's event, copies the per-image
//    results out of pinned memory, records the completion time and returns the
//    arena space.  Latency is measured from submission to that moment.
//
// Results do not depend on how requests are batched: per-image PQ folds each
// image's IoUs in slot order and prefix snapshots are per image.
#include <hip/hip_runtime.h>
#include <time.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "common.h"
#include "../../include/panoptic_hip.h"

namespace pe {
namespace {

int64_t now_ns() {
  timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return static_cast<int64_t>(ts.tv_sec) * 1000000000LL + ts.tv_nsec;
}

int64_t align256(int64_t v) { return (v + 255) & ~static_cast<int64_t>(255); }

}  // namespace

// First-fit allocator over offsets [0, capacity) in 256-byte granules, with
// coalescing of freed neighbours.  Host-only bookkeeping: the server maps
// offsets into its device arena and its pinned staging block.
class Arena {
 public:
  explicit Arena(int64_t capacity) : capacity_(capacity > 0 ? capacity / 256 * 256 : 0) {
    if (capacity_) free_[0] = capacity_;
  }
  bool alloc(int64_t size, int64_t* offset) {
    size = align256(std::max<int64_t>(size, 1));
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < size) continue;
      *offset = it->first;
      const int64_t start = it->first + size, rest = it->second - size;
      free_.erase(it);
      if (rest) free_[start] = rest;
      in_use_ += size;
      return true;
    }
    return false;
  }
  void release(int64_t offset, int64_t size) {
    size = align256(std::max<int64_t>(size, 1));
    in_use_ -= size;
    auto it = free_.emplace(offset, size).first;
    auto next = std::next(it);
    if (next != free_.end() && it->first + it->second == next->first) {
      it->second += next->second;
      free_.erase(next);
    }
    if (it != free_.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        free_.erase(it);
      }
    }
  }
  int64_t in_use() const { return in_use_; }
  int64_t capacity() const { return capacity_; }
  int64_t largest_free() const {
    int64_t best = 0;
    for (const auto& kv : free_) best = std::max(best, kv.second);
    return best;
  }

 private:
  int64_t capacity_;
  int64_t in_use_ = 0;
  std::map<int64_t, int64_t> free_;
};

// ---------------------------------------------------------------- gather copy

struct CopyEntry {
  uint64_t src;    // copy: source address; fills: unused
  uint64_t dst;
  int64_t bytes;
  int32_t mode;    // 0 copy, 1 fill 32-bit words with value, 2 fill bytes with value
  int32_t value;
};
static_assert(sizeof(CopyEntry) == 32, "copy entries are 32 bytes");

constexpr int64_t kCopyChunk = 256 << 10;  // bytes per entry (one workgroup each)

// One workgroup per entry: 16-byte vectors when every address and the size
// allow, else 4-byte words, else bytes.  Every access stays inside
// [dst, dst + bytes) and [src, src + bytes).
__global__ __launch_bounds__(256) void gather_copy(const CopyEntry* entries) {
  const CopyEntry e = entries[blockIdx.x];
  const int tid = threadIdx.x;
  char* dst = reinterpret_cast<char*>(e.dst);
  const char* src = reinterpret_cast<const char*>(e.src);
  if (e.mode == 0) {
    if (((e.src | e.dst | static_cast<uint64_t>(e.bytes)) & 15) == 0) {
      const i64 n = e.bytes / 16;
      for (i64 i = tid; i < n; i += blockDim.x) reinterpret_cast<uint4*>(dst)[i] = reinterpret_cast<const uint4*>(src)[i];
    } else if (((e.src | e.dst | static_cast<uint64_t>(e.bytes)) & 3) == 0) {
      const i64 n = e.bytes / 4;
      for (i64 i = tid; i < n; i += blockDim.x)
        reinterpret_cast<uint32_t*>(dst)[i] = reinterpret_cast<const uint32_t*>(src)[i];
    } else {
      for (i64 i = tid; i < e.bytes; i += blockDim.x) dst[i] = src[i];
    }
  } else if (e.mode == 1) {
    const i64 n = e.bytes / 4;
    for (i64 i = tid; i < n; i += blockDim.x) reinterpret_cast<int32_t*>(dst)[i] = e.value;
  } else {
    for (i64 i = tid; i < e.bytes; i += blockDim.x) dst[i] = static_cast<char>(e.value);
  }
}

// ------------------------------------------------------------------- requests

enum RequestKind { kPQ = 0, kPrefix = 1 };

// Per-image descriptor, as submitted (device addresses owned by the caller).
struct Image {
  uint64_t map;       // target slots, H x W int32
  uint64_t other;     // PQ: prediction slots (H x W int32); prefix: N x H x W bool masks
  uint64_t meta0;     // PQ: target classes (G int32); prefix: generated classes (N int32)
  uint64_t meta1;     // PQ: prediction classes (P int32); prefix: target classes (G int32)
  uint64_t crowd;     // target crowd flags (G bool) or 0
  int64_t H, W;
  int64_t G;          // target slots
  int64_t PN;         // PQ: prediction slots; prefix: generated masks
};

struct Request {
  int64_t id = 0;
  int kind = kPQ;
  std::vector<Image> images;
  std::vector<std::vector<int32_t>> points;  // prefix: checkpoints per image
  int64_t rows = 0;                          // prefix: checkpoint rows over all images
  int64_t arrival = 0, launched = 0, completed = 0;
  bool done = false;
  int error = 0;
  std::string message;
  std::vector<double> result;
};

// Device layout of one merged launch (byte offsets within its block).
struct Layout {
  int kind = kPQ;
  int64_t B = 0, G = 0, P = 0, X = 0, C = 0, T = 0;
  int64_t sum_n = 0, sum_tri = 0, max_n = 0, rows = 0;
  // offsets
  int64_t upload = 0, upload_bytes = 0;  // pix_off | (n_off | k_off | mask table | points) | entries
  int64_t pix_off = 0, n_off = 0, k_off = 0, mask_table = 0, points = 0, entries = 0;
  int64_t map = 0, pred = 0, gc = 0, pc = 0, crowd = 0, cls = 0;
  int64_t pairs = 0, areas = 0, matched = 0, tp = 0, fp = 0, fn = 0, sums = 0, confusion = 0, error = 0;
  int64_t itp = 0, ifp = 0, ifn = 0, precision = 0, recall = 0;
  int64_t per_class = 0, summaries = 0, counts = 0, miou = 0, out = 0;
  int64_t workspace = 0, workspace_bytes = 0;
  int64_t device_bytes = 0;
  int64_t host_bytes = 0;     // pinned: upload region, then results
  int64_t host_results = 0;   // offset of results in the pinned block
  int64_t result_doubles = 0;
  int64_t entry_count = 0;
};

struct Launch {
  int kind = kPQ;
  std::vector<Request*> requests;
  Layout layout;
  hipStream_t stream = nullptr;
  hipEvent_t event = nullptr;
  int64_t device_offset = 0, host_offset = 0;
  void* own_device = nullptr;  // a request larger than the arena: memory of its own
  void* own_host = nullptr;
};

class Server {
 public:
  Server(int device, int64_t budget, int streams, int64_t C, const uint8_t* things, const double* thresholds,
         int64_t T, int64_t pinned_bytes)
      : device_(device), C_(C), T_(T), arena_(budget), pinned_(pinned_bytes) {
    things_host_.assign(things, things + C);
    thresholds_host_.assign(thresholds, thresholds + T);
    stream_count_ = std::max(1, streams);
  }

  int start() {
    int caller = 0;
    PE_HIP(hipGetDevice(&caller));
    PE_HIP(hipSetDevice(device_));
    const int rc = allocate();
    PE_HIP(hipSetDevice(caller));
    if (rc) return rc;
    worker_ = std::thread([this] { run(); });
    return kOk;
  }

  int allocate() {
    // The budget is capped by the device memory free at creation (80% of it).
    size_t free_bytes = 0, total_bytes = 0;
    PE_HIP(hipMemGetInfo(&free_bytes, &total_bytes));
    const int64_t cap = static_cast<int64_t>(free_bytes / 10 * 8);
    if (arena_.capacity() > cap) arena_ = Arena(cap);
    if (arena_.capacity() > 0) PE_HIP(hipMalloc(&device_base_, arena_.capacity()));
    if (pinned_.capacity() > 0) PE_HIP(hipHostMalloc(&host_base_, pinned_.capacity(), hipHostMallocDefault));
    PE_HIP(hipMalloc(&constants_, align256(C_) + 8 * T_));
    PE_HIP(hipMemcpy(constants_, things_host_.data(), C_, hipMemcpyHostToDevice));
    PE_HIP(hipMemcpy(static_cast<char*>(constants_) + align256(C_), thresholds_host_.data(), 8 * T_,
                     hipMemcpyHostToDevice));
    for (int s = 0; s < stream_count_; ++s) {
      hipStream_t stream;
      PE_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
      streams_.push_back(stream);
    }
    return kOk;
  }

  ~Server() {
    {
      std::lock_guard<std::mutex> guard(mutex_);
      stop_ = true;
    }
    wake_.notify_all();
    if (worker_.joinable()) worker_.join();
    for (hipStream_t s : streams_) (void)hipStreamDestroy(s);
    if (device_base_) (void)hipFree(device_base_);
    if (host_base_) (void)hipHostFree(host_base_);
    if (constants_) (void)hipFree(constants_);
  }

  int64_t submit(std::unique_ptr<Request> request) {
    request->arrival = request->arrival ? request->arrival : now_ns();
    std::lock_guard<std::mutex> guard(mutex_);
    request->id = next_id_++;
    const int64_t id = request->id;
    incoming_.push_back(request.get());
    tickets_[id] = std::move(request);
    wake_.notify_all();
    return id;
  }

  int64_t poll(int64_t* ids, int64_t max_ids, int64_t wait_us) {
    std::unique_lock<std::mutex> lock(mutex_);
    if (completed_.empty() && wait_us > 0)
      done_cv_.wait_for(lock, std::chrono::microseconds(wait_us), [this] { return !completed_.empty(); });
    int64_t n = 0;
    while (n < max_ids && !completed_.empty()) {
      ids[n++] = completed_.front();
      completed_.pop_front();
    }
    return n;
  }

  // Copies a finished request's result and times and forgets it.
  int take(int64_t id, double* out, int64_t capacity, int64_t* length, int64_t* times, char* message) {
    std::lock_guard<std::mutex> guard(mutex_);
    auto it = tickets_.find(id);
    PE_REQUIRE(it != tickets_.end(), "unknown request id %lld", static_cast<long long>(id));
    Request& r = *it->second;
    PE_REQUIRE(r.done, "request %lld has not completed", static_cast<long long>(id));
    *length = static_cast<int64_t>(r.result.size());
    PE_REQUIRE(capacity >= *length || !out, "result buffer too small");
    if (out) std::memcpy(out, r.result.data(), 8 * r.result.size());
    if (times) {
      times[0] = r.arrival;
      times[1] = r.launched;
      times[2] = r.completed;
    }
    if (message) {
      std::strncpy(message, r.message.c_str(), 255);
      message[255] = 0;
    }
    const int error = r.error;
    tickets_.erase(it);
    return error;
  }

  void stats(int64_t* out) {
    std::lock_guard<std::mutex> guard(mutex_);
    out[0] = stat_in_use_;
    out[1] = arena_capacity_stat();
    out[2] = stat_waiting_;
    out[3] = stat_inflight_;
    out[4] = blocked_;
    out[5] = launches_[kPQ];
    out[6] = launches_[kPrefix];
    out[7] = done_count_;
    out[8] = launched_requests_;
  }

 private:
  int64_t arena_capacity_stat() const { return arena_.capacity(); }

  Layout plan(int kind, const std::vector<Request*>& group) const { return plan_layout(kind, group, C_, T_); }

 public:
  // ------------------------------------------------------------- layout
  static Layout plan_layout(int kind, const std::vector<Request*>& group, int64_t C_, int64_t T_) {
    Layout L;
    L.kind = kind;
    L.C = C_;
    L.T = T_;
    std::vector<const Image*> images;
    for (Request* r : group)
      for (const Image& im : r->images) images.push_back(&im);
    L.B = static_cast<int64_t>(images.size());
    // Copy-list entries: one per kCopyChunk bytes of every copy, one per fill
    // (exactly what issue() writes, so the staged list cannot overflow).
    auto chunks = [](int64_t bytes) { return (bytes + kCopyChunk - 1) / kCopyChunk; };
    int64_t entries = 0;
    for (const Image* im : images) {
      const int64_t px = im->H * im->W;
      L.X += px;
      L.G = std::max(L.G, im->G);
      entries += chunks(4 * px) + chunks(4 * im->G) + chunks(im->G) + 3;  // map, classes, crowd (+ pads)
      if (kind == kPQ) {
        L.P = std::max(L.P, im->PN);
        entries += chunks(4 * px) + chunks(4 * im->PN) + 1;             // prediction map, classes (+ pad)
      } else {
        L.sum_n += im->PN;
        L.sum_tri += im->PN * (im->PN + 1) / 2;
        L.max_n = std::max(L.max_n, im->PN);
        entries += chunks(4 * im->PN);                                    // generated classes
      }
    }
    if (kind == kPrefix)
      for (Request* r : group) L.rows += r->rows;
    entries += 4;  // statistics fills
    L.entry_count = entries;
    int64_t at = 0;
    auto take = [&](int64_t bytes) {
      const int64_t off = at;
      at = align256(at + std::max<int64_t>(bytes, 0));
      return off;
    };
    const int64_t B = L.B, C = L.C, T = L.T;
    // Upload region first: it mirrors the pinned staging layout.
    L.upload = take(0);
    int64_t up = 0;
    auto sub = [&](int64_t bytes) {
      const int64_t off = up;
      up = (up + bytes + 15) / 16 * 16;
      return off;
    };
    L.pix_off = sub(8 * (B + 1));
    if (kind == kPrefix) {
      L.n_off = sub(8 * (B + 1));
      L.k_off = sub(8 * (B + 1));
      L.mask_table = sub(8 * B);
      L.points = sub(4 * std::max<int64_t>(L.rows, 1));
    }
    L.entries = sub(32 * entries);
    L.upload_bytes = up;
    at = align256(up);
    L.map = take(4 * L.X);
    L.gc = take(4 * B * L.G);
    L.crowd = take(B * L.G);
    if (kind == kPQ) {
      L.pred = take(4 * L.X);
      L.pc = take(4 * B * L.P);
      L.pairs = take(8 * B * L.G * L.P);
      L.areas = take(8 * B * (L.G + L.P));
      L.matched = take(4 * B * L.G);
      L.tp = take(8 * B * C);
      L.fp = take(8 * B * C);
      L.fn = take(8 * B * C);
      L.sums = take(8 * B * C);
      L.error = take(4);
      L.workspace_bytes = pe_update_workspace_bytes(B, L.G, L.P, C, L.X);
      L.workspace = take(L.workspace_bytes);
      L.out = take(8 * B);
      L.result_doubles = B;
    } else {
      const int64_t R = L.rows;
      L.cls = take(4 * std::max<int64_t>(L.sum_n, 1));
      L.tp = take(8 * R * C);
      L.fp = take(8 * R * C);
      L.fn = take(8 * R * C);
      L.sums = take(8 * R * C);
      L.confusion = take(8 * R * C * (C + 1));
      L.itp = take(8 * R * T * C);
      L.ifp = take(8 * R * T * C);
      L.ifn = take(8 * R * T * C);
      L.precision = take(8 * R * T * C);
      L.recall = take(8 * R * T * C);
      L.error = take(4);
      L.workspace_bytes = pe_prefix_workspace_bytes(B, L.G, C, T, L.sum_n, L.sum_tri, L.X);
      L.workspace = take(L.workspace_bytes);
      L.per_class = take(8 * R * 4 * C);
      L.summaries = take(8 * R * 9);
      L.counts = take(8 * R * 3);
      L.miou = take(8 * R);
      L.result_doubles = 10 * R;
    }
    L.device_bytes = at;
    L.host_results = align256(L.upload_bytes);
    L.host_bytes = L.host_results + 8 * L.result_doubles + 8;  // results, then the error word
    return L;
  }

 private:

  // ---------------------------------------------------------------- loop
  void run() {
    (void)hipSetDevice(device_);
    for (;;) {
      {
        std::unique_lock<std::mutex> lock(mutex_);
        for (Request* r : incoming_) waiting_.push_back(r);
        incoming_.clear();
        if (stop_ && waiting_.empty() && inflight_.empty()) break;
        if (waiting_.empty() && inflight_.empty()) {
          wake_.wait(lock, [this] { return stop_ || !incoming_.empty(); });
          continue;
        }
      }
      bool progress = complete();
      progress |= launch_ready();
      publish_stats();
      if (!progress) {
        std::unique_lock<std::mutex> lock(mutex_);
        wake_.wait_for(lock, std::chrono::microseconds(20), [this] { return stop_ || !incoming_.empty(); });
      }
    }
  }

  void publish_stats() {
    std::lock_guard<std::mutex> guard(mutex_);
    stat_in_use_ = arena_.in_use();
    stat_waiting_ = static_cast<int64_t>(waiting_.size());
    stat_inflight_ = static_cast<int64_t>(inflight_.size());
  }

  // Launches every waiting group that fits; returns whether any launched.
  bool launch_ready() {
    bool any = false;
    while (!waiting_.empty()) {
      const int kind = waiting_.front()->kind;
      std::vector<Request*> group;
      int64_t images = 0;
      for (Request* r : waiting_) {
        if (r->kind != kind) continue;
        if (!group.empty() && images + static_cast<int64_t>(r->images.size()) > kMaxImages) break;
        group.push_back(r);
        images += static_cast<int64_t>(r->images.size());
      }
      // The longest first-in first-out run of this kind whose merged layout fits.
      size_t k = group.size();
      Layout L;
      int64_t dev_off = 0, host_off = 0;
      bool fits = false;
      while (k > 0) {
        std::vector<Request*> head(group.begin(), group.begin() + static_cast<long>(k));
        L = plan(kind, head);
        if (L.host_bytes <= pinned_.capacity() && arena_.alloc(L.device_bytes, &dev_off)) {
          if (pinned_.alloc(L.host_bytes, &host_off)) {
            fits = true;
            break;
          }
          arena_.release(dev_off, L.device_bytes);
        }
        k = k > 1 ? k / 2 : 0;
      }
      std::vector<Request*> chosen(group.begin(), group.begin() + static_cast<long>(std::max<size_t>(k, 1)));
      auto launch = std::make_unique<Launch>();
      launch->kind = kind;
      launch->requests = chosen;
      if (!fits) {
        // The head request alone does not fit in the free memory: it waits,
        // unless nothing is running and it exceeds the whole budget, in which
        // case it runs alone in memory of its own.
        L = plan(kind, chosen);
        const bool alone = inflight_.empty() && (L.device_bytes > arena_.capacity() || L.host_bytes > pinned_.capacity());
        if (!alone) {
          ++blocked_;
          break;
        }
        if (hipMalloc(&launch->own_device, L.device_bytes) != hipSuccess ||
            hipHostMalloc(&launch->own_host, L.host_bytes, hipHostMallocDefault) != hipSuccess) {
          fail_requests(chosen, kHipError, "memory for a request larger than the budget could not be allocated");
          if (launch->own_device) (void)hipFree(launch->own_device);
          if (launch->own_host) (void)hipHostFree(launch->own_host);
          remove_waiting(chosen);
          any = true;
          continue;
        }
      }
      launch->layout = L;
      launch->device_offset = dev_off;
      launch->host_offset = host_off;
      remove_waiting(chosen);
      const int rc = issue(*launch);
      any = true;
      if (rc) {
        fail_requests(chosen, rc, error_buffer());
        release(*launch);
        continue;
      }
      {
        std::lock_guard<std::mutex> guard(mutex_);
        ++launches_[kind];
        launched_requests_ += static_cast<int64_t>(chosen.size());
      }
      inflight_.push_back(std::move(launch));
    }
    return any;
  }

  void remove_waiting(const std::vector<Request*>& chosen) {
    for (Request* r : chosen) waiting_.erase(std::find(waiting_.begin(), waiting_.end(), r));
  }

  void release(Launch& launch) {
    if (launch.own_device) {
      (void)hipFree(launch.own_device);
      (void)hipHostFree(launch.own_host);
      launch.own_device = launch.own_host = nullptr;
    } else {
      arena_.release(launch.device_offset, launch.layout.device_bytes);
      pinned_.release(launch.host_offset, launch.layout.host_bytes);
    }
    if (launch.event) (void)hipEventDestroy(launch.event);
    launch.event = nullptr;
  }

  void fail_requests(const std::vector<Request*>& requests, int code, const char* message) {
    const int64_t t = now_ns();
    std::lock_guard<std::mutex> guard(mutex_);
    for (Request* r : requests) {
      r->done = true;
      r->error = code;
      r->message = message ? message : "launch failed";
      r->completed = t;
      completed_.push_back(r->id);
      ++done_count_;
    }
    done_cv_.notify_all();
  }

  // Polls the in-flight launches; finishes those whose event has completed.
  bool complete() {
    bool any = false;
    for (auto it = inflight_.begin(); it != inflight_.end();) {
      Launch& launch = **it;
      const hipError_t q = hipEventQuery(launch.event);
      if (q == hipErrorNotReady) {
        ++it;
        continue;
      }
      const int64_t t = now_ns();
      if (q != hipSuccess) {
        fail_requests(launch.requests, kHipError, hipGetErrorString(q));
      } else {
        deliver(launch, t);
      }
      release(launch);
      it = inflight_.erase(it);
      any = true;
    }
    return any;
  }

  void deliver(Launch& launch, int64_t t) {
    const Layout& L = launch.layout;
    const char* host = launch.own_host ? static_cast<const char*>(launch.own_host)
                                       : static_cast<const char*>(host_base_) + launch.host_offset;
    const double* results = reinterpret_cast<const double*>(host + L.host_results);
    const int32_t error = *reinterpret_cast<const int32_t*>(host + L.host_results + 8 * L.result_doubles);
    std::lock_guard<std::mutex> guard(mutex_);
    int64_t image = 0, row = 0;
    for (Request* r : launch.requests) {
      if (L.kind == kPQ) {
        r->result.assign(results + image, results + image + static_cast<int64_t>(r->images.size()));
      } else {
        // Row r of the batch: All pq, sq, rq (summaries[r][0][*]) and mIoU.
        r->result.resize(4 * r->rows);
        for (int64_t q = 0; q < r->rows; ++q) {
          const double* s = results + 9 * (row + q);
          r->result[4 * q] = s[0];
          r->result[4 * q + 1] = s[1];
          r->result[4 * q + 2] = s[2];
          r->result[4 * q + 3] = results[9 * L.rows + row + q];
        }
        row += r->rows;
      }
      image += static_cast<int64_t>(r->images.size());
      if (error && L.kind == kPrefix && (error & ~0x3E)) {
        r->error = 1;
        r->message = "target slot outside metadata capacity";
      }
      r->completed = t;
      r->done = true;
      completed_.push_back(r->id);
      ++done_count_;
    }
    done_cv_.notify_all();
  }

  // ------------------------------------------------------------ launch
  int issue(Launch& launch) {
    const Layout& L = launch.layout;
    char* dev = launch.own_device ? static_cast<char*>(launch.own_device)
                                  : static_cast<char*>(device_base_) + launch.device_offset;
    char* host = launch.own_host ? static_cast<char*>(launch.own_host)
                                 : static_cast<char*>(host_base_) + launch.host_offset;
    hipStream_t stream = streams_[next_stream_];
    next_stream_ = (next_stream_ + 1) % stream_count_;
    launch.stream = stream;
    const int64_t t = now_ns();
    for (Request* r : launch.requests) r->launched = t;
    auto D = [&](int64_t off) { return reinterpret_cast<uint64_t>(dev + off); };
    // Host staging of the upload region.
    char* up = host;
    auto* pix = reinterpret_cast<int64_t*>(up + L.pix_off);
    auto* entries = reinterpret_cast<CopyEntry*>(up + L.entries);
    int64_t ne = 0;
    bool overflow = false;
    auto push = [&](const CopyEntry& e) {
      if (ne < L.entry_count) entries[ne++] = e;
      else overflow = true;
    };
    auto copy = [&](uint64_t src, uint64_t dst, int64_t bytes) {
      for (int64_t o = 0; o < bytes; o += kCopyChunk) push(CopyEntry{src + o, dst + o, std::min(kCopyChunk, bytes - o), 0, 0});
    };
    auto fill32 = [&](uint64_t dst, int64_t words, int32_t value) {
      if (words > 0) push(CopyEntry{0, dst, 4 * words, 1, value});
    };
    auto fill8 = [&](uint64_t dst, int64_t bytes, int32_t value) {
      if (bytes > 0) push(CopyEntry{0, dst, bytes, 2, value});
    };
    int64_t b = 0, px = 0, n = 0, rows = 0;
    int64_t* n_off = L.kind == kPrefix ? reinterpret_cast<int64_t*>(up + L.n_off) : nullptr;
    int64_t* k_off = L.kind == kPrefix ? reinterpret_cast<int64_t*>(up + L.k_off) : nullptr;
    int64_t* table = L.kind == kPrefix ? reinterpret_cast<int64_t*>(up + L.mask_table) : nullptr;
    int32_t* pts = L.kind == kPrefix ? reinterpret_cast<int32_t*>(up + L.points) : nullptr;
    bool aligned = true;
    pix[0] = 0;
    if (n_off) n_off[0] = k_off[0] = 0;
    for (Request* r : launch.requests) {
      for (size_t i = 0; i < r->images.size(); ++i, ++b) {
        const Image& im = r->images[i];
        const int64_t hw = im.H * im.W;
        copy(im.map, D(L.map) + 4 * px, 4 * hw);
        copy(L.kind == kPQ ? im.meta0 : im.meta1, D(L.gc) + 4 * b * L.G, 4 * im.G);
        fill32(D(L.gc) + 4 * (b * L.G + im.G), L.G - im.G, -1);
        if (im.crowd) copy(im.crowd, D(L.crowd) + b * L.G, im.G);
        else fill8(D(L.crowd) + b * L.G, im.G, 0);
        fill8(D(L.crowd) + b * L.G + im.G, L.G - im.G, 0);
        if (L.kind == kPQ) {
          copy(im.other, D(L.pred) + 4 * px, 4 * hw);
          copy(im.meta1, D(L.pc) + 4 * b * L.P, 4 * im.PN);
          fill32(D(L.pc) + 4 * (b * L.P + im.PN), L.P - im.PN, -1);
        } else {
          copy(im.meta0, D(L.cls) + 4 * n, 4 * im.PN);
          table[b] = static_cast<int64_t>(im.other);
          aligned = aligned && hw % 16 == 0 && im.other % 16 == 0;
          n += im.PN;
          n_off[b + 1] = n;
          for (int32_t q : r->points[i]) pts[rows++] = q;
          k_off[b + 1] = rows;
        }
        px += hw;
        pix[b + 1] = px;
      }
    }
    if (L.kind == kPQ) {
      fill32(D(L.tp), 2 * L.B * L.C, 0);
      fill32(D(L.fp), 2 * L.B * L.C, 0);
      fill32(D(L.fn), 2 * L.B * L.C, 0);
      fill32(D(L.sums), 2 * L.B * L.C, 0);
    }
    PE_REQUIRE(!overflow, "copy list overflow");
    PE_HIP(hipMemcpyAsync(dev + L.upload, up, L.upload_bytes, hipMemcpyHostToDevice, stream));
    if (ne) {
      gather_copy<<<static_cast<unsigned>(ne), 256, 0, stream>>>(reinterpret_cast<const CopyEntry*>(dev + L.entries));
      PE_LAUNCH_CHECK();
    }
    int rc = kOk;
    double* results_dev = nullptr;
    int64_t result_bytes = 8 * L.result_doubles;
    if (L.kind == kPQ) {
      rc = pe_update(dev + L.map, 4, dev + L.pred, 4, L.B, 1, L.X, reinterpret_cast<const int64_t*>(dev + L.pix_off),
                     reinterpret_cast<const int32_t*>(dev + L.gc), reinterpret_cast<const int32_t*>(dev + L.pc),
                     reinterpret_cast<const uint8_t*>(dev + L.crowd), L.G, L.P, L.C,
                     reinterpret_cast<int64_t*>(dev + L.pairs), reinterpret_cast<int64_t*>(dev + L.areas),
                     reinterpret_cast<int32_t*>(dev + L.matched), reinterpret_cast<int64_t*>(dev + L.tp),
                     reinterpret_cast<int64_t*>(dev + L.fp), reinterpret_cast<int64_t*>(dev + L.fn),
                     reinterpret_cast<double*>(dev + L.sums), nullptr, nullptr, nullptr, nullptr, nullptr,
                     reinterpret_cast<int32_t*>(dev + L.error), nullptr, dev + L.workspace, L.workspace_bytes, 0, 1,
                     0, stream);
      if (!rc)
        rc = pe_image_pq(reinterpret_cast<const int64_t*>(dev + L.tp), reinterpret_cast<const int64_t*>(dev + L.fp),
                         reinterpret_cast<const int64_t*>(dev + L.fn), reinterpret_cast<const double*>(dev + L.sums),
                         L.B, L.C, reinterpret_cast<double*>(dev + L.out), nullptr, stream);
      results_dev = reinterpret_cast<double*>(dev + L.out);
      // No validation: the error word is reported as zero.
      *reinterpret_cast<int32_t*>(host + L.host_results + result_bytes) = 0;
    } else {
      const uint8_t* things = static_cast<const uint8_t*>(constants_);
      const double* thresholds = reinterpret_cast<const double*>(static_cast<char*>(constants_) + align256(C_));
      rc = pe_prefix(L.B, reinterpret_cast<const int32_t*>(dev + L.map),
                     reinterpret_cast<const int64_t*>(dev + L.pix_off), L.X, L.G, L.C,
                     reinterpret_cast<const uint8_t* const*>(dev + L.mask_table),
                     reinterpret_cast<const int64_t*>(dev + L.n_off), L.sum_n, L.sum_tri, L.max_n,
                     reinterpret_cast<const int32_t*>(dev + L.cls), reinterpret_cast<const int32_t*>(dev + L.gc),
                     reinterpret_cast<const uint8_t*>(dev + L.crowd), things, thresholds, L.T,
                     reinterpret_cast<const int64_t*>(dev + L.k_off), reinterpret_cast<const int32_t*>(dev + L.points),
                     L.rows, reinterpret_cast<int64_t*>(dev + L.tp), reinterpret_cast<int64_t*>(dev + L.fp),
                     reinterpret_cast<int64_t*>(dev + L.fn), reinterpret_cast<double*>(dev + L.sums),
                     reinterpret_cast<int64_t*>(dev + L.confusion), reinterpret_cast<int64_t*>(dev + L.itp),
                     reinterpret_cast<int64_t*>(dev + L.ifp), reinterpret_cast<int64_t*>(dev + L.ifn),
                     reinterpret_cast<double*>(dev + L.precision), reinterpret_cast<double*>(dev + L.recall),
                     reinterpret_cast<int32_t*>(dev + L.error), dev + L.workspace, L.workspace_bytes, aligned ? 1 : 0,
                     stream);
      if (!rc && L.rows)
        rc = pe_compute(reinterpret_cast<const int64_t*>(dev + L.tp), reinterpret_cast<const int64_t*>(dev + L.fp),
                        reinterpret_cast<const int64_t*>(dev + L.fn), reinterpret_cast<const double*>(dev + L.sums),
                        reinterpret_cast<const int64_t*>(dev + L.confusion), things, L.rows, L.C,
                        reinterpret_cast<double*>(dev + L.per_class), reinterpret_cast<double*>(dev + L.summaries),
                        reinterpret_cast<int64_t*>(dev + L.counts), reinterpret_cast<double*>(dev + L.miou), nullptr,
                        nullptr, nullptr, nullptr, nullptr, stream);
      // summaries (rows x 9) then mIoU (rows) then the error word.
      if (!rc && L.rows) {
        PE_HIP(hipMemcpyAsync(host + L.host_results, dev + L.summaries, 8 * 9 * L.rows, hipMemcpyDeviceToHost,
                              stream));
        PE_HIP(hipMemcpyAsync(host + L.host_results + 8 * 9 * L.rows, dev + L.miou, 8 * L.rows,
                              hipMemcpyDeviceToHost, stream));
      }
      if (!rc)
        PE_HIP(hipMemcpyAsync(host + L.host_results + result_bytes, dev + L.error, 4, hipMemcpyDeviceToHost, stream));
    }
    if (rc) return rc;
    if (L.kind == kPQ) PE_HIP(hipMemcpyAsync(host + L.host_results, results_dev, result_bytes, hipMemcpyDeviceToHost, stream));
    PE_HIP(hipEventCreateWithFlags(&launch.event, hipEventDisableTiming));
    PE_HIP(hipEventRecord(launch.event, stream));
    return kOk;
  }

  static constexpr int64_t kMaxImages = 4096;

  int device_;
  int64_t C_, T_;
  Arena arena_, pinned_;
  std::vector<uint8_t> things_host_;
  std::vector<double> thresholds_host_;
  int stream_count_ = 1;
  std::vector<hipStream_t> streams_;
  int next_stream_ = 0;
  void* device_base_ = nullptr;
  void* host_base_ = nullptr;
  void* constants_ = nullptr;
  std::thread worker_;
  std::mutex mutex_;
  std::condition_variable wake_, done_cv_;
  bool stop_ = false;
  int64_t next_id_ = 1;
  std::deque<Request*> incoming_;                          // guarded by mutex_
  std::deque<Request*> waiting_;                           // worker only
  std::vector<std::unique_ptr<Launch>> inflight_;          // worker only
  std::unordered_map<int64_t, std::unique_ptr<Request>> tickets_;  // guarded
  std::deque<int64_t> completed_;                          // guarded
  int64_t blocked_ = 0, launches_[2] = {0, 0}, done_count_ = 0, launched_requests_ = 0;
  int64_t stat_in_use_ = 0, stat_waiting_ = 0, stat_inflight_ = 0;
};

}  // namespace pe

using pe::Server;

extern "C" void* pe_server_create(int device, int64_t budget_bytes, int streams, int64_t C, const uint8_t* things,
                                  const double* thresholds, int64_t T, int64_t pinned_bytes) {
  if (budget_bytes <= 0 || C <= 0 || C >= INT_MAX || T <= 0 || !things || !thresholds || pinned_bytes <= 0) {
    pe::fail(pe::kBadArgument, "invalid server arguments");
    return nullptr;
  }
  auto server = std::make_unique<Server>(device, budget_bytes, streams, C, things, thresholds, T, pinned_bytes);
  if (server->start()) return nullptr;
  return server.release();
}

extern "C" int pe_server_destroy(void* server) {
  delete static_cast<Server*>(server);
  return pe::kOk;
}

namespace pe {
// A request from its per-image descriptors (9 int64 per image) and, for
// prefix requests, each image's checkpoints; returns kOk or an error.
static int parse_request(int kind, int64_t B, const int64_t* images, const int32_t* points,
                         const int64_t* point_counts, Request* request) {
  PE_REQUIRE((kind == kPQ || kind == kPrefix) && B >= 0 && (B == 0 || images), "invalid request");
  request->kind = kind;
  int64_t p = 0;
  for (int64_t b = 0; b < B; ++b) {
    const int64_t* d = images + 9 * b;
    pe::Image im{static_cast<uint64_t>(d[0]), static_cast<uint64_t>(d[1]), static_cast<uint64_t>(d[2]),
                 static_cast<uint64_t>(d[3]), static_cast<uint64_t>(d[4]), d[5], d[6], d[7], d[8]};
    const bool ok = kind == pe::kPQ
                        ? im.other && im.meta0 && im.meta1 && im.PN > 0 && im.PN < INT_MAX
                        : im.meta1 && im.PN >= 0 && im.PN < INT_MAX && (im.PN == 0 || (im.other && im.meta0));
    PE_REQUIRE(ok && im.map && im.H > 0 && im.W > 0 && im.G > 0 && im.G < INT_MAX,
               "invalid image %lld of a request", static_cast<long long>(b));
    request->images.push_back(im);
    if (kind == pe::kPrefix) {
      PE_REQUIRE(point_counts && point_counts[b] >= 1, "every image needs at least one checkpoint");
      std::vector<int32_t> pts(points + p, points + p + point_counts[b]);
      for (size_t q = 0; q < pts.size(); ++q)
        PE_REQUIRE(pts[q] >= 0 && pts[q] <= im.PN && (q == 0 || pts[q] > pts[q - 1]), "invalid checkpoints");
      p += point_counts[b];
      request->rows += static_cast<int64_t>(pts.size());
      request->points.push_back(std::move(pts));
    }
  }
  return kOk;
}
}  // namespace pe

extern "C" int64_t pe_server_submit(void* server, int kind, int64_t B, const int64_t* images, const int32_t* points,
                                    const int64_t* point_counts, int64_t arrival_ns) {
  PE_REQUIRE(server, "no server");
  auto request = std::make_unique<pe::Request>();
  const int rc = pe::parse_request(kind, B, images, points, point_counts, request.get());
  if (rc) return rc;
  request->arrival = arrival_ns;
  return static_cast<Server*>(server)->submit(std::move(request));
}

// Device bytes a request takes when it runs alone (its admission footprint);
// host only.
extern "C" int64_t pe_server_footprint(int64_t C, int64_t T, int kind, int64_t B, const int64_t* images,
                                       const int32_t* points, const int64_t* point_counts) {
  PE_REQUIRE(C > 0 && C < INT_MAX && T > 0, "invalid class or threshold count");
  pe::Request request;
  const int rc = pe::parse_request(kind, B, images, points, point_counts, &request);
  if (rc) return rc;
  return Server::plan_layout(kind, {&request}, C, T).device_bytes;
}

extern "C" int64_t pe_server_poll(void* server, int64_t* ids, int64_t max_ids, int64_t wait_us) {
  PE_REQUIRE(server && ids && max_ids > 0, "invalid poll arguments");
  return static_cast<Server*>(server)->poll(ids, max_ids, wait_us);
}

extern "C" int pe_server_take(void* server, int64_t id, double* out, int64_t capacity, int64_t* length,
                              int64_t* times, char* message) {
  PE_REQUIRE(server && length, "invalid arguments");
  return static_cast<Server*>(server)->take(id, out, capacity, length, times, message);
}

extern "C" int pe_server_stats(void* server, int64_t* out) {
  PE_REQUIRE(server && out, "invalid arguments");
  static_cast<Server*>(server)->stats(out);
  return pe::kOk;
}

// Host-only access to the server's first-fit arena (no device memory).
extern "C" void* pe_arena_create(int64_t capacity) { return new pe::Arena(capacity); }
extern "C" int pe_arena_destroy(void* arena) {
  delete static_cast<pe::Arena*>(arena);
  return pe::kOk;
}
extern "C" int64_t pe_arena_alloc(void* arena, int64_t bytes) {
  int64_t offset = -1;
  return static_cast<pe::Arena*>(arena)->alloc(bytes, &offset) ? offset : -1;
}
extern "C" int pe_arena_release(void* arena, int64_t offset, int64_t bytes) {
  static_cast<pe::Arena*>(arena)->release(offset, bytes);
  return pe::kOk;
}
extern "C" int64_t pe_arena_in_use(void* arena) { return static_cast<pe::Arena*>(arena)->in_use(); }
extern "C" int64_t pe_arena_largest_free(void* arena) { return static_cast<pe::Arena*>(arena)->largest_free(); }
