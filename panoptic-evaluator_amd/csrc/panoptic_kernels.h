// Kernel argument blocks and launchers of the panoptic path, shared by the
// plain entry points (panoptic.hip) and the batched / serving forms.
#pragma once

#include "common.h"

namespace pe {

constexpr int kHistThreads = 512;
constexpr int kScoreThreads = 1024;
constexpr int kMaxChunk = 65532;                 // packed 16-bit LDS counters stay exact
constexpr size_t kMaxHistLds = 128 * 1024;       // beyond this, global atomics
constexpr size_t kMaxStageLds = 120 * 1024;      // per-image scoring metadata in LDS

struct HistArgs {
  const void* gt;
  const void* pred;
  int B;
  i64 HW;                // pixels per image (uniform batches)
  const i64* pix_off;    // B + 1 pixel offsets of a packed batch, or nullptr
  int G, P;
  TileList tiles;        // chunks of <= kMaxChunk pixels, each inside one image
  bool persist;          // grid-stride walk of the chunk list
  int words4;            // LDS size in 16-byte units
  u64* pairs;            // B x G x P
  const i64* gt_ids;
  const int* gt_slots;
  const i64* pred_ids;
  const int* pred_slots;
  int* error;            // nullptr: no validation
  size_t lds_bytes;
  bool stage_tables;     // original-ID tables hashed into LDS after the bins
  int table_offset;      // their byte offset in LDS
  int gt_hash_bits;      // log2 of the table sizes: buckets of 4 keys (32-bit pixels), else entries; >= 2 entries per ID
  int pred_hash_bits;
};

struct ScoreArgs {
  const u64* pairs;
  const int* gc;
  const int* pc;
  const uint8_t* crowd;
  i64* areas;
  int* matched;
  u64* tp;
  u64* fp;
  u64* fn;
  double* iou_sum;
  u64* confusion;
  int* rec_lab;
  double* rec_val;
  u64* iou_exact;        // C x 2: exact 2^-53 fixed-point IoU sums (dataset mode)
  int iou_add;           // 1: iou_exact holds this update only; add it to iou_sum
  const int* guard;
  int B, G, P, C;
  int reward;
  int areas_ready;
  size_t lds_bytes;
};

struct FinalArgs {
  const i64* tp;
  const i64* fp;
  const i64* fn;
  const double* sums;
  const i64* confusion;
  const uint8_t* things;
  int C;
  double* per_class;
  double* summaries;
  i64* counts;
  double* miou;
  i64* out_tp;
  i64* out_fp;
  i64* out_fn;
  double* out_sums;
  i64* out_confusion;
};

int hist_default_chunk(i64 total_pixels);
// The pair table must be zero on entry; launch_score leaves it zeroed again.
int launch_joint_histogram(HistArgs a, int gt_bytes, int pred_bytes, int chunk, int* tile_ws, hipStream_t stream);
int launch_prepare(const ScoreArgs& a, int* error, hipStream_t stream);
int launch_score(ScoreArgs a, hipStream_t stream);
int launch_finalize(const FinalArgs& a, i64 K, hipStream_t stream);
// Checks pe_compute's arguments and fills `out` (0, or a negative status).
int make_final_args(const int64_t* tp, const int64_t* fp, const int64_t* fn, const double* sums,
                    const int64_t* confusion, const uint8_t* things, int64_t K, int64_t C, double* per_class,
                    double* summaries, int64_t* counts, double* miou, int64_t* out_tp, int64_t* out_fp,
                    int64_t* out_fn, double* out_sums, int64_t* out_confusion, FinalArgs* out);

}  // namespace pe
