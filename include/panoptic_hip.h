/* C ABI of libpanoptic_hip.so: the MI355X (gfx950) evaluation kernels behind
 * the panoptic_evaluator Python API.
 *
 * Every pointer is a device pointer unless stated otherwise; tensors are
 * dense, row-major and contiguous.  ``stream`` is a hipStream_t passed as
 * void*.  No entry point allocates device memory: outputs and scratch are
 * caller-owned, sized by the *_workspace_bytes queries, so every entry point
 * that does not validate (validate = 0) is safe to capture in a HIP graph.
 *
 * Return value: 0 = success; 1/2/4 = input-validation bitmask (1: unknown ID or
 * slot outside metadata capacity, 2: bad class / slot 0 not void / crowd on
 * slot 0, 4: active metadata and map pixels disagree); < 0 = argument or HIP
 * error with text from pe_last_error().  A validation error is reported before
 * any accumulator is modified.
 */
#ifndef PANOPTIC_HIP_H_
#define PANOPTIC_HIP_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PE_ABI_VERSION 1

const char* pe_last_error(void);
int pe_abi_version(void);

/* ------------------------------------------------------------ panoptic ---- */

/* Scratch bytes pe_update needs beyond the reference's pairs/areas/matched. */
int64_t pe_update_workspace_bytes(int64_t B, int64_t G, int64_t P, int64_t C);

/* Joint slot histogram, areas, validation, PQ matching and accumulation.
 * Replaces the reference's ``update`` binding
 * (panoptic_evaluator/csrc/bindings.cpp:4-9, :29-49; update_cuda at
 * histogram_kernel.cu:257-341).
 *   gt, pred        B x H x W maps; element size map_bytes 4 (int32) or 8
 *                   (int64).  Compact slots, or original IDs when the
 *                   matching ID/slot lookup tables (B x G / B x P) are given.
 *   gc, pc          B x G / B x P int32 classes; crowd B x G uint8 (bool).
 *   pairs           B x G x P int64 scratch (fully rewritten).
 *   areas           B x (G + P) int64 scratch: GT then predicted slot areas.
 *   matched         B x G int32 scratch.
 *   tp, fp, fn, iou_sum   C (or B x C when reward != 0) accumulators.
 *   confusion       C x (C + 1) int64 accumulator (ignored when reward != 0).
 *   error           1 int32 scratch.
 *   reward != 0 selects per-image statistics (panoptic_quality): empty
 *   prediction slots are ignored and no confusion is accumulated.
 * validate != 0 synchronizes the stream once to read the error bitmask.
 * iou_sum is accumulated in (image, prediction slot) order, deterministic. */
int pe_update(const void* gt, int gt_bytes, const void* pred, int pred_bytes,
              int64_t B, int64_t H, int64_t W,
              const int32_t* gc, const int32_t* pc, const uint8_t* crowd,
              int64_t G, int64_t P, int64_t C,
              int64_t* pairs, int64_t* areas, int32_t* matched,
              int64_t* tp, int64_t* fp, int64_t* fn, double* iou_sum, int64_t* confusion,
              const int64_t* gt_ids, const int32_t* gt_slots,
              const int64_t* pred_ids, const int32_t* pred_slots,
              int32_t* error, void* workspace, int64_t workspace_bytes,
              int validate, int reward, void* stream);

/* FP64 finaliser over K independent accumulator rows.  Replaces ``compute``
 * (bindings.cpp:10-11; compute_cuda at compute_kernel.cu:66-95; K > 1 is
 * compute_prefix_cuda, compute_kernel.cu:98-118).
 *   in:  tp, fp, fn (K x C int64), sums (K x C f64), confusion (K x C x (C+1)),
 *        things (C uint8)
 *   out: per_class (K x 4 x C f64: pq, sq, rq, iou), summaries (K x 3 x 3 f64:
 *        All/Things/Stuff x pq/sq/rq), counts (K x 3 int64), miou (K f64),
 *        and copies out_tp/out_fp/out_fn/out_sums/out_confusion (may be NULL). */
int pe_compute(const int64_t* tp, const int64_t* fp, const int64_t* fn, const double* sums,
               const int64_t* confusion, const uint8_t* things, int64_t K, int64_t C,
               double* per_class, double* summaries, int64_t* counts, double* miou,
               int64_t* out_tp, int64_t* out_fp, int64_t* out_fn, double* out_sums,
               int64_t* out_confusion, void* stream);

/* Per-image mean-class PQ from pe_update(reward=1) statistics: out[b] =
 * sum_c pq[b,c] / #active classes (evaluator.py:539-544); stats (may be NULL)
 * receives B x C x 4 f64 (iou_sum, tp, fp, fn) as return_stats=True
 * (evaluator.py:539-540). */
int pe_image_pq(const int64_t* tp, const int64_t* fp, const int64_t* fn, const double* sums,
                int64_t B, int64_t C, double* out, double* stats, void* stream);

/* Original segment IDs -> compact int32 slots (0 on failure, error |= 1).
 * Replaces ``pack`` (bindings.cpp:12; pack_cuda at histogram_kernel.cu:343-366).
 * ids: B x S sorted int64 (slot 0 = id 0), slots: B x S int32. */
int pe_pack(const void* maps, int map_bytes, int64_t B, int64_t HW,
            const int64_t* ids, const int32_t* slots, int64_t S,
            int32_t* out, int32_t* error, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* PANOPTIC_HIP_H_ */
