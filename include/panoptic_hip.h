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

#define PE_ABI_VERSION 3

const char* pe_last_error(void);
int pe_abi_version(void);

/* Streaming-read ceiling (no counterpart in the reference): reads `bytes`
 * of device memory at `data` (16-byte aligned, a multiple of 16) with
 * non-temporal 16-byte loads on `blocks` 256-thread workgroups and nothing
 * else; `sink` is a 4-byte device buffer, written only if the folded data
 * happens to equal a fixed sentinel (the write keeps the loads alive).
 * Timed by the profiling hooks as "stream_read": bench.py reports it as the
 * measured read bandwidth of the same byte count beside each HBM roofline. */
int pe_stream_read(const void* data, int64_t bytes, int blocks, int* sink, void* stream);

/* Stream fork and join (no counterpart in the reference: its extension runs
 * every kernel on the current stream).  pe_stream_wait records `event` on
 * `signalling` and makes `waiting` wait for it; recorded inside HIP graph
 * capture it forks a side stream into the capture or joins it back.  Events
 * come from pe_event_create (timing disabled) and are reusable once the wait
 * has been issued.  Used by the combined evaluator to overlap its panoptic
 * and instance branches. */
int pe_event_create(void** event);
int pe_event_destroy(void* event);
int pe_stream_wait(void* waiting, void* signalling, void* event);

/* Marks (on != 0) or unmarks a stream whose launches run beside a long
 * kernel on another stream (no counterpart in the reference).  The
 * library's latency-bound kernels (segment scoring, iou conversion) then
 * launch 256-thread workgroups on it instead of 1024-thread ones: a
 * 1024-thread workgroup needs 16 free wave slots on one compute unit and
 * waits until the other kernel drains, a narrow one fits beside it.
 * Results do not depend on the mark.  Used for the combined evaluator's
 * side stream. */
int pe_stream_set_concurrent(void* stream, int on);

/* Benchmark timing hooks (no counterpart in the reference): while enabled,
 * HIP events are recorded on the launch stream around every launch of the
 * library's dominant kernels (by kernel name, e.g. "joint_histogram");
 * pe_profile_read synchronizes on them and returns the summed milliseconds and
 * the launch count.  For eager runs only (not inside HIP graph capture). */
int pe_profile_enable(int on);
int pe_profile_reset(void);
int pe_profile_read(const char* kernel, double* total_ms, int64_t* launches);

/* Launch-shape tuning (no counterpart in the reference; used by the
 * autotuner of the batched form): each tuning knob ("joint_histogram",
 * "detection_histogram", "mask_overlap_mfma": launch shapes of those kernels;
 * "packed_grid": one workgroup per work item or a persistent grid for packed
 * batches; "detection_group": detections per detection-histogram workgroup)
 * has pe_tune_count variants that all compute identical results.  pe_tune_set selects one for every later
 * launch in the process (-1 restores the default); pe_tune_get returns the
 * variant in use and pe_tune_describe a short description of one. */
int pe_tune_count(const char* kernel);
int pe_tune_set(const char* kernel, int variant);
int pe_tune_get(const char* kernel);
const char* pe_tune_describe(const char* kernel, int variant);

/* ------------------------------------------------------------ panoptic ---- */

/* Scratch bytes pe_update needs beyond the reference's pairs/areas/matched;
 * packed_pixels = the pixel total of a packed batch (0 for B x H x W maps). */
int64_t pe_update_workspace_bytes(int64_t B, int64_t G, int64_t P, int64_t C, int64_t packed_pixels);

/* Joint slot histogram, areas, validation, PQ matching and accumulation.
 * Replaces the reference's ``update`` binding
 * (panoptic_evaluator/csrc/bindings.cpp:4-9, :29-49; update_cuda at
 * histogram_kernel.cu:257-341).
 *   gt, pred        B x H x W maps; element size map_bytes 4 (int32) or 8
 *                   (int64).  Compact slots, or original IDs when the
 *                   matching ID/slot lookup tables (B x G / B x P) are given.
 *                   map_bytes 3: decoded COCO panoptic PNGs, B x H x W x 3
 *                   uint8 whose ID is R + 256 G + 65536 B (README.md:67-68),
 *                   decoded in the kernel; needs the lookup table and pairs
 *                   with an RGB or int32 map (no reference counterpart: its
 *                   callers decode on the host).
 *   pix_off         NULL, or B + 1 device pixel offsets of a packed batch of
 *                   uneven images stored one after another (image b covers
 *                   pix_off[b] .. pix_off[b + 1]; then H * W is the pixel
 *                   total).  No counterpart in the reference, whose batches
 *                   share one size (README.md:57-58).
 *   gc, pc          B x G / B x P int32 classes; crowd B x G uint8 (bool).
 *   pairs           B x G x P int64 scratch.  pairs_zeroed != 0 promises it
 *                   is all zero on entry (as every call leaves it), which
 *                   saves a fill; otherwise it is zeroed first.
 *   areas           B x (G + P) int64 scratch: GT then predicted slot areas.
 *   matched         B x G int32 scratch.
 *   tp, fp, fn, iou_sum   C (or B x C when reward != 0) accumulators.
 *   confusion       C x (C + 1) int64 accumulator (ignored when reward != 0).
 *   error           1 int32 scratch.
 *   iou_exact       C x 2 uint64 accumulator (dataset mode) or NULL: the
 *                   exact sum of matched IoUs in units of 2^-53 (every
 *                   matched IoU is an exact multiple), as low-32-bit and
 *                   high-bit partial sums; iou_sum is rewritten as its
 *                   correctly rounded FP64 value after every update, so the
 *                   caller zeroes these words together with iou_sum.  NULL
 *                   (the reference's state, which has no such words): this
 *                   update's exact sums live in the workspace and their
 *                   correctly rounded value is added to iou_sum once per
 *                   class, so a caller that zeroes only its own tensors on
 *                   reset gets the right totals.  Ignored when reward != 0.
 *   reward != 0 selects per-image statistics (panoptic_quality): empty
 *   prediction slots are ignored and no confusion is accumulated.
 * validate != 0 synchronizes the stream once to read the error bitmask.
 * Dataset iou_sum is thus independent of batch composition and launch
 * order; per-image (reward) sums fold in prediction-slot order. */
int pe_update(const void* gt, int gt_bytes, const void* pred, int pred_bytes,
              int64_t B, int64_t H, int64_t W, const int64_t* pix_off,
              const int32_t* gc, const int32_t* pc, const uint8_t* crowd,
              int64_t G, int64_t P, int64_t C,
              int64_t* pairs, int64_t* areas, int32_t* matched,
              int64_t* tp, int64_t* fp, int64_t* fn, double* iou_sum, int64_t* confusion,
              const int64_t* gt_ids, const int32_t* gt_slots,
              const int64_t* pred_ids, const int32_t* pred_slots,
              int32_t* error, uint64_t* iou_exact, void* workspace, int64_t workspace_bytes,
              int validate, int reward, int pairs_zeroed, void* stream);

/* pe_update without a workspace argument and without any allocation.  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same arguments, behaviour and results as pe_update,
 * minus workspace/workspace_bytes: the workspace is a static device array of
 * pe_update_noalloc_max_workspace_bytes() bytes (8 MiB) per GPU, held until
 * the process ends; batches whose pe_update_workspace_bytes exceed it are
 * rejected (kWorkspaceTooSmall, < 0) before anything runs.  Call it with the
 * stream's device current.  All calls on one GPU share that array: calls
 * that overlap on one GPU (on different streams or host threads) race, so
 * serialise them. */
int64_t pe_update_noalloc_max_workspace_bytes(void);
int pe_update_noalloc(const void* gt, int gt_bytes, const void* pred, int pred_bytes,
                      int64_t B, int64_t H, int64_t W, const int64_t* pix_off,
                      const int32_t* gc, const int32_t* pc, const uint8_t* crowd,
                      int64_t G, int64_t P, int64_t C,
                      int64_t* pairs, int64_t* areas, int32_t* matched,
                      int64_t* tp, int64_t* fp, int64_t* fn, double* iou_sum, int64_t* confusion,
                      const int64_t* gt_ids, const int32_t* gt_slots,
                      const int64_t* pred_ids, const int32_t* pred_slots,
                      int32_t* error, uint64_t* iou_exact, int validate, int reward, int pairs_zeroed,
                      void* stream);

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
 * maps: map_bytes 4 / 8 integer IDs, or 3: RGB-encoded IDs as in pe_update.
 * ids: B x S sorted int64 (slot 0 = id 0), slots: B x S int32. */
int pe_pack(const void* maps, int map_bytes, int64_t B, int64_t HW,
            const int64_t* ids, const int32_t* slots, int64_t S,
            int32_t* out, int32_t* error, void* stream);

/* ------------------------------------------------------------ instance ---- */

/* Mask x mask intersections with int8 MFMA (bool bytes are 0/1 int8):
 * inter[d, g] = #pixels set in both pm[d] and gm[g] (D x G int32), plus the
 * mask areas pa (D int32) and ga (G int32) from the same pass (either may be
 * NULL).  pm: D x HW bool, gm: G x HW bool.  Replaces pack_masks + overlap of
 * instance_update_cuda (instance_kernel.cu:10-39, :282-309; CPU _match_cpu,
 * instance.py:255-264). */
int pe_mask_overlap(const uint8_t* pm, int64_t D, const uint8_t* gm, int64_t G, int64_t HW,
                    int32_t* inter, int32_t* pa, int32_t* ga, void* stream);

/* The same contraction for a batch of B images in one launch: pm/gm are
 * device arrays of B pointers (D_b x pixels_b and G_b x pixels_b bool
 * masks); det_off, gt_off (B + 1) give the rows of each image's areas in
 * pa/ga and inter_off (B + 1) the start of its D_b x G_b block in inter.
 * pixels (B, device) holds each image's pixel count, or is NULL when every
 * image has HW (then max_hw == HW).  max_d/max_g/max_hw bound the images
 * (they size the grid); total_* size the outputs (zeroed first).  aligned != 0
 * promises 16-byte aligned mask pointers and pixel counts multiple of 16.
 * Replaces the per-image instance_update_cuda contraction loop
 * (instance.py:182-186 calls it per image). */
int pe_mask_overlap_batch(int64_t B, const uint8_t* const* pm, const uint8_t* const* gm,
                          const int64_t* det_off, const int64_t* gt_off, const int64_t* inter_off,
                          const int64_t* pixels, int64_t HW, int64_t max_d, int64_t max_g, int64_t max_hw,
                          int64_t total_d, int64_t total_g, int64_t total_inter, int32_t* inter, int32_t* pa,
                          int32_t* ga, int aligned, void* stream);

/* Bit-packed masks (no reference counterpart: the reference takes bool
 * bytes; its compressed-mask input is a listed extension).  A mask of
 * `pixels` pixels is a row of pe_mask_words(pixels) 32-bit words, multiple
 * of 4: bit p % 32 of word p / 32 is pixel p, bits past `pixels` are zero,
 * rows are 16-byte aligned.  pe_pack_masks packs N rows of HW bool bytes
 * (any nonzero byte is set) into bits (N x pe_mask_words(HW) words). */
int64_t pe_mask_words(int64_t pixels);
int pe_pack_masks(const uint8_t* masks, int64_t N, int64_t HW, uint32_t* bits, void* stream);

/* COCO run-length encoded masks decoded into bit rows (the layout above).
 * No counterpart in the reference, which takes bool bytes and hands its masks
 * to pycocotools as RLE only in its parity tests (tests/test_instance.py:110,
 * ``mask_utils.encode(np.asfortranarray(pixels))``); this is the compressed
 * mask input of its listed extensions.  Format: pycocotools 2.0's
 * (common/maskApi.c) -- column-major runs of 0s then 1s summing to H * W,
 * either as the compressed string or as int32 counts.
 *
 * `table` is a device array of N x 8 int64, one row per mask: byte offset of
 * its characters / int32 counts in `data`, their number (`units`), word
 * offset of its bit row in `bits`, offset of its W entries in the column
 * table, H, W, format (0: string, 1: int32 counts; counts 4-byte aligned)
 * and offset of its run table (`units` entries).  `units` and `columns` are
 * the totals of the units and W entries, `max_width` the largest W.  Each
 * mask's row (pe_mask_words(H * W) words at its offset in `bits`) is
 * rewritten whole, so `bits` need not be cleared.  A malformed mask (bad
 * character, a count of more than seven 5-bit groups, a negative count or a
 * total other than H * W) is left all-zero and ORs a bit into the device
 * word *error (1 character, 2 group length, 4 negative count, 8 total).
 * Workspace: pe_rle_workspace_bytes(N, units, columns). */
/* extents (may be NULL): N x 2 int32, set to each mask's nonzero-word range
 * [first, end) (INT32_MAX, 0 for an empty or malformed mask), as
 * pe_mask_extents would compute it. */
int64_t pe_rle_workspace_bytes(int64_t N, int64_t units, int64_t columns);
int pe_rle_decode(int64_t N, const int64_t* table, const uint8_t* data, int64_t units, int64_t columns,
                  int64_t max_width, uint32_t* bits, int32_t* error, int32_t* extents, void* workspace,
                  int64_t workspace_bytes, void* stream);

/* pe_mask_overlap_batch over bit-packed masks: pm/gm are device arrays of B
 * pointers to bit rows (16-byte aligned); the intersections are exact counts
 * (FP4 matrix cores by default).  Same arguments, outputs and replaced
 * reference interface as pe_mask_overlap_batch (no `aligned`: bit rows always
 * are), plus optional mask extents pm_ext / gm_ext (device arrays of B
 * pointers, image b's to 2 int32 per mask: a word range [first, end) outside
 * which the row is zero, as pe_mask_extents writes; NULL for none): rows
 * that are zero over a work item's pixel range are then skipped. */
int pe_mask_overlap_bits_batch(int64_t B, const uint32_t* const* pm, const uint32_t* const* gm,
                               const int64_t* det_off, const int64_t* gt_off, const int64_t* inter_off,
                               const int64_t* pixels, int64_t HW, int64_t max_d, int64_t max_g, int64_t max_hw,
                               int64_t total_d, int64_t total_g, int64_t total_inter, int32_t* inter,
                               int32_t* pa, int32_t* ga, const int32_t* const* pm_ext,
                               const int32_t* const* gm_ext, void* stream);

/* Scratch bytes of pe_detection_histogram for a packed batch of
 * packed_pixels pixels (0 for a batch of one image size: no scratch). */
int64_t pe_detection_histogram_workspace_bytes(int64_t B, int64_t packed_pixels);

/* Detection x target-slot intersections: pairs[row, s] = #pixels of mask
 * order[row] (within its image; order NULL: mask row) whose target slot is
 * s; rows are grouped per image by det_off (B + 1, device).  masks: device array of B pointers to
 * D_b x pixels_b bool tensors; target: the images' int32 slots one after
 * another.  pix_off NULL: every image has HW pixels (target is B x HW);
 * otherwise B + 1 device pixel offsets of a packed batch of uneven images
 * and HW is the pixel total (workspace from the query above).  pairs:
 * (sum D_b) x S int64; pairs_zeroed != 0 promises it is all zero on entry
 * (as pe_coco_match with zero_inter leaves it), which saves a fill.
 * aligned != 0 promises that every image's pixel count is a multiple of 16
 * and the mask pointers are 16-byte aligned.  Replaces
 * ``instance_histogram`` (bindings.cpp:18-19; instance_histogram_cuda,
 * instance_kernel.cu:335-363), which takes one image at a time. */
int pe_detection_histogram(int64_t B, const uint8_t* const* masks, const int64_t* det_off,
                           const int64_t* order, const int32_t* target, const int64_t* pix_off,
                           int64_t HW, int64_t S, int64_t max_dets, int64_t total_dets, int64_t* pairs,
                           int pairs_zeroed, int aligned, void* workspace, int64_t workspace_bytes, void* stream);

/* pe_detection_histogram over bit-packed detection masks (rows as for
 * pe_pack_masks, 16-byte aligned): same arguments and outputs, no
 * `aligned`.  extents (may be NULL): B device pointers, image b's to 2 int32
 * per mask in mask order, a word range [first, end) outside which the mask's
 * row is zero (pe_mask_extents); a workgroup then reads only the detections
 * whose range meets its tile.  Replaces the same ``instance_histogram``
 * interface. */
int pe_detection_histogram_bits(int64_t B, const uint32_t* const* masks, const int64_t* det_off,
                                const int64_t* order, const int32_t* target, const int64_t* pix_off,
                                int64_t HW, int64_t S, int64_t max_dets, int64_t total_dets, int64_t* pairs,
                                int pairs_zeroed, const int32_t* const* extents, void* workspace,
                                int64_t workspace_bytes, void* stream);

/* Extents of N bit rows of row_words words (a multiple of 4, 16-byte aligned
 * rows): extents[2n], extents[2n + 1] = the first nonzero word of row n and
 * one past its last (0, 0 for an empty row).  No reference counterpart. */
int pe_mask_extents(const uint32_t* bits, int64_t N, int64_t row_words, int32_t* extents, void* stream);

/* COCO greedy matching of a batch of images.  Per image b, detections
 * det_off[b]..det_off[b+1] are in score order (det_cls int32, det_area f64
 * pixel areas) and ground truth gt_off[b]..gt_off[b+1] in metadata order
 * (gt_cls int32, -1 = not a candidate; gt_area f64 pixel areas; gt_crowd;
 * gt_ann f64 annotation areas for the area ranges; det_area may be NULL for
 * int64 intersections, then each detection's area is the sum of its
 * intersection row, as for the shared-target table whose rows span every
 * target slot).  Intersections of image b
 * start at inter_off[b] as a D_b x G_b block (int32 or int64, inter_bytes);
 * det_row (may be NULL) gives the block row of each detection.  Outputs:
 * status (sum D) x 40 int8 [area*10 + threshold], ranks (sum D) int32 (1..100,
 * else 0), counts C x 4 int64 (accumulated).  guard (may be NULL): nothing is
 * written when *guard != 0.  gt_things (C uint8, may be NULL) keeps only GT
 * of thing classes as candidates; gt_area_i64 (may be NULL, then gt_area is
 * used) reads int64 pixel areas at b * gt_area_stride + j; gt_ann may be NULL
 * (pixel areas drive the area ranges).  zero_inter != 0 (int64 rows only)
 * zeroes every detection's intersection row after matching, so a caller that
 * keeps the table can pass it to pe_detection_histogram with pairs_zeroed.
 * Replaces ``instance_update`` /
 * ``instance_match_histogram`` matching (bindings.cpp:13-15, :20-22;
 * match_instances, instance_kernel.cu:41-174; CPU _match_ious_cpu,
 * instance.py:267-297). */
int pe_coco_match(int64_t B, int64_t C, const int64_t* det_off, const int64_t* gt_off,
                  const int64_t* inter_off, const int32_t* det_cls, const double* det_area,
                  const int64_t* det_row, const int32_t* gt_cls, const double* gt_area,
                  const uint8_t* gt_crowd, const double* gt_ann, const void* inter, int inter_bytes,
                  int64_t total_dets, int64_t max_gt, int8_t* status, int32_t* ranks, int64_t* counts,
                  const int32_t* guard, const uint8_t* gt_things, const int64_t* gt_area_i64,
                  int64_t gt_area_stride, int zero_inter, void* stream);

/* pe_coco_match for the shared-target combined path, split in two so that
 * the part needing no intersections runs beside the detection histogram.
 * Replaces the same matching interface as pe_coco_match (the reference's
 * combined.py:90-161 matching over the target's thing segments;
 * instance_kernel.cu:41-174), restricted to its combined-path case: GT are
 * the S target slots of each image (gt_cls B x S int32, -1 = not a
 * candidate), intersections are the detection x slot int64 rows of
 * pe_detection_histogram (image b's rows start at row det_off[b]; row of a
 * ranked detection = det_row[i], its mask index), and the rows are zeroed
 * after matching.
 *
 * pe_match_groups (no intersections needed): per image, groups the
 * detections (score order within a class; det_cls int32 score order) and
 * the candidate slots (class in range and, if gt_things is given, a thing
 * class) by class into `groups` (pe_match_groups_bytes), and writes every
 * detection's rank (1..100 within its class, else 0) and the flags (-1) of
 * the unranked ones.  At most 4096 detections and slots per image.
 *
 * pe_coco_match_grouped (after the histogram): one wave per (image, class)
 * reads its lists, counts the class's positives per area range into
 * counts (C x 4, accumulated), matches the ranked detections (flags into
 * status, (sum D) x 40) and zeroes their rows.  Detection areas are the
 * row sums; GT pixel areas are gt_area_i64[b * gt_area_stride + j]; gt_ann
 * (B x S f64, may be NULL) replaces them for the area ranges. */
int64_t pe_match_groups_bytes(int64_t B, int64_t C, int64_t total_dets, int64_t S);
int pe_match_groups(int64_t B, int64_t C, int64_t S, const int64_t* det_off, int64_t total_dets,
                    int64_t max_dets, const int32_t* det_cls, const int64_t* det_row, const int32_t* gt_cls,
                    const uint8_t* gt_things, int8_t* status, int32_t* ranks, void* groups,
                    int64_t groups_bytes, void* stream);
int pe_coco_match_grouped(int64_t B, int64_t C, int64_t S, const int64_t* det_off, int64_t total_dets,
                          const void* groups, int64_t groups_bytes, const uint8_t* gt_crowd,
                          const double* gt_ann, const int64_t* gt_area_i64, int64_t gt_area_stride,
                          int64_t* pairs, int8_t* status, int64_t* counts, void* stream);

/* Scratch bytes of pe_score_order for N detections in B images. */
int64_t pe_score_order_workspace_bytes(int64_t N, int64_t B);

/* Per-image stable descending score order (ties keep input order) of N
 * detections grouped by det_off (B + 1, device): order[i] = within-image
 * index of sorted position i; scores/classes gathered into that order
 * (classes also as int32).  max_dets: the largest image's detection count
 * (at most 4096: one in-LDS sort per image; 0 or more: a segmented radix
 * sort).  max_dets must be at least every image's count: an image with more
 * detections than the in-LDS sort was sized for is not sorted and gets
 * order entries of -1.  Replaces the per-image
 * torch.argsort(scores, descending=True, stable=True) of instance.py:140. */
int pe_score_order(int64_t B, int64_t N, int64_t max_dets, const int64_t* det_off, const double* scores,
                   const int64_t* classes, int64_t* order, double* sorted_scores,
                   int64_t* sorted_classes, int32_t* sorted_classes32, void* workspace,
                   int64_t workspace_bytes, void* stream);

/* The tables InstanceEvaluator.update hands to pe_coco_match, in one launch
 * for at most 64 images: det_area[i] = pa[det_off[b] + order[i]] as f64 for
 * the i-th detection in score order (pa in mask order, as pe_mask_overlap*
 * writes it), and for the GT masks of every image concatenated (image b at
 * gt_off[b]) their f64 pixel areas (from ga), classes as int64 and int32 and
 * crowd flags.  gt_classes and gt_crowd are HOST arrays of B device pointers
 * (image b's int32 or int64 classes, per class_bytes, and its bool flags),
 * copied into the kernel's arguments.  Replaces the per-image pixel areas,
 * their score-order permutation and the concatenation of classes, crowd flags
 * and areas that the reference's InstanceEvaluator update performs
 * (instance.py:121, :127-130, :143, :186-197). */
int pe_instance_tables(int64_t B, const int64_t* det_off, const int64_t* order, const int32_t* pa, int64_t N,
                       const int64_t* gt_off, const int32_t* ga, int64_t total_g, const void* const* gt_classes,
                       int class_bytes, const uint8_t* const* gt_crowd, double* det_area, double* gt_area,
                       int64_t* gt_cls, int32_t* gt_cls32, uint8_t* crowd, void* stream);

/* pe_score_order for at most 64 images of at most 4096 detections each, read
 * in place: scores and classes are HOST arrays of B device pointers (image b's
 * f64 scores and its int32 or int64 classes, per class_bytes), copied into the
 * kernel's arguments, so no concatenated copy and no upload is needed.  No
 * workspace. */
int pe_score_order_images(int64_t B, int64_t max_dets, const int64_t* det_off, const double* const* scores,
                          const void* const* classes, int class_bytes, int64_t* order, double* sorted_scores,
                          int64_t* sorted_classes, int32_t* sorted_classes32, void* stream);

/* Scratch bytes of pe_ap_compute for N detection records and C classes. */
int64_t pe_ap_workspace_bytes(int64_t N, int64_t C);

/* COCO AP/AR accumulation and summaries.  Records (scores f64, classes int32,
 * ranks int32, status N x 40 int8) in any order; they are stably sorted by
 * descending score on the device.  counts: C x 4 positives.  Outputs:
 * precision 10 x 101 x C x 4 x 3 f64, recall 10 x C x 4 x 3 f64, and summary
 * (12 f64: ap, ap50, ap75, ap_small, ap_medium, ap_large, ar1, ar10, ar100,
 * ar_small, ar_medium, ar_large; -1 when undefined).  status must be 8-byte
 * aligned.  Replaces
 * ``instance_compute`` (bindings.cpp:16-17; instance_compute_cuda,
 * instance_kernel.cu:176-209, :311-333; CPU instance.py:205-248, :300-331). */
int pe_ap_compute(int64_t N, const double* scores, const int32_t* classes, const int32_t* ranks,
                  const int8_t* status, const int64_t* counts, int64_t C, double* precision,
                  double* recall, double* summary, void* workspace, int64_t workspace_bytes,
                  void* stream);

/* pe_ap_compute with the 12 summaries folded by the AP launch's last class
 * workgroup instead of a launch of their own (for at most 4096 records; larger
 * sets fold them in a separate launch as pe_ap_compute does).  done is one
 * device int32 word, zero before the first call; every call leaves it zero, and
 * calls that share it must not overlap.  Replaces the same instance_compute
 * binding (bindings.cpp:16-17, :54; instance_kernel.cu:176-209, :311-333). */
int pe_ap_compute_done(int64_t N, const double* scores, const int32_t* classes, const int32_t* ranks,
                       const int8_t* status, const int64_t* counts, int64_t C, double* precision,
                       double* recall, double* summary, void* workspace, int64_t workspace_bytes,
                       int32_t* done, void* stream);

/* pe_compute and pe_ap_compute in one call (no reference counterpart; used by
 * the combined evaluator's compute()).  The first 17 arguments are
 * pe_compute's without its stream, the rest pe_ap_compute's.  For at most
 * 4096 records the finaliser rows run as extra workgroups of the AP launch,
 * beside the AP classes; otherwise the finaliser is launched first.  Same
 * outputs as the two calls. */
int pe_compute_with_ap(const int64_t* tp, const int64_t* fp, const int64_t* fn, const double* sums,
                       const int64_t* confusion, const uint8_t* things, int64_t K, int64_t C,
                       double* per_class, double* summaries, int64_t* counts, double* miou,
                       int64_t* out_tp, int64_t* out_fp, int64_t* out_fn, double* out_sums,
                       int64_t* out_confusion, int64_t N, const double* scores, const int32_t* classes,
                       const int32_t* ranks, const int8_t* status, const int64_t* ap_counts, int64_t ap_C,
                       double* precision, double* recall, double* summary, void* workspace,
                       int64_t workspace_bytes, void* stream);

/* pe_compute_with_ap with the grouped matching of the last update done in
 * the same launch (no reference counterpart; used by the combined
 * evaluator's compute() after a graph-replayed update, whose matching it
 * defers).  The records (N, scores, classes, ranks, status) are that
 * update's: ranks and the unranked flags come from pe_match_groups, and the
 * ranked flags are written into `status` here; positives are added into
 * ap_counts.  The arguments after `workspace_bytes` are
 * pe_coco_match_grouped's (B, S, det_off, groups, groups_bytes, gt_crowd,
 * gt_ann, gt_area_i64, gt_area_stride, pairs; total_dets = N).  With at
 * most 4096 records each AP class workgroup first matches its class in every
 * image, then accumulates; otherwise the matcher is launched first.  `done`
 * (optional): a device int32 that is zero between calls; the last AP class
 * workgroup to finish then folds the summaries in the same launch and zeroes
 * it again (calls sharing it must not overlap).  Same outputs as
 * pe_coco_match_grouped followed by pe_compute_with_ap. */
int pe_compute_with_ap_match(const int64_t* tp, const int64_t* fp, const int64_t* fn, const double* sums,
                             const int64_t* confusion, const uint8_t* things, int64_t K, int64_t C,
                             double* per_class, double* summaries, int64_t* counts, double* miou,
                             int64_t* out_tp, int64_t* out_fp, int64_t* out_fn, double* out_sums,
                             int64_t* out_confusion, int64_t N, const double* scores, const int32_t* classes,
                             const int32_t* ranks, int8_t* status, int64_t* ap_counts, int64_t ap_C,
                             double* precision, double* recall, double* summary, void* workspace,
                             int64_t workspace_bytes, int64_t B, int64_t S, const int64_t* det_off,
                             const void* groups, int64_t groups_bytes, const uint8_t* gt_crowd,
                             const double* gt_ann, const int64_t* gt_area_i64, int64_t gt_area_stride,
                             int64_t* pairs, int32_t* done, void* stream);

/* out[r] = sum of int64 row r as f64 (detection areas of a shared target
 * histogram, combined.py:107-108). */
int pe_row_sums(const int64_t* pairs, int64_t rows, int64_t cols, double* out, void* stream);

/* Instance input validation into an int32 error bitmask (zeroed first):
 * 1 non-finite score / annotation area or negative annotation area; 2 class
 * outside 0..C-1; 4 mask without pixels; 8 detection class not a thing
 * (things may be NULL); 16 non-finite or negative instance_areas.  Any array
 * may be NULL to skip its check (instance.py:104-126; combined.py:47-61,
 * :92-93, :109-110). */
int pe_instance_checks(int64_t n_det, const int64_t* det_cls, const double* scores, const double* det_area,
                       int64_t n_gt, const int64_t* gt_cls, const double* gt_area, const double* ann,
                       int64_t C, const uint8_t* things, int64_t n_inst, const double* inst_areas,
                       int32_t* error, void* stream);

/* -------------------------------------------------------------- prefix ---- */

/* Scratch bytes of pe_prefix: sum_n = total masks, sum_tri = sum over images
 * of N_b (N_b + 1) / 2, packed_pixels = pixel total of a packed batch (0 for
 * images of one size). */
int64_t pe_prefix_workspace_bytes(int64_t B, int64_t G, int64_t C, int64_t T, int64_t sum_n, int64_t sum_tri,
                                  int64_t packed_pixels);

/* Generation-order prefix rewards for B images.  pix_off NULL: the images
 * share one size and target is B x HW int32 slots; otherwise B + 1 device
 * pixel offsets of a packed batch of uneven images (target holds them one
 * after another, HW is the pixel total, mask b is N_b x pixels_b; no
 * counterpart in the reference).  max_n = the largest N_b.  masks: device
 * array of B pointers to N_b x HW
 * bool masks in generation order (n_off: B + 1 device offsets; cls: sum N
 * int32 classes); gc/crowd: B x G target metadata; things: C uint8;
 * thresholds: T f64 IoU thresholds; k_off (B + 1) / points (sum K int32):
 * strictly increasing checkpoints per image, rows of every output.
 * Outputs per checkpoint row r < total_rows: tp/fp/fn (C int64), sums (C
 * f64), confusion (C x (C + 1) int64) of the visible canvas (feed pe_compute
 * with K = total_rows), and out_itp/out_ifp/out_ifn (T x C int64) and
 * precision/recall (T x C f64) of greedy instance matching.  error bits: 1
 * target slot outside 0..G-1, 2 bad target classes / slot 0, 4 active target
 * metadata and pixels disagree, 8 generated class outside 0..C-1, 16 empty
 * mask, 32 checkpoints not increasing or beyond N_b.  Replaces ``prefix`` (bindings.cpp:24-26; prefix_cuda,
 * prefix_kernel.cu:142-199; CPU segmentation_prefixes, prefix.py:11-122). */
int pe_prefix(int64_t B, const int32_t* target, const int64_t* pix_off, int64_t HW, int64_t G, int64_t C,
              const uint8_t* const* masks, const int64_t* n_off, int64_t sum_n, int64_t sum_tri, int64_t max_n,
              const int32_t* cls, const int32_t* gc, const uint8_t* crowd, const uint8_t* things,
              const double* thresholds, int64_t T, const int64_t* k_off, const int32_t* points,
              int64_t total_rows, int64_t* tp, int64_t* fp, int64_t* fn, double* sums,
              int64_t* confusion, int64_t* out_itp, int64_t* out_ifp, int64_t* out_ifn,
              double* precision, double* recall, int32_t* error, void* workspace,
              int64_t workspace_bytes, int aligned, void* stream);

/* ------------------------------------------------------------- serving ---- */

/* Reward server (no counterpart in the reference: the serving form of
 * pe_update(reward) + pe_image_pq and pe_prefix + pe_compute).  A worker
 * thread admits requests first-in first-out into a device arena of
 * budget_bytes (the HBM budget, allocated once), merges every waiting request
 * of a kind that fits into one packed launch on the next of `streams` streams
 * of default priority, and completes them from their events; a request larger
 * than the whole budget runs alone in memory of its own.  Inputs stay in the
 * caller's device memory until the request completes; they are not validated.
 *   pe_server_create: things (C uint8) and thresholds (T f64) are HOST
 *     arrays; pinned_bytes sizes the pinned staging block; the budget is
 *     capped at 80% of the device memory free at creation.  NULL on error.
 *   pe_server_submit: kind 0 = per-image PQ, 1 = prefix rewards; `images` is a
 *     HOST array of B x 9 int64 per image: target map, then PQ: prediction map,
 *     target classes, prediction classes / prefix: N x H x W bool masks,
 *     generated classes (int32), target classes; then target crowd flags (or
 *     0), H, W, G (target slots), P (prediction slots) or N (masks).  Maps are
 *     H x W int32 compact slots.  Prefix requests give each image's strictly
 *     increasing checkpoints (HOST `points`, `point_counts` per image).
 *     arrival_ns: CLOCK_MONOTONIC submission time (0: now).  Returns the
 *     request id (> 0) or < 0.
 *   pe_server_poll: copies up to max_ids ids of completed requests (waiting up
 *     to wait_us for one); returns the count.
 *   pe_server_take: copies a completed request's result into `out` (PQ: one
 *     f64 per image; prefix: pq, sq, rq, mIoU per checkpoint row, images in
 *     order) and `times` (arrival, launch, completion, CLOCK_MONOTONIC ns),
 *     and forgets the request; returns 0, > 0 when its launch failed (text in
 *     `message`, 256 bytes), < 0 on argument errors.
 *   pe_server_stats: 9 int64: arena bytes in use, arena capacity, requests
 *     waiting, launches in flight, admission passes that waited for memory,
 *     PQ launches, prefix launches, requests completed, requests launched.
 *   pe_server_footprint: device bytes a request takes when it runs alone
 *     (host only; no server needed). */
void* pe_server_create(int device, int64_t budget_bytes, int streams, int64_t C, const uint8_t* things,
                       const double* thresholds, int64_t T, int64_t pinned_bytes);
int pe_server_destroy(void* server);
int64_t pe_server_submit(void* server, int kind, int64_t B, const int64_t* images, const int32_t* points,
                         const int64_t* point_counts, int64_t arrival_ns);
int64_t pe_server_poll(void* server, int64_t* ids, int64_t max_ids, int64_t wait_us);
int pe_server_take(void* server, int64_t id, double* out, int64_t capacity, int64_t* length, int64_t* times,
                   char* message);
int pe_server_stats(void* server, int64_t* out);
int64_t pe_server_footprint(int64_t C, int64_t T, int kind, int64_t B, const int64_t* images,
                            const int32_t* points, const int64_t* point_counts);

/* The server's first-fit arena as host-only bookkeeping (offsets in 256-byte
 * granules, freed neighbours coalesced), for callers and tests that plan
 * memory without a device: alloc returns the offset or -1. */
void* pe_arena_create(int64_t capacity);
int pe_arena_destroy(void* arena);
int64_t pe_arena_alloc(void* arena, int64_t bytes);
int pe_arena_release(void* arena, int64_t offset, int64_t bytes);
int64_t pe_arena_in_use(void* arena);
int64_t pe_arena_largest_free(void* arena);

#ifdef __cplusplus
}
#endif

#endif /* PANOPTIC_HIP_H_ */
