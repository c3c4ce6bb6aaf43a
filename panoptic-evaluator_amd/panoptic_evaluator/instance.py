"""Individual mask inputs and COCO instance segmentation evaluation on AMD GPUs.

Public surface of the reference's panoptic_evaluator/instance.py
(InstanceBatch :13-26, InstanceEvaluator :46-227).  Mask intersections run as
int8 MFMA contractions, matching and AP/AR accumulation in the gfx950 kernels;
only scores, classes, ranks and match flags are kept between updates.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import NamedTuple, Sequence

import torch

from . import _native

_i64, _int, _ptr = ctypes.c_int64, ctypes.c_int, ctypes.c_void_p
for _name, _sig in {
    "pe_mask_overlap": (_int, [_ptr, _i64, _ptr, _i64, _i64, _ptr, _ptr, _ptr, _ptr]),
    "pe_mask_overlap_batch": (_int, [_i64, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _i64, _i64, _i64, _i64, _i64, _i64,
                                     _i64, _ptr, _ptr, _ptr, _int, _ptr]),
    "pe_mask_words": (_i64, [_i64]),
    "pe_pack_masks": (_int, [_ptr, _i64, _i64, _ptr, _ptr]),
    "pe_mask_overlap_bits_batch": (_int, [_i64, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _i64, _i64, _i64, _i64, _i64,
                                          _i64, _i64, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr]),
    "pe_detection_histogram_bits": (_int, [_i64, _ptr, _ptr, _ptr, _ptr, _ptr, _i64, _i64, _i64, _i64, _ptr,
                                           _int, _ptr, _ptr, _i64, _ptr]),
    "pe_mask_extents": (_int, [_ptr, _i64, _i64, _ptr, _ptr]),
    "pe_detection_histogram_workspace_bytes": (_i64, [_i64, _i64]),
    "pe_detection_histogram": (_int, [_i64, _ptr, _ptr, _ptr, _ptr, _ptr, _i64, _i64, _i64, _i64, _ptr, _int,
                                      _int, _ptr, _i64, _ptr]),
    "pe_coco_match": (_int, [_i64, _i64, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _int,
                             _i64, _i64, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _i64, _int, _ptr]),
    "pe_match_groups_bytes": (_i64, [_i64, _i64, _i64, _i64]),
    "pe_match_groups": (_int, [_i64, _i64, _i64, _ptr, _i64, _i64, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _i64,
                               _ptr]),
    "pe_coco_match_grouped": (_int, [_i64, _i64, _i64, _ptr, _i64, _ptr, _i64, _ptr, _ptr, _ptr, _i64, _ptr, _ptr,
                                     _ptr, _ptr]),
    "pe_score_order_workspace_bytes": (_i64, [_i64, _i64]),
    "pe_score_order": (_int, [_i64, _i64, _i64, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _i64, _ptr]),
    "pe_score_order_images": (_int, [_i64, _i64, _ptr, _ptr, _ptr, _int, _ptr, _ptr, _ptr, _ptr, _ptr]),
    "pe_instance_tables": (_int, [_i64, _ptr, _ptr, _ptr, _i64, _ptr, _ptr, _i64, _ptr, _int, _ptr, _ptr, _ptr, _ptr,
                                  _ptr, _ptr, _ptr]),
    "pe_ap_workspace_bytes": (_i64, [_i64, _i64]),
    "pe_ap_compute": (_int, [_i64, _ptr, _ptr, _ptr, _ptr, _ptr, _i64, _ptr, _ptr, _ptr, _ptr, _i64, _ptr]),
    "pe_ap_compute_done": (_int, [_i64, _ptr, _ptr, _ptr, _ptr, _ptr, _i64, _ptr, _ptr, _ptr, _ptr, _i64, _ptr,
                                  _ptr]),
    "pe_row_sums": (_int, [_ptr, _i64, _i64, _ptr, _ptr]),
    "pe_instance_checks": (_int, [_i64, _ptr, _ptr, _ptr, _i64, _ptr, _ptr, _ptr, _i64, _ptr, _i64, _ptr, _ptr,
                                  _ptr]),
}.items():
    _native.register(_name, *_sig)

SUMMARY_KEYS = ("ap", "ap50", "ap75", "ap_small", "ap_medium", "ap_large", "ar1", "ar10", "ar100", "ar_small",
                "ar_medium", "ar_large")


def mask_words(pixels: int) -> int:
    """32-bit words per bit-packed mask row of ``pixels`` pixels (rows are
    zero-padded to a multiple of 128 pixels, so every row is 16-byte aligned)."""
    return (int(pixels) + 127) // 128 * 4


@dataclass(frozen=True)
class PackedMasks:
    """N x H x W instance masks stored as bits, one bit per pixel (no
    counterpart in the reference, whose masks are bool bytes): ``bits`` is a
    contiguous int32 N x mask_words(H * W) device tensor in which bit p % 32
    of word p // 32 of row n is pixel p = y * W + x of mask n; bits past H * W
    are zero.  InstanceBatch accepts it wherever it accepts a bool mask
    tensor, and the evaluators then read 1/8 of the bytes.

    ``extents`` (optional, from_bool computes it) is an N x 2 int32 device
    tensor: for mask n, a word range [first, end) outside which its row is
    zero.  The combined evaluator's detection histogram then reads only the
    masks whose range meets each pixel tile.  A caller that builds the bits
    itself may pass extents that bound the nonzero words (wider is allowed,
    narrower gives wrong counts) or leave them out."""

    bits: torch.Tensor
    height: int
    width: int
    extents: torch.Tensor | None = None

    def __post_init__(self):
        if self.height <= 0 or self.width <= 0:
            raise ValueError("packed masks need positive image dimensions")
        words = mask_words(self.height * self.width)
        bits = self.bits
        if (bits.ndim != 2 or bits.dtype != torch.int32 or bits.shape[1] != words or not bits.is_contiguous()
                or bits.data_ptr() % 16):
            raise ValueError("packed mask bits must be a contiguous 16-byte aligned int32 N x mask_words(H * W) "
                             "tensor")
        ext = self.extents
        if ext is not None and (ext.shape != (bits.shape[0], 2) or ext.dtype != torch.int32
                                or not ext.is_contiguous() or ext.device != bits.device):
            raise ValueError("packed mask extents must be a contiguous int32 N x 2 tensor on the bits' device")

    @classmethod
    def from_bool(cls, masks: torch.Tensor) -> PackedMasks:
        """Pack an N x H x W bool device tensor with one streaming kernel."""
        if masks.ndim != 3 or masks.dtype != torch.bool or min(masks.shape[1:]) <= 0:
            raise ValueError("masks must be bool N x H x W with positive image dimensions")
        device = _native.require_gpu(masks.device)
        masks = masks.contiguous()
        N, H, W = masks.shape
        bits = torch.empty((N, mask_words(H * W)), dtype=torch.int32, device=device)
        stream = _native.stream(device)
        _native.call("pe_pack_masks", masks.data_ptr(), N, H * W, bits.data_ptr(), stream)
        return cls(bits, H, W, mask_extents(bits))

    def to_bool(self) -> torch.Tensor:
        """The N x H x W bool masks (torch ops; for inspection and tests)."""
        shifts = torch.arange(32, dtype=torch.int32, device=self.bits.device)
        flags = (self.bits[:, :, None] >> shifts) & 1
        N, HW = self.bits.shape[0], self.height * self.width
        return flags.reshape(N, 32 * self.bits.shape[1])[:, :HW].bool().reshape(N, self.height, self.width)

    @property
    def shape(self) -> torch.Size:
        return torch.Size((self.bits.shape[0], self.height, self.width))

    @property
    def ndim(self) -> int:
        return 3

    @property
    def dtype(self) -> torch.dtype:
        return self.bits.dtype

    @property
    def device(self) -> torch.device:
        return self.bits.device

    def data_ptr(self) -> int:
        return self.bits.data_ptr()

    def contiguous(self) -> PackedMasks:
        return self


def mask_extents(bits: torch.Tensor) -> torch.Tensor:
    """[first, end) of the nonzero words of every row of an N x words int32
    bit tensor (0, 0 for an empty row), as an N x 2 int32 device tensor."""
    extents = torch.empty((bits.shape[0], 2), dtype=torch.int32, device=bits.device)
    _native.call("pe_mask_extents", bits.data_ptr(), bits.shape[0], bits.shape[1], extents.data_ptr(),
                 _native.stream(bits.device))
    return extents


def pack_masks(masks: torch.Tensor) -> PackedMasks:
    """Bit-pack an N x H x W bool device tensor (PackedMasks.from_bool)."""
    return PackedMasks.from_bool(masks)


def as_bits(masks) -> PackedMasks:
    return masks if isinstance(masks, PackedMasks) else PackedMasks.from_bool(masks)


@dataclass(frozen=True)
class InstanceBatch:
    """One N x H x W bool tensor per image, with N category indices.

    A PackedMasks may stand in for any image's bool tensor (one bit per pixel).
    Predictions require floating point scores.  Targets accept crowd flags and
    annotation areas (COCO's area filtering uses annotation area for targets).
    Masks may overlap.  Empty images use tensors with N = 0.
    """

    masks: Sequence[torch.Tensor]
    classes: Sequence[torch.Tensor]
    scores: Sequence[torch.Tensor] | None = None
    crowd: Sequence[torch.Tensor] | None = None
    areas: Sequence[torch.Tensor] | None = None


class _PreparedImage(NamedTuple):
    masks: torch.Tensor
    classes: torch.Tensor         # input order
    crowd: torch.Tensor | None    # targets only
    annotation_areas: torch.Tensor | None
    scores: torch.Tensor | None   # input order
    order: torch.Tensor | None    # unused (kept for the record layout)


class ScoreOrder(NamedTuple):
    """Stable descending score order of every detection of a batch."""

    counts: list                  # detections per image
    det_off: torch.Tensor         # B + 1 (device)
    order: torch.Tensor           # within-image mask index per sorted position (int64)
    scores: torch.Tensor          # float64, sorted
    classes: torch.Tensor         # int64, sorted (validation)
    classes32: torch.Tensor       # int32, sorted (matching, records)


ORDER_ARG_IMAGES = 64      # images whose arrays pe_score_order_images reads in place
ORDER_LDS_DETS = 4096      # detections per image of the in-LDS sort


def score_order(images, device) -> ScoreOrder:
    """One segmented radix sort over all images (instance.py:140 per image)."""
    counts = [image.masks.shape[0] for image in images]
    B, N = len(images), sum(counts)
    det_off = offsets(counts, device)
    class_dtype = images[0].classes.dtype if images else torch.int64
    if (N and B <= ORDER_ARG_IMAGES and max(counts) <= ORDER_LDS_DETS
            and all(image.scores.dtype == torch.float64 and image.scores.is_contiguous()
                    and image.classes.dtype == class_dtype and image.classes.is_contiguous() for image in images)):
        # Every image's arrays read in place: their addresses travel in the
        # kernel's arguments (no concatenation, no upload).
        order = torch.empty(N, dtype=torch.int64, device=device)
        sorted_scores = torch.empty(N, dtype=torch.float64, device=device)
        sorted_classes = torch.empty(N, dtype=torch.int64, device=device)
        sorted_classes32 = torch.empty(N, dtype=torch.int32, device=device)
        pointers = ctypes.c_void_p * B
        _native.call("pe_score_order_images", B, max(counts), det_off.data_ptr(),
                     pointers(*(image.scores.data_ptr() for image in images)),
                     pointers(*(image.classes.data_ptr() for image in images)),
                     8 if class_dtype == torch.int64 else 4, order.data_ptr(), sorted_scores.data_ptr(),
                     sorted_classes.data_ptr(), sorted_classes32.data_ptr(), _native.stream(device))
        return ScoreOrder(counts, det_off, order, sorted_scores, sorted_classes, sorted_classes32)
    if B == 1:
        # One image: its own arrays (converted only if they are not float64 /
        # int64 already), no concatenated copy.
        scores = images[0].scores.reshape(-1).to(torch.float64).contiguous()
        classes = images[0].classes.reshape(-1).to(torch.int64).contiguous()
    else:
        scores = _cat([image.scores for image in images], torch.float64, device)
        classes = _cat([image.classes for image in images], torch.int64, device)
    order = torch.empty(N, dtype=torch.int64, device=device)
    sorted_scores = torch.empty(N, dtype=torch.float64, device=device)
    sorted_classes = torch.empty(N, dtype=torch.int64, device=device)
    sorted_classes32 = torch.empty(N, dtype=torch.int32, device=device)
    if N:
        size = max(_native.library().pe_score_order_workspace_bytes(N, B), 256)
        workspace = torch.empty(size, dtype=torch.uint8, device=device)
        _native.call("pe_score_order", B, N, max(counts), det_off.data_ptr(), scores.data_ptr(),
                     classes.data_ptr(), order.data_ptr(), sorted_scores.data_ptr(), sorted_classes.data_ptr(),
                     sorted_classes32.data_ptr(), workspace.data_ptr(), workspace.numel(), _native.stream(device))
    return ScoreOrder(counts, det_off, order, sorted_scores, sorted_classes, sorted_classes32)


def match_tables(ranked, pa, ga, gt_off, targets, device, stream):
    """Detection pixel areas in score order and the GT areas, classes and crowd
    flags of every image concatenated (instance.py:121, :127-130, :143): one
    launch (pe_instance_tables) reading each image's arrays in place."""
    B, N, total_g = len(targets), pa.numel(), ga.numel()
    class_dtype = targets[0].classes.dtype if targets else torch.int64
    if B <= ORDER_ARG_IMAGES and all(g.classes.dtype == class_dtype and g.classes.is_contiguous()
                                     and g.crowd.is_contiguous() for g in targets):
        det_area = torch.empty(N, dtype=torch.float64, device=device)
        gt_area = torch.empty(total_g, dtype=torch.float64, device=device)
        gt_cls = torch.empty(total_g, dtype=torch.int64, device=device)
        gt_cls32 = torch.empty(total_g, dtype=torch.int32, device=device)
        crowd = torch.empty(total_g, dtype=torch.bool, device=device)
        pointers = ctypes.c_void_p * B
        _native.call("pe_instance_tables", B, ranked.det_off.data_ptr(), ranked.order.data_ptr(), pa.data_ptr(), N,
                     gt_off.data_ptr(), ga.data_ptr(), total_g, pointers(*(g.classes.data_ptr() for g in targets)),
                     8 if class_dtype == torch.int64 else 4, pointers(*(g.crowd.data_ptr() for g in targets)),
                     det_area.data_ptr(), gt_area.data_ptr(), gt_cls.data_ptr(), gt_cls32.data_ptr(),
                     crowd.data_ptr(), stream)
        return det_area, gt_area, gt_cls, gt_cls32, crowd
    first, base = 0, []
    for count in ranked.counts:
        base.extend([first] * count)
        first += count
    det_area = pa.double()[ranked.order + _native.device_table(base, device)]
    gt_cls = _cat([g.classes for g in targets], torch.int64, device)
    return det_area, ga.double(), gt_cls, gt_cls.int(), _cat([g.crowd for g in targets], torch.bool, device)


def raise_instance_error(code: int, *, combined: bool = False):
    """Map the instance validation bitmask to the reference's messages."""
    if code & 1:
        raise ValueError("scores must be finite; areas must be nonnegative")
    if code & 2:
        raise ValueError("classes must be valid category indices")
    if code & 16:
        raise ValueError("instance_areas must be finite and nonnegative")
    if code & 8:
        raise ValueError("detection categories must be thing classes")
    raise ValueError("each instance mask must have pixels")


def _cat(tensors, dtype, device):
    tensors = [t.reshape(-1).to(dtype) for t in tensors]
    return torch.cat(tensors) if tensors else torch.empty(0, dtype=dtype, device=device)


def offsets(counts, device):
    """Device int64 exclusive prefix sums (len(counts) + 1) of host counts."""
    total, values = 0, [0]
    for count in counts:
        total += int(count)
        values.append(total)
    return _native.device_table(values, device)


class InstanceEvaluator:
    """COCO mask AP/AR with MFMA intersections and on-device greedy matching
    (instance.py:46-227).  Only scores, matches and counts are retained."""

    def __init__(self, num_classes: int, *, device: str | torch.device = "cpu", validate: bool = True):
        """The reference's signature, default device included (instance.py:53-58);
        this build evaluates on AMD GPUs only, so pass device="cuda"."""
        if not isinstance(num_classes, int) or isinstance(num_classes, bool) or num_classes <= 0:
            raise ValueError("num_classes must be a positive integer")
        self.num_classes = num_classes
        self.device = _native.require_gpu(device)
        self.validate = validate
        self._error = torch.zeros(1, dtype=torch.int32, device=self.device)
        # Completion word of compute()'s AP launch (zero between launches).
        self._done = torch.zeros(1, dtype=torch.int32, device=self.device)
        self.reset()
        self.device = self._counts.device  # e.g. "cuda" -> cuda:0, as instance.py:70

    def reset(self):
        """Drop the records and zero the positive counts in place (their
        storage is stable, so a captured update keeps accumulating into it)."""
        self._records = []
        if getattr(self, "_counts", None) is None:
            self._counts = torch.zeros((self.num_classes, 4), dtype=torch.int64, device=self.device)
        else:
            self._counts.zero_()

    def _prepare(self, batch: InstanceBatch, *, prediction: bool):
        """Shape/dtype/device checks and the stable per-image score order
        (instance.py:78-159).  Value checks run on the device."""
        count = len(batch.masks)
        for field in (batch.classes, batch.scores, batch.crowd, batch.areas):
            if field is not None and len(field) != count:
                raise ValueError("all fields must contain one tensor per image")
        if prediction and batch.scores is None:
            raise ValueError("predictions require scores")
        prepared = []
        for b, (masks, classes) in enumerate(zip(batch.masks, batch.classes)):
            if not isinstance(masks, PackedMasks) and (
                    masks.ndim != 3 or masks.dtype != torch.bool or min(masks.shape[1:]) <= 0):
                raise ValueError("masks must be bool N x H x W with positive image dimensions")
            n = masks.shape[0]
            if classes.shape != (n,) or classes.dtype not in (torch.int32, torch.int64):
                raise ValueError("classes must be int32/int64 N tensors")
            for name in ("scores", "crowd", "areas"):
                field = getattr(batch, name)
                if field is None:
                    continue
                value = field[b]
                if value.shape != (n,) or value.device != self.device:
                    raise ValueError(f"{name} must have shape N and share evaluator device")
                if name == "crowd":
                    if value.dtype != torch.bool:
                        raise ValueError("crowd must be bool")
                elif not value.is_floating_point():
                    raise ValueError(f"{name} must be floating point")
            if masks.device != self.device or classes.device != self.device:
                raise ValueError("inputs must share evaluator device")
            crowd = None  # detections carry no crowd flags
            if not prediction:
                crowd = (batch.crowd[b].contiguous() if batch.crowd is not None
                         else torch.zeros(n, dtype=torch.bool, device=self.device))
            annotation = batch.areas[b].double() if not prediction and batch.areas is not None else None
            # Predictions stay in input order here; ScoreOrder sorts every
            # image of the batch in one segmented radix sort.
            scores = batch.scores[b] if prediction else None
            prepared.append(_PreparedImage(masks.contiguous(), classes, crowd, annotation, scores, None))
        return prepared

    def _commit(self, record):
        self._records.append(record)

    @torch.no_grad()
    def update(self, prediction: InstanceBatch, target: InstanceBatch):
        predictions = self._prepare(prediction, prediction=True)
        targets = self._prepare(target, prediction=False)
        if len(predictions) != len(targets) or any(
                p.masks.shape[1:] != g.masks.shape[1:] for p, g in zip(predictions, targets)):
            raise ValueError("prediction and target must have matching image counts and dimensions")
        device, C = self.device, self.num_classes
        stream = _native.stream(device)
        ranked = score_order(predictions, device)
        det_cls, scores, det_cls32 = ranked.classes, ranked.scores, ranked.classes32
        # Intersections and pixel areas of every image in one MFMA launch, in
        # input (mask) order.
        det_counts = ranked.counts
        gt_counts = [g.masks.shape[0] for g in targets]
        pixels = [p.masks.shape[1] * p.masks.shape[2] for p in predictions]
        inter_counts = [d * g for d, g in zip(det_counts, gt_counts)]
        N, total_g, total_inter = sum(det_counts), sum(gt_counts), sum(inter_counts)
        gt_off, inter_off = offsets(gt_counts, device), offsets(inter_counts, device)
        uniform = all(n == pixels[0] for n in pixels)
        inter = torch.empty(total_inter, dtype=torch.int32, device=device)
        pa = torch.empty(N, dtype=torch.int32, device=device)
        ga = torch.empty(total_g, dtype=torch.int32, device=device)
        shape_args = (ranked.det_off.data_ptr(), gt_off.data_ptr(), inter_off.data_ptr(),
                      None if uniform else _native.device_table(pixels, device).data_ptr(),
                      pixels[0] if pixels else 0, max(det_counts, default=0), max(gt_counts, default=0),
                      max(pixels, default=0), N, total_g, total_inter, inter.data_ptr(), pa.data_ptr(),
                      ga.data_ptr())
        if any(isinstance(x.masks, PackedMasks) for x in predictions + targets):
            # Bit rows on both sides (bool images are packed first): the FP4
            # matrix-core contraction instead of the int8 one over bytes.
            pred_bits = [as_bits(p.masks) for p in predictions]
            gt_bits = [as_bits(g.masks) for g in targets]
            # Mask extents (when every mask set has them) let each work item
            # skip the rows that are zero over its pixel range.
            ext = [None, None]
            for side, bits in enumerate((pred_bits, gt_bits)):
                if all(m.extents is not None for m in bits):
                    ext[side] = _native.device_table([m.extents.data_ptr() for m in bits], device).data_ptr()
            _native.call("pe_mask_overlap_bits_batch", len(predictions),
                         _native.device_table([m.data_ptr() for m in pred_bits], device).data_ptr(),
                         _native.device_table([m.data_ptr() for m in gt_bits], device).data_ptr(),
                         *shape_args, ext[0], ext[1], stream)
        else:
            aligned = int(all(n % 16 == 0 for n in pixels)
                          and all(x.masks.data_ptr() % 16 == 0 for x in predictions + targets))
            _native.call("pe_mask_overlap_batch", len(predictions),
                         _native.device_table([p.masks.data_ptr() for p in predictions], device).data_ptr(),
                         _native.device_table([g.masks.data_ptr() for g in targets], device).data_ptr(),
                         *shape_args, aligned, stream)
        det_area, gt_area, gt_cls, gt_cls32, crowd = match_tables(ranked, pa, ga, gt_off, targets, device, stream)
        # Annotation areas drive the area ranges; images without them use the
        # pixel areas (the matcher reads gt_area when none are given at all).
        ann = None
        if any(g.annotation_areas is not None for g in targets):
            parts, start = [], 0
            for g, count in zip(targets, gt_counts):
                parts.append(g.annotation_areas if g.annotation_areas is not None else gt_area[start:start + count])
                start += count
            ann = _cat(parts, torch.float64, device)
        if self.validate:
            _native.call("pe_instance_checks", det_cls.numel(), det_cls.data_ptr(), scores.data_ptr(),
                         det_area.data_ptr(), gt_cls.numel(), gt_cls.data_ptr(), gt_area.data_ptr(), _native.ptr(ann),
                         C, None, 0, None, self._error.data_ptr(), stream)
            code = int(self._error.item())
            if code:
                raise_instance_error(code)
        det_off = ranked.det_off
        det_row = ranked.order  # intersection rows are in mask order
        status = torch.empty((N, 40), dtype=torch.int8, device=device)
        ranks = torch.empty(N, dtype=torch.int32, device=device)
        _native.call("pe_coco_match", len(predictions), C, det_off.data_ptr(), gt_off.data_ptr(),
                     inter_off.data_ptr(), det_cls32.data_ptr(), det_area.data_ptr(), det_row.data_ptr(),
                     gt_cls32.data_ptr(), gt_area.data_ptr(), crowd.data_ptr(), _native.ptr(ann), inter.data_ptr(), 4,
                     N, max(gt_counts, default=0), status.data_ptr(), ranks.data_ptr(), self._counts.data_ptr(),
                     None, None, None, 0, 0, stream)
        self._commit((scores, det_cls32, ranks, status))
        return self

    @torch.no_grad()
    def compute(self):
        """COCO AP/AR summaries and the full precision/recall arrays."""
        return accumulate(self._records, self._counts, self.num_classes, self.device, done=self._done)


def accumulate(records, counts, num_classes, device, *, done=None):
    """Sort all records by score and accumulate AP/AR on the device
    (instance.py:205-248).  With `done` (a zeroed int32 word that every call
    leaves zero) the AP launch folds the summaries itself."""
    outputs, args, keep = ap_launch_args(records, counts, num_classes, device)
    if done is None:
        _native.call("pe_ap_compute", *args, _native.stream(device))
    else:
        _native.call("pe_ap_compute_done", *args, done.data_ptr(), _native.stream(device))
    return ap_result(*outputs)


def ap_result(summary, precision, recall):
    result = {key: summary[i] for i, key in enumerate(SUMMARY_KEYS)}
    result.update(precision=precision, recall=recall)
    return result


def ap_launch_args(records, counts, num_classes, device):
    """AP outputs (summary, precision, recall), pe_ap_compute's arguments
    without the stream, and the tensors those arguments point into."""
    C = num_classes
    if len(records) == 1:
        scores, classes, ranks, status = records[0]
    elif records:
        scores = torch.cat([r[0] for r in records])
        classes = torch.cat([r[1] for r in records])
        ranks = torch.cat([r[2] for r in records])
        status = torch.cat([r[3] for r in records])
    else:
        scores = torch.empty(0, dtype=torch.float64, device=device)
        classes = torch.empty(0, dtype=torch.int32, device=device)
        ranks = torch.empty(0, dtype=torch.int32, device=device)
        status = torch.empty((0, 40), dtype=torch.int8, device=device)
    N = scores.numel()
    precision = torch.empty((10, 101, C, 4, 3), dtype=torch.float64, device=device)
    recall = torch.empty((10, C, 4, 3), dtype=torch.float64, device=device)
    summary = torch.empty(12, dtype=torch.float64, device=device)
    size = max(_native.library().pe_ap_workspace_bytes(N, C), 256)
    workspace = torch.empty(size, dtype=torch.uint8, device=device)
    args = (N, scores.data_ptr(), classes.data_ptr(), ranks.data_ptr(), status.data_ptr(), counts.data_ptr(), C,
            precision.data_ptr(), recall.data_ptr(), summary.data_ptr(), workspace.data_ptr(), workspace.numel())
    return (summary, precision, recall), args, (scores, classes, ranks, status, workspace)
