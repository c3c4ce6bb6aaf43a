"""Panoptic and semantic evaluation on AMD GPUs.

Public surface identical to the reference's panoptic_evaluator/evaluator.py
(PanopticBatch :23-158, PanopticEvaluator :246-468, panoptic_quality
:471-544).  Every pixel- and segment-proportional step runs in the gfx950
kernels of libpanoptic_hip.so (joint histogram, areas/validation, PQ matching,
confusion, FP64 finaliser); this module only checks shapes, owns the
accumulator tensors and maps error codes to the reference's exceptions.
"""

from __future__ import annotations

from dataclasses import dataclass, replace
from numbers import Integral
from typing import Sequence

import numpy as np
import torch

from . import _native

INT64_MAX = torch.iinfo(torch.int64).max

MSG_UNKNOWN = "map contains an unknown segment ID or a slot outside metadata capacity"
MSG_CLASSES = "classes must be -1 or valid category indices; slot 0 must be void"
MSG_ACTIVE = "nonvoid map slots and active metadata must correspond exactly"


def is_rgb(maps: torch.Tensor) -> bool:
    """Decoded COCO panoptic PNGs: B x H x W x 3 uint8, segment ID
    R + 256 * G + 65536 * B (README.md:67-68)."""
    return maps.dtype == torch.uint8 and maps.ndim == 4 and maps.shape[-1] == 3


def id_map_ok(maps: torch.Tensor) -> bool:
    """Original-ID maps: int32/int64 B x H x W, or RGB B x H x W x 3 uint8."""
    return (maps.ndim == 3 and maps.dtype in (torch.int32, torch.int64)) or is_rgb(maps)


def map_bytes(maps: torch.Tensor) -> int:
    """Bytes per pixel as the kernels read them (3 for RGB)."""
    return 3 if is_rgb(maps) else maps.element_size()


def raise_validation(code: int):
    """Map the kernels' validation bitmask to the reference's messages
    (evaluator.py:365-376)."""
    if code & 1:
        raise ValueError(MSG_UNKNOWN)
    if code & 2:
        raise ValueError(MSG_CLASSES)
    raise ValueError(MSG_ACTIVE)


@dataclass(frozen=True)
class PanopticBatch:
    """Segment maps and metadata; optional lookup tables allow original IDs.

    segment_map  B x H x W; compact int32 slots, or original int32/int64 IDs,
                 or original IDs as decoded RGB PNGs (B x H x W x 3 uint8)
    classes      B x S int32; slot 0 is void, class -1 is padding
    crowd        optional B x S bool (ground truth only)
    lookup_ids   optional B x S sorted int64 original IDs
    lookup_slots optional B x S int32 original metadata slots
    """

    segment_map: torch.Tensor
    classes: torch.Tensor
    crowd: torch.Tensor | None = None
    lookup_ids: torch.Tensor | None = None
    lookup_slots: torch.Tensor | None = None
    image_sizes: tuple | None = None

    @classmethod
    def pack(cls, maps: Sequence[torch.Tensor], classes: Sequence[torch.Tensor],
             crowd: Sequence[torch.Tensor] | None = None) -> PanopticBatch:
        """A batch of images of uneven size (no counterpart in the reference,
        whose batches share one size, README.md:57-58): compact int32 H_i x W_i
        maps are stored one after another in a flat ``segment_map`` and the
        per-image slot metadata (S_i int32 classes, S_i bool crowd) is padded
        with class -1 to the largest S.  ``image_sizes`` keeps (H_i, W_i)."""
        if len(maps) == 0 or len(maps) != len(classes) or (crowd is not None and len(crowd) != len(maps)):
            raise ValueError("pack needs one map, one classes tensor (and one crowd tensor) per image")
        device = maps[0].device
        for m, c in zip(maps, classes):
            if m.ndim != 2 or m.dtype != torch.int32 or c.ndim != 1 or c.dtype != torch.int32 or c.numel() < 1:
                raise ValueError("packed maps must be int32 H x W and classes int32 S tensors with S >= 1")
            if m.device != device or c.device != device:
                raise ValueError("batch tensors must share a device")
        if crowd is not None:
            for c, k in zip(classes, crowd):
                if k.shape != c.shape or k.dtype != torch.bool or k.device != device:
                    raise ValueError("crowd must be bool with the same shape as classes")
        lengths = np.fromiter((c.numel() for c in classes), dtype=np.int64, count=len(classes))
        S, B = int(lengths.max()), len(maps)
        flat = torch.cat([m.reshape(-1) for m in maps])
        # Every image's metadata lands in its padded row with one scatter:
        # element j of image b goes to b * S + j.
        starts = np.arange(B, dtype=np.int64) * S - (np.cumsum(lengths) - lengths)
        index = torch.from_numpy(np.arange(int(lengths.sum()), dtype=np.int64) + np.repeat(starts, lengths))
        index = index.to(device)
        padded = torch.full((B * S,), -1, dtype=torch.int32, device=device)
        padded.index_copy_(0, index, torch.cat(list(classes)))
        padded_crowd = None
        if crowd is not None:
            padded_crowd = torch.zeros(B * S, dtype=torch.bool, device=device)
            padded_crowd.index_copy_(0, index, torch.cat(list(crowd)))
            padded_crowd = padded_crowd.view(B, S)
        sizes = tuple((int(m.shape[0]), int(m.shape[1])) for m in maps)
        return cls(flat, padded.view(B, S), padded_crowd, image_sizes=sizes)

    @property
    def packed(self) -> bool:
        return self.image_sizes is not None

    @classmethod
    def from_coco(cls, maps: torch.Tensor, segments_info: Sequence[Sequence[dict]],
                  category_ids: Sequence[int], *, compact: bool = True) -> PanopticBatch:
        """Build slot metadata from COCO ``segments_info`` with one host-to-device
        upload (evaluator.py:33-106).  Metadata keeps input order, so the
        official "last crowd segment per category" rule applies.  With
        compact=False the maps keep their original IDs and the lookup is fused
        into the histogram kernel; pixel validation then happens in update().
        ``maps`` may also be decoded COCO panoptic PNGs (B x H x W x 3 uint8 on
        the device): the ID R + 256 * G + 65536 * B is decoded inside the same
        kernels, so an RGB map is read once at 3 bytes per pixel.
        """
        if not id_map_ok(maps):
            raise ValueError("maps must be integer B x H x W tensors")
        if len(segments_info) != maps.shape[0] or len(set(category_ids)) != len(category_ids):
            raise ValueError("batch length or category IDs are invalid")
        index_of = {category: index for index, category in enumerate(category_ids)}
        capacity = 1 + max((len(info) for info in segments_info), default=0)
        ids, slots, classes, crowd = [], [], [], []
        for info in segments_info:
            segment_ids = [entry["id"] for entry in info]
            if len(set(segment_ids)) != len(segment_ids) or any(
                    not isinstance(i, Integral) or not 0 < i < INT64_MAX for i in segment_ids):
                raise ValueError(
                    "segment IDs must be unique positive int64 values below INT64_MAX; 0 is void")
            try:
                labels = [index_of[entry["category_id"]] for entry in info]
            except KeyError as error:
                raise ValueError("unknown category_id") from error
            # Sorted (id, metadata slot) table for the binary search; slot
            # numbers keep metadata order.
            table = sorted(zip(segment_ids, range(1, len(info) + 1)))
            padding = capacity - 1 - len(info)
            ids += [0, *(i for i, _ in table)] + [INT64_MAX] * padding
            slots += [0, *(s for _, s in table)] + [0] * padding
            classes += [-1, *labels] + [-1] * padding
            crowd += [False, *(bool(entry.get("iscrowd", 0)) for entry in info)] + [False] * padding
        host = [torch.tensor(ids, dtype=torch.int64), torch.tensor(slots, dtype=torch.int32),
                torch.tensor(classes, dtype=torch.int32), torch.tensor(crowd, dtype=torch.bool)]
        # One upload: the typed fields share a single byte buffer (bool last
        # keeps the 8- and 4-byte fields aligned).
        blob = torch.cat([t.view(torch.uint8) for t in host]).to(maps.device)
        views, offset = [], 0
        for t in host:
            size = t.numel() * t.element_size()
            views.append(blob[offset:offset + size].view(t.dtype).reshape(maps.shape[0], capacity))
            offset += size
        batch = cls(maps, views[2], views[3], views[0], views[1])
        return batch.compact() if compact else batch

    def with_maps(self, maps: torch.Tensor) -> PanopticBatch:
        """Reuse prepared metadata for another map batch (evaluator.py:108-118)."""
        if self.packed:
            if maps.shape != self.segment_map.shape or maps.device != self.classes.device:
                raise ValueError("new packed maps must have the same pixel count and metadata device")
            return replace(self, segment_map=maps)
        if (maps.ndim != 3 and not is_rgb(maps)) or maps.shape[0] != self.classes.shape[0] \
                or maps.device != self.classes.device:
            raise ValueError("new maps must have the same batch size and metadata device")
        return replace(self, segment_map=maps)

    def compact(self) -> PanopticBatch:
        """Materialize compact slots with one pixel kernel (evaluator.py:120-158)."""
        if self.lookup_ids is None:
            return self
        maps = self.segment_map
        if not id_map_ok(maps):
            raise ValueError("maps must be integer B x H x W tensors")
        if (self.lookup_slots is None or self.lookup_ids.shape != self.classes.shape
                or self.lookup_slots.shape != self.classes.shape):
            raise ValueError("lookup IDs and slots must match classes shape")
        if (self.lookup_ids.device != maps.device or self.lookup_slots.device != maps.device
                or maps.shape[0] != self.classes.shape[0]):
            raise ValueError("lookup tables and maps must share device and batch size")
        _native.require_gpu(maps.device)
        maps = maps.contiguous()
        ids = self.lookup_ids.contiguous()
        slots = self.lookup_slots.contiguous()
        B, H, W = maps.shape[:3]
        packed = torch.empty((B, H, W), dtype=torch.int32, device=maps.device)
        error = torch.empty(1, dtype=torch.int32, device=maps.device)
        _native.call("pe_pack", maps.data_ptr(), map_bytes(maps), B, H * W, ids.data_ptr(),
                     slots.data_ptr(), ids.shape[1], packed.data_ptr(), error.data_ptr(),
                     _native.stream(maps.device))
        if error.item():
            raise ValueError("map contains a segment absent from segments_info")
        return PanopticBatch(packed, self.classes, self.crowd)


def _prepare(batch: PanopticBatch, num_classes: int, validate: bool, *, ignore_crowd: bool = False,
             allow_empty_slots: bool = False):
    """Shape, dtype and device checks (evaluator.py:161-207).  Value checks run
    inside the kernels (``validate``) and never on the host."""
    maps, classes = batch.segment_map, batch.classes
    if batch.packed:
        if (maps.ndim != 1 or classes.ndim != 2 or classes.shape[0] != len(batch.image_sizes)
                or classes.shape[1] < 1 or maps.numel() != sum(h * w for h, w in batch.image_sizes)):
            raise ValueError("a packed batch needs flat maps of sum(H_i * W_i) pixels and classes B x S")
        if batch.lookup_ids is not None:
            raise ValueError("packed batches take compact slots (compact() them image by image first)")
    elif ((maps.ndim != 3 and not is_rgb(maps)) or classes.ndim != 2 or classes.shape[0] != maps.shape[0]
          or classes.shape[1] < 1):
        raise ValueError("expected maps B x H x W and classes B x S with S >= 1")
    original_ids = batch.lookup_ids is not None
    if (batch.lookup_slots is not None) != original_ids:
        raise ValueError("lookup IDs and slots must be supplied together")
    if not (id_map_ok(maps) if original_ids else maps.dtype == torch.int32) or classes.dtype != torch.int32:
        raise ValueError("classes and compact maps must be int32; original ID maps may also be int64")
    if original_ids:
        for lookup, dtype in ((batch.lookup_ids, torch.int64), (batch.lookup_slots, torch.int32)):
            if lookup.shape != classes.shape or lookup.dtype != dtype or lookup.device != maps.device:
                raise ValueError("lookup tables must match metadata shape, device and expected dtypes")
    crowd = None
    if not ignore_crowd:
        crowd = batch.crowd if batch.crowd is not None else torch.zeros_like(classes, dtype=torch.bool)
        if crowd.shape != classes.shape or crowd.dtype != torch.bool:
            raise ValueError("crowd must be bool with the same shape as classes")
    if classes.device != maps.device or (crowd is not None and crowd.device != maps.device):
        raise ValueError("batch tensors must share a device")
    return maps.contiguous(), classes.contiguous(), None if crowd is None else crowd.contiguous()


def _lookup(batch: PanopticBatch):
    if batch.lookup_ids is None:
        return None, None
    return batch.lookup_ids.contiguous(), batch.lookup_slots.contiguous()


def _geometry(prediction: PanopticBatch, target: PanopticBatch, gt: torch.Tensor, pred: torch.Tensor):
    """(B, H, W, pix_off, packed_pixels) of a prediction/target pair: B x H x W
    maps, or packed images (H = 1, W = pixel total, device pixel offsets)."""
    if target.packed or prediction.packed:
        if target.image_sizes != prediction.image_sizes:
            raise ValueError("prediction and target must pack the same image sizes")
        if gt.shape != pred.shape or gt.device != pred.device:
            raise ValueError("prediction and target must share shape and device")
        offsets, total = [0], 0
        for h, w in target.image_sizes:
            total += h * w
            offsets.append(total)
        return len(target.image_sizes), 1, total, _native.device_table(offsets, gt.device), total
    if gt.shape[:3] != pred.shape[:3] or gt.device != pred.device:
        raise ValueError("prediction and target must share shape and device")
    if is_rgb(gt) and pred.dtype == torch.int64 or is_rgb(pred) and gt.dtype == torch.int64:
        raise ValueError("an RGB map pairs with an RGB or int32 map")
    B, H, W = gt.shape[:3]
    return B, H, W, None, 0


class PanopticEvaluator:
    """Accumulate dataset PQ/SQ/RQ and semantic IoU without copying pixels to
    the host (evaluator.py:246-468)."""

    def __init__(self, num_classes: int, *, isthing: Sequence[bool] | None = None,
                 device: str | torch.device = "cuda", validate: bool = True):
        if not isinstance(num_classes, int) or isinstance(num_classes, bool) or num_classes <= 0:
            raise ValueError("num_classes must be a positive integer")
        if isthing is not None and len(isthing) != num_classes:
            raise ValueError("isthing must contain one flag per category")
        self.num_classes = num_classes
        self.device = _native.require_gpu(device)
        self.validate = validate
        self.isthing = torch.tensor(list(isthing) if isthing is not None else [True] * num_classes,
                                    dtype=torch.bool, device=self.device)
        self._scratch = None
        self._scratch_key = None
        self._workspace = None
        self._default_crowd = None
        self._validation_error = torch.zeros(1, dtype=torch.int32, device=self.device)
        self.reset()

    def reset(self):
        """Zero the accumulators in place (their storage is stable)."""
        if hasattr(self, "tp"):
            self._state.zero_()  # one fill for all accumulators
            return
        # One 8-byte-element buffer holds every accumulator, including the
        # exact 2^-53 fixed-point IoU sums behind iou_sum (2 words per class).
        self._bind_state(torch.zeros(self.state_words(self.num_classes), dtype=torch.int64, device=self.device))

    @staticmethod
    def state_words(num_classes: int) -> int:
        """int64 words of the accumulator buffer for ``num_classes``."""
        return 6 * num_classes + num_classes * (num_classes + 1)

    def _bind_state(self, state: torch.Tensor):
        """Use ``state`` (state_words int64 words, zero or a copy of the
        current totals) as the accumulator buffer."""
        C = self.num_classes
        self._state = state
        self.tp = self._state[0:C]
        self.fp = self._state[C:2 * C]
        self.fn = self._state[2 * C:3 * C]
        self.iou_sum = self._state[3 * C:4 * C].view(torch.float64)
        # Column C counts void (or inactive) predictions on valid GT pixels.
        self.confusion = self._state[4 * C:4 * C + C * (C + 1)].view(C, C + 1)
        self._iou_exact = self._state[4 * C + C * (C + 1):]

    @torch.no_grad()
    def update(self, prediction: PanopticBatch, target: PanopticBatch):
        return self._update(prediction, target)

    def _scratch_for(self, B, G, P, packed_pixels=0):
        shape = (B, G, P)
        if self._scratch is not None and self._scratch_key == (B, G, P, packed_pixels):
            return self._scratch  # same layout as the previous update
        self._scratch_key = (B, G, P, packed_pixels)
        size = _native.library().pe_update_workspace_bytes(B, G, P, self.num_classes, packed_pixels)
        if self._workspace is not None and self._workspace.numel() < size:
            self._scratch = None
        if self._scratch is None or self._scratch[0].shape != shape:
            device = self.device
            # The pair table starts zeroed and every update leaves it zeroed.
            self._scratch = (torch.zeros(shape, dtype=torch.int64, device=device),
                             torch.empty((B, G + P), dtype=torch.int64, device=device),
                             torch.empty((B, G), dtype=torch.int32, device=device))
            self._workspace = torch.zeros(max(size, 256), dtype=torch.uint8, device=device)
        return self._scratch

    @torch.no_grad()
    def _update(self, prediction: PanopticBatch, target: PanopticBatch, *, return_areas=False, side=None):
        """``side`` (a ``_native.SideStream``): fork it from the current stream
        once every input and scratch buffer is ready and launch there; the
        caller joins it back."""
        if target.crowd is None:
            if (self._default_crowd is None or self._default_crowd.shape != target.classes.shape
                    or self._default_crowd.device != target.classes.device):
                self._default_crowd = torch.zeros_like(target.classes, dtype=torch.bool)
            target = replace(target, crowd=self._default_crowd)
        gt, gc, crowd = _prepare(target, self.num_classes, False)
        pred, pc, _ = _prepare(prediction, self.num_classes, False, ignore_crowd=True)
        if gt.device != self.tp.device:
            raise ValueError("prediction and target must share shape and evaluator device")
        B, H, W, pix_off, packed_pixels = _geometry(prediction, target, gt, pred)
        G, P = gc.shape[1], pc.shape[1]
        pairs, areas, matched = self._scratch_for(B, G, P, packed_pixels)
        gt_ids, gt_slots = _lookup(target)
        pred_ids, pred_slots = _lookup(prediction)
        stream = _native.stream(self.device)
        if side is not None:
            side.fork(stream)
            stream = side.handle
        code = _native.call(
            "pe_update", gt.data_ptr(), map_bytes(gt), pred.data_ptr(), map_bytes(pred), B, H, W,
            _native.ptr(pix_off), gc.data_ptr(), pc.data_ptr(), crowd.data_ptr(), G, P, self.num_classes,
            pairs.data_ptr(), areas.data_ptr(), matched.data_ptr(), self.tp.data_ptr(),
            self.fp.data_ptr(), self.fn.data_ptr(), self.iou_sum.data_ptr(), self.confusion.data_ptr(),
            _native.ptr(gt_ids), _native.ptr(gt_slots), _native.ptr(pred_ids), _native.ptr(pred_slots),
            self._validation_error.data_ptr(), self._iou_exact.data_ptr(), self._workspace.data_ptr(),
            self._workspace.numel(), int(self.validate), 0, 1, stream)
        if code:
            raise_validation(code)
        return areas if return_areas else self

    @torch.no_grad()
    def compute(self) -> dict[str, torch.Tensor | dict]:
        """Device tensors in [0, 1]; absent classes have zero PQ and NaN IoU.
        The returned state tensors are copies (valid after reset/update)."""
        return finalize(self.tp[None], self.fp[None], self.fn[None], self.iou_sum[None],
                        self.confusion[None], self.isthing)[0]


def final_launch_args(tp, fp, fn, sums, confusion, things, *, copies=True):
    """Outputs of the FP64 finaliser over K accumulator rows, allocated:
    ((per_class K x 4 x C, summaries K x 3 x 3, counts K x 3, miou K, copies
    of the five state tensors or None), pe_compute's arguments without the
    stream)."""
    K, C = tp.shape
    device = tp.device
    doubles = dict(dtype=torch.float64, device=device)
    per_class = torch.empty((K, 4, C), **doubles)
    summaries = torch.empty((K, 3, 3), **doubles)
    counts = torch.empty((K, 3), dtype=torch.int64, device=device)
    miou = torch.empty(K, **doubles)
    state = [torch.empty_like(t) for t in (tp, fp, fn, sums, confusion)] if copies else [None] * 5
    tensors = [t.contiguous() for t in (tp, fp, fn, sums, confusion)]
    args = (*(t.data_ptr() for t in tensors), things.data_ptr(), K, C, per_class.data_ptr(),
            summaries.data_ptr(), counts.data_ptr(), miou.data_ptr(), *(_native.ptr(t) for t in state))
    # The inputs stay referenced until the call is issued.
    return (per_class, summaries, counts, miou, state), (args, tensors)


def finalize_arrays(tp, fp, fn, sums, confusion, things, *, copies=True):
    """FP64 finaliser over K accumulator rows as device arrays: per_class
    (K x 4 x C), summaries (K x 3 x 3: All/Things/Stuff x pq/sq/rq), counts
    (K x 3), miou (K), and copies of the five state tensors (or None)."""
    outputs, (args, _) = final_launch_args(tp, fp, fn, sums, confusion, things, copies=copies)
    _native.call("pe_compute", *args, _native.stream(tp.device))
    return outputs


def finalize(tp, fp, fn, sums, confusion, things):
    """Run the FP64 finaliser over K accumulator rows; return K result dicts."""
    return final_results(*finalize_arrays(tp, fp, fn, sums, confusion, things))


def final_results(per_class, summaries, counts, miou, copies):
    """K result dictionaries over the finaliser's output arrays."""
    K = per_class.shape[0]
    # Result views come from a few unbind calls (one dispatch each) rather
    # than one indexing op per entry.
    summary_rows = summaries.reshape(K, 9).unbind(0)
    count_rows, class_rows, miou_rows = counts.unbind(0), per_class.unbind(0), miou.unbind(0)
    out_tp, out_fp, out_fn, out_sums, out_confusion = (t.unbind(0) for t in copies)
    results = []
    for k in range(K):
        s, n = summary_rows[k].unbind(0), count_rows[k].unbind(0)
        result = {name: {"pq": s[3 * i], "sq": s[3 * i + 1], "rq": s[3 * i + 2], "n": n[i]}
                  for i, name in enumerate(("All", "Things", "Stuff"))}
        result.update(per_class=dict(zip(("pq", "sq", "rq", "iou"), class_rows[k].unbind(0))),
                      miou=miou_rows[k], tp=out_tp[k], fp=out_fp[k], fn=out_fn[k], iou_sum=out_sums[k],
                      confusion=out_confusion[k])
        results.append(result)
    return results


@torch.no_grad()
def panoptic_quality(prediction: PanopticBatch, target: PanopticBatch, num_classes: int,
                     reduction: str = "none", *, validate: bool = True,
                     return_stats: bool = False) -> torch.Tensor:
    """Stateless mean-class PQ per image, float64 on the input device
    (evaluator.py:471-544).  Prediction slots without pixels are ignored.
    With return_stats=True return [image, class, (IoU sum, TP, FP, FN)]."""
    if not isinstance(num_classes, int) or isinstance(num_classes, bool) or num_classes <= 0:
        raise ValueError("num_classes must be a positive integer")
    if reduction != "none":
        raise ValueError('reduction must be "none"')
    gt, gc, crowd = _prepare(target, num_classes, False)
    pred, pc, _ = _prepare(prediction, num_classes, False, ignore_crowd=True, allow_empty_slots=True)
    device = _native.require_gpu(gt.device)
    B, H, W, pix_off, packed_pixels = _geometry(prediction, target, gt, pred)
    G, P, C = gc.shape[1], pc.shape[1], num_classes
    tp = torch.zeros((B, C), dtype=torch.int64, device=device)
    fp, fn = torch.zeros_like(tp), torch.zeros_like(tp)
    sums = torch.zeros((B, C), dtype=torch.float64, device=device)
    pairs = torch.empty((B, G, P), dtype=torch.int64, device=device)
    areas = torch.empty((B, G + P), dtype=torch.int64, device=device)
    matched = torch.empty((B, G), dtype=torch.int32, device=device)
    error = torch.empty(1, dtype=torch.int32, device=device)
    size = max(_native.library().pe_update_workspace_bytes(B, G, P, C, packed_pixels), 256)
    workspace = torch.zeros(size, dtype=torch.uint8, device=device)
    gt_ids, gt_slots = _lookup(target)
    pred_ids, pred_slots = _lookup(prediction)
    stream = _native.stream(device)
    code = _native.call(
        "pe_update", gt.data_ptr(), map_bytes(gt), pred.data_ptr(), map_bytes(pred), B, H, W,
        _native.ptr(pix_off), gc.data_ptr(), pc.data_ptr(), crowd.data_ptr(), G, P, C, pairs.data_ptr(),
        areas.data_ptr(),
        matched.data_ptr(), tp.data_ptr(), fp.data_ptr(), fn.data_ptr(), sums.data_ptr(), None,
        _native.ptr(gt_ids), _native.ptr(gt_slots), _native.ptr(pred_ids), _native.ptr(pred_slots),
        error.data_ptr(), None, workspace.data_ptr(), workspace.numel(), int(validate), 1, 0, stream)
    if code:
        raise_validation(code)
    if return_stats:
        stats = torch.empty((B, C, 4), dtype=torch.float64, device=device)
        _native.call("pe_image_pq", tp.data_ptr(), fp.data_ptr(), fn.data_ptr(), sums.data_ptr(), B, C,
                     None, stats.data_ptr(), stream)
        return stats
    out = torch.empty(B, dtype=torch.float64, device=device)
    _native.call("pe_image_pq", tp.data_ptr(), fp.data_ptr(), fn.data_ptr(), sums.data_ptr(), B, C,
                 out.data_ptr(), None, stream)
    return out
