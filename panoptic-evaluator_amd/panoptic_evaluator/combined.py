"""Joint panoptic, semantic and instance evaluation with a shared target map.

Public surface of the reference's panoptic_evaluator/combined.py
(SegmentationEvaluator :18-167).  One update() runs, for the whole batch:
the detection x target-slot histogram (one launch), the device-side input
checks (one host read when validating), the panoptic update, and one COCO
matching launch; AP ground truth is exactly the target's thing segments.

The detection x slot histogram counts rows in mask order, so it needs
neither the score order nor the panoptic branch (joint histogram and segment
scoring); the matcher joins them.  While an update is captured into a HIP
graph (validate=False), the panoptic branch and the score order are launched
on a second stream of default priority that joins the caller's stream before
the matcher, so the graph holds them beside the detection histogram.  Eager
updates keep one stream: a cross-stream event wait costs more on the host and
device than the overlap saves at these sizes.  The detection x slot table is
kept between updates and returned all-zero by the matcher, so no fill runs
per update.
"""

from __future__ import annotations

from typing import NamedTuple, Sequence

import torch

from . import _native
from .evaluator import PanopticBatch, PanopticEvaluator, _prepare, final_launch_args, final_results
from .instance import (InstanceBatch, InstanceEvaluator, PackedMasks, ap_launch_args, ap_result, as_bits, offsets,
                       raise_instance_error, score_order)


# Images up to this many detections and target slots use the grouped matcher.
GROUP_LIMIT = 4096


class _PendingMatch(NamedTuple):
    """The grouped matching of an update, deferred: pe_coco_match_grouped's
    arguments (without the stream), the record whose flags it completes, and
    the tensors those arguments point into."""
    args: tuple
    record: tuple
    keep: tuple


class SegmentationEvaluator:
    """PQ/mIoU and mask AP/AR with separate predictions and one shared target."""

    def __init__(self, num_classes: int, *, isthing: Sequence[bool] | None = None,
                 device: str | torch.device = "cuda", validate: bool = True):
        self.panoptic = PanopticEvaluator(num_classes, isthing=isthing, device=device, validate=validate)
        self.device = self.panoptic.tp.device
        self.instance = InstanceEvaluator(num_classes, device=self.device, validate=validate)
        self.num_classes = num_classes
        self.validate = validate
        self._error = torch.zeros(1, dtype=torch.int32, device=self.device)
        self._side = None
        # Set while a whole step (reset, update, compute) is captured into one
        # graph: reset() then clears the accumulators on the side stream, and
        # the update's grouped matching runs inside compute()'s launch instead
        # of at the end of the update.
        self._capturing_step = False
        # Set while a step graph records its mask-making work on a branch of
        # its own: makes the current stream wait for that branch.
        self._masks_ready = None
        self._pending = None
        self._pairs = None
        self._pairs_clean = False
        # The panoptic accumulators and the AP positive counts share one
        # buffer, so reset() is a single fill.  One more word: the completion
        # counter of compute()'s launch (zero between launches; the launch
        # leaves it zero, and reset() zeroes it too).
        words = PanopticEvaluator.state_words(num_classes)
        shared = torch.zeros(words + 4 * num_classes + 1, dtype=torch.int64, device=self.device)
        self.panoptic._bind_state(shared[:words])
        self.instance._counts = shared[words:words + 4 * num_classes].view(num_classes, 4)
        self._done = shared[words + 4 * num_classes:].view(torch.int32)[:1]
        self._shared = shared

    def _pair_table(self, N, S):
        """The N x S detection x slot table, a view of a buffer kept between
        updates (grown when a batch needs more rows): every update leaves its
        rows all-zero (the matcher zeroes the rows it consumed), so no fill
        and no allocation run per update."""
        if self._pairs is None or self._pairs.shape[1] != S or self._pairs.shape[0] < N:
            rows = N if self._pairs is None or self._pairs.shape[1] != S else max(N, 2 * self._pairs.shape[0])
            self._pairs = torch.zeros((max(rows, 1), S), dtype=torch.int64, device=self.device)
            self._pairs_clean = True
        return self._pairs[:N]

    def _side_stream(self, which=0):
        # Created on first use: the second only for bit-packed detections.
        if self._side is None:
            self._side = [None, None]
        if self._side[which] is None:
            self._side[which] = _native.SideStream(self.device)
        return self._side[which]

    def _wait_for_masks(self):
        ready, self._masks_ready = self._masks_ready, None
        if ready is not None:
            ready()

    def _flush(self):
        """Run the deferred matching of the last update, if any."""
        pending, self._pending = self._pending, None
        if pending is not None:
            _native.call("pe_coco_match_grouped", *pending.args, _native.stream(self.device))

    def reset(self):
        # A deferred match still zeroes the rows it owns (the table stays clean).
        self._flush()
        if self._capturing_step:
            # The side stream's panoptic branch writes the accumulators first;
            # the detection histogram need not wait for the fill.  The update
            # captured next joins the side stream back.
            side = self._side_stream()
            side.fork(_native.stream(self.device))
            with torch.cuda.stream(side.stream):
                self._shared.zero_()
        else:
            self._shared.zero_()
        self.instance._records = []

    def _validate_areas(self, areas, classes):
        if areas is None:
            return
        if areas.shape != classes.shape or areas.device != self.device or not areas.is_floating_point():
            raise ValueError("instance_areas must be floating point B x S on evaluator device")

    @torch.no_grad()
    def update(self, prediction: PanopticBatch, detections: InstanceBatch, target: PanopticBatch, *,
               instance_areas: torch.Tensor | None = None):
        """Accumulate all metrics (combined.py:63-161); optional annotation areas
        have shape B x S in target slot order and only drive AP area ranges."""
        self._flush()
        target = target.compact()
        maps, classes, crowd = _prepare(target, self.num_classes, False)
        if maps.device != self.device:
            raise ValueError("target must share evaluator device")
        prepared = self.instance._prepare(detections, prediction=True)
        sizes = target.image_sizes if target.packed else [tuple(maps.shape[1:])] * maps.shape[0]
        if len(prepared) != len(sizes) or any(tuple(image.masks.shape[1:]) != tuple(size)
                                              for image, size in zip(prepared, sizes)):
            raise ValueError("detections and target must have matching image counts and dimensions")
        self._validate_areas(instance_areas, classes)
        device, C = self.device, self.num_classes
        B, S = len(sizes), classes.shape[1]
        counts = [image.masks.shape[0] for image in prepared]
        N = sum(counts)
        det_off = offsets(counts, device)
        fork = not self.validate and torch.cuda.is_current_stream_capturing()
        stream = _native.stream(device)
        max_d = max(counts, default=0)
        # Captured updates match in two parts (images of at most 4096
        # detections and slots): the per-class grouping needs only the score
        # order and the slot classes, so it runs beside the histogram.  Eager
        # updates keep one stream, where the one-part matcher is shorter.
        grouped = fork and max_d <= GROUP_LIMIT and S <= GROUP_LIMIT
        status = torch.empty((N, 40), dtype=torch.int8, device=device)
        ranks = torch.empty(N, dtype=torch.int32, device=device)
        groups = None
        if grouped:
            size = _native.library().pe_match_groups_bytes(B, C, N, S)
            groups = torch.empty(max(size, 256), dtype=torch.uint8, device=device)

        def order_and_group():
            order = score_order(prepared, device)
            if grouped:
                _native.call("pe_match_groups", B, C, S, det_off.data_ptr(), N, max_d, order.classes32.data_ptr(),
                             order.order.data_ptr(), classes.data_ptr(), self.panoptic.isthing.data_ptr(),
                             status.data_ptr(), ranks.data_ptr(), groups.data_ptr(), groups.numel(),
                             _native.stream(device))
            return order

        ranked = None
        bits = any(isinstance(image.masks, PackedMasks) for image in prepared)
        if not all(isinstance(image.masks, PackedMasks) for image in prepared):
            self._wait_for_masks()  # bool masks are read (packed) below
        masks = [as_bits(image.masks) for image in prepared] if bits else [image.masks for image in prepared]
        pointers = _native.device_table([m.data_ptr() for m in masks], device)
        aligned = int(all((h * w) % 16 == 0 for h, w in sizes) and all(m.data_ptr() % 16 == 0 for m in masks))
        pairs = self._pair_table(N, S)
        if target.packed:
            # Packed images: pixel offsets, the pixel total and a tile-list workspace.
            pix_off = offsets([h * w for h, w in sizes], device)
            HW = maps.numel()
            size = _native.library().pe_detection_histogram_workspace_bytes(B, HW)
            workspace = torch.empty(max(size, 256), dtype=torch.uint8, device=device)
        else:
            pix_off, HW, workspace = None, sizes[0][0] * sizes[0][1], None
        # Rows in mask order (order NULL): the matcher maps each ranked
        # detection to its row, so the histogram does not wait for the sort.
        common = (B, pointers.data_ptr(), det_off.data_ptr(), None, maps.data_ptr(), _native.ptr(pix_off), HW, S,
                  max(counts, default=0), N, pairs.data_ptr(), int(self._pairs_clean))
        ws = (_native.ptr(workspace), 0 if workspace is None else workspace.numel(), stream)
        self._pairs_clean = False
        if fork:
            # The panoptic branch, the score order and the grouping need
            # nothing the detection histogram makes: they run on a second
            # stream beside it (the panoptic scratch is allocated once per
            # shape and kept).  They are launched first, so that their
            # workgroups are placed before the histogram's fill the device,
            # and in narrow workgroups, which fit where the histogram's retire.
            # With bit rows the histogram is shorter than the panoptic branch,
            # so the score order and the grouping (which need nothing from
            # that branch) take a third stream; beside the byte-mask histogram
            # a third queue only slows the step.
            side = self._side_stream()
            areas = self.panoptic._update(prediction, target, return_areas=True, side=side)
            order_side = self._side_stream(1) if bits else side
            if bits:
                order_side.fork(stream)
            with torch.cuda.stream(order_side.stream):
                ranked = order_and_group()
        self._wait_for_masks()
        if bits:
            # Bit rows (bool images are packed first): 1/8 of the mask bytes.
            # With every image's extents, a tile reads only the masks that can
            # meet it.
            extents = None
            if all(m.extents is not None for m in masks):
                extents = _native.device_table([m.extents.data_ptr() for m in masks], device).data_ptr()
            _native.call("pe_detection_histogram_bits", *common, extents, *ws)
        else:
            _native.call("pe_detection_histogram", *common, aligned, *ws)
        if ranked is None:
            ranked = order_and_group()
        det_cls, scores, det_cls32 = ranked.classes, ranked.scores, ranked.classes32
        inst = None if instance_areas is None else instance_areas.double().contiguous()
        if self.validate:
            # Full detection areas (row sums, mask order: only emptiness is
            # checked here; the matcher sums the rows it ranks itself).
            det_area = torch.empty(N, dtype=torch.float64, device=device)
            _native.call("pe_row_sums", pairs.data_ptr(), N, S, det_area.data_ptr(), stream)
            _native.call("pe_instance_checks", N, det_cls.data_ptr(), scores.data_ptr(), det_area.data_ptr(), 0,
                         None, None, None, C, self.panoptic.isthing.data_ptr(),
                         0 if inst is None else inst.numel(), _native.ptr(inst), self._error.data_ptr(), stream)
            code = int(self._error.item())
            if code:
                raise_instance_error(code, combined=True)
        # Panoptic totals change only after the instance inputs passed.  The
        # matcher reads the target slot areas straight from the panoptic
        # update's int64 area table and keeps only thing-class slots.
        if fork:
            side.join(stream)
            if order_side is not side:
                order_side.join(stream)
        else:
            areas = self.panoptic._update(prediction, target, return_areas=True)
        record = (scores, det_cls32, ranks, status)
        if grouped and not self._capturing_step:
            _native.call("pe_coco_match_grouped", B, C, S, det_off.data_ptr(), N, groups.data_ptr(), groups.numel(),
                         _native.ptr(crowd), _native.ptr(inst), areas.data_ptr(), areas.stride(0), pairs.data_ptr(),
                         status.data_ptr(), self.instance._counts.data_ptr(), stream)
        elif grouped:
            # The matching is deferred: compute() does it in its own launch
            # (or the next update, reset or compute flushes it first).
            # Detection areas are the matcher's own row sums, and it zeroes
            # every row it consumed, so the table is clean for the next update.
            args = (B, C, S, det_off.data_ptr(), N, groups.data_ptr(), groups.numel(), _native.ptr(crowd),
                    _native.ptr(inst), areas.data_ptr(), areas.stride(0), pairs.data_ptr(), status.data_ptr(),
                    self.instance._counts.data_ptr())
            self._pending = _PendingMatch(args, record, (det_off, groups, crowd, inst, areas, pairs))
        else:
            gt_off = _native.device_table(range(0, (B + 1) * S, S), device)
            inter_off = offsets([d * S for d in counts], device)
            _native.call("pe_coco_match", B, C, det_off.data_ptr(), gt_off.data_ptr(), inter_off.data_ptr(),
                         det_cls32.data_ptr(), None, ranked.order.data_ptr(), classes.data_ptr(), None,
                         _native.ptr(crowd), _native.ptr(inst), pairs.data_ptr(), 8, N, S, status.data_ptr(),
                         ranks.data_ptr(), self.instance._counts.data_ptr(), None, self.panoptic.isthing.data_ptr(),
                         areas.data_ptr(), areas.stride(0), 1, stream)
        self._pairs_clean = True
        self.instance._commit(record)
        return self

    def compute(self):
        """Panoptic/semantic results plus an 'instance' AP/AR dictionary."""
        # One launch runs the PQ finaliser beside the AP classes.
        pan = self.panoptic
        final_outputs, (final_args, _inputs) = final_launch_args(
            pan.tp[None], pan.fp[None], pan.fn[None], pan.iou_sum[None], pan.confusion[None], pan.isthing)
        records = self.instance._records
        pending = self._pending
        if pending is not None and not (len(records) == 1 and records[0] is pending.record):
            self._flush()
            pending = None
        ap_outputs, ap_args, _keep = ap_launch_args(records, self.instance._counts, self.num_classes, self.device)
        if pending is not None:
            # The last update's matching runs in the same launch: each AP
            # class workgroup matches its class in every image first.
            self._pending = None
            B, _C, S, det_off, _N, groups, groups_bytes, crowd, inst, areas, stride, pairs, _status, _counts = \
                pending.args
            _native.call("pe_compute_with_ap_match", *final_args, *ap_args, B, S, det_off, groups, groups_bytes,
                         crowd, inst, areas, stride, pairs, self._done.data_ptr(), _native.stream(self.device))
        else:
            _native.call("pe_compute_with_ap", *final_args, *ap_args, _native.stream(self.device))
        result = final_results(*final_outputs)[0]
        result["instance"] = ap_result(*ap_outputs)
        return result
