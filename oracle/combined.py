"""Oracle for SegmentationEvaluator: PQ/mIoU plus mask AP with the panoptic
target's thing segments as instance ground truth.

Test infrastructure only (see oracle/__init__.py).  Restates
panoptic_evaluator/combined.py of the reference.
"""

from __future__ import annotations

import numpy as np

from . import instance as inst
from . import panoptic as pan


def detection_pairs(masks, target_map, order, S):
    """Detection x target-slot intersections, row r = masks[order[r]]
    (combined.py:98-106)."""
    pairs = np.zeros((len(order), S), np.int64)
    flat = target_map.ravel().astype(np.int64)
    for rank, index in enumerate(order):
        pairs[rank] = np.bincount(flat[masks[index].ravel()], minlength=S)[:S]
    return pairs


class SegmentationState:
    """combined.py:18-167 in numpy form (compact target maps)."""

    def __init__(self, num_classes, isthing=None, validate=True):
        self.panoptic = pan.PanopticState(num_classes, isthing)
        self.instance = inst.InstanceState(num_classes, validate)
        self.num_classes = num_classes
        self.validate = validate

    def reset(self):
        self.panoptic.reset()
        self.instance.reset()

    def update(self, pred_map, pred_classes, detections, target_map, target_classes,
               target_crowd=None, instance_areas=None):
        """combined.py:63-161.  ``detections`` is a list of dicts with masks,
        classes and scores per image."""
        C = self.num_classes
        gc = target_classes
        crowd = target_crowd if target_crowd is not None else np.zeros(gc.shape, bool)
        isthing = self.panoptic.isthing
        if self.validate:
            pan.validate_values(target_map, gc, crowd, C)
        if instance_areas is not None and self.validate:
            if not np.isfinite(instance_areas).all() or (instance_areas < 0).any():
                raise ValueError("instance_areas must be finite and nonnegative")
        prepared = []
        for det in detections:
            classes = np.asarray(det["classes"], np.int64)
            scores = np.asarray(det["scores"], np.float64)
            if self.validate:
                if not np.isfinite(scores).all():
                    raise ValueError("scores must be finite; areas must be nonnegative")
                if ((classes < 0) | (classes >= C)).any():
                    raise ValueError("classes must be valid category indices")
            order = inst.score_order(scores)
            prepared.append((np.asarray(det["masks"], bool), classes[order], scores[order], order))
        histograms, det_areas = [], []
        S = gc.shape[1]
        for (masks, classes, scores, order), tmap in zip(prepared, target_map):
            if self.validate and (~isthing[classes]).any():
                raise ValueError("detection categories must be thing classes")
            pairs = detection_pairs(masks, tmap, order, S)
            areas = pairs.sum(1).astype(np.float64)
            if self.validate and (areas == 0).any():
                raise ValueError("each instance mask must have pixels")
            histograms.append(pairs)
            det_areas.append(areas)
        target_areas = self.panoptic.update(target_map, pred_map, gc, pred_classes, crowd,
                                            validate=self.validate)
        valid_classes = np.clip(gc, 0, C - 1)
        gt_classes = np.where((gc >= 0) & isthing[valid_classes], gc, -1)
        records, counts = [], np.zeros_like(self.instance.counts)
        for b, ((masks, classes, scores, order), pairs, pa) in enumerate(
                zip(prepared, histograms, det_areas)):
            ga = target_areas[b].astype(np.float64)
            tcrowd = crowd[b]
            ann = ga if instance_areas is None else instance_areas[b].astype(np.float64)
            inter = pairs.astype(np.float64)
            union = np.where(tcrowd[None], pa[:, None], pa[:, None] + ga[None] - inter)
            status, ranks, positive = inst.match_ious(inter / np.maximum(union, 1), classes,
                                                      gt_classes[b], pa, tcrowd, ann, C)
            records.append((scores.copy(), classes.astype(np.int32), ranks, status))
            counts += positive
        self.instance.commit(records, counts)

    def compute(self):
        result = self.panoptic.compute()
        result["instance"] = self.instance.compute()
        return result
