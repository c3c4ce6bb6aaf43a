"""The reference's CPU SegmentationEvaluator path restated in PyTorch CPU ops.

TEST INFRASTRUCTURE ONLY, like everything under ``oracle/``: the bench's
``cpu_baseline`` leg times it at the host's threads and at one thread, and
tests/test_oracle.py checks that its results equal the numpy oracle's.

The reference evaluates on the CPU with torch ops, so torch's intra-op threads
apply to its pixel passes.  Those run as torch ops here:

* the joint GT x prediction slot histogram and the semantic confusion
  (evaluator.py:228-243), one ``bincount`` each over the whole batch;
* the detection x target-slot histogram (combined.py:98-106), here one
  ``bincount`` over the nonzero pixels of all of an image's masks at once,
  keyed by (score rank, target slot), instead of one per detection;
* the PQ accumulation (evaluator.py:381-417) and its finaliser
  (evaluator.py:448-468).

The greedy COCO matching and the AP accumulation are sequential per-segment
loops in the reference (instance.py:267-331); they run as the numpy oracle's
restatements of those loops (``oracle.instance``), which are faster than the
same loops over torch scalars, so this baseline errs fast.
"""

from __future__ import annotations

import numpy as np
import torch

from . import instance as inst


def joint_histogram(gt, pred, gc, pc, crowd, C):
    """pairs[b, g, p] and semantic[tc, pc or C] of evaluator.py:228-243 for
    int64 B x H x W slot maps and B x S metadata."""
    B = gt.shape[0]
    G, P = gc.shape[1], pc.shape[1]
    g = gt.reshape(B, -1)
    p = pred.reshape(B, -1)
    image = torch.arange(B, dtype=torch.int64)[:, None] * (G * P)
    pairs = torch.bincount((image + g * P + p).reshape(-1), minlength=B * G * P).view(B, G, P)
    target = torch.gather(gc, 1, g)
    predicted = torch.gather(pc, 1, p)
    counted = (target >= 0) & ~torch.gather(crowd, 1, g)
    predicted = torch.where(predicted >= 0, predicted, C)
    semantic = torch.bincount((target * (C + 1) + predicted)[counted], minlength=C * (C + 1)).view(C, C + 1)
    return pairs, semantic


def detection_rows(masks, target_map, rank, S):
    """rows[r, s] = pixels of the detection of score rank r on target slot s
    (combined.py:98-106).  ``rank[i]`` is detection i's position in score order."""
    D = masks.shape[0]
    if D == 0:
        return torch.zeros((0, S), dtype=torch.int64)
    hits = masks.reshape(D, -1).nonzero()
    keys = rank[hits[:, 0]] * S + target_map.reshape(-1)[hits[:, 1]]
    return torch.bincount(keys, minlength=D * S).view(D, S)


class SegmentationState:
    """SegmentationEvaluator (combined.py:18-167) with torch CPU tensors:
    int64 compact slot maps, int64 classes, bool crowd."""

    def __init__(self, num_classes, isthing=None):
        self.num_classes = num_classes
        self.isthing = torch.tensor(isthing if isthing is not None else [True] * num_classes, dtype=torch.bool)
        self.instance = inst.InstanceState(num_classes, validate=False)
        self.reset()

    def reset(self):
        C = self.num_classes
        self.tp = torch.zeros(C, dtype=torch.int64)
        self.fp = torch.zeros_like(self.tp)
        self.fn = torch.zeros_like(self.tp)
        self.iou_sum = torch.zeros(C, dtype=torch.float64)
        self.confusion = torch.zeros((C, C + 1), dtype=torch.int64)
        self.instance.reset()

    def _accumulate_pq(self, pairs, semantic, gc, pc, crowd):
        """evaluator.py:381-417; returns the GT slot areas."""
        ga, pa = pairs.sum(2), pairs.sum(1)
        union = ga[:, :, None] + pa[:, None, :] - pairs - pairs[:, :1, :]
        iou = pairs.double() / union.clamp_min(1)
        same_class = (gc[:, :, None] == pc[:, None, :]) & (gc[:, :, None] >= 0)
        matched = same_class & ~crowd[:, :, None] & (iou > 0.5)
        gt_matched, pred_matched = matched.any(2), matched.any(1)
        # The last crowd GT slot of the prediction's class (metadata order).
        slot = torch.arange(gc.shape[1])[None, :, None]
        crowd_slot = torch.where(same_class & crowd[:, :, None], slot, -1).amax(1) if gc.shape[1] else \
            torch.full(pc.shape, -1, dtype=torch.int64)
        crowd_overlap = torch.gather(pairs, 1, crowd_slot.clamp_min(0)[:, None, :])[:, 0, :]
        ignored = pairs[:, 0, :] + torch.where(crowd_slot >= 0, crowd_overlap, 0)
        false_pos = (pc >= 0) & ~pred_matched & (ignored.double() <= 0.5 * pa.double())
        false_neg = (gc >= 0) & ~crowd & ~gt_matched
        for labels, flags, total in ((pc, pred_matched, self.tp), (pc, false_pos, self.fp),
                                     (gc, false_neg, self.fn)):
            total.scatter_add_(0, labels.clamp_min(0).reshape(-1), flags.reshape(-1).long())
        self.iou_sum.scatter_add_(0, pc.clamp_min(0).reshape(-1), (iou * matched).sum(1).reshape(-1))
        self.confusion += semantic
        return ga

    def update(self, pred_map, pred_classes, detections, target_map, target_classes, target_crowd,
               instance_areas=None):
        """combined.py:63-161 without input validation.  ``detections`` holds
        one (masks N x H x W bool, classes N int64, scores N f64) per image."""
        C = self.num_classes
        gc, crowd = target_classes, target_crowd
        pairs, semantic = joint_histogram(target_map, pred_map, gc, pred_classes, crowd, C)
        target_areas = self._accumulate_pq(pairs, semantic, gc, pred_classes, crowd)
        gt_classes = torch.where((gc >= 0) & self.isthing[gc.clamp(0, C - 1)], gc, -1)
        S = gc.shape[1]
        records, counts = [], np.zeros((C, 4), np.int64)
        for b, (masks, classes, scores) in enumerate(detections):
            order = torch.argsort(scores, descending=True, stable=True)
            rank = torch.empty_like(order)
            rank[order] = torch.arange(len(order))
            rows = detection_rows(masks, target_map[b], rank, S)
            det_area = rows.sum(1).double()
            gt_area = target_areas[b].double()
            annotation = gt_area if instance_areas is None else instance_areas[b].double()
            inter = rows.double()
            union = torch.where(crowd[b][None], det_area[:, None], det_area[:, None] + gt_area[None] - inter)
            status, ranks, positive = inst.match_ious(
                (inter / union.clamp_min(1)).numpy(), classes[order].numpy(), gt_classes[b].numpy(),
                det_area.numpy(), crowd[b].numpy(), annotation.numpy(), C)
            records.append((scores[order].numpy().copy(), classes[order].numpy().astype(np.int32), ranks, status))
            counts += positive
        self.instance.commit(records, counts)

    def compute(self):
        """evaluator.py:448-468 plus the instance summaries (instance.py:205-227)."""
        C = self.num_classes
        denom = self.tp.double() + 0.5 * (self.fp + self.fn).double()
        pq = self.iou_sum / denom.clamp_min(1)
        sq = self.iou_sum / self.tp.clamp_min(1)
        rq = self.tp / denom.clamp_min(1)
        active = denom > 0

        def mean_over(mask):
            chosen = mask & active
            n = chosen.sum()
            return {"pq": (pq * chosen).sum() / n.clamp_min(1), "sq": (sq * chosen).sum() / n.clamp_min(1),
                    "rq": (rq * chosen).sum() / n.clamp_min(1), "n": n}

        hits = self.confusion[:, :C].diagonal()
        union = self.confusion.sum(1) + self.confusion[:, :C].sum(0) - hits
        class_iou = torch.where(union > 0, hits.double() / union.clamp_min(1), torch.nan)
        return {"All": mean_over(torch.ones_like(active)), "Things": mean_over(self.isthing),
                "Stuff": mean_over(~self.isthing),
                "per_class": {"pq": pq, "sq": sq, "rq": rq, "iou": class_iou},
                "miou": torch.nan_to_num(class_iou).sum() / (union > 0).sum().clamp_min(1),
                "tp": self.tp.clone(), "fp": self.fp.clone(), "fn": self.fn.clone(),
                "iou_sum": self.iou_sum.clone(), "confusion": self.confusion.clone(),
                "instance": self.instance.compute()}
