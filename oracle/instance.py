"""Oracle for COCO mask AP/AR: mask IoU, greedy matching, accumulation and
summaries.

Test infrastructure only (see oracle/__init__.py).  Restates
panoptic_evaluator/instance.py of the reference.  The matcher is the
reference's pure-Python loop, so it is also the CPU baseline's dominant cost.
"""

from __future__ import annotations

import numpy as np

RANGES = ((0.0, 1e10), (0.0, 1024.0), (1024.0, 9216.0), (9216.0, 1e10))  # instance.py:251
IOU_THRESHOLDS = (0.5, 0.55, 0.6, 0.65, 0.7, 0.75, 0.8, 0.85, 0.8999999999999999, 0.95)  # :252
EPS = np.finfo(np.float64).eps


def score_order(scores):
    """Stable descending argsort (instance.py:140, :213)."""
    return np.argsort(-np.asarray(scores, np.float64), kind="stable")


def mask_ious(pm, gm, pa, ga, crowd):
    """_match_cpu intersections and IoU (instance.py:255-264)."""
    d, g = len(pm), len(gm)
    inter = np.zeros((d, g), np.float64)
    pixels = pm.shape[1] * pm.shape[2] if pm.ndim == 3 else 0
    gflat = gm.reshape(g, gm.shape[1] * gm.shape[2])
    for i in range(d):
        inter[i] = (pm[i].reshape(1, pixels) & gflat).sum(1)
    union = np.where(crowd[None], pa[:, None], pa[:, None] + ga[None] - inter)
    return inter / np.maximum(union, 1)


def match_ious(ious, pc, gc, pa, crowd, areas, classes):
    """Greedy COCO matching (instance.py:267-297).

    Returns status int8 (40, D) [a*10+t], ranks int32 (D,), counts int64 (C, 4).
    """
    pc = np.asarray(pc)
    gc = np.asarray(gc)
    d = len(pc)
    status = np.full((40, d), -1, np.int8)
    ranks = np.zeros(d, np.int32)
    counts = np.zeros((classes, 4), np.int64)
    crowd = np.asarray(crowd, bool)
    areas = np.asarray(areas, np.float64)
    pa = np.asarray(pa, np.float64)
    for c in range(classes):
        ds = np.nonzero(pc == c)[0].tolist()[:100]
        for rank, i in enumerate(ds):
            ranks[i] = rank + 1
        gs = np.nonzero(gc == c)[0].tolist()
        for a, (low, high) in enumerate(RANGES):
            ignored = crowd | (areas < low) | (areas > high)
            ordered = sorted(gs, key=lambda j: bool(ignored[j]))
            counts[c, a] = sum(not bool(ignored[j]) for j in gs)
            for t in range(10):
                used = set()
                for i in ds:
                    best, overlap = -1, IOU_THRESHOLDS[t]
                    for j in ordered:
                        if j in used and not crowd[j]:
                            continue
                        if best >= 0 and not ignored[best] and ignored[j]:
                            break
                        if ious[i, j] >= overlap:
                            best, overlap = j, float(ious[i, j])
                    if best >= 0:
                        used.add(best)
                        status[a * 10 + t, i] = -1 if ignored[best] else 1
                    elif low <= pa[i] <= high:
                        status[a * 10 + t, i] = 0
    return status, ranks, counts


def accumulate(classes, ranks, status, counts):
    """Precision/recall arrays (instance.py:300-331)."""
    C = counts.shape[0]
    precision = np.full((10, 101, C, 4, 3), -1.0)
    recall = np.full((10, C, 4, 3), -1.0)
    thresholds = np.arange(101, dtype=np.float64) * 0.01
    for c in range(C):
        for a in range(4):
            if not counts[c, a]:
                continue
            for m, maximum in enumerate((1, 10, 100)):
                selected = (classes == c) & (ranks > 0) & (ranks <= maximum)
                flags = status[a * 10: a * 10 + 10][:, selected]
                n = flags.shape[1]
                if n == 0:
                    precision[:, :, c, a, m] = 0
                    recall[:, c, a, m] = 0
                    continue
                tp = np.cumsum(flags == 1, 1).astype(np.float64)
                fp = np.cumsum(flags == 0, 1).astype(np.float64)
                rc = tp / counts[c, a]
                pr = tp / (tp + fp + EPS)
                envelope = np.flip(np.maximum.accumulate(np.flip(pr, 1), 1), 1)
                for t in range(10):
                    pos = np.searchsorted(rc[t], thresholds, side="left")
                    precision[t, :, c, a, m] = np.where(
                        pos < n, envelope[t][np.minimum(pos, n - 1)], 0)
                recall[:, c, a, m] = rc[:, -1]
    return precision, recall


def mean_valid(values):
    """instance.py:230-233."""
    valid = values >= 0
    if not valid.any():
        return np.float64(-1.0)
    return np.where(valid, values, 0).sum() / max(valid.sum(), 1)


def summarize(precision, recall):
    """instance.py:236-248."""
    result = {"ap": mean_valid(precision[:, :, :, 0, 2]),
              "ap50": mean_valid(precision[0, :, :, 0, 2]),
              "ap75": mean_valid(precision[5, :, :, 0, 2])}
    for area, name in enumerate(("small", "medium", "large"), 1):
        result[f"ap_{name}"] = mean_valid(precision[:, :, :, area, 2])
        result[f"ar_{name}"] = mean_valid(recall[:, :, area, 2])
    for index, maximum in enumerate((1, 10, 100)):
        result[f"ar{maximum}"] = mean_valid(recall[:, :, 0, index])
    result.update(precision=precision, recall=recall)
    return result


class InstanceState:
    """InstanceEvaluator records and counts (instance.py:46-227), numpy form.

    Images are dicts with masks (N,H,W bool), classes (N,), and scores (N,)
    for predictions or crowd/areas (optional) for targets.
    """

    def __init__(self, num_classes, validate=True):
        self.num_classes = num_classes
        self.validate = validate
        self.reset()

    def reset(self):
        self.records = []
        self.counts = np.zeros((self.num_classes, 4), np.int64)

    def commit(self, records, counts):
        self.records.extend(records)
        self.counts += counts

    def update(self, predictions, targets):
        """instance.py:171-203 with _prepare (:78-159)."""
        prepared, positives = [], np.zeros_like(self.counts)
        for pred, tgt in zip(predictions, targets):
            pm, pcl = np.asarray(pred["masks"], bool), np.asarray(pred["classes"], np.int64)
            gm, gcl = np.asarray(tgt["masks"], bool), np.asarray(tgt["classes"], np.int64)
            scores = np.asarray(pred["scores"], np.float64)
            pa = pm.reshape(len(pm), pm[0].size if len(pm) else 0).sum(1).astype(np.float64)
            ga = gm.reshape(len(gm), gm[0].size if len(gm) else 0).sum(1).astype(np.float64)
            if self.validate:
                for cl, ar in ((pcl, pa), (gcl, ga)):
                    if ((cl < 0) | (cl >= self.num_classes)).any():
                        raise ValueError("classes must be valid category indices")
                    if (ar == 0).any():
                        raise ValueError("each instance mask must have pixels")
                if not np.isfinite(scores).all():
                    raise ValueError("scores must be finite; areas must be nonnegative")
            crowd = np.asarray(tgt.get("crowd") if tgt.get("crowd") is not None
                               else np.zeros(len(gm), bool), bool)
            areas = (np.asarray(tgt["areas"], np.float64) if tgt.get("areas") is not None else ga)
            order = score_order(scores)
            pm, pcl, scores, pa = pm[order], pcl[order], scores[order], pa[order]
            prepared.append((pm, gm, pcl, gcl, pa, ga, crowd, areas, scores))
        records = []
        for pm, gm, pcl, gcl, pa, ga, crowd, areas, scores in prepared:
            ious = mask_ious(pm, gm, pa, ga, crowd)
            status, ranks, positive = match_ious(ious, pcl, gcl, pa, crowd, areas, self.num_classes)
            records.append((scores.copy(), pcl.astype(np.int32), ranks, status))
            positives += positive
        self.commit(records, positives)

    def compute(self):
        """instance.py:205-227."""
        if self.records:
            scores = np.concatenate([r[0] for r in self.records])
            classes = np.concatenate([r[1] for r in self.records])
            ranks = np.concatenate([r[2] for r in self.records])
            status = np.concatenate([r[3] for r in self.records], 1)
            order = score_order(scores)
            classes, ranks, status = classes[order], ranks[order], status[:, order]
        else:
            classes = np.zeros(0, np.int32)
            ranks = np.zeros(0, np.int32)
            status = np.zeros((40, 0), np.int8)
        precision, recall = accumulate(classes, ranks, status, self.counts)
        return summarize(precision, recall)
