"""Oracle for generation-order prefix rewards (segmentation_prefixes).

Test infrastructure only (see oracle/__init__.py).  Restates the CPU branch of
panoptic_evaluator/prefix.py of the reference (prefix.py:11-122).
"""

from __future__ import annotations

import math

import numpy as np

from . import panoptic as pan


def check_arguments(iou_thresholds, checkpoints):
    """prefix.py:28-36."""
    thresholds = tuple(iou_thresholds)
    if not thresholds or any(isinstance(t, bool) or not math.isfinite(t) or not 0 < t <= 1
                             for t in thresholds):
        raise ValueError("iou_thresholds must be finite values in (0, 1]")
    requested = None if checkpoints is None else tuple(checkpoints)
    if requested is not None and (any(type(k) is not int or k < 0 for k in requested)
                                  or any(a >= b for a, b in zip(requested, requested[1:]))):
        raise ValueError("checkpoints must be strictly increasing nonnegative integers")
    return np.array(thresholds, np.float64), requested


def prefixes(masks_list, classes_list, target_map, gc_all, crowd_all, num_classes, *,
             isthing=None, checkpoints=None, iou_thresholds=(0.5, 0.75), validate=True):
    """segmentation_prefixes CPU branch on compact target maps (prefix.py:11-122).

    Returns one list per image of snapshot dicts (numpy values).
    """
    C = num_classes
    thresholds, requested = check_arguments(iou_thresholds, checkpoints)
    isthing = np.array(isthing if isthing is not None else [True] * C, bool)
    crowd_all = crowd_all if crowd_all is not None else np.zeros(gc_all.shape, bool)
    if validate:
        pan.validate_values(target_map, gc_all, crowd_all, C)
        for masks, classes in zip(masks_list, classes_list):
            classes = np.asarray(classes)
            if ((classes < 0) | (classes >= C)).any():
                raise ValueError("classes must be valid category indices")
            if len(masks) and (np.asarray(masks).reshape(len(masks), -1).sum(1) == 0).any():
                raise ValueError("each instance mask must have pixels")
    if requested and any(requested[-1] > len(m) for m in masks_list):
        raise ValueError("checkpoint exceeds prediction count")
    evaluator = pan.PanopticState(C, isthing)
    results = []
    T = len(thresholds)
    for b, (masks, classes) in enumerate(zip(masks_list, classes_list)):
        masks = np.asarray(masks, bool)
        classes = np.asarray(classes, np.int64)
        n, g = len(masks), gc_all.shape[1]
        gc, crowd = gc_all[b], crowd_all[b]
        points = set(requested if requested is not None else (range(1, n + 1) if n else [0]))
        labels = np.concatenate(([-1], classes)).astype(np.int32)
        gt = target_map[b].astype(np.int64)
        canvas = np.zeros(gt.shape, np.int64)
        ga = np.bincount(gt.ravel(), minlength=g)[:g].astype(np.int64)
        pairs = np.zeros((g, n + 1), np.int64)
        pairs[:, 0] = ga
        tp = np.zeros((T, C), np.int64)
        fp = np.zeros_like(tp)
        used = np.zeros((T, g), bool)
        things = (gc >= 0) & isthing[np.maximum(gc, 0)]
        positives = np.zeros(C, np.int64)
        np.add.at(positives, np.maximum(gc, 0), (things & ~crowd).astype(np.int64))
        pixel_areas = masks.reshape(n, gt.size).sum(1).astype(np.float64)
        snapshots = []

        def snapshot(k):
            evaluator.reset()
            semantic = np.zeros((C, C + 1), np.int64)
            valid = (gc >= 0) & ~crowd
            active_labels = np.where(pairs.sum(0) > 0, labels, -1).astype(np.int32)
            columns = np.where(active_labels >= 0, active_labels, C).astype(np.int64)
            indices = gc[valid, None].astype(np.int64) * (C + 1) + columns[None]
            np.add.at(semantic.reshape(-1), indices.ravel(), pairs[valid].ravel())
            evaluator.update_histogram(pairs[None], semantic, gc[None], active_labels[None],
                                       crowd[None])
            result = evaluator.compute()
            result["prefix"] = k
            result["instance"] = dict(
                iou_thresholds=thresholds.copy(), tp=tp.copy(), fp=fp.copy(),
                fn=positives[None] - tp,
                precision=tp.astype(np.float64) / np.maximum(tp + fp, 1),
                recall=tp.astype(np.float64) / np.maximum(positives, 1)[None])
            snapshots.append(result)

        if 0 in points:
            snapshot(0)
        for i in range(n):
            mask = masks[i]
            covered_gt, previous = gt[mask], canvas[mask]
            intersections = np.bincount(covered_gt, minlength=g)[:g].astype(np.int64)
            np.add.at(pairs.reshape(-1), covered_gt * (n + 1) + previous, -1)
            pairs[:, i + 1] = intersections
            canvas[mask] = i + 1
            c = int(classes[i])
            if isthing[c]:
                area = pixel_areas[i]
                union = np.where(crowd, area, area + ga - intersections)
                ious = intersections.astype(np.float64) / np.maximum(union, 1)
                same = gc == c
                eligible = (same[None] & ~crowd[None] & ~used
                            & (ious[None] >= thresholds[:, None]))
                cand = np.where(eligible, ious[None], -1.0)
                best = cand.argmax(1)  # first maximum, as torch.max on CPU
                best_iou = cand[np.arange(T), best]
                matched = best_iou >= 0
                used[np.arange(T), best] |= matched
                ignored = (same[None] & crowd[None] & (ious[None] >= thresholds[:, None])).any(1)
                tp[:, c] += matched.astype(np.int64)
                fp[:, c] += (~matched & ~ignored).astype(np.int64)
            if i + 1 in points:
                snapshot(i + 1)
        results.append(snapshots)
    return results
