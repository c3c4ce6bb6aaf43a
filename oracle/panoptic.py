"""Oracle for the panoptic path: ID compaction, validation, joint histogram,
semantic confusion, PQ matching/accumulation, compute() and per-image PQ.

Test infrastructure only (see oracle/__init__.py).  Restates
panoptic_evaluator/evaluator.py of the reference.
"""

from __future__ import annotations

import numpy as np

INT64_MAX = np.iinfo(np.int64).max

MSG_UNKNOWN = "map contains an unknown segment ID or a slot outside metadata capacity"
MSG_CLASSES = "classes must be -1 or valid category indices; slot 0 must be void"
MSG_ACTIVE = "nonvoid map slots and active metadata must correspond exactly"
MSG_ABSENT = "map contains a segment absent from segments_info"


def from_coco(maps, segments_info, category_ids):
    """Host metadata of PanopticBatch.from_coco (evaluator.py:33-106).

    Returns (ids int64 B x S, slots int32 B x S, classes int32 B x S,
    crowd bool B x S) with S = max segments + 1; ids sorted per image with
    slot 0 = id 0 and padding id INT64_MAX (evaluator.py:69-82).
    """
    if len(segments_info) != maps.shape[0] or len(set(category_ids)) != len(category_ids):
        raise ValueError("batch length or category IDs are invalid")
    lookup = {c: i for i, c in enumerate(category_ids)}
    capacity = max((len(info) for info in segments_info), default=0) + 1
    ids, slots, classes, crowd = [], [], [], []
    for info in segments_info:
        seg = [e["id"] for e in info]
        if any(int(i) <= 0 or int(i) >= INT64_MAX for i in seg) or len(set(seg)) != len(seg):
            raise ValueError("segment IDs must be unique positive int64 values below INT64_MAX; 0 is void")
        try:
            labels = [lookup[e["category_id"]] for e in info]
        except KeyError as error:
            raise ValueError("unknown category_id") from error
        pad = capacity - len(info) - 1
        ordered = sorted((int(i), s) for s, i in enumerate(seg, 1))
        ids.append([0] + [i for i, _ in ordered] + [INT64_MAX] * pad)
        slots.append([0] + [s for _, s in ordered] + [0] * pad)
        classes.append([-1] + labels + [-1] * pad)
        crowd.append([False] + [bool(e.get("iscrowd", 0)) for e in info] + [False] * pad)
    B = len(segments_info)
    shape = (B, capacity)
    return (np.array(ids, np.int64).reshape(shape), np.array(slots, np.int32).reshape(shape),
            np.array(classes, np.int32).reshape(shape), np.array(crowd, bool).reshape(shape))


def rgb_to_id(rgb):
    """Decoded COCO panoptic PNG (... x 3 uint8) -> segment IDs: R + 256 * G
    + 65536 * B (README.md:67-68)."""
    rgb = np.asarray(rgb).astype(np.int64)
    return rgb[..., 0] + 256 * rgb[..., 1] + 65536 * rgb[..., 2]


def id_to_rgb(ids):
    """Segment IDs (< 2**24) -> the RGB bytes a COCO panoptic PNG stores; the
    encoder of the reference's tests (test_evaluator.py:39)."""
    ids = np.asarray(ids).astype(np.int64)
    return np.stack((ids % 256, ids // 256 % 256, ids // 65536 % 256), -1).astype(np.uint8)


def compact(maps, lookup_ids, lookup_slots):
    """Original IDs -> compact slots (evaluator.py:147-157, CPU branch).

    searchsorted (left) in the sorted per-image id table, clamped to the last
    entry; a pixel is valid iff the id matches and (id == 0 or slot > 0).
    """
    packed = np.empty(maps.shape, np.int32)
    for b in range(maps.shape[0]):
        ids = lookup_ids[b]
        m = maps[b].astype(np.int64)
        pos = np.minimum(np.searchsorted(ids, m, side="left"), ids.size - 1)
        slots = lookup_slots[b][pos]
        valid = (ids[pos] == m) & ((m == 0) | (slots > 0))
        if not valid.all():
            raise ValueError(MSG_ABSENT)
        packed[b] = slots
    return packed


def validate_values(maps, classes, crowd, num_classes, *, allow_empty_slots=False):
    """Value checks of _prepare (evaluator.py:208-224)."""
    S = classes.shape[1]
    if ((maps < 0) | (maps >= S)).any():
        raise ValueError("segment map slot outside metadata capacity")
    if (((classes < -1) | (classes >= num_classes)).any() or (classes[:, 0] != -1).any()
            or (crowd is not None and crowd[:, 0].any())):
        raise ValueError(MSG_CLASSES)
    B = maps.shape[0]
    areas = np.zeros(classes.shape, np.int64)
    for b in range(B):
        areas[b] = np.bincount(maps[b].ravel().astype(np.int64), minlength=S)
    if allow_empty_slots:
        invalid = (areas[:, 1:] > 0) & (classes[:, 1:] < 0)
    else:
        invalid = (areas[:, 1:] > 0) != (classes[:, 1:] >= 0)
    if invalid.any():
        raise ValueError(MSG_ACTIVE)


def histogram(gt, pred, gc, pc, crowd, C, *, pq_only=False):
    """Joint slot histogram and semantic confusion (evaluator.py:228-243).

    pairs[b, g, p] = #pixels with GT slot g and predicted slot p;
    semantic[tc, pcls or C] over pixels whose GT slot is active and not crowd.
    """
    B = gt.shape[0]
    G, P = gc.shape[1], pc.shape[1]
    offsets = np.arange(B, dtype=np.int64)[:, None, None] * G * P
    keys = (offsets + gt.astype(np.int64) * P + pred.astype(np.int64)).ravel()
    pairs = np.bincount(keys, minlength=B * G * P).reshape(B, G, P).astype(np.int64)
    if pq_only:
        return pairs
    g_flat = gt.reshape(B, -1).astype(np.int64)
    p_flat = pred.reshape(B, -1).astype(np.int64)
    target = np.take_along_axis(gc, g_flat, 1)
    prediction = np.take_along_axis(pc, p_flat, 1)
    valid = (target >= 0) & ~np.take_along_axis(crowd, g_flat, 1)
    prediction = np.where(prediction >= 0, prediction, C)
    semantic = np.bincount((target.astype(np.int64) * (C + 1) + prediction)[valid],
                           minlength=C * (C + 1)).reshape(C, C + 1).astype(np.int64)
    return pairs, semantic


def _match_stats(pairs, gc, pc, crowd):
    """Shared PQ matching of _update_histogram / panoptic_quality
    (evaluator.py:383-404 and :522-534).

    Returns ga, pa, iou, matches, gm, pm, ignored_area.
    """
    ga, pa = pairs.sum(2), pairs.sum(1)
    union = ga[:, :, None] + pa[:, None, :] - pairs - pairs[:, :1, :]
    iou = pairs.astype(np.float64) / np.maximum(union, 1)
    matches = ((gc[:, :, None] == pc[:, None, :]) & (gc[:, :, None] >= 0)
               & ~crowd[:, :, None] & (iou > 0.5))
    gm, pm = matches.any(2), matches.any(1)
    G = gc.shape[1]
    slots = np.arange(G)[None, :, None]
    same_crowd = crowd[:, :, None] & (gc[:, :, None] == pc[:, None, :]) & (gc[:, :, None] >= 0)
    last = np.where(same_crowd, slots, -1).max(1) if G else np.full(pc.shape, -1)
    overlap = np.take_along_axis(pairs, np.maximum(last, 0)[:, None, :], 1)[:, 0, :]
    ignored_area = pairs[:, 0, :] + np.where(last >= 0, overlap, 0)
    return ga, pa, iou, matches, gm, pm, ignored_area


def _half(pa):
    # evaluator.py:403 and :533 compute ``0.5 * pa`` for an int64 tensor, which
    # torch promotes to float32; it is exact while pa < 2**24 (all supported
    # image sizes), so float64 gives the identical comparison.
    return 0.5 * pa.astype(np.float64)


class PanopticState:
    """Accumulators of PanopticEvaluator (evaluator.py:274-285)."""

    def __init__(self, num_classes, isthing=None):
        C = num_classes
        self.num_classes = C
        self.isthing = np.array(isthing if isthing is not None else [True] * C, bool)
        self.reset()

    def reset(self):
        C = self.num_classes
        self.iou_sum = np.zeros(C, np.float64)
        self.tp = np.zeros(C, np.int64)
        self.fp = np.zeros(C, np.int64)
        self.fn = np.zeros(C, np.int64)
        self.confusion = np.zeros((C, C + 1), np.int64)

    def update_histogram(self, pairs, semantic, gc, pc, crowd):
        """evaluator.py:381-417: accumulate sufficient statistics."""
        ga, pa, iou, matches, gm, pm, ignored_area = _match_stats(pairs, gc, pc, crowd)
        false_positive = (pc >= 0) & ~pm & (ignored_area.astype(np.float64) <= _half(pa))
        false_negative = (gc >= 0) & ~crowd & ~gm
        for labels, mask, output in ((pc, pm, self.tp), (pc, false_positive, self.fp),
                                     (gc, false_negative, self.fn)):
            np.add.at(output, np.maximum(labels, 0).ravel(), mask.ravel().astype(np.int64))
        # In-order left fold, as the reference's 1-D scatter_add_ (:413-415).
        np.add.at(self.iou_sum, np.maximum(pc, 0).ravel(), (iou * matches).sum(1).ravel())
        self.confusion += semantic
        return ga

    def update(self, gt, pred, gc, pc, crowd, *, validate=True):
        """PanopticEvaluator._update CPU branch (evaluator.py:291-379) on compact maps."""
        if crowd is None:
            crowd = np.zeros(gc.shape, bool)
        if validate:
            validate_inputs = (gt, gc, crowd), (pred, pc, None)
            for maps, classes, cr in validate_inputs:
                validate_values(maps, classes, cr, self.num_classes)
        pairs, semantic = histogram(gt, pred, gc, pc, crowd, self.num_classes)
        return self.update_histogram(pairs, semantic, gc, pc, crowd)

    def compute(self):
        """evaluator.py:448-468 (CPU formulas)."""
        C = self.num_classes
        tp, fp, fn, s = self.tp, self.fp, self.fn, self.iou_sum
        denom = tp.astype(np.float64) + 0.5 * (fp + fn).astype(np.float64)
        pq = s / np.maximum(denom, 1)
        sq = s / np.maximum(tp, 1)
        rq = tp / np.maximum(denom, 1)
        active = denom > 0

        def average(mask):
            selected = mask & active
            n = selected.sum()
            d = max(n, 1)
            return {"pq": (pq * selected).sum() / d, "sq": (sq * selected).sum() / d,
                    "rq": (rq * selected).sum() / d, "n": np.int64(n)}

        conf = self.confusion
        diagonal = np.diagonal(conf[:, :C]).copy()
        union = conf.sum(1) + conf[:, :C].sum(0) - diagonal
        with np.errstate(invalid="ignore", divide="ignore"):
            semantic_iou = np.where(union > 0, diagonal / np.maximum(union, 1), np.nan)
        miou = np.nan_to_num(semantic_iou, nan=0.0).sum() / max((union > 0).sum(), 1)
        return {"All": average(np.ones(C, bool)), "Things": average(self.isthing),
                "Stuff": average(~self.isthing),
                "per_class": {"pq": pq, "sq": sq, "rq": rq, "iou": semantic_iou},
                "miou": np.float64(miou), "tp": tp.copy(), "fp": fp.copy(), "fn": fn.copy(),
                "iou_sum": s.copy(), "confusion": conf.copy()}


def panoptic_quality(gt, pred, gc, pc, crowd, num_classes, *, validate=True, return_stats=False):
    """Per-image mean-class PQ (evaluator.py:471-544, CPU branch) on compact maps."""
    if crowd is None:
        crowd = np.zeros(gc.shape, bool)
    if validate:
        validate_values(gt, gc, crowd, num_classes)
        validate_values(pred, pc, None, num_classes, allow_empty_slots=True)
    B = gt.shape[0]
    C = num_classes
    tp = np.zeros((B, C), np.int64)
    fp = np.zeros_like(tp)
    fn = np.zeros_like(tp)
    sums = np.zeros((B, C), np.float64)
    pairs = histogram(gt, pred, gc, pc, crowd, C, pq_only=True)
    ga, pa, iou, matches, gm, pm, ignored_area = _match_stats(pairs, gc, pc, crowd)
    false_positive = (pc >= 0) & (pa > 0) & ~pm & (ignored_area.astype(np.float64) <= _half(pa))
    false_negative = (gc >= 0) & ~crowd & ~gm
    rows = np.arange(B)[:, None]
    for labels, mask, output in ((pc, pm, tp), (pc, false_positive, fp), (gc, false_negative, fn)):
        lab = np.maximum(labels, 0)
        np.add.at(output, (np.broadcast_to(rows, lab.shape).ravel(), lab.ravel()),
                  mask.ravel().astype(np.int64))
    lab = np.maximum(pc, 0)
    np.add.at(sums, (np.broadcast_to(rows, lab.shape).ravel(), lab.ravel()),
              (iou * matches).sum(1).ravel())
    if return_stats:
        return np.stack((sums, tp.astype(np.float64), fp.astype(np.float64), fn.astype(np.float64)), -1)
    denominator = tp.astype(np.float64) + 0.5 * (fp + fn).astype(np.float64)
    active = denominator > 0
    pq = sums / np.maximum(denominator, 1)
    return pq.sum(1) / np.maximum(active.sum(1), 1)
