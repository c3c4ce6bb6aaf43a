"""The combined evaluator's grouped matching (class lists built beside the
detection histogram) and its deferred form (matching inside compute()'s
launch, in a whole-step graph), on inputs that reach their edge paths:

* more than 100 detections of one class in an image (ranks past 100 are 0,
  their flags -1, and their rows are still zeroed for the next update);
* detections of stuff classes (no candidates: every one unmatched);
* crowd ground truth and annotation areas on the range edges;
* more target slots than the matcher stages in LDS (pairs read from HBM);
* repeated replays, with the detection x slot table reused between them.

Every result equals the oracle (integer statistics and AP/AR arrays exact,
FP64 values within 1e-12 relative) and the eager one-part matcher exactly."""

import numpy as np
import pytest
import torch

from helpers import assert_close, assert_exact, compare_panoptic, to_numpy
from oracle import combined as ocomb

pytestmark = pytest.mark.gpu

SUMMARY = ("ap", "ap50", "ap75", "ap_small", "ap_medium", "ap_large", "ar1", "ar10", "ar100", "ar_small",
           "ar_medium", "ar_large")


@pytest.fixture(scope="module")
def pe():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import panoptic_evaluator
    return panoptic_evaluator


def cuda(x):
    return torch.as_tensor(np.asarray(x)).cuda()


def compare_instance(actual, expected, exact=False):
    assert_exact(actual["precision"], expected["precision"], "precision")
    assert_exact(actual["recall"], expected["recall"], "recall")
    for key in SUMMARY:
        if exact:
            assert_exact(actual[key], expected[key], key)
        else:
            assert_close(actual[key], expected[key], what=key)


def make_case(seed, B=2, H=48, W=64, segments=40, dets=200, C=6, things=(True, True, True, False, True, False)):
    """Maps with `segments` slots per image (classes cycling over C, crowd on
    some thing slots), predictions shifted by a pixel, and `dets` detections
    per image: most of class 0 (past 100 of one class), some of a stuff
    class, the rest spread; masks are slot unions and random blobs."""
    rng = np.random.default_rng(seed)
    S = segments + 1
    # blocky segments: a coarse map repeated 4 x 4
    coarse = rng.integers(1, S, (B, H // 4, W // 4)).astype(np.int32)
    maps = np.repeat(np.repeat(coarse, 4, axis=1), 4, axis=2)
    maps[:, :2, :] = 0
    classes = np.full((B, S), -1, np.int32)
    classes[:, 1:] = np.arange(segments) % C
    present = np.zeros((B, S), bool)
    for b in range(B):
        present[b, np.unique(maps[b])] = True
    classes[~present] = -1
    classes[:, 0] = -1
    crowd = np.zeros((B, S), bool)
    thing_slot = np.isin(classes, [c for c in range(C) if things[c]])
    crowd[:, 1:] = (rng.random((B, S - 1)) < 0.15) & thing_slot[:, 1:]
    pred_maps = np.roll(maps, 1, axis=2).copy()
    pred_classes = np.full((B, S), -1, np.int32)
    for b in range(B):
        for s in np.unique(pred_maps[b]):
            if s:
                pred_classes[b, s] = classes[b, s] if classes[b, s] >= 0 else 0
    masks, labels, scores = [], [], []
    stuff = [c for c in range(C) if not things[c]]
    for b in range(B):
        m = np.zeros((dets, H, W), bool)
        lab = np.zeros(dets, np.int64)
        for d in range(dets):
            if d % 3 == 0:
                slot = rng.integers(1, S)
                m[d] = maps[b] == slot
                lab[d] = max(classes[b, slot], 0)
            else:
                y, x = rng.integers(0, H - 8), rng.integers(0, W - 8)
                m[d, y:y + rng.integers(2, 9), x:x + rng.integers(2, 9)] = True
                lab[d] = 0 if d % 7 else rng.integers(0, C)
            if d % 17 == 0 and stuff:
                lab[d] = stuff[d % len(stuff)]
            if not m[d].any():
                m[d, rng.integers(0, H), rng.integers(0, W)] = True
        masks.append(m)
        labels.append(lab)
        scores.append(rng.choice([0.25, 0.5, 0.75, 0.9], dets) + rng.integers(0, 3, dets) * 1e-3)
    areas = rng.choice([0.0, 100.0, 1024.0, 1025.0, 9216.0, 9217.0], (B, S))
    return maps, classes, crowd, pred_maps, pred_classes, masks, labels, scores, areas, C, things


@pytest.mark.parametrize("segments,ann", [(40, False), (40, True), (150, False), (150, True)])
def test_grouped_and_deferred_match_oracle(pe, segments, ann):
    from panoptic_evaluator.graphs import CapturedStep, CapturedUpdate
    (maps, classes, crowd, pred_maps, pred_classes, masks, labels, scores, areas, C,
     things) = make_case(7 + segments, segments=segments)
    assert max((lab == 0).sum() for lab in labels) > 100
    inst = areas if ann else None
    state = ocomb.SegmentationState(C, list(things), validate=False)
    odets = [dict(masks=m, classes=x, scores=s) for m, x, s in zip(masks, labels, scores)]
    state.update(pred_maps, pred_classes, odets, maps, classes, crowd, instance_areas=inst)
    expected = state.compute()
    args = (pe.PanopticBatch(cuda(pred_maps), cuda(pred_classes)),
            pe.InstanceBatch([cuda(m) for m in masks], [cuda(x) for x in labels], [cuda(s) for s in scores]),
            pe.PanopticBatch(cuda(maps), cuda(classes), cuda(crowd)))
    kwargs = {} if inst is None else dict(instance_areas=cuda(inst))
    make = lambda: pe.SegmentationEvaluator(C, isthing=list(things), validate=False)  # noqa: E731
    eager = make()
    eager.update(*args, **kwargs)
    want = to_numpy(eager.compute())
    compare_panoptic(want, expected)
    compare_instance(want["instance"], expected["instance"])
    # Update graph: grouped matching at the end of every replay.
    ev = make()
    captured = CapturedUpdate(ev, *args, **kwargs)
    for _ in range(3):
        ev.reset()
        captured.replay()
        got = to_numpy(ev.compute())
        compare_panoptic(got, want, exact_iou_sum=True)
        compare_instance(got["instance"], want["instance"], exact=True)
    # Two replays without reset: the first replay's rows were zeroed.
    ev.reset()
    captured.replay()
    captured.replay()
    twice = make()
    twice.update(*args, **kwargs)
    twice.update(*args, **kwargs)
    got = to_numpy(ev.compute())
    twice_result = to_numpy(twice.compute())
    compare_panoptic(got, twice_result, exact_iou_sum=True)
    compare_instance(got["instance"], twice_result["instance"], exact=True)
    # Step graph: matching and summaries inside compute()'s launch.
    stepped = make()
    step = CapturedStep(stepped, *args, **kwargs)
    for _ in range(3):
        got = to_numpy(step.replay())
        compare_panoptic(got, want, exact_iou_sum=True)
        compare_instance(got["instance"], want["instance"], exact=True)
    # The evaluator's own compute() after a replay (eager path, same state).
    got = to_numpy(stepped.compute())
    compare_instance(got["instance"], want["instance"], exact=True)
