"""GPU parity of the instance and combined paths (InstanceEvaluator,
SegmentationEvaluator, the int8 MFMA mask contraction, the detection x slot
histogram and the COCO matcher) against the CPU oracle and golden vectors.

Restates the reference's tests/test_instance.py and tests/test_combined.py
behaviours: integer outputs (intersections, match flags, ranks, counts) are
bit-exact; precision/recall arrays are compared exactly and the summaries to
1e-12 relative.
"""

import numpy as np
import pytest
import torch

from golden_data import group, load
from helpers import assert_close, assert_exact, compare_panoptic, to_numpy
from oracle import combined as ocomb
from oracle import instance as oinst

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


def compare_instance(actual, expected):
    assert_exact(actual["precision"], expected["precision"], "precision")
    assert_exact(actual["recall"], expected["recall"], "recall")
    for key in SUMMARY:
        assert_close(actual[key], expected[key], what=key)


# ---------------------------------------------------------- MFMA contraction

@pytest.mark.parametrize("D,G,H,W", [(1, 1, 4, 4), (17, 33, 24, 37), (100, 48, 64, 96), (5, 130, 32, 64),
                                     (0, 7, 8, 8), (9, 0, 8, 8), (40, 20, 800, 1088)])
def test_mask_overlap_mfma_exact(pe, D, G, H, W):
    from panoptic_evaluator import _native
    rng = np.random.default_rng(D * 1000 + G)
    pm = rng.random((D, H, W)) < 0.3
    gm = rng.random((G, H, W)) < 0.4
    if D and G:
        gm[0] = True  # a full mask: intersection == detection area
    inter = torch.empty(D * G, dtype=torch.int32, device="cuda")
    pa = torch.empty(D, dtype=torch.int32, device="cuda")
    ga = torch.empty(G, dtype=torch.int32, device="cuda")
    pm_dev, gm_dev = cuda(pm), cuda(gm)  # keep the device copies alive during the launch
    _native.call("pe_mask_overlap", pm_dev.data_ptr(), D, gm_dev.data_ptr(), G, H * W, inter.data_ptr(),
                 pa.data_ptr(), ga.data_ptr(), _native.stream())
    pflat = pm.reshape(D, H * W).astype(np.int64)
    gflat = gm.reshape(G, H * W).astype(np.int64)
    assert_exact(inter.view(D, G).long(), pflat @ gflat.T)
    assert_exact(pa.long(), pflat.sum(1))
    assert_exact(ga.long(), gflat.sum(1))


# ------------------------------------------------------------ golden vectors

def test_golden_coco_matcher(pe):
    """Direct matcher parity on the reference's adversarial cases (ties at
    exactly 0.5, groups wider than a wavefront, the 100-detection cap,
    eligible-vs-ignored preference, reusable crowds)."""
    from panoptic_evaluator import _native
    data = load("coco_matcher")
    for index in range(int(data["num_cases"])):
        inp = group(data, f"in{index}")
        D, G, C = len(inp["pc"]), len(inp["gc"]), int(inp["C"])
        status = torch.empty((D, 40), dtype=torch.int8, device="cuda")
        ranks = torch.empty(D, dtype=torch.int32, device="cuda")
        counts = torch.zeros((C, 4), dtype=torch.int64, device="cuda")
        off = lambda n: torch.tensor([0, n], dtype=torch.int64, device="cuda")  # noqa: E731
        det_off, gt_off, inter_off = off(D), off(G), off(0)  # alive for the launch
        pairs = cuda(inp["pairs"]).long().contiguous()
        args = [cuda(inp["pc"]).int(), cuda(inp["pa"]).double(), cuda(inp["gc"]).int(), cuda(inp["ga"]).double(),
                cuda(inp["crowd"]), cuda(inp["areas"]).double()]
        _native.call("pe_coco_match", 1, C, det_off.data_ptr(), gt_off.data_ptr(), inter_off.data_ptr(),
                     args[0].data_ptr(), args[1].data_ptr(), None, args[2].data_ptr(), args[3].data_ptr(),
                     args[4].data_ptr(), args[5].data_ptr(), pairs.data_ptr(), 8, D, G, status.data_ptr(),
                     ranks.data_ptr(), counts.data_ptr(), None, None, None, 0, 0, _native.stream())
        assert_exact(status.t(), data[f"out{index}.status"], f"status case {index}")
        assert_exact(ranks, data[f"out{index}.ranks"], f"ranks case {index}")
        assert_exact(counts, data[f"out{index}.counts"], f"counts case {index}")


def test_golden_instance_evaluator(pe):
    data = load("instance_eval")
    ev = pe.InstanceEvaluator(4, device="cuda")
    for image in range(int(data["num_images"])):
        inp = group(data, f"in{image}")
        ev.update(pe.InstanceBatch([cuda(inp["pm"])], [cuda(inp["pc"])], [cuda(inp["scores"])]),
                  pe.InstanceBatch([cuda(inp["gm"])], [cuda(inp["gc"])], crowd=[cuda(inp["crowd"])],
                                   areas=[cuda(inp["areas"])]))
    result = to_numpy(ev.compute())
    out = group(data, "out")
    assert_exact(result["precision"], out["precision"])
    assert_exact(result["recall"], out["recall"])
    for key in SUMMARY:
        assert_close(result[key], out[key], what=key)


def test_golden_combined(pe):
    data = load("combined")
    inp = group(data, "in")
    joint = pe.SegmentationEvaluator(3, isthing=[True, True, False])
    dets = pe.InstanceBatch([cuda(inp[f"masks{b}"]) for b in range(2)], [cuda(inp[f"labels{b}"]) for b in range(2)],
                            [cuda(inp[f"scores{b}"]) for b in range(2)])
    for areas in (None, cuda(inp["areas"])):
        joint.update(pe.PanopticBatch(cuda(inp["pred_maps"]), cuda(inp["classes"])), dets,
                     pe.PanopticBatch(cuda(inp["maps"]), cuda(inp["classes"]), cuda(inp["crowd"])),
                     instance_areas=areas)
    result = to_numpy(joint.compute())
    out = group(data, "out")
    for key in ("tp", "fp", "fn", "confusion"):
        assert_exact(result[key], out[key], key)
    assert_close(result["iou_sum"], out["iou_sum"], what="iou_sum")
    assert_exact(result["instance"]["precision"], out["instance.precision"])
    assert_exact(result["instance"]["recall"], out["instance.recall"])
    for key in SUMMARY:
        assert_close(result["instance"][key], out[f"instance.{key}"], what=key)


# ------------------------------------------------- reference behaviours (restated)

def instance_batch(pe, scores=None, crowd=None):
    masks = torch.zeros((2, 16, 16), dtype=torch.bool, device="cuda")
    masks[0, :10, :10] = True
    masks[1, 5:15, 5:15] = True
    return pe.InstanceBatch([masks], [torch.zeros(2, dtype=torch.int64, device="cuda")],
                            scores=[torch.tensor(scores, device="cuda")] if scores is not None else None,
                            crowd=[torch.tensor(crowd, device="cuda")] if crowd is not None else None)


def test_instance_known_answers(pe):  # test_instance.py:27-69
    ev = pe.InstanceEvaluator(2, device="cuda")
    ev.update(instance_batch(pe, [0.9, 0.8]), instance_batch(pe))
    result = ev.compute()
    assert abs(result["ap"].item() - 1.0) < 1e-12 and result["precision"].shape == (10, 101, 2, 4, 3)
    assert (result["precision"][:, :, 1] == -1).all()
    ev.reset()
    assert ev.compute()["ap"].item() == -1.0 and abs(result["ap"].item() - 1.0) < 1e-12
    empty = pe.InstanceBatch([torch.empty((0, 16, 16), dtype=torch.bool, device="cuda")],
                             [torch.empty(0, dtype=torch.int64, device="cuda")], [torch.empty(0, device="cuda")])
    ev = pe.InstanceEvaluator(1, device="cuda")
    ev.update(empty, instance_batch(pe))
    assert ev.compute()["ap"].item() == 0.0
    ev = pe.InstanceEvaluator(1, device="cuda")
    ev.update(instance_batch(pe, [0.9, 0.8]), instance_batch(pe, crowd=[True, False]))
    ev.update(instance_batch(pe, [0.9, 0.8]), instance_batch(pe))
    assert abs(ev.compute()["ap"].item() - 1.0) < 1e-12
    ev = pe.InstanceEvaluator(1, device="cuda")
    with pytest.raises(ValueError):
        ev.update(instance_batch(pe), instance_batch(pe))
    assert ev.compute()["ap"].item() == -1.0
    target = instance_batch(pe)
    target = pe.InstanceBatch([target.masks[0][:1]], [target.classes[0][:1]])
    ev = pe.InstanceEvaluator(1, device="cuda")
    ev.update(instance_batch(pe, [0.8, 0.9]), target)
    assert abs(ev.compute()["ap"].item() - 0.5) < 1e-12


def test_max_detections_per_category(pe):  # test_instance.py:179-197
    masks = torch.ones((106, 1, 1), dtype=torch.bool, device="cuda")
    labels = torch.zeros(106, dtype=torch.int64, device="cuda")
    labels[-1] = 1
    ev = pe.InstanceEvaluator(2, device="cuda")
    ev.update(pe.InstanceBatch([masks], [labels], [torch.ones(106, device="cuda")]), pe.InstanceBatch([masks], [labels]))
    r = ev.compute()
    assert abs(r["ar1"].item() - (1 / 105 + 1) / 2) < 1e-12
    assert abs(r["ar10"].item() - (10 / 105 + 1) / 2) < 1e-12
    assert abs(r["ar100"].item() - (100 / 105 + 1) / 2) < 1e-12


def test_exact_iou_threshold_and_empty_target(pe):  # test_instance.py:199-218
    pm = torch.tensor([[[True, False]]], device="cuda")
    gm = torch.tensor([[[True, True]]], device="cuda")
    label = torch.tensor([0], device="cuda")
    pred = pe.InstanceBatch([pm], [label], [torch.tensor([0.9], device="cuda")])
    ev = pe.InstanceEvaluator(1, device="cuda")
    ev.update(pred, pe.InstanceBatch([gm], [label]))
    r = ev.compute()
    assert abs(r["ap50"].item() - 1.0) < 1e-12 and abs(r["ap"].item() - 0.1) < 1e-12
    ev.reset()
    ev.update(pred, pe.InstanceBatch([gm[:0]], [label[:0]]))
    assert ev.compute()["ap"].item() == -1.0
    ev.reset()
    ev.update(pe.InstanceBatch([pm[:0]], [label[:0]], [torch.empty(0, device="cuda")]), pe.InstanceBatch([gm], [label]))
    assert ev.compute()["ap"].item() == 0.0


def test_nondefault_stream_snapshots_and_validation(pe):  # test_instance.py:220-237
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        prediction = instance_batch(pe, [0.9, 0.8])
        target = instance_batch(pe)
        ev = pe.InstanceEvaluator(1, device="cuda", validate=False)
        ev.update(prediction, target)
        prediction.masks[0].zero_()
        prediction.scores[0].zero_()
        result = ev.compute()
        ev.reset()
    stream.synchronize()
    assert abs(result["ap"].item() - 1.0) < 1e-12
    ev = pe.InstanceEvaluator(1, device="cuda")
    with pytest.raises(ValueError):
        ev.update(prediction, target)
    assert ev.compute()["ap"].item() == -1.0


def test_randomized_instance_parity(pe):
    """Random overlapping masks with duplicates, exact ties, crowds and area
    edges (the COCOeval parity inputs of test_instance.py:80-177) against the
    oracle, over several updates."""
    rng = np.random.default_rng(47)
    ev = pe.InstanceEvaluator(4, device="cuda")
    state = oinst.InstanceState(4)
    for _ in range(12):
        h, w = 24, 37
        gm = rng.random((6, h, w)) > 0.65
        pm = np.concatenate([gm[:4].copy(), rng.random((5, h, w)) > 0.65])
        pm[1] &= rng.random((h, w)) > 0.2
        pm[2] = gm[0]
        gcl = rng.integers(0, 3, 6)
        pcl = np.concatenate([gcl[:4], rng.integers(0, 3, 5)])
        pcl[2] = gcl[0]
        scores = rng.choice([0.1, 0.5, 0.9], 9)
        crowd = rng.random(6) < 0.3
        areas = rng.choice([100.0, 1024.0, 1025.0, 9216.0, 9217.0], 6)
        ev.update(pe.InstanceBatch([cuda(pm)], [cuda(pcl)], [cuda(scores)]),
                  pe.InstanceBatch([cuda(gm)], [cuda(gcl)], crowd=[cuda(crowd)], areas=[cuda(areas)]))
        state.update([dict(masks=pm, classes=pcl, scores=scores)],
                     [dict(masks=gm, classes=gcl, crowd=crowd, areas=areas)])
    compare_instance(to_numpy(ev.compute()), state.compute())


def combined_fixture():
    gen = np.random.default_rng(91)
    maps = gen.integers(0, 6, (2, 13, 35)).astype(np.int32)
    classes = np.array([[-1, 0, 1, 0, 2, -1], [-1, 1, 0, 2, 0, 1]], np.int32)
    maps[0][maps[0] == 5] = 0
    crowd = np.array([[False, False, True, False, False, False], [False, False, False, False, True, False]])
    pred_maps = maps.copy()
    pred_maps[:, :3] = 0
    masks = [gen.random((8, 13, 35)) > 0.7 for _ in range(2)]
    for b in range(2):
        masks[b][:3] = maps[b][None] == np.array([1, 2, 3])[:, None, None]
    masks[0][3] = True
    labels = [np.array([0, 1, 0, 0, 1, 0, 1, 0]), np.array([1, 0, 0, 1, 0, 1, 0, 1])]
    scores = [np.array([0.9, 0.9, 0.8, 0.5, 0.5, 0.5, 0.1, 0.1]) for _ in range(2)]
    return maps, classes, crowd, pred_maps, masks, labels, scores


@pytest.mark.parametrize("validate", [True, False])
def test_combined_equals_oracle(pe, validate):  # test_combined.py:93-120
    maps, classes, crowd, pred_maps, masks, labels, scores = combined_fixture()
    areas = np.array([[0.0, 1024.0, 1025.0, 9216.0, 9217.0, 0.0], [0.0, 100.0, 1024.0, 200.0, 9217.0, 9216.0]])
    joint = pe.SegmentationEvaluator(3, isthing=[True, True, False], validate=validate)
    state = ocomb.SegmentationState(3, [True, True, False], validate=validate)
    dets = pe.InstanceBatch([cuda(m) for m in masks], [cuda(x) for x in labels], [cuda(s) for s in scores])
    odets = [dict(masks=m, classes=x, scores=s) for m, x, s in zip(masks, labels, scores)]
    for ann in (None, areas):
        joint.update(pe.PanopticBatch(cuda(pred_maps), cuda(classes)), dets,
                     pe.PanopticBatch(cuda(maps), cuda(classes), cuda(crowd)),
                     instance_areas=None if ann is None else cuda(ann))
        state.update(pred_maps, classes, odets, maps, classes, crowd, instance_areas=ann)
    result, expected = to_numpy(joint.compute()), state.compute()
    compare_panoptic(result, expected)
    compare_instance(result["instance"], expected["instance"])
    snapshot = joint.compute()
    joint.reset()
    assert joint.compute()["instance"]["ap"].item() == -1.0
    compare_instance(to_numpy(snapshot["instance"]), expected["instance"])


def test_combined_failed_validation_preserves_all_metrics(pe):  # test_combined.py:155-178
    maps, classes, crowd, pred_maps, masks, labels, scores = combined_fixture()
    p = pe.PanopticBatch(cuda(pred_maps), cuda(classes))
    g = pe.PanopticBatch(cuda(maps), cuda(classes), cuda(crowd))
    d = pe.InstanceBatch([cuda(m) for m in masks], [cuda(x) for x in labels], [cuda(s) for s in scores])
    joint = pe.SegmentationEvaluator(3, isthing=[True, True, False])
    joint.update(p, d, g)
    before = to_numpy(joint.compute())
    bad_labels = [x.clone() for x in d.classes]
    bad_labels[1][-1] = 2
    with pytest.raises(ValueError, match="thing"):
        joint.update(p, pe.InstanceBatch(d.masks, bad_labels, d.scores), g)
    bad_masks = [x.clone() for x in d.masks]
    bad_masks[1][-1].zero_()
    with pytest.raises(ValueError, match="pixels"):
        joint.update(p, pe.InstanceBatch(bad_masks, d.classes, d.scores), g)
    with pytest.raises(ValueError, match="instance_areas"):
        joint.update(p, d, g, instance_areas=torch.full(g.classes.shape, float("nan"), device="cuda"))
    after = to_numpy(joint.compute())
    compare_panoptic(after, before)
    compare_instance(after["instance"], before["instance"])


def test_combined_original_ids_and_empty_detections(pe):  # test_combined.py:123-150
    maps, classes, crowd, pred_maps, masks, labels, scores = combined_fixture()
    info = [[dict(id=int(s) * 10000000001, category_id=int(classes[b, s]), iscrowd=int(crowd[b, s]))
             for s in range(1, classes.shape[1]) if classes[b, s] >= 0] for b in range(2)]
    target = pe.PanopticBatch.from_coco(cuda(maps).long() * 10000000001, info, [0, 1, 2], compact=False)
    empty = pe.InstanceBatch([torch.empty((0, 13, 35), dtype=torch.bool, device="cuda") for _ in range(2)],
                             [torch.empty(0, dtype=torch.int32, device="cuda") for _ in range(2)],
                             [torch.empty(0, device="cuda") for _ in range(2)])
    joint = pe.SegmentationEvaluator(3, isthing=[True, True, False])
    joint.update(pe.PanopticBatch(cuda(pred_maps), cuda(classes)), empty, target)
    state = ocomb.SegmentationState(3, [True, True, False])
    state.update(pred_maps, classes, [dict(masks=np.zeros((0, 13, 35), bool), classes=np.zeros(0, int),
                                           scores=np.zeros(0))] * 2, maps, classes, crowd)
    result, expected = to_numpy(joint.compute()), state.compute()
    compare_panoptic(result, expected)
    compare_instance(result["instance"], expected["instance"])


def test_combined_nondefault_stream(pe):  # test_combined.py:180-206
    maps, classes, crowd, pred_maps, masks, labels, scores = combined_fixture()
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        joint = pe.SegmentationEvaluator(3, isthing=[True, True, False], validate=False)
        g = pe.PanopticBatch(cuda(maps), cuda(classes), cuda(crowd))
        d = pe.InstanceBatch([cuda(m) for m in masks], [cuda(x) for x in labels], [cuda(s) for s in scores])
        joint.update(pe.PanopticBatch(cuda(pred_maps), cuda(classes)), d, g)
        actual = joint.compute()
        g.segment_map.zero_()
        for x in list(d.masks) + list(d.scores):
            x.zero_()
        joint.reset()
    stream.synchronize()
    state = ocomb.SegmentationState(3, [True, True, False], validate=False)
    state.update(pred_maps, classes, [dict(masks=m, classes=x, scores=s) for m, x, s in zip(masks, labels, scores)],
                 maps, classes, crowd)
    expected = state.compute()
    result = to_numpy(actual)
    compare_panoptic(result, expected)
    compare_instance(result["instance"], expected["instance"])


def test_config1_parity_subset(pe):
    """c1 image size and detection count (800 x 1088, 100 ellipses) on 2 of the
    16 images: every statistic against the oracle."""
    import workloads
    pair, masks, classes, scores = workloads.config1(device="cpu", B=2)
    isthing = workloads.isthing_vector(133, 80)
    joint = pe.SegmentationEvaluator(133, isthing=isthing)
    joint.update(pe.PanopticBatch(pair.pred.cuda(), pair.pc.cuda()),
                 pe.InstanceBatch([m.cuda() for m in masks], [c.cuda() for c in classes], [s.cuda() for s in scores]),
                 pe.PanopticBatch(pair.gt.cuda(), pair.gc.cuda(), pair.crowd.cuda()))
    state = ocomb.SegmentationState(133, isthing)
    state.update(pair.pred.numpy(), pair.pc.numpy(),
                 [dict(masks=m.numpy(), classes=c.numpy(), scores=s.numpy()) for m, c, s in zip(masks, classes, scores)],
                 pair.gt.numpy(), pair.gc.numpy(), pair.crowd.numpy())
    result, expected = to_numpy(joint.compute()), state.compute()
    compare_panoptic(result, expected)
    compare_instance(result["instance"], expected["instance"])


@pytest.mark.parametrize("N,C", [(3000, 7), (4000, 2), (9000, 7)])
def test_ap_accumulation_paths(pe, N, C):
    """Synthetic match records (tied scores, ranks 0..100, all flag values) on
    both AP paths: one-launch per-class LDS sort (N <= 4096; with C=2 a class
    holds more records than one LDS chunk) and the two global radix sorts
    (larger N)."""
    from panoptic_evaluator.instance import accumulate
    rng = np.random.default_rng(N)
    scores = rng.integers(0, 50, N).astype(np.float64) / 50.0   # many ties
    classes = rng.integers(0, C, N).astype(np.int32)
    ranks = rng.integers(0, 101, N).astype(np.int32)
    status = rng.integers(-1, 2, (N, 40)).astype(np.int8)
    counts = rng.integers(0, 40, (C, 4)).astype(np.int64)
    if C > 2:
        counts[2] = 0                                            # a class without positives
    half = N // 2                                                # two update records
    records = [(cuda(scores[:half]), cuda(classes[:half]), cuda(ranks[:half]), cuda(status[:half])),
               (cuda(scores[half:]), cuda(classes[half:]), cuda(ranks[half:]), cuda(status[half:]))]
    got = accumulate(records, cuda(counts), C, torch.device("cuda", 0))
    order = oinst.score_order(scores)
    precision, recall = oinst.accumulate(classes[order], ranks[order], status[order].T, counts)
    expected = oinst.summarize(precision, recall)
    assert_exact(got["precision"], precision, "precision")
    assert_exact(got["recall"], recall, "recall")
    for key in SUMMARY:
        assert_close(got[key], expected[key], what=key)


def test_instance_batch_of_uneven_images(pe):
    """One update over images of different sizes (odd, not multiples of 16),
    with an image without detections and one without ground truth: the single
    batched MFMA launch against the oracle."""
    rng = np.random.default_rng(5)
    sizes = [(24, 37), (40, 64), (17, 29), (33, 50)]
    dets, gts = [9, 0, 21, 6], [6, 4, 0, 11]
    pm, gm, pcl, gcl, sc, cr = [], [], [], [], [], []
    for (h, w), d, g in zip(sizes, dets, gts):
        gmask = rng.random((g, h, w)) > 0.6
        pmask = rng.random((d, h, w)) > 0.6
        if d and g:
            pmask[0] = gmask[0]
        gm.append(gmask)
        pm.append(pmask)
        gcl.append(rng.integers(0, 3, g))
        pcl.append(rng.integers(0, 3, d))
        sc.append(rng.choice([0.2, 0.6, 0.9], d))
        cr.append(rng.random(g) < 0.2)
    ev = pe.InstanceEvaluator(3, device="cuda")
    ev.update(pe.InstanceBatch([cuda(m) for m in pm], [cuda(c) for c in pcl], [cuda(s) for s in sc]),
              pe.InstanceBatch([cuda(m) for m in gm], [cuda(c) for c in gcl], crowd=[cuda(c) for c in cr]))
    state = oinst.InstanceState(3)
    state.update([dict(masks=m, classes=c, scores=s) for m, c, s in zip(pm, pcl, sc)],
                 [dict(masks=m, classes=c, crowd=k) for m, c, k in zip(gm, gcl, cr)])
    compare_instance(to_numpy(ev.compute()), state.compute())


def test_more_than_100_detections_of_one_class(pe):
    """The matcher ranks only the first 100 detections of a class (in score
    order); the rest get rank 0 and flags -1 from the same kernel."""
    rng = np.random.default_rng(11)
    h, w, d, g = 20, 24, 130, 9
    pm = rng.random((d, h, w)) > 0.7
    gm = rng.random((g, h, w)) > 0.6
    pm[:g] = gm                                  # exact matches among the detections
    pcl = np.where(np.arange(d) < 120, 1, 0)     # 120 detections of class 1, 10 of class 0
    gcl = rng.integers(0, 2, g)
    sc = rng.random(d)
    ev = pe.InstanceEvaluator(2, device="cuda")
    ev.update(pe.InstanceBatch([cuda(pm)], [cuda(pcl)], [cuda(sc)]),
              pe.InstanceBatch([cuda(gm)], [cuda(gcl)]))
    state = oinst.InstanceState(2)
    state.update([dict(masks=pm, classes=pcl, scores=sc)], [dict(masks=gm, classes=gcl)])
    compare_instance(to_numpy(ev.compute()), state.compute())


def test_score_order_input_dtypes(pe):
    """float64 or float32 scores and int32 or int64 classes (also mixed across
    images, and a single image passed without a copy) give the oracle's
    results."""
    rng = np.random.default_rng(12)
    sizes, dets, gts = [(16, 21), (16, 21), (16, 21)], [14, 0, 9], [5, 3, 6]
    pm = [rng.random((d, 16, 21)) > 0.65 for d in dets]
    gm = [rng.random((g, 16, 21)) > 0.6 for g in gts]
    pcl = [rng.integers(0, 3, d) for d in dets]
    gcl = [rng.integers(0, 3, g) for g in gts]
    sc = [rng.integers(0, 8, d) / 8.0 for d in dets]  # ties; exact in float32
    state = oinst.InstanceState(3)
    state.update([dict(masks=m, classes=c, scores=s) for m, c, s in zip(pm, pcl, sc)],
                 [dict(masks=m, classes=c) for m, c in zip(gm, gcl)])
    expected = state.compute()
    targets = pe.InstanceBatch([cuda(m) for m in gm], [cuda(c) for c in gcl])
    forms = {
        "int64 classes": ([cuda(c) for c in pcl], [cuda(s) for s in sc]),
        "int32 classes": ([cuda(c).int() for c in pcl], [cuda(s) for s in sc]),
        "float32 scores": ([cuda(c) for c in pcl], [cuda(s).float() for s in sc]),
        "mixed class dtypes": ([cuda(c).int() if i == 0 else cuda(c) for i, c in enumerate(pcl)],
                               [cuda(s) for s in sc]),
    }
    for name, (classes, scores) in forms.items():
        ev = pe.InstanceEvaluator(3, device="cuda")
        ev.update(pe.InstanceBatch([cuda(m) for m in pm], classes, scores), targets)
        compare_instance(to_numpy(ev.compute()), expected)
        # The same images one update at a time (single-image batches).
        ev = pe.InstanceEvaluator(3, device="cuda")
        for b in range(len(pm)):
            ev.update(pe.InstanceBatch([cuda(pm[b])], [classes[b]], [scores[b]]),
                      pe.InstanceBatch([targets.masks[b]], [targets.classes[b]]))
        compare_instance(to_numpy(ev.compute()), expected)


def test_match_tables_gt_dtypes_and_image_counts(pe):
    """The matcher's GT tables (areas, classes, crowd flags) built in one launch
    from each image's arrays in place, or by concatenation when the batch has
    mixed GT class dtypes or more than 64 images: every form gives the oracle's
    results."""
    rng = np.random.default_rng(13)
    B = 70
    dets = rng.integers(0, 6, B)
    gts = rng.integers(0, 5, B)
    pm = [rng.random((d, 12, 15)) > 0.6 for d in dets]
    gm = [rng.random((g, 12, 15)) > 0.55 for g in gts]
    for p, g in zip(pm, gm):
        if len(p) and len(g):
            p[0] = g[0]
    pcl = [rng.integers(0, 3, d) for d in dets]
    gcl = [rng.integers(0, 3, g) for g in gts]
    sc = [rng.random(d) for d in dets]
    cr = [rng.random(g) < 0.2 for g in gts]
    for images in (5, B):
        state = oinst.InstanceState(3)
        state.update([dict(masks=m, classes=c, scores=s) for m, c, s in zip(pm[:images], pcl[:images], sc[:images])],
                     [dict(masks=m, classes=c, crowd=k) for m, c, k in zip(gm[:images], gcl[:images], cr[:images])])
        expected = state.compute()
        forms = {
            "int64": [cuda(c) for c in gcl[:images]],
            "int32": [cuda(c).int() for c in gcl[:images]],
            "mixed": [cuda(c).int() if i % 2 else cuda(c) for i, c in enumerate(gcl[:images])],
        }
        for name, classes in forms.items():
            ev = pe.InstanceEvaluator(3, device="cuda")
            ev.update(pe.InstanceBatch([cuda(m) for m in pm[:images]], [cuda(c) for c in pcl[:images]],
                                       [cuda(s) for s in sc[:images]]),
                      pe.InstanceBatch([cuda(m) for m in gm[:images]], classes,
                                       crowd=[cuda(k) for k in cr[:images]]))
            compare_instance(to_numpy(ev.compute()), expected)


def test_matcher_with_many_pairs_of_one_class(pe):
    """40 detections x 20 ground-truth masks of one class exceed the matcher's
    LDS staging (512 pairs), so its chains read global memory; a second image
    stays within the staging.  Both against the oracle."""
    rng = np.random.default_rng(13)
    h, w = 18, 22
    pm = [rng.random((40, h, w)) > 0.6, rng.random((6, h, w)) > 0.6]
    gm = [rng.random((20, h, w)) > 0.55, rng.random((4, h, w)) > 0.55]
    pm[0][:20] = gm[0]
    pm[0][20:25] &= gm[0][:5]
    pcl = [np.zeros(40, int), rng.integers(0, 2, 6)]
    gcl = [np.zeros(20, int), rng.integers(0, 2, 4)]
    sc = [rng.random(40), rng.random(6)]
    cr = [rng.random(20) < 0.1, np.zeros(4, bool)]
    ev = pe.InstanceEvaluator(2, device="cuda")
    ev.update(pe.InstanceBatch([cuda(m) for m in pm], [cuda(c) for c in pcl], [cuda(s) for s in sc]),
              pe.InstanceBatch([cuda(m) for m in gm], [cuda(c) for c in gcl], crowd=[cuda(c) for c in cr]))
    state = oinst.InstanceState(2)
    state.update([dict(masks=m, classes=c, scores=s) for m, c, s in zip(pm, pcl, sc)],
                 [dict(masks=m, classes=c, crowd=k) for m, c, k in zip(gm, gcl, cr)])
    compare_instance(to_numpy(ev.compute()), state.compute())


def test_score_order_beyond_the_lds_sort(pe):
    """An image with more detections than the in-LDS per-image sort holds
    (4096) takes the global merge sort; both images against the oracle."""
    rng = np.random.default_rng(14)
    h, w = 6, 7
    dets = [5000, 37]
    pm = [rng.random((d, h, w)) > 0.5 for d in dets]
    gm = [rng.random((3, h, w)) > 0.5 for _ in dets]
    pcl = [rng.integers(0, 2, d) for d in dets]
    gcl = [rng.integers(0, 2, 3) for _ in dets]
    sc = [rng.integers(0, 300, d) / 300.0 for d in dets]  # many ties
    ev = pe.InstanceEvaluator(2, device="cuda")
    ev.update(pe.InstanceBatch([cuda(m) for m in pm], [cuda(c) for c in pcl], [cuda(s) for s in sc]),
              pe.InstanceBatch([cuda(m) for m in gm], [cuda(c) for c in gcl]))
    state = oinst.InstanceState(2)
    state.update([dict(masks=m, classes=c, scores=s) for m, c, s in zip(pm, pcl, sc)],
                 [dict(masks=m, classes=c) for m, c in zip(gm, gcl)])
    compare_instance(to_numpy(ev.compute()), state.compute())


def test_sorts_over_many_merge_passes(pe):
    """Three images of 14000 detections: the per-image score order and the AP
    record order each sort 42000 items (11 tiles, four merge passes with a
    partial last pair); against the oracle."""
    rng = np.random.default_rng(16)
    h, w = 3, 4
    dets = [14000, 14000, 14000]
    pm = [rng.random((d, h, w)) > 0.5 for d in dets]
    gm = [rng.random((4, h, w)) > 0.5 for _ in dets]
    for m in pm + gm:
        m[:, 0, 0] = True  # no empty mask
    pcl = [rng.integers(0, 3, d) for d in dets]
    gcl = [rng.integers(0, 3, 4) for _ in dets]
    sc = [rng.integers(0, 1000, d) / 1000.0 for d in dets]  # many ties
    ev = pe.InstanceEvaluator(3, device="cuda")
    ev.update(pe.InstanceBatch([cuda(m) for m in pm], [cuda(c) for c in pcl], [cuda(s) for s in sc]),
              pe.InstanceBatch([cuda(m) for m in gm], [cuda(c) for c in gcl]))
    state = oinst.InstanceState(3)
    state.update([dict(masks=m, classes=c, scores=s) for m, c, s in zip(pm, pcl, sc)],
                 [dict(masks=m, classes=c) for m, c in zip(gm, gcl)])
    compare_instance(to_numpy(ev.compute()), state.compute())


def test_combined_compute_beyond_the_small_ap_path(pe):
    """More than 4096 detection records: compute() launches the finaliser on
    its own before the sorted AP path (below that, both run in one
    launch, as every other combined test checks)."""
    rng = np.random.default_rng(15)
    B, h, w, D = 2, 10, 12, 2100
    maps = rng.integers(0, 6, (B, h, w)).astype(np.int32)
    classes = np.array([[-1, 0, 1, 2, 0, 3]] * B, dtype=np.int32)
    crowd = np.zeros((B, 6), bool)
    pred_maps = np.where(rng.random((B, h, w)) < 0.8, maps, rng.integers(0, 6, (B, h, w))).astype(np.int32)
    masks = [rng.random((D, h, w)) > 0.6 for _ in range(B)]
    labels = [rng.integers(0, 3, D) for _ in range(B)]
    scores = [rng.integers(0, 500, D) / 500.0 for _ in range(B)]
    isthing = [True, True, True, False]
    joint = pe.SegmentationEvaluator(4, isthing=isthing, validate=False)
    joint.update(pe.PanopticBatch(cuda(pred_maps), cuda(classes)),
                 pe.InstanceBatch([cuda(m) for m in masks], [cuda(x) for x in labels], [cuda(s) for s in scores]),
                 pe.PanopticBatch(cuda(maps), cuda(classes), cuda(crowd)))
    state = ocomb.SegmentationState(4, isthing, validate=False)
    state.update(pred_maps, classes, [dict(masks=m, classes=x, scores=s) for m, x, s in zip(masks, labels, scores)],
                 maps, classes, crowd)
    result, expected = to_numpy(joint.compute()), state.compute()
    compare_panoptic(result, expected)
    compare_instance(result["instance"], expected["instance"])
