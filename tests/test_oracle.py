"""Pins for the CPU oracle: golden vectors from the reference, the reference's
known-answer tests restated, brute force on tiny inputs and invariants.

CPU only.  A plausible mistake in the oracle (dropped void term, wrong crowd
rule, transposed histogram, wrong tie-break) fails at least one of these.
"""

import itertools

import numpy as np
import pytest

import oracle
from golden_data import group, load
from oracle import combined as ocomb
from oracle import instance as oinst
from oracle import panoptic as opan
from oracle import prefix as opre

CATS = (7, 42, 91)


def assert_exact(actual, expected):
    actual, expected = np.asarray(actual), np.asarray(expected)
    assert actual.shape == expected.shape, (actual.shape, expected.shape)
    np.testing.assert_array_equal(actual, expected)


def assert_close(actual, expected, rtol=1e-12):
    actual, expected = np.asarray(actual, np.float64), np.asarray(expected, np.float64)
    assert actual.shape == expected.shape
    np.testing.assert_allclose(actual, expected, rtol=rtol, atol=1e-15, equal_nan=True)


def packed(maps, info, categories=CATS):
    maps = np.asarray(maps, np.int64)
    ids, slots, classes, crowd = opan.from_coco(maps, info, categories)
    return opan.compact(maps, ids, slots), classes, crowd


def evaluate(gt, pred, gi, pi, isthing=(True, False, True)):
    state = opan.PanopticState(3, isthing)
    g, gc, crowd = packed(gt, gi)
    p, pc, _ = packed(pred, pi)
    state.update(g, p, gc, pc, crowd)
    return state


# ---------------------------------------------------------------- golden vectors

def test_golden_panoptic_accumulate():
    data = load("panoptic_accumulate")
    state = opan.PanopticState(5, data["isthing"])
    for index in range(int(data["num_cases"])):
        inp = group(data, f"in{index}")
        state.update(inp["gt"], inp["pred"], inp["gc"], inp["pc"], inp["crowd"], validate=False)
        out = group(data, f"out{index}")
        result = state.compute()
        for key in ("tp", "fp", "fn", "confusion"):
            assert_exact(result[key], out[key])
        # in-order fold: bit-exact with the reference's scatter_add_
        assert_exact(result["iou_sum"], out["iou_sum"])
        for name in ("All", "Things", "Stuff"):
            for key in ("pq", "sq", "rq"):
                assert_close(result[name][key], out[f"{name}.{key}"])
            assert_exact(result[name]["n"], out[f"{name}.n"])
        for key in ("pq", "sq", "rq", "iou"):
            assert_close(result["per_class"][key], out[f"per_class.{key}"])
        assert_close(result["miou"], out["miou"])


def test_golden_panoptic_quality():
    data = load("panoptic_quality")
    for index in range(int(data["num_cases"])):
        inp = group(data, f"in{index}")
        args = (inp["gt"], inp["pred"], inp["gc"], inp["pc"], inp["crowd"], 5)
        assert_close(opan.panoptic_quality(*args), data[f"out{index}.pq"])
        assert_exact(opan.panoptic_quality(*args, return_stats=True), data[f"out{index}.stats"])


def test_golden_compact():
    data = load("compact")
    large = 2**40 + 17
    gi = [[{"id": 1234, "category_id": 7, "iscrowd": 1}, {"id": large, "category_id": 7, "iscrowd": 1},
           {"id": 7, "category_id": 42}], [{"id": 99, "category_id": 91}]]
    ids, slots, classes, crowd = opan.from_coco(data["in.maps"], gi, CATS)
    assert_exact(ids, data["out.lookup_ids"])
    assert_exact(slots, data["out.lookup_slots"])
    assert_exact(classes, data["out.classes"])
    assert_exact(crowd, data["out.crowd"])
    assert_exact(opan.compact(data["in.maps"], ids, slots), data["out.packed"])


def test_golden_coco_matcher():
    data = load("coco_matcher")
    for index in range(int(data["num_cases"])):
        inp = group(data, f"in{index}")
        pairs = inp["pairs"].astype(np.float64)
        union = np.where(inp["crowd"][None], inp["pa"][:, None], inp["pa"][:, None] + inp["ga"][None] - pairs)
        ious = pairs / np.maximum(union, 1)
        status, ranks, counts = oinst.match_ious(ious, inp["pc"], inp["gc"], inp["pa"], inp["crowd"],
                                                 inp["areas"], int(inp["C"]))
        assert_exact(status, data[f"out{index}.status"])
        assert_exact(ranks, data[f"out{index}.ranks"])
        assert_exact(counts, data[f"out{index}.counts"])


def test_golden_instance_evaluator():
    data = load("instance_eval")
    state = oinst.InstanceState(4)
    for image in range(int(data["num_images"])):
        inp = group(data, f"in{image}")
        state.update([dict(masks=inp["pm"], classes=inp["pc"], scores=inp["scores"])],
                     [dict(masks=inp["gm"], classes=inp["gc"], crowd=inp["crowd"], areas=inp["areas"])])
    result = state.compute()
    out = group(data, "out")
    assert_exact(result["precision"], out["precision"])
    assert_exact(result["recall"], out["recall"])
    for key, value in out.items():
        if key not in ("precision", "recall"):
            assert_close(result[key], value)


def test_golden_combined():
    data = load("combined")
    inp = group(data, "in")
    state = ocomb.SegmentationState(3, [True, True, False])
    dets = [dict(masks=inp[f"masks{b}"], classes=inp[f"labels{b}"], scores=inp[f"scores{b}"]) for b in range(2)]
    for areas in (None, inp["areas"]):
        state.update(inp["pred_maps"], inp["classes"], dets, inp["maps"], inp["classes"], inp["crowd"],
                     instance_areas=areas)
    result = state.compute()
    out = group(data, "out")
    for key in ("tp", "fp", "fn", "confusion", "iou_sum"):
        assert_exact(result[key], out[key])
    assert_exact(result["instance"]["precision"], out["instance.precision"])
    assert_exact(result["instance"]["recall"], out["instance.recall"])
    for key, value in out.items():
        if key.startswith("instance.") and key not in ("instance.precision", "instance.recall"):
            assert_close(result["instance"][key[9:]], value)
    assert_close(result["miou"], out["miou"])


def test_golden_prefix():
    data = load("prefix")
    for index in range(int(data["num_cases"])):
        inp = group(data, f"in{index}")
        out = group(data, f"out{index}")
        result = opre.prefixes([inp["masks"]], [inp["classes"]], inp["maps"], inp["labels"], inp["crowd"], 3,
                               isthing=[True, False, True], iou_thresholds=tuple(inp["thresholds"]))[0]
        assert len(result) == int(out["count"])
        for k, snap in enumerate(result):
            assert snap["prefix"] == out["prefix"][k]
            for key in ("tp", "fp", "fn", "confusion", "iou_sum"):
                assert_exact(snap[key], out[key][k])
            for key in ("tp", "fp", "fn"):
                assert_exact(snap["instance"][key], out[f"instance.{key}"][k])
            for key in ("precision", "recall"):
                assert_close(snap["instance"][key], out[f"instance.{key}"][k])
            for key in ("pq", "sq", "rq", "iou"):
                assert_close(snap["per_class"][key], out[f"per_class.{key}"][k])
            assert_close(snap["miou"], out["miou"][k])


# ------------------------------------------------- known-answer tests (restated)

def test_kat_perfect_and_reset():  # test_evaluator.py:55-67
    maps = [[[0, 1, 1, 2]]]
    info = [[{"id": 1, "category_id": 7}, {"id": 2, "category_id": 42}]]
    state = evaluate(maps, maps, info, info)
    result = state.compute()
    assert result["All"]["pq"] == 1 and result["Things"]["n"] == 1 and result["Stuff"]["pq"] == 1
    assert np.isnan(result["per_class"]["iou"][2]) and result["miou"] == 1
    state.reset()
    assert state.compute()["All"]["pq"] == 0 and state.compute()["miou"] == 0


def test_kat_strict_half_and_void_union():  # test_evaluator.py:69-81
    gi = [[{"id": 1, "category_id": 7}]]
    pi = [[{"id": 9, "category_id": 7}]]
    r = evaluate([[[1, 1, 1, 1]]], [[[9, 9, 0, 0]]], gi, pi).compute()
    assert (r["tp"].sum(), r["fp"].sum(), r["fn"].sum(), r["miou"]) == (0, 1, 1, 0.5)
    r = evaluate([[[1, 1, 1]]], [[[9, 9, 0]]], gi, pi).compute()
    assert r["tp"].sum() == 1 and abs(r["iou_sum"].sum() - 2 / 3) < 1e-15
    r = evaluate([[[1, 0, 0, 0]]], [[[9, 9, 9, 9]]], gi, pi).compute()
    assert r["All"]["pq"] == 1


def test_kat_crowd_and_half_ignore():  # test_evaluator.py:83-94
    gi = [[{"id": 1, "category_id": 7, "iscrowd": 1}]]
    pi = [[{"id": 9, "category_id": 7}]]
    r = evaluate([[[1, 1, 0]]], [[[9, 9, 9]]], gi, pi).compute()
    assert r["fp"].sum() == 0 and r["fn"].sum() == 0
    r = evaluate([[[1, 1, 1, 1]]], [[[9, 9, 9, 9]]], gi, [[{"id": 9, "category_id": 42}]]).compute()
    assert r["fp"][1] == 1
    r = evaluate([[[0, 0, 1, 1]]], [[[9, 9, 9, 9]]], [[{"id": 1, "category_id": 42}]], pi).compute()
    assert r["fp"][0] == 1


def test_kat_confusion():  # test_evaluator.py:96-101
    r = evaluate([[[1, 1, 2, 2]]], [[[9, 0, 9, 10]]],
                 [[{"id": 1, "category_id": 7}, {"id": 2, "category_id": 42}]],
                 [[{"id": 9, "category_id": 7}, {"id": 10, "category_id": 42}]]).compute()
    assert_exact(r["confusion"], [[1, 0, 0, 1], [1, 1, 0, 0], [0, 0, 0, 0]])
    assert abs(r["miou"] - (1 / 3 + 1 / 2) / 2) < 1e-15


def test_kat_rewards():  # test_rewards.py:12-49
    gt = np.array([[[1, 1, 1, 1]], [[1, 1, 1, 0]], [[1, 0, 0, 0]], [[2, 2, 0, 0]], [[0, 0, 0, 0]],
                   [[1, 1, 1, 1]], [[1, 1, 1, 1]], [[1, 1, 2, 2]]], np.int32)
    pred = np.array([[[1, 1, 0, 0]], [[1, 1, 0, 0]], [[1, 1, 1, 1]], [[2, 2, 2, 2]], [[0, 0, 0, 0]],
                     [[0, 0, 0, 0]], [[1, 1, 1, 1]], [[1, 1, 2, 2]]], np.int32)
    gc = np.array([[-1, 0, -1], [-1, 0, -1], [-1, 0, -1], [-1, -1, 1], [-1, -1, -1], [-1, 0, -1],
                   [-1, 0, -1], [-1, 0, 2]], np.int32)
    pc = np.array([[-1, 0, 2]] * 8, np.int32)
    pc[3, 2] = 1
    crowd = np.zeros(gc.shape, bool)
    crowd[3, 2] = True
    pq = opan.panoptic_quality(gt, pred, gc, pc, crowd, 4)
    assert pq[0] == 0 and abs(pq[1] - 2 / 3) < 1e-15 and pq[2] == 1
    # invariant: equals a per-image evaluator with empty prediction slots masked
    for b in range(8):
        present = np.bincount(pred[b].ravel(), minlength=3) > 0
        labels = np.where(present, pc[b], -1)[None].astype(np.int32)
        state = opan.PanopticState(4)
        state.update(gt[b:b + 1], pred[b:b + 1], gc[b:b + 1], labels, crowd[b:b + 1])
        assert_close(pq[b], state.compute()["All"]["pq"])


def instance_batch(scores=None, crowd=None):
    masks = np.zeros((2, 16, 16), bool)
    masks[0, :10, :10] = True
    masks[1, 5:15, 5:15] = True
    item = dict(masks=masks, classes=np.zeros(2, np.int64))
    if scores is not None:
        item["scores"] = np.array(scores)
    if crowd is not None:
        item["crowd"] = np.array(crowd)
    return item


def test_kat_instance():  # test_instance.py:27-69
    state = oinst.InstanceState(2)
    state.update([instance_batch([0.9, 0.8])], [instance_batch()])
    result = state.compute()
    assert abs(result["ap"] - 1) < 1e-12 and (result["precision"][:, :, 1] == -1).all()
    state.reset()
    assert state.compute()["ap"] == -1
    state = oinst.InstanceState(1)
    empty = dict(masks=np.zeros((0, 16, 16), bool), classes=np.zeros(0, np.int64), scores=np.zeros(0))
    state.update([empty], [instance_batch()])
    assert state.compute()["ap"] == 0
    state = oinst.InstanceState(1)
    state.update([instance_batch([0.9, 0.8])], [instance_batch(crowd=[True, False])])
    state.update([instance_batch([0.9, 0.8])], [instance_batch()])
    assert abs(state.compute()["ap"] - 1) < 1e-12
    target = instance_batch()
    target = dict(masks=target["masks"][:1], classes=target["classes"][:1])
    state = oinst.InstanceState(1)
    state.update([instance_batch([0.8, 0.9])], [target])
    assert abs(state.compute()["ap"] - 0.5) < 1e-12


def test_kat_max_detections():  # test_instance.py:179-197
    masks = np.ones((106, 1, 1), bool)
    labels = np.zeros(106, np.int64)
    labels[-1] = 1
    state = oinst.InstanceState(2)
    state.update([dict(masks=masks, classes=labels, scores=np.ones(106))], [dict(masks=masks, classes=labels)])
    r = state.compute()
    assert abs(r["ar1"] - (1 / 105 + 1) / 2) < 1e-12
    assert abs(r["ar10"] - (10 / 105 + 1) / 2) < 1e-12
    assert abs(r["ar100"] - (100 / 105 + 1) / 2) < 1e-12


def test_kat_exact_iou_threshold():  # test_instance.py:199-218
    pm = np.array([[[True, False]]])
    gm = np.array([[[True, True]]])
    label = np.array([0])
    state = oinst.InstanceState(1)
    state.update([dict(masks=pm, classes=label, scores=np.array([0.9]))], [dict(masks=gm, classes=label)])
    r = state.compute()
    assert abs(r["ap50"] - 1) < 1e-12 and abs(r["ap"] - 0.1) < 1e-12
    state.reset()
    state.update([dict(masks=pm, classes=label, scores=np.array([0.9]))], [dict(masks=gm[:0], classes=label[:0])])
    assert state.compute()["ap"] == -1
    state.reset()
    state.update([dict(masks=pm[:0], classes=label[:0], scores=np.zeros(0))], [dict(masks=gm, classes=label)])
    assert state.compute()["ap"] == 0


def prefix_target():
    return (np.array([[[1, 1, 2, 2, 3, 0]]], np.int32), np.array([[-1, 0, 0, 1]], np.int32))


def test_kat_prefix_independent_masks_and_crowd():  # test_prefix.py:43-61
    maps, labels = prefix_target()
    masks = np.array([[[1, 1, 0, 0, 0, 0]], [[1, 1, 0, 0, 0, 0]], [[0, 0, 1, 1, 0, 0]],
                      [[0, 0, 0, 0, 1, 0]]], bool)
    classes = np.array([0, 0, 0, 1])
    result = opre.prefixes([masks], [classes], maps, labels, None, 2, isthing=[True, False])[0]
    assert [r["instance"]["tp"][0, 0] for r in result] == [1, 1, 2, 2]
    assert [r["instance"]["fp"][0, 0] for r in result] == [0, 1, 1, 1]
    assert abs(result[-1]["instance"]["precision"][0, 0] - 2 / 3) < 1e-15
    crowd = np.array([[False, True, False, False]])
    result = opre.prefixes([masks], [classes], maps, labels, crowd, 2)[0]
    assert result[1]["instance"]["tp"].sum() == 0 and result[1]["instance"]["fp"].sum() == 0
    assert result[2]["instance"]["recall"][0, 0] == 1


def test_prefix_equals_canvas_evaluation():  # invariant of test_prefix.py:17-41
    maps, labels = prefix_target()
    rng = np.random.default_rng(17)
    for _ in range(5):
        masks = rng.random((7, 1, 6)) > 0.4
        masks[:, :, 0] = True
        classes = rng.integers(0, 2, 7)
        actual = opre.prefixes([masks], [classes], maps, labels, None, 3, isthing=[True, False, True],
                               checkpoints=range(8))[0]
        canvas = np.zeros_like(maps)
        for k, result in enumerate(actual):
            if k:
                canvas[0][masks[k - 1]] = k
            lab = np.concatenate(([-1], classes))
            present = np.bincount(canvas.ravel(), minlength=8) > 0
            lab = np.where(present, lab, -1)[None].astype(np.int32)
            state = opan.PanopticState(3, [True, False, True])
            state.update(maps, canvas, labels, lab, None)
            expected = state.compute()
            for key in ("tp", "fp", "fn", "confusion"):
                assert_exact(result[key], expected[key])
            assert_close(result["iou_sum"], expected["iou_sum"])


# ------------------------------------------------------- brute force and invariants

def test_histogram_brute_force():
    rng = np.random.default_rng(5)
    B, H, W, G, P, C = 2, 5, 7, 4, 6, 3
    gt = rng.integers(0, G, (B, H, W)).astype(np.int32)
    pred = rng.integers(0, P, (B, H, W)).astype(np.int32)
    gc = rng.integers(-1, C, (B, G)).astype(np.int32)
    pc = rng.integers(-1, C, (B, P)).astype(np.int32)
    crowd = rng.random((B, G)) < 0.3
    pairs, semantic = opan.histogram(gt, pred, gc, pc, crowd, C)
    want = np.zeros((B, G, P), np.int64)
    sem = np.zeros((C, C + 1), np.int64)
    for b, y, x in itertools.product(range(B), range(H), range(W)):
        g, p = gt[b, y, x], pred[b, y, x]
        want[b, g, p] += 1
        if gc[b, g] >= 0 and not crowd[b, g]:
            sem[gc[b, g], pc[b, p] if pc[b, p] >= 0 else C] += 1
    assert_exact(pairs, want)
    assert_exact(semantic, sem)
    assert pairs.sum() == B * H * W
    assert_exact(pairs.sum(2)[0], np.bincount(gt[0].ravel(), minlength=G))


def test_pq_brute_force_textbook():
    """Textbook panopticapi per-segment loop on tiny random inputs."""
    rng = np.random.default_rng(143)
    for _ in range(20):
        H, W, G, P, C = 6, 7, 5, 6, 3
        gt = rng.integers(0, G, (1, H, W)).astype(np.int32)
        pred = np.where(rng.random((1, H, W)) < 0.6, gt, rng.integers(0, P, (1, H, W))).astype(np.int32)
        gc = rng.integers(0, C, (1, G)).astype(np.int32)
        pc = np.concatenate([gc[:, :G], rng.integers(0, C, (1, P - G))], 1).astype(np.int32)
        gc[:, 0] = pc[:, 0] = -1
        crowd = rng.random((1, G)) < 0.3
        crowd[:, 0] = False
        ga = np.bincount(gt.ravel(), minlength=G)
        pa = np.bincount(pred.ravel(), minlength=P)
        gc[0, ga == 0] = -1
        pc[0, pa == 0] = -1
        state = opan.PanopticState(C)
        state.update(gt, pred, gc, pc, crowd)
        tp, fn, s = np.zeros(C, int), np.zeros(C, int), np.zeros(C)
        gmatched, pmatched = set(), set()
        inter = {}
        for g, p in zip(gt.ravel(), pred.ravel()):
            inter[(g, p)] = inter.get((g, p), 0) + 1
        for (g, p), i in inter.items():
            if g == 0 or p == 0 or gc[0, g] < 0 or pc[0, p] < 0 or crowd[0, g] or gc[0, g] != pc[0, p]:
                continue
            union = ga[g] + pa[p] - i - inter.get((0, p), 0)
            if i / union > 0.5:
                tp[gc[0, g]] += 1
                s[gc[0, g]] += i / union
                gmatched.add(g)
                pmatched.add(p)
        for g in range(1, G):
            if gc[0, g] >= 0 and not crowd[0, g] and g not in gmatched:
                fn[gc[0, g]] += 1
        # unmatched predictions are FP unless void + last same-class crowd
        # overlap covers more than half of them (panopticapi's ignore rule)
        fp_true = np.zeros(C, int)
        for p in range(1, P):
            if pc[0, p] < 0 or p in pmatched:
                continue
            ignore = inter.get((0, p), 0)
            crowd_g = [g for g in range(1, G) if crowd[0, g] and gc[0, g] == pc[0, p]]
            if crowd_g:
                ignore += inter.get((crowd_g[-1], p), 0)
            if ignore / pa[p] <= 0.5:
                fp_true[pc[0, p]] += 1
        assert_exact(state.tp, tp)
        assert_exact(state.fn, fn)
        assert_exact(state.fp, fp_true)
        assert_close(state.iou_sum, s)


def test_accumulate_brute_force_single_class():
    """AP for one class by the COCOeval definition (interpolated precision at
    101 recall points) against an explicit loop."""
    rng = np.random.default_rng(3)
    for _ in range(10):
        n = 12
        status = np.full((40, n), -1, np.int8)
        flags = rng.choice([-1, 0, 1], n, p=[0.2, 0.4, 0.4]).astype(np.int8)
        status[0:10] = flags
        ranks = np.arange(1, n + 1, dtype=np.int32)
        classes = np.zeros(n, np.int32)
        npos = max(int((flags == 1).sum()) + 2, 1)
        counts = np.array([[npos, 0, 0, 0]])
        precision, recall = oinst.accumulate(classes, ranks, status, counts)
        tp = fp = 0
        prs, rcs = [], []
        for f in flags:
            tp += f == 1
            fp += f == 0
            rcs.append(tp / npos)
            prs.append(tp / (tp + fp + np.finfo(float).eps))
        for r in range(101):
            thr = r * 0.01
            idx = [i for i in range(n) if rcs[i] >= thr]
            want = max(prs[idx[0]:]) if idx else 0.0
            assert precision[0, r, 0, 0, 2] == want
        assert recall[0, 0, 0, 2] == rcs[-1]


def test_rgb_decode():
    """README.md:67-68 decode against the reference tests' encoder
    (test_evaluator.py:39) and hand values; each byte weight checked alone."""
    assert opan.rgb_to_id(np.array([1, 0, 0], np.uint8)) == 1
    assert opan.rgb_to_id(np.array([0, 1, 0], np.uint8)) == 256
    assert opan.rgb_to_id(np.array([0, 0, 1], np.uint8)) == 65536
    assert opan.rgb_to_id(np.array([1, 2, 3], np.uint8)) == 1 + 2 * 256 + 3 * 65536
    assert opan.rgb_to_id(np.array([255, 255, 255], np.uint8)) == 2**24 - 1
    ids = np.random.default_rng(4).integers(0, 2**24, (3, 5, 7))
    rgb = opan.id_to_rgb(ids)
    assert rgb.dtype == np.uint8 and rgb.shape == (3, 5, 7, 3)
    assert_exact(opan.rgb_to_id(rgb), ids)
    assert_exact(rgb[..., 0], ids % 256)  # R is the low byte


def test_oracle_is_test_only():
    assert oracle.__doc__ and "TEST INFRASTRUCTURE ONLY" in oracle.__doc__


# ------------------------------------------- torch-op restatement (CPU baseline)

def _torch_and_numpy(C, isthing, batches):
    import torch
    from oracle import torch_cpu
    ours, theirs = torch_cpu.SegmentationState(C, isthing), ocomb.SegmentationState(C, isthing, validate=False)
    for pred, pc, dets, gt, gc, crowd, areas in batches:
        theirs.update(pred, pc, [dict(masks=m, classes=c, scores=s) for m, c, s in dets], gt, gc, crowd,
                      instance_areas=areas)
        t = lambda a, dtype=torch.int64: torch.from_numpy(np.asarray(a)).to(dtype)  # noqa: E731
        ours.update(t(pred), t(pc), [(t(m, torch.bool), t(c), t(s, torch.float64)) for m, c, s in dets],
                    t(gt), t(gc), t(crowd, torch.bool), None if areas is None else t(areas, torch.float64))
    return ours.compute(), theirs.compute()


def _assert_same_results(got, want):
    for key in ("tp", "fp", "fn", "confusion"):
        assert_exact(got[key].numpy(), want[key])
    assert_close(got["iou_sum"].numpy(), want["iou_sum"])
    assert_close(got["miou"].item(), want["miou"])
    for group_name in ("All", "Things", "Stuff"):
        for key in ("pq", "sq", "rq", "n"):
            assert_close(got[group_name][key].item(), want[group_name][key])
    for key in ("pq", "sq", "rq", "iou"):
        assert_close(got["per_class"][key].numpy(), want["per_class"][key])
    assert_exact(got["instance"]["precision"], want["instance"]["precision"])
    assert_exact(got["instance"]["recall"], want["instance"]["recall"])


def test_torch_restatement_golden_combined():
    data = load("combined")
    inp = group(data, "in")
    dets = [(inp[f"masks{b}"], inp[f"labels{b}"], inp[f"scores{b}"]) for b in range(2)]
    batches = [(inp["pred_maps"], inp["classes"], dets, inp["maps"], inp["classes"], inp["crowd"], areas)
               for areas in (None, inp["areas"])]
    got, want = _torch_and_numpy(3, [True, True, False], batches)
    _assert_same_results(got, want)
    out = group(data, "out")
    for key in ("tp", "fp", "fn", "confusion", "iou_sum"):
        assert_exact(got[key].numpy(), out[key])


def test_torch_restatement_random_batches():
    rng = np.random.default_rng(31)
    C, isthing = 6, [True, True, True, False, True, False]
    batches = []
    for _ in range(3):
        B, H, W, S = 3, 20, 24, 7
        gt = rng.integers(0, S, (B, H, W))
        pred = np.where(rng.random((B, H, W)) < 0.3, rng.integers(0, S, (B, H, W)), gt)
        gc = np.where(rng.random((B, S)) < 0.85, rng.integers(0, C, (B, S)), -1)
        gc[:, 0] = -1
        pc = np.where(rng.random((B, S)) < 0.7, gc, rng.integers(0, C, (B, S)))
        pc[:, 0] = -1
        crowd = (rng.random((B, S)) < 0.2) & (gc >= 0)
        dets = []
        for b in range(B):
            D = int(rng.integers(0, 9))
            masks = (gt[b][None] == rng.integers(1, S, (D, 1, 1))) ^ (rng.random((D, H, W)) < 0.1)
            masks[:, 0, 0] = True
            classes = rng.choice([c for c in range(C) if isthing[c]], D)
            scores = np.round(rng.random(D), 1)  # ties exercise the stable order
            dets.append((masks, classes, scores))
        batches.append((pred, pc, dets, gt, gc, crowd, None))
    got, want = _torch_and_numpy(C, isthing, batches)
    _assert_same_results(got, want)
    assert int(got["tp"].sum()) > 0 and (got["instance"]["precision"] > 0).any()


# ---------------------------------------------------------------- COCO RLE

def test_rle_string_known_answers():
    """Hand-derived strings of pycocotools' compressed RLE: single groups,
    a multi-group count, the delta from count i - 2 (only for i > 2) and a
    negative delta with its sign bit."""
    from oracle import rle
    cases = [
        ([4], b"4"),              # all-zero 2 x 2
        ([0, 4], b"04"),          # all-one 2 x 2: an empty leading zero run
        ([40], b"X1"),            # 40 = 8 + 32 * 1: group 8 with the continuation bit, then 1
        ([3, 2, 5, 2], b"3250"),  # count 3 is stored as 2 - 2 = 0; count 2 as it is
        ([1, 5, 1, 1], b"151L"),  # count 3: 1 - 5 = -4, one group 28 whose bit 4 is the sign
        ([16], b"`0"),            # 16 has bit 4 set: it takes the continuation bit (96), then group 0
        ([0, 1, 0, 0, 1], b"010O1"),  # count 3: 0 - 1 = -1 is group 31 (79); count 4: 1 - 0
    ]
    for counts, text in cases:
        assert rle.counts_to_string(counts) == text, counts
        assert rle.counts_from_string(text) == counts, text


def test_rle_round_trip_and_identities():
    from oracle import rle
    rng = np.random.default_rng(11)
    for h, w in ((1, 1), (1, 7), (5, 1), (3, 4), (17, 33), (32, 64), (40, 50)):
        for density in (0.0, 0.05, 0.5, 0.97, 1.0):
            mask = rng.random((h, w)) < density
            code = rle.encode(mask)
            assert code["size"] == [h, w]
            counts = rle.counts_from_string(code["counts"])
            assert sum(counts) == h * w
            assert rle.area(code) == int(mask.sum())
            assert np.array_equal(rle.decode(code), mask)
            assert np.array_equal(rle.decode({"size": [h, w], "counts": counts}), mask)
            # Column-major: the first run covers column 0 from the top.
            first = mask[:, 0].argmax() if mask[:, 0].any() else h
            assert counts[0] == (first if first < h else counts[0])
    with pytest.raises(ValueError):
        rle.decode({"size": [2, 2], "counts": [1, 2]})
