"""GPU parity of segmentation_prefixes against the CPU oracle and golden
vectors (restates the reference's tests/test_prefix.py behaviours)."""

import numpy as np
import pytest
import torch

from golden_data import group, load
from helpers import assert_close, assert_exact, to_numpy
from oracle import prefix as opre

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pe():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import panoptic_evaluator
    return panoptic_evaluator


def cuda(x):
    return torch.as_tensor(np.asarray(x)).cuda()


def compare_snapshots(actual, expected):
    assert len(actual) == len(expected)
    for a, e in zip(actual, expected):
        assert a["prefix"] == e["prefix"]
        for key in ("tp", "fp", "fn", "confusion", "iou_sum"):
            assert_exact(a[key], e[key], f"prefix {e['prefix']} {key}")
        for key in ("tp", "fp", "fn"):
            assert_exact(a["instance"][key], e["instance"][key], f"instance {key}")
        for key in ("precision", "recall"):
            assert_close(a["instance"][key], e["instance"][key], what=f"instance {key}")
        for name in ("All", "Things", "Stuff"):
            for key in ("pq", "sq", "rq"):
                assert_close(a[name][key], e[name][key], what=f"{name}.{key}")
            assert_exact(a[name]["n"], e[name]["n"])
        for key in ("pq", "sq", "rq", "iou"):
            assert_close(a["per_class"][key], e["per_class"][key], what=f"per_class.{key}")
        assert_close(a["miou"], e["miou"], what="miou")


def test_golden_prefix(pe):
    data = load("prefix")
    for index in range(int(data["num_cases"])):
        inp = group(data, f"in{index}")
        out = group(data, f"out{index}")
        result = to_numpy(pe.segmentation_prefixes(
            pe.InstanceBatch([cuda(inp["masks"])], [cuda(inp["classes"])]),
            pe.PanopticBatch(cuda(inp["maps"]), cuda(inp["labels"]), cuda(inp["crowd"])), 3,
            isthing=[True, False, True], iou_thresholds=tuple(inp["thresholds"])))[0]
        assert len(result) == int(out["count"])
        for k, snap in enumerate(result):
            assert snap["prefix"] == out["prefix"][k]
            for key in ("tp", "fp", "fn", "confusion", "iou_sum"):
                assert_exact(snap[key], out[key][k], key)
            for key in ("tp", "fp", "fn"):
                assert_exact(snap["instance"][key], out[f"instance.{key}"][k], key)
            for key in ("precision", "recall"):
                assert_close(snap["instance"][key], out[f"instance.{key}"][k])
            for key in ("pq", "sq", "rq", "iou"):
                assert_close(snap["per_class"][key], out[f"per_class.{key}"][k])
            assert_close(snap["miou"], out["miou"][k])


def target(pe):
    return pe.PanopticBatch(torch.tensor([[[1, 1, 2, 2, 3, 0]]], dtype=torch.int32, device="cuda"),
                            torch.tensor([[-1, 0, 0, 1]], dtype=torch.int32, device="cuda"))


def test_independent_masks_duplicates_and_crowd(pe):  # test_prefix.py:43-61
    masks = torch.tensor([[[1, 1, 0, 0, 0, 0]], [[1, 1, 0, 0, 0, 0]], [[0, 0, 1, 1, 0, 0]], [[0, 0, 0, 0, 1, 0]]],
                         dtype=torch.bool, device="cuda")
    classes = torch.tensor([0, 0, 0, 1], dtype=torch.int32, device="cuda")
    result = pe.segmentation_prefixes(pe.InstanceBatch([masks], [classes]), target(pe), 2, isthing=[True, False])[0]
    assert [r["instance"]["tp"][0, 0].item() for r in result] == [1, 1, 2, 2]
    assert [r["instance"]["fp"][0, 0].item() for r in result] == [0, 1, 1, 1]
    assert abs(result[-1]["instance"]["precision"][0, 0].item() - 2 / 3) < 1e-15
    t = target(pe)
    crowded = pe.PanopticBatch(t.segment_map, t.classes, torch.tensor([[False, True, False, False]], device="cuda"))
    result = pe.segmentation_prefixes(pe.InstanceBatch([masks], [classes]), crowded, 2)[0]
    assert result[1]["instance"]["tp"].sum().item() == 0 and result[1]["instance"]["fp"].sum().item() == 0
    assert result[2]["instance"]["recall"][0, 0].item() == 1


def test_every_prefix_matches_oracle_canvas(pe):  # test_prefix.py:17-41
    gen = np.random.default_rng(17)
    maps = np.array([[[1, 1, 2, 2, 3, 0]]], np.int32)
    labels = np.array([[-1, 0, 0, 1]], np.int32)
    for _ in range(5):
        masks = gen.random((7, 1, 6)) > 0.4
        masks[:, :, 0] = True
        classes = gen.integers(0, 2, 7).astype(np.int32)
        actual = to_numpy(pe.segmentation_prefixes(pe.InstanceBatch([cuda(masks)], [cuda(classes)]),
                                                   pe.PanopticBatch(cuda(maps), cuda(labels)), 3,
                                                   isthing=[True, False, True], checkpoints=range(8)))[0]
        expected = opre.prefixes([masks], [classes], maps, labels, None, 3, isthing=[True, False, True],
                                 checkpoints=range(8))[0]
        compare_snapshots(actual, expected)


def test_variable_lengths_original_ids_and_thresholds(pe):  # test_prefix.py:63-80
    ids = torch.tensor([0, 2**40, 71, 9], device="cuda")
    t = target(pe)
    tgt = pe.PanopticBatch.from_coco(ids[t.segment_map.long()].repeat(2, 1, 1),
                                     [[dict(id=2**40, category_id=0), dict(id=71, category_id=0),
                                       dict(id=9, category_id=1)]] * 2, [0, 1], compact=False)
    mask = torch.tensor([[[1, 0, 0, 0, 0, 0]]], dtype=torch.bool, device="cuda")
    classes = torch.zeros(1, dtype=torch.int32, device="cuda")
    prediction = pe.InstanceBatch([mask, mask[:0]], [classes, classes[:0]])
    result = pe.segmentation_prefixes(prediction, tgt, 2)
    assert [r[0]["prefix"] for r in result] == [1, 0]
    assert result[0][0]["instance"]["tp"][:, 0].tolist() == [1, 0]
    assert result[0][0]["instance"]["fp"][:, 0].tolist() == [0, 1]
    selected = pe.segmentation_prefixes(prediction, tgt, 2, checkpoints=[0])
    assert [r[0]["prefix"] for r in selected] == [0, 0]
    assert pe.segmentation_prefixes(prediction, tgt, 2, checkpoints=[]) == [[], []]


def test_empty_checkpoints_and_validation(pe):  # test_prefix.py:82-93
    prediction = pe.InstanceBatch([torch.empty((0, 1, 6), dtype=torch.bool, device="cuda")],
                                  [torch.empty(0, dtype=torch.int32, device="cuda")])
    result = pe.segmentation_prefixes(prediction, target(pe), 2)[0]
    assert len(result) == 1 and result[0]["prefix"] == 0
    assert result[0]["instance"]["fn"][0].sum().item() == 3
    for kwargs in ({"checkpoints": [1]}, {"checkpoints": [0, 0]}, {"iou_thresholds": [0]},
                   {"iou_thresholds": [float("nan")]}):
        with pytest.raises(ValueError):
            pe.segmentation_prefixes(prediction, target(pe), 2, **kwargs)
    empty_mask = pe.InstanceBatch([torch.zeros((1, 1, 6), dtype=torch.bool, device="cuda")],
                                  [torch.zeros(1, dtype=torch.int32, device="cuda")])
    with pytest.raises(ValueError):
        pe.segmentation_prefixes(empty_mask, target(pe), 2)


def test_native_parity_streams_and_snapshots(pe):  # test_prefix.py:100-144
    gen = np.random.default_rng(123)
    for g, n in ((3, 0), (9, 17), (65, 129), (257, 300)):
        maps = gen.integers(0, g, (1, 19, 23)).astype(np.int32)
        maps.reshape(-1)[:g] = np.arange(g)
        labels = (np.arange(g) % 3).astype(np.int32)[None]
        labels[:, 0] = -1
        crowd = gen.random((1, g)) < 0.3
        crowd[:, 0] = False
        slots = gen.integers(1, g, n)
        masks = maps[0][None] == slots[:, None, None]
        if n:
            masks |= gen.random(masks.shape) < 0.01
        classes = labels[0, slots].astype(np.int32)
        kwargs = dict(isthing=[True, False, True], iou_thresholds=(0.5,) * 300 if n == 17 else (0.5, 0.75, 0.95),
                      checkpoints=[0, n] if n else [0])
        expected = opre.prefixes([masks], [classes], maps, labels, crowd, 3, **kwargs)
        stream = torch.cuda.Stream()
        stream.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(stream):
            actual = pe.segmentation_prefixes(pe.InstanceBatch([cuda(masks)], [cuda(classes)]),
                                              pe.PanopticBatch(cuda(maps), cuda(labels), cuda(crowd)), 3, **kwargs)
        stream.synchronize()
        snapshot = to_numpy(actual)
        compare_snapshots(snapshot[0], expected[0])
        pe.segmentation_prefixes(pe.InstanceBatch([cuda(masks)], [cuda(classes)]),
                                 pe.PanopticBatch(cuda(maps), cuda(labels), cuda(crowd)), 3, **kwargs)
        compare_snapshots(to_numpy(actual)[0], expected[0])  # retained outputs are not mutated


def test_multi_image_every_prefix(pe):
    """Several images of different sequence lengths in one call, stride-1 and
    stride-8 checkpoints, at a 96 x 128 resolution with ellipse masks."""
    import workloads
    gen = torch.Generator().manual_seed(5)
    pair = workloads.panoptic_batch(5, 3, 96, 128, 12, 5, 3)
    masks = [workloads.ellipse_masks(gen, n, 96, 128, max_axis=40.0) for n in (9, 0, 17)]
    classes = [torch.randint(0, 5, (len(m),), generator=gen).int() for m in masks]
    for stride in (1, 8):
        points = None if stride == 1 else [0, 8, 9]
        if points and any(p > len(m) for m in masks for p in points):
            points = [0]
        expected = opre.prefixes([m.numpy() for m in masks], [c.numpy() for c in classes], pair.gt.numpy(),
                                 pair.gc.numpy(), pair.crowd.numpy(), 5, isthing=[True, True, True, False, False],
                                 checkpoints=points)
        actual = to_numpy(pe.segmentation_prefixes(
            pe.InstanceBatch([m.cuda() for m in masks], [c.cuda() for c in classes]),
            pe.PanopticBatch(pair.gt.cuda(), pair.gc.cuda(), pair.crowd.cuda()), 5,
            isthing=[True, True, True, False, False], checkpoints=points))
        for a, e in zip(actual, expected):
            compare_snapshots(a, e)


def test_scoring_state_beyond_lds():
    """200 target segments x 160 masks: the G x (N + 1) canvas table no
    longer fits LDS, so scoring keeps its state in global memory (the
    unstaged path, matching inside the mask loop) -- same results."""
    import workloads
    import panoptic_evaluator as pe
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    gen = torch.Generator().manual_seed(21)
    pair = workloads.panoptic_batch(21, 1, 64, 64, 200, 5, 3)
    masks = [workloads.ellipse_masks(gen, 160, 64, 64, min_axis=2.0, max_axis=12.0)]
    classes = [torch.randint(0, 5, (160,), generator=gen).int()]
    isthing = [True, True, True, False, False]
    points = list(range(0, 161, 8))
    expected = opre.prefixes([m.numpy() for m in masks], [c.numpy() for c in classes], pair.gt.numpy(),
                             pair.gc.numpy(), pair.crowd.numpy(), 5, isthing=isthing, checkpoints=points)
    actual = to_numpy(pe.segmentation_prefixes(
        pe.InstanceBatch([m.cuda() for m in masks], [c.cuda() for c in classes]),
        pe.PanopticBatch(pair.gt.cuda(), pair.gc.cuda(), pair.crowd.cuda()), 5, isthing=isthing,
        checkpoints=points))
    assert pair.gc.shape[1] * 161 * 4 > 120 * 1024  # beyond the LDS stage
    for a, e in zip(actual, expected):
        compare_snapshots(a, e)


def test_checkpoint_slices_of_several_rows(pe):
    """1031 checkpoint rows over 8 images (one without masks): more rows than
    two per compute unit, so every scoring workgroup takes a slice of
    several consecutive checkpoints and moves its canvas between them."""
    import workloads
    gen = torch.Generator().manual_seed(9)
    ns = (150, 0, 140, 160, 130, 150, 145, 155)
    pair = workloads.panoptic_batch(9, len(ns), 48, 64, 14, 5, 3)
    masks = [workloads.ellipse_masks(gen, n, 48, 64, min_axis=2.0, max_axis=20.0) for n in ns]
    classes = [torch.randint(0, 5, (len(m),), generator=gen).int() for m in masks]
    isthing = [True, True, True, False, False]
    rows = sum(n if n else 1 for n in ns)
    assert rows > 2 * torch.cuda.get_device_properties(0).multi_processor_count
    expected = opre.prefixes([m.numpy() for m in masks], [c.numpy() for c in classes], pair.gt.numpy(),
                             pair.gc.numpy(), pair.crowd.numpy(), 5, isthing=isthing)
    actual = to_numpy(pe.segmentation_prefixes(
        pe.InstanceBatch([m.cuda() for m in masks], [c.cuda() for c in classes]),
        pe.PanopticBatch(pair.gt.cuda(), pair.gc.cuda(), pair.crowd.cuda()), 5, isthing=isthing))
    assert sum(len(a) for a in actual) == rows
    for a, e in zip(actual, expected):
        compare_snapshots(a, e)


def test_native_checkpoint_beyond_masks_is_reported(pe):
    """At the native boundary a checkpoint past the image's mask count is
    caught by the scoring kernel's own range check (error bit 32); the
    Python entry point rejects it before any launch."""
    from panoptic_evaluator.evaluator import _prepare
    from panoptic_evaluator.prefix import launch_prefix
    tgt = target(pe).compact()
    maps, gc, crowd = _prepare(tgt, 2, False)
    masks = [torch.zeros((3,) + tuple(maps.shape[1:]), dtype=torch.bool, device="cuda")]
    masks[0][:, 0, :2] = True
    classes = [torch.zeros(3, dtype=torch.int32, device="cuda")]
    things = torch.ones(2, dtype=torch.bool, device="cuda")
    sizes = [tuple(maps.shape[1:])] * maps.shape[0]
    out = launch_prefix(masks, classes, maps, gc, crowd, sizes, False, [[1, 5]], things, (0.5,), 2,
                        torch.device("cuda"))
    assert int(out["error"].item()) & 32
    ok = launch_prefix(masks, classes, maps, gc, crowd, sizes, False, [[1, 3]], things, (0.5,), 2,
                       torch.device("cuda"))
    assert int(ok["error"].item()) == 0


def test_events_beyond_lds_table():
    """260 target segments x 130 masks: G x (N + 1) exceeds the events
    kernel's LDS step table, so every pixel run goes to global atomics (and
    scoring keeps its state in global memory) -- same results."""
    import workloads
    import panoptic_evaluator as pe
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    gen = torch.Generator().manual_seed(23)
    pair = workloads.panoptic_batch(23, 1, 64, 64, 260, 5, 3)
    masks = [workloads.ellipse_masks(gen, 130, 64, 64, min_axis=2.0, max_axis=14.0)]
    classes = [torch.randint(0, 5, (130,), generator=gen).int()]
    isthing = [True, True, True, False, False]
    points = [0, 1, 2, 3] + list(range(10, 131, 10))
    assert pair.gc.shape[1] * 131 > 32768
    expected = opre.prefixes([m.numpy() for m in masks], [c.numpy() for c in classes], pair.gt.numpy(),
                             pair.gc.numpy(), pair.crowd.numpy(), 5, isthing=isthing, checkpoints=points)
    actual = to_numpy(pe.segmentation_prefixes(
        pe.InstanceBatch([m.cuda() for m in masks], [c.cuda() for c in classes]),
        pe.PanopticBatch(pair.gt.cuda(), pair.gc.cuda(), pair.crowd.cuda()), 5, isthing=isthing,
        checkpoints=points))
    for a, e in zip(actual, expected):
        compare_snapshots(a, e)
