"""Every launch variant of the tunable kernels computes the same statistics.

The autotuner may pick any variant, so each one is checked against the CPU
oracle (joint histogram, detection x slot histogram) or exact integer
products (MFMA mask contraction), on shapes that exercise partial tiles and
unaligned rows.
"""

import numpy as np
import pytest
import torch

from helpers import assert_exact, compare_panoptic, to_numpy
from oracle import combined as ocomb
from oracle import panoptic as opan

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pe():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import panoptic_evaluator
    return panoptic_evaluator


@pytest.fixture
def tuning(pe):
    from panoptic_evaluator import tuning
    yield tuning
    for kernel in tuning.KERNELS:
        tuning.select(kernel, -1)


def count(tuning, kernel):
    return len(tuning.variants(kernel))


@pytest.mark.parametrize("H,W", [(300, 517), (256, 512)])
def test_joint_histogram_variants(pe, tuning, H, W):
    import workloads
    pair = workloads.panoptic_batch(11, 3, H, W, 30, 12, 6)
    isthing = workloads.isthing_vector(12, 6)
    arrays = [t.numpy() for t in (pair.gt, pair.pred, pair.gc, pair.pc, pair.crowd)]
    state = opan.PanopticState(12, isthing)
    state.update(*arrays, validate=True)
    expected = state.compute()
    for variant in range(count(tuning, "joint_histogram")):
        tuning.select("joint_histogram", variant)
        assert tuning.selected("joint_histogram") == variant
        ev = pe.PanopticEvaluator(12, isthing=isthing)
        ev.update(pe.PanopticBatch(pair.pred.cuda(), pair.pc.cuda()),
                  pe.PanopticBatch(pair.gt.cuda(), pair.gc.cuda(), pair.crowd.cuda()))
        compare_panoptic(to_numpy(ev.compute()), expected)


@pytest.mark.parametrize("H,W", [(200, 333), (160, 256)])
def test_detection_histogram_variants(pe, tuning, H, W):
    import workloads
    pair, masks, classes, scores = workloads.config1(device="cpu", seed=5, B=2, H=H, W=W, detections=37)
    isthing = workloads.isthing_vector(133, 80)
    state = ocomb.SegmentationState(133, isthing)
    state.update(pair.pred.numpy(), pair.pc.numpy(),
                 [dict(masks=m.numpy(), classes=c.numpy(), scores=s.numpy()) for m, c, s in zip(masks, classes, scores)],
                 pair.gt.numpy(), pair.gc.numpy(), pair.crowd.numpy())
    expected = state.compute()
    dets = pe.InstanceBatch([m.cuda() for m in masks], [c.cuda() for c in classes], [s.cuda() for s in scores])
    for variant in range(count(tuning, "detection_histogram")):
        tuning.select("detection_histogram", variant)
        joint = pe.SegmentationEvaluator(133, isthing=isthing)
        joint.update(pe.PanopticBatch(pair.pred.cuda(), pair.pc.cuda()), dets,
                     pe.PanopticBatch(pair.gt.cuda(), pair.gc.cuda(), pair.crowd.cuda()))
        result = to_numpy(joint.compute())
        compare_panoptic(result, expected)
        for key in ("precision", "recall"):
            assert_exact(result["instance"][key], expected["instance"][key], f"{key} variant {variant}")


@pytest.mark.parametrize("D,G,H,W", [(17, 140, 24, 37), (33, 20, 160, 300), (150, 70, 40, 52), (0, 9, 30, 31)])
def test_mask_overlap_variants(pe, tuning, D, G, H, W):
    from panoptic_evaluator import _native
    rng = np.random.default_rng(D + G)
    pm = rng.random((D, H, W)) < 0.3
    gm = rng.random((G, H, W)) < 0.4
    pflat = pm.reshape(D, H * W).astype(np.int64)
    gflat = gm.reshape(G, H * W).astype(np.int64)
    pm_dev, gm_dev = torch.from_numpy(pm).cuda(), torch.from_numpy(gm).cuda()
    for variant in range(count(tuning, "mask_overlap_mfma")):
        tuning.select("mask_overlap_mfma", variant)
        inter = torch.empty(D * G, dtype=torch.int32, device="cuda")
        pa = torch.empty(D, dtype=torch.int32, device="cuda")
        ga = torch.empty(G, dtype=torch.int32, device="cuda")
        _native.call("pe_mask_overlap", pm_dev.data_ptr(), D, gm_dev.data_ptr(), G, H * W, inter.data_ptr(),
                     pa.data_ptr(), ga.data_ptr(), _native.stream())
        assert_exact(inter.view(D, G).long(), pflat @ gflat.T, f"variant {variant}")
        assert_exact(pa.long(), pflat.sum(1))
        assert_exact(ga.long(), gflat.sum(1))


def test_autotune_keeps_results(pe, tuning):
    import workloads
    pair = workloads.config0(B=2)
    isthing = workloads.isthing_vector(133, 80)
    ev = pe.PanopticEvaluator(133, isthing=isthing, validate=False)
    p = pe.PanopticBatch(pair.pred.cuda(), pair.pc.cuda())
    g = pe.PanopticBatch(pair.gt.cuda(), pair.gc.cuda(), pair.crowd.cuda())

    def step():
        ev.reset()
        ev.update(p, g)

    step()
    before = to_numpy(ev.compute())
    report = tuning.autotune(step, ["joint_histogram", "detection_histogram"], repeats=2)
    assert 0 <= report["selection"]["joint_histogram"] < count(tuning, "joint_histogram")
    assert all(ms is not None and ms > 0 for ms in report["timings"]["joint_histogram"])
    # The step never launches the detection histogram: one probe, no change.
    assert report["timings"]["detection_histogram"] == [None]
    step()
    compare_panoptic(to_numpy(ev.compute()), before, exact_iou_sum=True)


def test_packed_grid_variants(pe, tuning):
    """Persistent and one-workgroup-per-item grids over a packed batch."""
    import workloads
    batch = workloads.config2(seed=9, B=5, sizes=(2, 6), segments=(3, 25), detections=(0, 40))
    isthing = workloads.isthing_vector(133, 80)
    state = ocomb.SegmentationState(133, isthing)
    for i in range(len(batch.sizes)):
        state.update(batch.pred[i][None].numpy(), batch.pc[i][None].numpy(),
                     [dict(masks=batch.masks[i].numpy(), classes=batch.classes[i].numpy(),
                           scores=batch.scores[i].numpy())],
                     batch.gt[i][None].numpy(), batch.gc[i][None].numpy(), batch.crowd[i][None].numpy())
    expected = state.compute()
    cuda = lambda xs: [x.cuda() for x in xs]  # noqa: E731
    target = pe.PanopticBatch.pack(cuda(batch.gt), cuda(batch.gc), cuda(batch.crowd))
    prediction = pe.PanopticBatch.pack(cuda(batch.pred), cuda(batch.pc))
    dets = pe.InstanceBatch(cuda(batch.masks), cuda(batch.classes), cuda(batch.scores))
    for grid in range(count(tuning, "packed_grid")):
        for det, group in ((0, 0), (count(tuning, "detection_histogram") - 1, 3), (2, 2)):
            tuning.select("packed_grid", grid)
            tuning.select("detection_histogram", det)
            tuning.select("detection_group", group)
            joint = pe.SegmentationEvaluator(133, isthing=isthing)
            joint.update(prediction, dets, target)
            result = to_numpy(joint.compute())
            compare_panoptic(result, expected)
            for key in ("precision", "recall"):
                assert_exact(result["instance"][key], expected["instance"][key], f"{key} grid {grid}")
