"""Batched form: packed batches of uneven images and HIP-graph-captured updates.

A packed batch must give exactly the statistics of evaluating its images one
at a time (integers bit-exact, FP64 within 1e-12 relative; the dataset iou_sum
is exact, so it is identical), and every replay of a captured update must equal
the eager update it recorded.
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


def uneven(seed, B, sizes=(2, 5), segments=(3, 20), detections=(0, 12), odd=False):
    import workloads
    batch = workloads.config2(seed=seed, B=B, sizes=sizes, segments=segments, detections=detections)
    if odd:
        # Sizes that are not multiples of 4 or 16 (scalar head/tail paths).
        for i, (h, w) in enumerate(batch.sizes):
            h2, w2 = h - 3 - i % 5, w - 1 - i % 3
            batch.sizes[i] = (h2, w2)
            for name in ("gt", "pred"):
                getattr(batch, name)[i] = getattr(batch, name)[i][:h2, :w2].contiguous()
            batch.masks[i] = batch.masks[i][:, :h2, :w2].contiguous()
            # Slots that lost every pixel become inactive metadata.
            for maps, meta in ((batch.gt, batch.gc), (batch.pred, batch.pc)):
                area = torch.bincount(maps[i].flatten().long(), minlength=meta[i].numel())
                meta[i] = torch.where(area[:meta[i].numel()] > 0, meta[i], -1).int()
                meta[i][0] = -1
            batch.crowd[i] = batch.crowd[i] & (batch.gc[i] >= 0)
    return batch


def oracle_panoptic(batch, C, isthing):
    state = opan.PanopticState(C, isthing)
    for i in range(len(batch.sizes)):
        state.update(batch.gt[i][None].numpy(), batch.pred[i][None].numpy(), batch.gc[i][None].numpy(),
                     batch.pc[i][None].numpy(), batch.crowd[i][None].numpy(), validate=True)
    return state.compute()


def packed(pe, batch, device="cuda"):
    cuda = lambda xs: [x.to(device) for x in xs]  # noqa: E731
    target = pe.PanopticBatch.pack(cuda(batch.gt), cuda(batch.gc), cuda(batch.crowd))
    prediction = pe.PanopticBatch.pack(cuda(batch.pred), cuda(batch.pc))
    return prediction, target


@pytest.mark.parametrize("odd", [False, True])
def test_packed_panoptic_equals_per_image_oracle(pe, odd):
    batch = uneven(21, 7, odd=odd)
    isthing = [c < 80 for c in range(133)]
    prediction, target = packed(pe, batch)
    for validate in (True, False):
        ev = pe.PanopticEvaluator(133, isthing=isthing, validate=validate)
        ev.update(prediction, target)
        compare_panoptic(to_numpy(ev.compute()), oracle_panoptic(batch, 133, isthing))


def test_packed_panoptic_matches_plain_loop_exactly(pe):
    batch = uneven(5, 9, sizes=(3, 9))
    isthing = [c < 80 for c in range(133)]
    prediction, target = packed(pe, batch)
    batched = pe.PanopticEvaluator(133, isthing=isthing, validate=False)
    batched.update(prediction, target)
    plain = pe.PanopticEvaluator(133, isthing=isthing, validate=False)
    for i in range(len(batch.sizes)):
        plain.update(pe.PanopticBatch(batch.pred[i][None].cuda(), batch.pc[i][None].cuda()),
                     pe.PanopticBatch(batch.gt[i][None].cuda(), batch.gc[i][None].cuda(),
                                      batch.crowd[i][None].cuda()))
    compare_panoptic(to_numpy(batched.compute()), to_numpy(plain.compute()), exact_iou_sum=True)


def test_packed_panoptic_quality(pe):
    batch = uneven(8, 5, odd=True)
    prediction, target = packed(pe, batch)
    got = pe.panoptic_quality(prediction, target, 133).cpu().numpy()
    want = np.concatenate([opan.panoptic_quality(batch.gt[i][None].numpy(), batch.pred[i][None].numpy(),
                                                 batch.gc[i][None].numpy(), batch.pc[i][None].numpy(),
                                                 batch.crowd[i][None].numpy(), 133)
                           for i in range(len(batch.sizes))])
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-15)


def test_packed_validation_errors(pe):
    batch = uneven(3, 4)
    batch.gc[2] = batch.gc[2].clone()
    batch.gc[2][1] = 500  # class outside 0..C-1
    prediction, target = packed(pe, batch)
    ev = pe.PanopticEvaluator(133)
    with pytest.raises(ValueError, match="classes"):
        ev.update(prediction, target)
    assert ev.tp.sum().item() == 0


@pytest.mark.parametrize("odd", [False, True])
def test_packed_combined_equals_per_image_oracle(pe, odd):
    batch = uneven(13, 6, detections=(0, 25), odd=odd)
    isthing = [c < 80 for c in range(133)]
    prediction, target = packed(pe, batch)
    dets = pe.InstanceBatch([m.cuda() for m in batch.masks], [c.cuda() for c in batch.classes],
                            [s.cuda() for s in batch.scores])
    joint = pe.SegmentationEvaluator(133, isthing=isthing, validate=True)
    joint.update(prediction, dets, target)
    state = ocomb.SegmentationState(133, isthing)
    for i in range(len(batch.sizes)):
        state.update(batch.pred[i][None].numpy(), batch.pc[i][None].numpy(),
                     [dict(masks=batch.masks[i].numpy(), classes=batch.classes[i].numpy(),
                           scores=batch.scores[i].numpy())],
                     batch.gt[i][None].numpy(), batch.gc[i][None].numpy(), batch.crowd[i][None].numpy())
    result, expected = to_numpy(joint.compute()), state.compute()
    compare_panoptic(result, expected)
    for key in ("precision", "recall"):
        assert_exact(result["instance"][key], expected["instance"][key], key)


def _replay_matches_eager(pe, make, args, replays=3):
    from panoptic_evaluator.graphs import CapturedUpdate
    eager = make()
    eager.update(*args)
    want = to_numpy(eager.compute())
    captured_ev = make()
    captured = CapturedUpdate(captured_ev, *args)
    # Capture leaves the state untouched.
    assert to_numpy(captured_ev.compute())["tp"].sum() == 0
    for _ in range(replays):
        captured_ev.reset()
        # eager work and a synchronisation between replays
        scratch = torch.randn(1 << 20, device="cuda").sum()
        torch.cuda.synchronize()
        captured.replay()
        got = to_numpy(captured_ev.compute())
        assert scratch.isfinite()
        yield got, want


def test_graph_replays_equal_eager_panoptic(pe):
    import workloads
    pair = workloads.config0(B=3)
    isthing = workloads.isthing_vector(133, 80)
    args = (pe.PanopticBatch(pair.pred.cuda(), pair.pc.cuda()),
            pe.PanopticBatch(pair.gt.cuda(), pair.gc.cuda(), pair.crowd.cuda()))
    make = lambda: pe.PanopticEvaluator(133, isthing=isthing, validate=False)  # noqa: E731
    for got, want in _replay_matches_eager(pe, make, args):
        compare_panoptic(got, want, exact_iou_sum=True)


def test_graph_replays_equal_eager_packed_combined(pe):
    batch = uneven(17, 5, detections=(0, 30))
    isthing = [c < 80 for c in range(133)]
    prediction, target = packed(pe, batch)
    dets = pe.InstanceBatch([m.cuda() for m in batch.masks], [c.cuda() for c in batch.classes],
                            [s.cuda() for s in batch.scores])
    make = lambda: pe.SegmentationEvaluator(133, isthing=isthing, validate=False)  # noqa: E731
    for got, want in _replay_matches_eager(pe, make, (prediction, dets, target)):
        compare_panoptic(got, want, exact_iou_sum=True)
        for key in ("precision", "recall"):
            assert_exact(got["instance"][key], want["instance"][key], key)


def test_graph_replay_accumulates_like_repeated_updates(pe):
    """Two replays without reset == two eager updates (records appended twice)."""
    from panoptic_evaluator.graphs import CapturedUpdate
    batch = uneven(19, 4, detections=(1, 20))
    isthing = [c < 80 for c in range(133)]
    prediction, target = packed(pe, batch)
    dets = pe.InstanceBatch([m.cuda() for m in batch.masks], [c.cuda() for c in batch.classes],
                            [s.cuda() for s in batch.scores])
    eager = pe.SegmentationEvaluator(133, isthing=isthing, validate=False)
    eager.update(prediction, dets, target)
    eager.update(prediction, dets, target)
    ev = pe.SegmentationEvaluator(133, isthing=isthing, validate=False)
    captured = CapturedUpdate(ev, prediction, dets, target)
    captured.replay()
    captured.replay()
    got, want = to_numpy(ev.compute()), to_numpy(eager.compute())
    compare_panoptic(got, want, exact_iou_sum=True)
    for key in ("precision", "recall"):
        assert_exact(got["instance"][key], want["instance"][key], key)


def test_graph_replay_keeps_records_of_refilled_inputs(pe):
    """Inputs refilled in place between replays: the first replay's detection
    records must survive the second replay (they are copied out of graph
    memory before it overwrites them)."""
    from panoptic_evaluator.graphs import CapturedUpdate
    batch = uneven(23, 4, detections=(2, 20))
    isthing = [c < 80 for c in range(133)]
    prediction, target = packed(pe, batch)
    scores = [s.cuda() for s in batch.scores]
    dets = pe.InstanceBatch([m.cuda() for m in batch.masks], [c.cuda() for c in batch.classes], scores)
    ev = pe.SegmentationEvaluator(133, isthing=isthing, validate=False)
    captured = CapturedUpdate(ev, prediction, dets, target)
    eager = pe.SegmentationEvaluator(133, isthing=isthing, validate=False)
    captured.replay()
    eager.update(prediction, dets, target)
    for s in scores:
        s.copy_(1.0 - s)  # reverse every image's score order in place
    captured.replay()
    eager.update(prediction, dets, target)
    got, want = to_numpy(ev.compute()), to_numpy(eager.compute())
    for key in ("precision", "recall"):
        assert_exact(got["instance"][key], want["instance"][key], key)


def test_capture_rejects_validation(pe):
    from panoptic_evaluator.graphs import CapturedUpdate
    import workloads
    pair = workloads.config0(B=1)
    ev = pe.PanopticEvaluator(133, validate=True)
    with pytest.raises(ValueError, match="validate=False"):
        CapturedUpdate(ev, pe.PanopticBatch(pair.pred.cuda(), pair.pc.cuda()),
                       pe.PanopticBatch(pair.gt.cuda(), pair.gc.cuda()))


def test_packed_prefix_equals_per_image_oracle(pe):
    """Prefix rewards over a packed batch of uneven images (odd sizes too)."""
    import workloads
    from oracle import prefix as opre
    from test_prefix_gpu import compare_snapshots
    batch = uneven(23, 4, sizes=(2, 4), segments=(3, 10), odd=True)
    gen = torch.Generator().manual_seed(3)
    masks = [workloads.ellipse_masks(gen, n, h, w, max_axis=40.0) for n, (h, w) in zip((7, 0, 12, 5), batch.sizes)]
    classes = [torch.randint(0, 133, (len(m),), generator=gen).int() for m in masks]
    isthing = [c < 80 for c in range(133)]
    target = pe.PanopticBatch.pack([g.cuda() for g in batch.gt], [c.cuda() for c in batch.gc],
                                   [c.cuda() for c in batch.crowd])
    for points in (None, [0]):
        actual = to_numpy(pe.segmentation_prefixes(pe.InstanceBatch([m.cuda() for m in masks],
                                                                    [c.cuda() for c in classes]),
                                                   target, 133, isthing=isthing, checkpoints=points))
        for i in range(len(masks)):
            expected = opre.prefixes([masks[i].numpy()], [classes[i].numpy()], batch.gt[i][None].numpy(),
                                     batch.gc[i][None].numpy(), batch.crowd[i][None].numpy(), 133,
                                     isthing=isthing, checkpoints=points)
            compare_snapshots(actual[i], expected[0])


def test_device_tables_survive_eviction_across_streams(pe):
    """Tables handed out on two streams keep their values while the cache
    evicts and reallocates around them (evicted entries are marked as used by
    every stream that requested them)."""
    from panoptic_evaluator import _native
    side = torch.cuda.Stream()
    held = []
    for i in range(3 * _native._TABLE_CACHE):
        values = [i, i + 1, 7 * i]
        if i % 2:
            with torch.cuda.stream(side):
                table = _native.device_table(values, "cuda")
        else:
            table = _native.device_table(values, "cuda")
        if i % 97 == 0:
            held.append((values, table))
    torch.cuda.synchronize()
    for values, table in held:
        assert table.cpu().tolist() == values


def test_combined_fork_equals_sequential_at_c1_shape(pe):
    """A captured update holds the panoptic branch on a second stream; two
    replays equal two updates of the validating (one-stream) evaluator."""
    import workloads
    from panoptic_evaluator.graphs import CapturedUpdate
    pair, masks, classes, scores = workloads.config1("cuda", seed=3, B=2)
    isthing = [c < 80 for c in range(133)]
    args = (pe.PanopticBatch(pair.pred, pair.pc), pe.InstanceBatch(masks, classes, scores),
            pe.PanopticBatch(pair.gt, pair.gc, pair.crowd))
    sequential = pe.SegmentationEvaluator(133, isthing=isthing, validate=True)
    sequential.update(*args)
    sequential.update(*args)
    forked = pe.SegmentationEvaluator(133, isthing=isthing, validate=False)
    captured = CapturedUpdate(forked, *args)
    captured.replay()
    torch.cuda.synchronize()
    captured.replay()
    got, want = to_numpy(forked.compute()), to_numpy(sequential.compute())
    compare_panoptic(got, want, exact_iou_sum=True)
    for key in ("precision", "recall"):
        assert_exact(got["instance"][key], want["instance"][key], key)


def test_captured_update_survives_table_cache_turnover(pe):
    """A captured update reads device metadata tables by address.  After the
    cache has handed out and evicted more than three times its capacity of new
    tables (whose memory the allocator reuses), replays still equal eager
    updates: the graph keeps its own tables."""
    from panoptic_evaluator import _native
    from panoptic_evaluator.graphs import CapturedUpdate
    batch = uneven(31, 5, detections=(1, 9))
    isthing = [c < 80 for c in range(133)]
    prediction, target = packed(pe, batch)
    dets = pe.InstanceBatch([m.cuda() for m in batch.masks], [c.cuda() for c in batch.classes],
                            [s.cuda() for s in batch.scores])
    eager = pe.SegmentationEvaluator(133, isthing=isthing, validate=False)
    eager.update(prediction, dets, target)
    want = to_numpy(eager.compute())
    replayed = pe.SegmentationEvaluator(133, isthing=isthing, validate=False)
    captured = CapturedUpdate(replayed, prediction, dets, target)
    churn = []
    for i in range(3 * _native._TABLE_CACHE + 7):
        churn.append(torch.full((4096,), -1, dtype=torch.int64, device="cuda"))  # reuse freed blocks
        _native.device_table([10 ** 9 + i, i, 2 * i, 5], "cuda")
        if len(churn) > 8:
            churn.pop(0)
    torch.cuda.synchronize()
    for _ in range(2):
        replayed.reset()
        captured.replay()
        got = to_numpy(replayed.compute())
        compare_panoptic(got, want, exact_iou_sum=True)
        for key in ("precision", "recall"):
            assert_exact(got["instance"][key], want["instance"][key], key)


# ---------------------------------------------------------------- whole step

def _step_replays_match_eager(pe, make, args, compare, replays=3):
    """CapturedStep: every replay (eager work and a synchronisation between
    replays) returns exactly the eager reset/update/compute result; after a
    replay the evaluator's own compute() and a further eager update agree
    with the eager sequence too."""
    from panoptic_evaluator.graphs import CapturedStep
    eager = make()
    eager.update(*args)
    want = to_numpy(eager.compute())
    ev = make()
    step = CapturedStep(ev, *args)
    for _ in range(replays):
        scratch = torch.randn(1 << 20, device="cuda").sum()
        torch.cuda.synchronize()
        got = to_numpy(step.replay())
        assert scratch.isfinite()
        compare(got, want)
    compare(to_numpy(ev.compute()), want)
    ev.update(*args)
    eager.update(*args)
    compare(to_numpy(ev.compute()), to_numpy(eager.compute()))
    # A replay starts from reset, whatever the evaluator held.
    compare(to_numpy(step.replay()), want)


def test_step_graph_panoptic(pe):
    import workloads
    pair = workloads.config0(B=3)
    isthing = workloads.isthing_vector(133, 80)
    args = (pe.PanopticBatch(pair.pred.cuda(), pair.pc.cuda()),
            pe.PanopticBatch(pair.gt.cuda(), pair.gc.cuda(), pair.crowd.cuda()))
    make = lambda: pe.PanopticEvaluator(133, isthing=isthing, validate=False)  # noqa: E731
    _step_replays_match_eager(pe, make, args, lambda got, want: compare_panoptic(got, want, exact_iou_sum=True))


def _compare_combined_exact(got, want):
    compare_panoptic(got, want, exact_iou_sum=True)
    for key in ("precision", "recall", "ap", "ar100"):
        assert_exact(got["instance"][key], want["instance"][key], key)


def test_step_graph_combined_uniform(pe):
    import workloads
    pair, masks, classes, scores = workloads.config1(device="cuda", seed=4, B=3, H=160, W=224, detections=40)
    isthing = workloads.isthing_vector(133, 80)
    args = (pe.PanopticBatch(pair.pred, pair.pc), pe.InstanceBatch(masks, classes, scores),
            pe.PanopticBatch(pair.gt, pair.gc, pair.crowd))
    make = lambda: pe.SegmentationEvaluator(133, isthing=isthing, validate=False)  # noqa: E731
    _step_replays_match_eager(pe, make, args, _compare_combined_exact)


def test_step_graph_combined_packed(pe):
    batch = uneven(29, 5, detections=(0, 30))
    isthing = [c < 80 for c in range(133)]
    prediction, target = packed(pe, batch)
    dets = pe.InstanceBatch([m.cuda() for m in batch.masks], [c.cuda() for c in batch.classes],
                            [s.cuda() for s in batch.scores])
    make = lambda: pe.SegmentationEvaluator(133, isthing=isthing, validate=False)  # noqa: E731
    _step_replays_match_eager(pe, make, (prediction, dets, target), _compare_combined_exact)


def test_step_graph_instance(pe):
    import workloads
    pair, masks, classes, scores = workloads.config1(device="cuda", seed=6, B=2, H=96, W=128, detections=25)
    gts = [(pair.gt[b][None] == torch.arange(1, 9, device="cuda")[:, None, None]) for b in range(2)]
    keep = [m.flatten(1).any(1) for m in gts]
    gts = [m[k] for m, k in zip(gts, keep)]
    gcl = [torch.arange(len(m), device="cuda") % 7 for m in gts]
    args = (pe.InstanceBatch(masks, classes, scores), pe.InstanceBatch(gts, gcl))
    make = lambda: pe.InstanceEvaluator(133, device="cuda", validate=False)  # noqa: E731

    def compare(got, want):
        for key in ("precision", "recall", "ap", "ar100"):
            assert_exact(got[key], want[key], key)
    _step_replays_match_eager(pe, make, args, compare)


def test_step_graph_with_refilled_inputs(pe):
    """Inputs refilled in place between replays: each replay evaluates the
    buffers' current contents."""
    from panoptic_evaluator.graphs import CapturedStep
    import workloads
    pair, masks, classes, scores = workloads.config1(device="cuda", seed=8, B=2, H=128, W=160, detections=30)
    isthing = workloads.isthing_vector(133, 80)
    args = (pe.PanopticBatch(pair.pred, pair.pc), pe.InstanceBatch(masks, classes, scores),
            pe.PanopticBatch(pair.gt, pair.gc, pair.crowd))
    ev = pe.SegmentationEvaluator(133, isthing=isthing, validate=False)
    step = CapturedStep(ev, *args)
    for round_ in range(3):
        for s in scores:
            s.copy_(torch.rand_like(s))
        pair.pred.copy_(torch.roll(pair.pred, shifts=round_ + 1, dims=2))
        eager = pe.SegmentationEvaluator(133, isthing=isthing, validate=False)
        eager.update(*args)
        _compare_combined_exact(to_numpy(step.replay()), to_numpy(eager.compute()))
