"""The drop-in module ``panoptic_evaluator._cuda_evaluator`` driven the way the
reference's Python files drive their pybind11 extension.

Each test restates one call sequence of the reference (quoted by file:line)
with plain torch tensors, calls the shim's function of the same name, and
checks the result against the oracle: integers bit-exact, FP64 within 1e-12
relative.  The reference itself is not imported.
"""

import numpy as np
import pytest
import torch

from helpers import assert_close, assert_exact, compare_panoptic, to_numpy
from oracle import combined as ocomb
from oracle import instance as oinst
from oracle import panoptic as opan
from oracle import prefix as opre

pytestmark = pytest.mark.gpu

C = 7
ISTHING = [True, True, True, False, True, False, False]


@pytest.fixture(scope="module")
def shim():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from panoptic_evaluator import _cuda_evaluator
    return _cuda_evaluator


def cuda(x):
    return torch.as_tensor(np.asarray(x)).cuda()


def panoptic_inputs(seed, B=3, H=41, W=57, G=9, P=12):
    rng = np.random.default_rng(seed)
    gt = rng.integers(0, G, (B, H, W)).astype(np.int32)
    pred = np.where(rng.random((B, H, W)) < 0.75, gt, rng.integers(0, P, (B, H, W))).astype(np.int32)
    gc = rng.integers(0, C, (B, G)).astype(np.int32)
    pc = np.concatenate([gc, rng.integers(0, C, (B, P - G))], 1).astype(np.int32)
    gc[:, 0] = pc[:, 0] = -1
    crowd = (rng.random((B, G)) < 0.2) & (gc >= 0)
    for maps, classes in ((gt, gc), (pred, pc)):
        for b in range(B):
            present = np.bincount(maps[b].ravel(), minlength=classes.shape[1]) > 0
            classes[b, 1:] = np.where(present[1:], classes[b, 1:], -1)
    crowd &= gc >= 0
    return gt, pred, gc, pc, crowd


class ReferenceState:
    """The reference's PanopticEvaluator state and scratch on the GPU
    (evaluator.py:280-285, :322-329)."""

    def __init__(self):
        device = "cuda"
        self.tp = torch.zeros(C, dtype=torch.int64, device=device)
        self.fp = torch.zeros_like(self.tp)
        self.fn = torch.zeros_like(self.tp)
        self.iou_sum = torch.zeros(C, dtype=torch.float64, device=device)
        self.confusion = torch.zeros((C, C + 1), dtype=torch.int64, device=device)
        self.error = torch.zeros(1, dtype=torch.int32, device=device)
        self.isthing = torch.tensor(ISTHING, dtype=torch.bool, device=device)

    def reset(self):
        # evaluator.py:274-278: the five state tensors are zeroed in place.
        for t in (self.tp, self.fp, self.fn, self.iou_sum, self.confusion):
            t.zero_()

    def update(self, shim, gt, pred, gc, pc, crowd, validate):
        B, G, P = gt.shape[0], gc.shape[1], pc.shape[1]
        # Fresh uninitialised scratch, as the first update of a shape gets.
        scratch = (torch.full((B, G, P), 7, dtype=torch.int64, device="cuda"),
                   torch.empty((B, G + P), dtype=torch.int64, device="cuda"),
                   torch.empty((B, G), dtype=torch.int32, device="cuda"))
        return shim.update(cuda(gt), cuda(pred), cuda(gc), cuda(pc), cuda(crowd), *scratch, self.tp, self.fp,
                           self.fn, self.iou_sum, self.confusion, None, None, None, None, self.error, validate)

    def compute(self, shim):
        return shim.compute(self.tp, self.fp, self.fn, self.iou_sum, self.confusion, self.isthing)


def result_dict(outputs):
    """evaluator.py:424-447: the compute binding's nine outputs as a result."""
    per_class, summaries, counts, miou, tp, fp, fn, sums, confusion = to_numpy(outputs)
    result = {name: {"pq": summaries[i, 0], "sq": summaries[i, 1], "rq": summaries[i, 2], "n": counts[i]}
              for i, name in enumerate(("All", "Things", "Stuff"))}
    result.update(per_class=dict(zip(("pq", "sq", "rq", "iou"), per_class)), miou=miou, tp=tp, fp=fp, fn=fn,
                  iou_sum=sums, confusion=confusion)
    return result


@pytest.mark.parametrize("validate", [True, False])
def test_update_reset_update(shim, validate):
    """update -> reset (zero_ of the caller's tensors) -> update -> compute:
    the totals are those of the second batch alone."""
    first, second = panoptic_inputs(1), panoptic_inputs(2)
    state = ReferenceState()
    assert state.update(shim, *first, validate) == 0
    state.reset()
    assert state.update(shim, *second, validate) == 0
    oracle = opan.PanopticState(C, ISTHING)
    oracle.update(*second, validate=True)
    compare_panoptic(result_dict(state.compute(shim)), oracle.compute())


def test_updates_accumulate_and_snapshots_are_copies(shim):
    batches = [panoptic_inputs(seed) for seed in (3, 4, 5)]
    state = ReferenceState()
    oracle = opan.PanopticState(C, ISTHING)
    for batch in batches:
        assert state.update(shim, *batch, True) == 0
        oracle.update(*batch, validate=True)
    outputs = state.compute(shim)
    compare_panoptic(result_dict(outputs), oracle.compute())
    state.reset()
    assert int(outputs[4].sum().item()) == int(oracle.tp.sum())  # the snapshot survives reset


def test_validation_error_leaves_totals(shim):
    gt, pred, gc, pc, crowd = panoptic_inputs(6)
    state = ReferenceState()
    assert state.update(shim, gt, pred, gc, pc, crowd, True) == 0
    before = [t.clone() for t in (state.tp, state.fp, state.fn, state.iou_sum, state.confusion)]
    bad = gc.copy()
    bad[0, 1] = C + 3
    assert state.update(shim, gt, pred, bad, pc, crowd, True) == 2
    for old, new in zip(before, (state.tp, state.fp, state.fn, state.iou_sum, state.confusion)):
        assert torch.equal(old, new)


def test_argument_errors_raise_runtime_error(shim):
    gt, pred, gc, pc, crowd = panoptic_inputs(7)
    state = ReferenceState()
    with pytest.raises(RuntimeError):
        shim.update(cuda(gt), cuda(pred).long().float(), cuda(gc), cuda(pc), cuda(crowd),
                    torch.empty(1, device="cuda"), None, None, state.tp, state.fp, state.fn, state.iou_sum,
                    state.confusion, None, None, None, None, state.error, True)


def test_reward_mode(shim):
    """panoptic_quality's call (evaluator.py:503-513): per-image statistics."""
    gt, pred, gc, pc, crowd = panoptic_inputs(8, B=4)
    B, G, P = gt.shape[0], gc.shape[1], pc.shape[1]
    tp = torch.zeros((B, C), dtype=torch.int64, device="cuda")
    fp, fn = torch.zeros_like(tp), torch.zeros_like(tp)
    sums = torch.zeros_like(tp, dtype=torch.float64)
    code = shim.update(cuda(gt), cuda(pred), cuda(gc), cuda(pc), cuda(crowd),
                       torch.empty((B, G, P), dtype=torch.int64, device="cuda"),
                       torch.empty((B, G + P), dtype=torch.int64, device="cuda"),
                       torch.empty((B, G), dtype=torch.int32, device="cuda"), tp, fp, fn, sums,
                       torch.empty(0, dtype=torch.int64, device="cuda"), None, None, None, None,
                       torch.empty(1, dtype=torch.int32, device="cuda"), True, True)
    assert code == 0
    want = opan.panoptic_quality(gt, pred, gc, pc, crowd, C, return_stats=True)
    assert_close(sums, want[..., 0], what="sums")
    assert_exact(tp, want[..., 1].astype(np.int64), "tp")
    assert_exact(fp, want[..., 2].astype(np.int64), "fp")
    assert_exact(fn, want[..., 3].astype(np.int64), "fn")


def test_pack(shim):
    """PanopticBatch.compact (evaluator.py:140-145)."""
    rng = np.random.default_rng(9)
    ids = np.array([[0, 5, 17, 400], [0, 3, 9, 1 << 40]], np.int64)
    slots = np.array([[0, 2, 1, 3], [0, 1, 3, 2]], np.int32)
    maps = np.stack([rng.choice(ids[b], (13, 21)) for b in range(2)]).astype(np.int64)
    packed, error = shim.pack(cuda(maps), cuda(ids), cuda(slots))
    assert error.item() == 0
    assert_exact(packed, opan.compact(maps, ids, slots), "packed")
    maps[1, 2, 3] = 12345
    packed, error = shim.pack(cuda(maps), cuda(ids), cuda(slots))
    assert error.item() == 1


def detections(seed, D=23, G=11, H=33, W=47):
    rng = np.random.default_rng(seed)
    pm = rng.random((D, H, W)) < 0.2
    gm = rng.random((G, H, W)) < 0.25
    pm[:, 0, 0] = True
    gm[:, 0, 1] = True
    gm[1] = pm[2]  # exact matches exist
    gm[4] = pm[5] | pm[6]
    pc = rng.integers(0, C, D).astype(np.int32)
    pc[2] = rng.integers(0, C)
    gc = rng.integers(0, C, G).astype(np.int32)
    gc[1] = pc[2]
    scores = rng.random(D)
    crowd = rng.random(G) < 0.2
    annotation = rng.random(G) * 12000.0
    return pm, gm, pc, gc, scores, crowd, annotation


def test_instance_update_and_compute(shim):
    """InstanceEvaluator.update/compute (instance.py:171-227): masks in score
    order, f64 areas, then accumulation over records in global score order."""
    records, counts_total = [], np.zeros((C, 4), np.int64)
    oracle_records, oracle_counts = [], np.zeros((C, 4), np.int64)
    for seed in (10, 11):
        pm, gm, pc, gc, scores, crowd, annotation = detections(seed)
        order = oinst.score_order(scores)
        pm, pc, scores = pm[order], pc[order], scores[order]
        pa = pm.reshape(len(pm), -1).sum(1).astype(np.float64)
        ga = gm.reshape(len(gm), -1).sum(1).astype(np.float64)
        flags, ranks, counts = shim.instance_update(cuda(pm), cuda(gm), cuda(pc), cuda(gc), cuda(pa), cuda(ga),
                                                    cuda(crowd), cuda(annotation), C)
        ious = oinst.mask_ious(pm, gm, pa, ga, crowd)
        want = oinst.match_ious(ious, pc, gc, pa, crowd, annotation, C)
        assert_exact(flags, want[0], "flags")
        assert_exact(ranks, want[1], "ranks")
        assert_exact(counts, want[2], "counts")
        records.append((scores, pc, ranks, flags))
        counts_total += to_numpy(counts)
        oracle_records.append((scores, pc, want[1], want[0]))
        oracle_counts += want[2]
    # instance.py:209-226: one global stable score sort, then accumulation.
    scores = np.concatenate([r[0] for r in oracle_records])
    order = oinst.score_order(scores)
    classes = np.concatenate([r[1] for r in oracle_records])[order]
    ranks = np.concatenate([r[2] for r in oracle_records])[order]
    status = np.concatenate([r[3] for r in oracle_records], 1)[:, order]
    index = cuda(order)
    precision, recall = shim.instance_compute(
        torch.cat([cuda(r[1]) for r in records])[index].contiguous(),
        torch.cat([r[2] for r in records])[index].contiguous(),
        torch.cat([r[3] for r in records], 1)[:, index].contiguous(), cuda(counts_total))
    want_p, want_r = oinst.accumulate(classes, ranks, status, oracle_counts)
    assert_exact(precision, want_p, "precision")
    assert_exact(recall, want_r, "recall")


def test_instance_histogram_and_match(shim):
    """SegmentationEvaluator.update (combined.py:90-141): detection x slot
    histogram, then matching against the target's thing slots."""
    rng = np.random.default_rng(12)
    H, W, S, D = 37, 53, 10, 19
    target = rng.integers(0, S, (H, W)).astype(np.int32)
    classes = rng.integers(0, C, S).astype(np.int32)
    classes[0] = -1
    crowd = (rng.random(S) < 0.2) & (classes >= 0)
    masks = rng.random((D, H, W)) < 0.3
    masks[3] = target == 4
    det_classes = rng.choice([c for c in range(C) if ISTHING[c]], D).astype(np.int32)
    det_classes[3] = classes[4] if classes[4] >= 0 and ISTHING[classes[4]] else det_classes[3]
    order = oinst.score_order(rng.random(D))
    pairs = shim.instance_histogram(cuda(masks), cuda(target), cuda(order.astype(np.int64)), S)
    want = ocomb.detection_pairs(masks, target, order, S)
    assert_exact(pairs, want, "pairs")
    pc = det_classes[order]
    pa = want.sum(1).astype(np.float64)
    ga = np.bincount(target.ravel(), minlength=S).astype(np.float64)
    thing = np.array([c >= 0 and ISTHING[c] for c in classes])
    gc = np.where(thing, classes, -1).astype(np.int32)
    flags, ranks, counts = shim.instance_match_histogram(pairs, cuda(pc), cuda(gc), cuda(pa), cuda(ga), cuda(crowd),
                                                         cuda(ga), C)
    inter = want.astype(np.float64)
    union = np.where(crowd[None], pa[:, None], pa[:, None] + ga[None] - inter)
    expected = oinst.match_ious(inter / np.maximum(union, 1), pc, gc, pa, crowd, ga, C)
    assert_exact(flags, expected[0], "flags")
    assert_exact(ranks, expected[1], "ranks")
    assert_exact(counts, expected[2], "counts")


@pytest.mark.parametrize("points", [None, [0, 2, 5]])
def test_prefix(shim, points):
    """segmentation_prefixes' call (prefix.py:51-60) and _unpack_cuda's view
    of the fifteen outputs (prefix.py:125-142)."""
    rng = np.random.default_rng(13)
    H, W, G, N = 29, 31, 8, 6
    target = rng.integers(0, G, (H, W)).astype(np.int32)
    gc = rng.integers(0, C, G).astype(np.int32)
    gc[0] = -1
    crowd = (rng.random(G) < 0.25) & (gc >= 0)
    masks = rng.random((N, H, W)) < 0.3
    masks[1] = target == 3
    classes = rng.integers(0, C, N).astype(np.int32)
    thresholds = (0.5, 0.75)
    ordered = list(range(1, N + 1)) if points is None else points
    outputs = shim.prefix(cuda(target), cuda(masks), cuda(classes), cuda(gc), cuda(crowd),
                          torch.tensor(ISTHING, dtype=torch.bool, device="cuda"),
                          torch.tensor(thresholds, dtype=torch.float64, device="cuda"), ordered, True)
    assert len(outputs) == 15 and outputs[-1].item() == 0
    want = opre.prefixes([masks], [classes], target[None], gc[None], crowd[None], C, isthing=ISTHING,
                         checkpoints=points, iou_thresholds=thresholds)[0]
    values, summaries, counts, miou, tp, fp, fn, sums, confusion = to_numpy(outputs[:9])
    itp, ifp, ifn, precision, recall = to_numpy(outputs[9:14])
    assert len(want) == len(ordered)
    for q, snap in enumerate(want):
        assert snap["prefix"] == ordered[q]
        for key, got in (("tp", tp), ("fp", fp), ("fn", fn), ("confusion", confusion)):
            assert_exact(got[q], snap[key], key)
        assert_close(sums[q], snap["iou_sum"], what="iou_sum")
        for i, name in enumerate(("All", "Things", "Stuff")):
            for j, key in enumerate(("pq", "sq", "rq")):
                assert_close(summaries[q, i, j], snap[name][key], what=f"{name}.{key}")
        assert_close(miou[q], snap["miou"], what="miou")
        for key, got in (("tp", itp), ("fp", ifp), ("fn", ifn)):
            assert_exact(got[q], snap["instance"][key], f"instance {key}")
        assert_close(precision[q], snap["instance"]["precision"], what="precision")
        assert_close(recall[q], snap["instance"]["recall"], what="recall")


def test_prefix_error_word(shim):
    """A target slot outside 0..G-1 sets bit 1; an empty mask sets bit 2 only
    when validating (prefix_kernel.cu:34, :74)."""
    target = np.zeros((8, 8), np.int32)
    target[2:4, 2:4] = 1
    gc = np.array([-1, 0], np.int32)
    crowd = np.zeros(2, bool)
    masks = np.zeros((2, 8, 8), bool)
    masks[0, 2:4, 2:4] = True
    things = torch.ones(C, dtype=torch.bool, device="cuda")
    thresholds = torch.tensor([0.5], dtype=torch.float64, device="cuda")
    args = (cuda(masks), cuda(np.array([0, 1], np.int32)), cuda(gc), cuda(crowd), things, thresholds, [1, 2])
    assert shim.prefix(cuda(target), *args, True)[-1].item() == 2
    assert shim.prefix(cuda(target), *args, False)[-1].item() == 0
    bad = target.copy()
    bad[0, 0] = 5
    assert shim.prefix(cuda(bad), *args, False)[-1].item() & 1
