"""Parity at the benchmark configurations' own sizes.

* c1 -- SegmentationEvaluator, 16 x 800 x 1088, 100 ellipse detections per
  image: all 16 images against the oracle, in the eager and the
  graph-replayed forms.
* c2 -- 256 images of uneven size (H, W multiples of 32 in [320, 1344]),
  5..120 segments, 0..300 detections: the packed batch (graph-replayed, as
  timed) equals the plain per-image loop bit for bit; eight images, the
  largest and the one with the most detections among them, against the
  oracle; and size-independent invariants over the whole batch (slot areas
  equal bincounts and sum to H x W, detection-histogram rows sum to the mask
  areas and equal per-image bincounts).
* c4 -- reward serving of requests of 1..64 images of 256^2..1024^2 with
  10..100 segments and prefix requests over 8..128 masks: served under an
  HBM budget that forces admission to wait, every result equals the plain
  API alone, and a subset equals the oracle.

Inputs are generated on the GPU from the configurations' seeds, as the
benchmark does, and copied to the host for the oracle.
"""

import numpy as np
import pytest
import torch

from helpers import assert_close, assert_exact, compare_panoptic, to_numpy
from oracle import combined as ocomb
from oracle import panoptic as opan
from oracle import prefix as opre

pytestmark = pytest.mark.gpu

SUMMARY = ("ap", "ap50", "ap75", "ap_small", "ap_medium", "ap_large", "ar1", "ar10", "ar100", "ar_small",
           "ar_medium", "ar_large")
ISTHING = [c < 80 for c in range(133)]


@pytest.fixture(scope="module")
def pe():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import panoptic_evaluator
    return panoptic_evaluator


def compare_combined(actual, expected, *, exact=False):
    compare_panoptic(actual, expected, exact_iou_sum=exact)
    assert_exact(actual["instance"]["precision"], expected["instance"]["precision"], "precision")
    assert_exact(actual["instance"]["recall"], expected["instance"]["recall"], "recall")
    for key in SUMMARY:
        if exact:
            assert_exact(actual["instance"][key], expected["instance"][key], key)
        else:
            assert_close(actual["instance"][key], expected["instance"][key], what=key)


def oracle_combined(gts, preds, gcs, pcs, crowds, masks, classes, scores):
    """The oracle SegmentationEvaluator over images given one at a time."""
    state = ocomb.SegmentationState(133, ISTHING)
    for i in range(len(gts)):
        state.update(preds[i][None], pcs[i][None],
                     [dict(masks=masks[i], classes=classes[i], scores=scores[i])],
                     gts[i][None], gcs[i][None], crowds[i][None])
    return state.compute()


# ------------------------------------------------------------------------ c1

def test_config1_all_images(pe):
    import workloads
    from panoptic_evaluator.graphs import CapturedStep, CapturedUpdate
    pair, masks, classes, scores = workloads.config1(device="cuda")
    print(f"c1: {pair.gt.shape[0]} images {tuple(pair.gt.shape[1:])}, {masks[0].shape[0]} detections", flush=True)
    args = (pe.PanopticBatch(pair.pred, pair.pc), pe.InstanceBatch(masks, classes, scores),
            pe.PanopticBatch(pair.gt, pair.gc, pair.crowd))
    eager = pe.SegmentationEvaluator(133, isthing=ISTHING, validate=True)
    eager.update(*args)
    got_eager = to_numpy(eager.compute())
    replayed = pe.SegmentationEvaluator(133, isthing=ISTHING, validate=False)
    captured = CapturedUpdate(replayed, *args)
    replayed.reset()
    captured.replay()
    got_graph = to_numpy(replayed.compute())
    stepped = pe.SegmentationEvaluator(133, isthing=ISTHING, validate=False)
    got_step = to_numpy(CapturedStep(stepped, *args).replay())
    host = lambda ts: [t.cpu().numpy() for t in ts]  # noqa: E731
    expected = oracle_combined(pair.gt.cpu().numpy(), pair.pred.cpu().numpy(), pair.gc.cpu().numpy(),
                               pair.pc.cpu().numpy(), pair.crowd.cpu().numpy(), host(masks), host(classes),
                               host(scores))
    compare_combined(got_eager, expected)
    compare_combined(got_graph, got_eager, exact=True)
    compare_combined(got_step, got_eager, exact=True)


# ------------------------------------------------------------------------ c2

@pytest.fixture(scope="module")
def c2(pe):
    import workloads
    batch = workloads.config2(device="cuda")
    torch.cuda.synchronize()
    print(f"c2: {len(batch.sizes)} images, {sum(h * w for h, w in batch.sizes)} px, "
          f"{sum(m.shape[0] for m in batch.masks)} detections", flush=True)
    return batch


def test_config2_batched_equals_plain_loop(pe, c2):
    from panoptic_evaluator.graphs import CapturedStep, CapturedUpdate
    target = pe.PanopticBatch.pack(c2.gt, c2.gc, c2.crowd)
    prediction = pe.PanopticBatch.pack(c2.pred, c2.pc)
    detections = pe.InstanceBatch(c2.masks, c2.classes, c2.scores)
    batched = pe.SegmentationEvaluator(133, isthing=ISTHING, validate=False)
    captured = CapturedUpdate(batched, prediction, detections, target)
    batched.reset()
    captured.replay()
    got = to_numpy(batched.compute())
    # The whole step as one graph (matching inside compute()'s launches; the
    # batch has more AP records than one launch accumulates in LDS).
    stepped = pe.SegmentationEvaluator(133, isthing=ISTHING, validate=False)
    step = CapturedStep(stepped, prediction, detections, target)
    for _ in range(2):
        compare_combined(to_numpy(step.replay()), got, exact=True)
    plain = pe.SegmentationEvaluator(133, isthing=ISTHING, validate=True)
    for i in range(len(c2.sizes)):
        plain.update(pe.PanopticBatch(c2.pred[i][None], c2.pc[i][None]),
                     pe.InstanceBatch([c2.masks[i]], [c2.classes[i]], [c2.scores[i]]),
                     pe.PanopticBatch(c2.gt[i][None], c2.gc[i][None], c2.crowd[i][None]))
    compare_combined(got, to_numpy(plain.compute()), exact=True)
    assert int(got["tp"].sum() + got["fn"].sum()) > 0


def test_config2_subset_against_oracle(pe, c2):
    pixels = [h * w for h, w in c2.sizes]
    dets = [m.shape[0] for m in c2.masks]
    chosen = {int(np.argmax(pixels)), int(np.argmax(dets)), int(np.argmin(pixels))}
    if 0 in dets:
        chosen.add(dets.index(0))  # an image without detections
    for index in np.linspace(0, len(pixels) - 1, 8).astype(int):
        if len(chosen) >= 8:
            break
        chosen.add(int(index))
    chosen = sorted(chosen)
    print("c2 oracle images:", [(i, c2.sizes[i], dets[i]) for i in chosen], flush=True)
    pick = lambda ts: [ts[i] for i in chosen]  # noqa: E731
    ev = pe.SegmentationEvaluator(133, isthing=ISTHING, validate=True)
    ev.update(pe.PanopticBatch.pack(pick(c2.pred), pick(c2.pc)),
              pe.InstanceBatch(pick(c2.masks), pick(c2.classes), pick(c2.scores)),
              pe.PanopticBatch.pack(pick(c2.gt), pick(c2.gc), pick(c2.crowd)))
    host = lambda ts: [ts[i].cpu().numpy() for i in chosen]  # noqa: E731
    expected = oracle_combined(host(c2.gt), host(c2.pred), host(c2.gc), host(c2.pc), host(c2.crowd),
                               host(c2.masks), host(c2.classes), host(c2.scores))
    compare_combined(to_numpy(ev.compute()), expected)


def test_config2_invariants_whole_batch(pe, c2):
    from panoptic_evaluator import _native
    from panoptic_evaluator.instance import offsets
    target = pe.PanopticBatch.pack(c2.gt, c2.gc, c2.crowd)
    prediction = pe.PanopticBatch.pack(c2.pred, c2.pc)
    ev = pe.PanopticEvaluator(133, isthing=ISTHING, validate=True)
    areas = ev._update(prediction, target, return_areas=True)
    G, P = target.classes.shape[1], prediction.classes.shape[1]
    for i, (h, w) in enumerate(c2.sizes):
        ga = torch.bincount(c2.gt[i].flatten().long(), minlength=G)
        pa = torch.bincount(c2.pred[i].flatten().long(), minlength=P)
        assert torch.equal(areas[i, :G], ga), f"GT areas of image {i}"
        assert torch.equal(areas[i, G:], pa), f"prediction areas of image {i}"
        assert int(areas[i, :G].sum()) == h * w
    # Detection x slot histogram of the whole packed batch through the C ABI,
    # rows in input (mask) order.
    device = torch.device("cuda")
    counts = [m.shape[0] for m in c2.masks]
    N, S = sum(counts), G
    order = torch.cat([torch.arange(n, device=device) for n in counts])
    det_off = offsets(counts, device)
    pix_off = offsets([h * w for h, w in c2.sizes], device)
    pointers = _native.device_table([m.data_ptr() for m in c2.masks], device)
    HW = target.segment_map.numel()
    size = max(_native.library().pe_detection_histogram_workspace_bytes(len(counts), HW), 256)
    workspace = torch.empty(size, dtype=torch.uint8, device=device)
    pairs = torch.empty((N, S), dtype=torch.int64, device=device)
    _native.call("pe_detection_histogram", len(counts), pointers.data_ptr(), det_off.data_ptr(), order.data_ptr(),
                 target.segment_map.data_ptr(), pix_off.data_ptr(), HW, S, max(counts), N, pairs.data_ptr(), 0, 1,
                 workspace.data_ptr(), workspace.numel(), _native.stream(device))
    mask_areas = torch.cat([m.flatten(1).sum(1) for m in c2.masks])
    assert torch.equal(pairs.sum(1), mask_areas)
    first = 0
    for i, n in enumerate(counts):
        if n and i % 16 == 0:
            flat = c2.gt[i].flatten().long()
            for d in range(0, n, max(1, n // 5)):
                want = torch.bincount(flat[c2.masks[i][d].flatten()], minlength=S)
                assert torch.equal(pairs[first + d], want), f"image {i} detection {d}"
        first += n


# ------------------------------------------------------------------------ c4

def alone(pe, request):
    from panoptic_evaluator.serving import PQRequest
    if isinstance(request, PQRequest):
        return pe.panoptic_quality(request.prediction, request.target, 133, validate=False).cpu().numpy()
    snaps = pe.segmentation_prefixes(request.prediction, request.target, 133, isthing=ISTHING,
                                     checkpoints=request.checkpoints, validate=False)
    out = []
    for image in snaps:
        summaries = torch.stack([torch.stack([s["All"]["pq"], s["miou"]]) for s in image]).cpu().numpy()
        out.append({"pq": summaries[:, 0], "miou": summaries[:, 1]})
    return out


def test_config4_serving_at_size(pe):
    import workloads
    from panoptic_evaluator.serving import PQRequest, PrefixRequest, RewardServer, footprint
    pq, prefix = workloads.config4(device="cuda", pool=3)
    requests = []
    for images in pq:
        requests.append(PQRequest(
            pe.PanopticBatch.pack([i.pred for i in images], [i.pc for i in images]),
            pe.PanopticBatch.pack([i.gt for i in images], [i.gc for i in images], [i.crowd for i in images])))
    for images, points in prefix:
        requests.append(PrefixRequest(
            pe.InstanceBatch([i.masks for i in images], [i.classes for i in images]),
            pe.PanopticBatch.pack([i.gt for i in images], [i.gc for i in images], [i.crowd for i in images]),
            points))
    requests = [r for pair in zip(requests[:len(pq)], requests[len(pq):]) for r in pair]
    print("c4 requests:", [(type(r).__name__, len(r.target.image_sizes)) for r in requests], flush=True)
    expected = [alone(pe, r) for r in requests]
    footprints = [footprint(r, 133) for r in requests]
    # A budget that holds the largest request but not every two: admission waits.
    server = RewardServer(133, isthing=ISTHING, hbm_budget=max(footprints) + min(footprints) // 2, streams=3)
    tickets = []
    for i, request in enumerate(requests):
        tickets.append(server.submit(request))
        if i % 2:
            server.poll()
    server.drain()
    assert server.idle and server.admission.in_use == 0
    assert server.admission.blocked > 0, "the budget never made admission wait"
    print("c4 launches:", dict(server.launches), "blocked passes:", server.admission.blocked, flush=True)
    for ticket, want in zip(tickets, expected):
        assert ticket.done and ticket.error is None
        if isinstance(want, np.ndarray):
            np.testing.assert_array_equal(ticket.result, want)
        else:
            assert len(ticket.result) == len(want)
            for g, w in zip(ticket.result, want):
                np.testing.assert_array_equal(g["pq"], w["pq"])
                np.testing.assert_array_equal(g["miou"], w["miou"])
    # Oracle: every image of the first PQ request, and the smallest image of
    # the first prefix request.
    images = pq[0]
    want = np.concatenate([opan.panoptic_quality(*(t.cpu().numpy()[None] for t in (i.gt, i.pred, i.gc, i.pc, i.crowd)),
                                                 133, validate=False) for i in images])
    assert_close(tickets[0].result, want, what="per-image PQ")
    images, points = prefix[0]
    small = int(np.argmin([i.size[0] for i in images]))
    image = images[small]
    snaps = opre.prefixes([image.masks.cpu().numpy()], [image.classes.cpu().numpy()], image.gt.cpu().numpy()[None],
                          image.gc.cpu().numpy()[None], image.crowd.cpu().numpy()[None], 133, isthing=ISTHING,
                          checkpoints=points, validate=False)[0]
    got = tickets[1].result[small]
    assert list(got["prefix"]) == [s["prefix"] for s in snaps]
    assert_close(got["pq"], np.array([s["All"]["pq"] for s in snaps]), what="prefix pq")
    assert_close(got["miou"], np.array([s["miou"] for s in snaps]), what="prefix miou")
