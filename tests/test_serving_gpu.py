"""Reward server: each request's result is the same alone and in the mix.

Requests of both kinds and of uneven sizes are served with a small HBM budget
(so admission has to wait for memory) and several streams (so batches
overlap); every result must equal the one the plain API gives for that
request alone -- bit-identical, since per-image sums fold in a fixed order.
"""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pe():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import panoptic_evaluator
    return panoptic_evaluator


def make_requests(pe, seed, count):
    import workloads
    from panoptic_evaluator.serving import PQRequest, PrefixRequest
    gen = torch.Generator().manual_seed(seed)
    requests = []
    for r in range(count):
        batch = workloads.config2(seed=seed * 100 + r, B=int(torch.randint(1, 5, (1,), generator=gen)),
                                  sizes=(2, 5), segments=(3, 15), detections=(0, 0))
        cuda = lambda xs: [x.cuda() for x in xs]  # noqa: E731
        target = pe.PanopticBatch.pack(cuda(batch.gt), cuda(batch.gc), cuda(batch.crowd))
        if r % 2 == 0:
            prediction = pe.PanopticBatch.pack(cuda(batch.pred), cuda(batch.pc))
            requests.append(PQRequest(prediction, target))
        else:
            masks = [workloads.ellipse_masks(gen, int(torch.randint(0, 9, (1,), generator=gen)), h, w,
                                             max_axis=30.0).cuda() for h, w in batch.sizes]
            classes = [torch.randint(0, 133, (len(m),), generator=gen).int().cuda() for m in masks]
            checkpoints = None if r % 4 == 1 else [0]
            requests.append(PrefixRequest(pe.InstanceBatch(masks, classes), target, checkpoints))
    return requests


def alone(pe, request, isthing):
    from panoptic_evaluator.serving import PQRequest
    if isinstance(request, PQRequest):
        return pe.panoptic_quality(request.prediction, request.target, 133).cpu().numpy()
    snaps = pe.segmentation_prefixes(request.prediction, request.target, 133, isthing=isthing,
                                     checkpoints=request.checkpoints)
    return [{"pq": np.array([s["All"]["pq"].item() for s in image]),
             "miou": np.array([s["miou"].item() for s in image])} for image in snaps]


def assert_same(got, want):
    if isinstance(want, np.ndarray):
        np.testing.assert_array_equal(got, want)
        return
    assert len(got) == len(want)
    for g, w in zip(got, want):
        np.testing.assert_array_equal(g["pq"], w["pq"])
        np.testing.assert_array_equal(g["miou"], w["miou"])


@pytest.mark.parametrize("streams,budget", [(1, 1 << 40), (3, 40 << 20)])
def test_results_same_alone_and_in_the_mix(pe, streams, budget):
    from panoptic_evaluator.serving import RewardServer
    isthing = [c < 80 for c in range(133)]
    requests = make_requests(pe, 3, 12)
    expected = [alone(pe, r, isthing) for r in requests]
    server = RewardServer(133, isthing=isthing, hbm_budget=budget, streams=streams)
    tickets = []
    for i, request in enumerate(requests):
        tickets.append(server.submit(request))
        if i % 3 == 2:
            server.poll()   # later requests arrive while earlier ones run
    server.drain()
    assert server.idle and server.admission.in_use == 0
    for ticket, want in zip(tickets, expected):
        assert ticket.done and ticket.latency >= 0
        assert_same(ticket.result, want)
    if budget < (1 << 30):
        assert server.launches["pq"] + server.launches["prefix"] > 2  # memory forced several batches


def test_checkpoint_beyond_sequence_is_rejected(pe):
    from panoptic_evaluator.serving import PrefixRequest, RewardServer
    requests = make_requests(pe, 5, 2)
    request = requests[1]
    bad = PrefixRequest(request.prediction, request.target, [0, 99])
    server = RewardServer(133)
    with pytest.raises(ValueError, match="checkpoint"):
        server.submit(bad)


def test_single_request_batches_packed_and_uniform(pe):
    """A batch of one request is evaluated as submitted (packed, or a
    uniform B x H x W batch without crowd flags); results equal the plain API."""
    import workloads
    from panoptic_evaluator.serving import PQRequest, PrefixRequest, RewardServer
    isthing = [c < 80 for c in range(133)]
    pair = workloads.config0("cuda", seed=4, B=3, H=40, W=56)
    gen = torch.Generator(device="cuda").manual_seed(9)
    masks = [workloads.ellipse_masks(gen, 7, 40, 56, device="cuda", max_axis=20.0) for _ in range(3)]
    classes = [torch.randint(0, 133, (7,), device="cuda", generator=gen).int() for _ in range(3)]
    uniform_target = pe.PanopticBatch(pair.gt, pair.gc)                  # no crowd flags
    requests = [PQRequest(pe.PanopticBatch(pair.pred, pair.pc), pe.PanopticBatch(pair.gt, pair.gc, pair.crowd)),
                PQRequest(pe.PanopticBatch(pair.pred, pair.pc), uniform_target),
                PrefixRequest(pe.InstanceBatch(masks, classes), uniform_target, None),
                PrefixRequest(pe.InstanceBatch(masks, classes), pe.PanopticBatch(pair.gt, pair.gc, pair.crowd),
                              [0, 3, 7])]
    requests += make_requests(pe, 8, 4)                                  # packed requests
    server = RewardServer(133, isthing=isthing, streams=2)
    for request in requests:
        ticket = server.submit(request)
        server.drain()                                                   # one request per batch
        assert_same(ticket.result, alone(pe, request, isthing))


def test_malformed_request_rejected_and_server_keeps_serving(pe):
    """A request the server cannot evaluate is refused at submission (nothing
    queued, no memory held); later requests are served and drain returns."""
    from panoptic_evaluator.serving import PQRequest, RewardServer
    isthing = [c < 80 for c in range(133)]
    requests = make_requests(pe, 6, 4)
    server = RewardServer(133, isthing=isthing, hbm_budget=64 << 20)
    good = requests[0]
    bad = PQRequest(pe.PanopticBatch(good.prediction.segment_map.long(), good.prediction.classes,
                                     image_sizes=good.prediction.image_sizes), good.target)
    with pytest.raises(ValueError):
        server.submit(bad)
    tickets = [server.submit(r) for r in requests]
    server.drain()
    assert server.idle and server.admission.in_use == 0 and server.admission.waiting == 0
    for ticket, request in zip(tickets, requests):
        assert ticket.error is None
        assert_same(ticket.result, alone(pe, request, isthing))


def test_request_larger_than_budget_runs_alone(pe):
    from panoptic_evaluator.serving import RewardServer, footprint
    isthing = [c < 80 for c in range(133)]
    requests = make_requests(pe, 7, 3)
    sizes = [footprint(r, 133) for r in requests]
    server = RewardServer(133, isthing=isthing, hbm_budget=min(sizes) // 2)
    tickets = [server.submit(r) for r in requests]
    server.drain()
    assert server.admission.in_use == 0
    for ticket, request in zip(tickets, requests):
        assert_same(ticket.result, alone(pe, request, isthing))
