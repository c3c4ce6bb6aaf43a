"""GPU parity of the panoptic path (PanopticEvaluator, panoptic_quality,
PanopticBatch.from_coco/compact) against the CPU oracle and golden vectors.

Restates the reference's tests/test_evaluator.py and tests/test_rewards.py
behaviours on the HIP path; every comparison is bit-exact for integer
statistics and iou_sum, <= 1e-12 relative for the final ratios.
"""

import numpy as np
import pytest
import torch

from golden_data import group, load
from helpers import assert_close, assert_exact, compare_panoptic, to_numpy
from oracle import panoptic as opan

pytestmark = pytest.mark.gpu

CATS = (7, 42, 91)


@pytest.fixture(scope="module")
def pe():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import panoptic_evaluator
    return panoptic_evaluator


def cuda(x, dtype=None):
    return torch.as_tensor(np.asarray(x), dtype=dtype).cuda()


def pack(pe, values, info, categories=CATS):
    return pe.PanopticBatch.from_coco(torch.tensor(values, dtype=torch.int64, device="cuda"), info, categories)


def evaluate(pe, gt, pred, gi, pi, validate=True):
    ev = pe.PanopticEvaluator(3, isthing=[True, False, True], validate=validate)
    ev.update(pack(pe, pred, pi), pack(pe, gt, gi))
    return ev


def oracle_state(gt, pred, gc, pc, crowd, C, isthing=None):
    state = opan.PanopticState(C, isthing)
    state.update(gt, pred, gc, pc, crowd, validate=False)
    return state


# ------------------------------------------------------------ golden vectors

def test_golden_accumulate(pe):
    data = load("panoptic_accumulate")
    ev = pe.PanopticEvaluator(5, isthing=list(data["isthing"]), validate=False)
    for index in range(int(data["num_cases"])):
        inp = group(data, f"in{index}")
        ev.update(pe.PanopticBatch(cuda(inp["pred"]), cuda(inp["pc"])),
                  pe.PanopticBatch(cuda(inp["gt"]), cuda(inp["gc"]), cuda(inp["crowd"])))
        out = group(data, f"out{index}")
        result = to_numpy(ev.compute())
        for key in ("tp", "fp", "fn", "confusion"):
            assert_exact(result[key], out[key], key)
        assert_close(result["iou_sum"], out["iou_sum"], what="iou_sum")
        for name in ("All", "Things", "Stuff"):
            for key in ("pq", "sq", "rq"):
                assert_close(result[name][key], out[f"{name}.{key}"])
        for key in ("pq", "sq", "rq", "iou"):
            assert_close(result["per_class"][key], out[f"per_class.{key}"])
        assert_close(result["miou"], out["miou"])


def test_golden_panoptic_quality(pe):
    data = load("panoptic_quality")
    for index in range(int(data["num_cases"])):
        inp = group(data, f"in{index}")
        args = (pe.PanopticBatch(cuda(inp["pred"]), cuda(inp["pc"])),
                pe.PanopticBatch(cuda(inp["gt"]), cuda(inp["gc"]), cuda(inp["crowd"])), 5)
        for validate in (True, False):
            assert_close(pe.panoptic_quality(*args, validate=validate), data[f"out{index}.pq"])
            assert_exact(pe.panoptic_quality(*args, return_stats=True, validate=validate),
                         data[f"out{index}.stats"])


def test_golden_compact(pe):
    data = load("compact")
    large = 2**40 + 17
    gi = [[{"id": 1234, "category_id": 7, "iscrowd": 1}, {"id": large, "category_id": 7, "iscrowd": 1},
           {"id": 7, "category_id": 42}], [{"id": 99, "category_id": 91}]]
    batch = pe.PanopticBatch.from_coco(cuda(data["in.maps"]), gi, CATS, compact=False)
    for name in ("classes", "crowd", "lookup_ids", "lookup_slots"):
        assert_exact(getattr(batch, name), data[f"out.{name}"], name)
    assert_exact(batch.compact().segment_map, data["out.packed"])


# ------------------------------------------- reference behaviours (restated)

def test_perfect_and_reset(pe):  # test_evaluator.py:55-67
    maps = [[[0, 1, 1, 2]]]
    info = [[{"id": 1, "category_id": 7}, {"id": 2, "category_id": 42}]]
    ev = evaluate(pe, maps, maps, info, info)
    r = ev.compute()
    assert r["All"]["pq"].item() == 1 and r["Things"]["n"].item() == 1 and r["Stuff"]["pq"].item() == 1
    assert torch.isnan(r["per_class"]["iou"][2]) and r["miou"].item() == 1
    ev.reset()
    assert ev.compute()["All"]["pq"].item() == 0 and ev.compute()["miou"].item() == 0
    assert r["All"]["pq"].item() == 1  # snapshots survive reset


def test_strict_half_void_union_crowd(pe):  # test_evaluator.py:69-94
    gi = [[{"id": 1, "category_id": 7}]]
    pi = [[{"id": 9, "category_id": 7}]]
    r = evaluate(pe, [[[1, 1, 1, 1]]], [[[9, 9, 0, 0]]], gi, pi).compute()
    assert (r["tp"].sum().item(), r["fp"].sum().item(), r["fn"].sum().item()) == (0, 1, 1)
    assert r["miou"].item() == 0.5
    r = evaluate(pe, [[[1, 1, 1]]], [[[9, 9, 0]]], gi, pi).compute()
    assert r["tp"].sum().item() == 1 and r["iou_sum"].sum().item() == 2 / 3
    assert evaluate(pe, [[[1, 0, 0, 0]]], [[[9, 9, 9, 9]]], gi, pi).compute()["All"]["pq"].item() == 1
    gi = [[{"id": 1, "category_id": 7, "iscrowd": 1}]]
    r = evaluate(pe, [[[1, 1, 0]]], [[[9, 9, 9]]], gi, pi).compute()
    assert r["fp"].sum().item() == 0 and r["fn"].sum().item() == 0
    r = evaluate(pe, [[[1, 1, 1, 1]]], [[[9, 9, 9, 9]]], gi, [[{"id": 9, "category_id": 42}]]).compute()
    assert r["fp"][1].item() == 1
    r = evaluate(pe, [[[0, 0, 1, 1]]], [[[9, 9, 9, 9]]], [[{"id": 1, "category_id": 42}]], pi).compute()
    assert r["fp"][0].item() == 1


def test_confusion_and_void_prediction(pe):  # test_evaluator.py:96-101
    r = evaluate(pe, [[[1, 1, 2, 2]]], [[[9, 0, 9, 10]]],
                 [[{"id": 1, "category_id": 7}, {"id": 2, "category_id": 42}]],
                 [[{"id": 9, "category_id": 7}, {"id": 10, "category_id": 42}]]).compute()
    assert_exact(r["confusion"], [[1, 0, 0, 1], [1, 1, 0, 0], [0, 0, 0, 0]])
    assert abs(r["miou"].item() - (1 / 3 + 1 / 2) / 2) < 1e-15


def test_arbitrary_ids_and_batch_accumulation(pe):  # test_evaluator.py:103-112
    maps = [[[0, 16777215, 16777215]], [[16777215, 0, 0]]]
    info = [[{"id": 16777215, "category_id": 7}], [{"id": 16777215, "category_id": 42}]]
    batched = evaluate(pe, maps, maps, info, info).compute()
    ev = pe.PanopticEvaluator(3)
    for b in range(2):
        item = pack(pe, maps[b:b + 1], info[b:b + 1])
        ev.update(item, item)
    single = ev.compute()
    for key in ("tp", "fp", "fn", "iou_sum", "confusion", "miou"):
        assert_exact(batched[key], to_numpy(single[key]), key)  # batching-invariant, bit for bit


def test_original_ids_fused_lookup(pe):  # test_evaluator.py:124-166
    large = 2**40 + 17
    gt = np.array([[[0, large, large, 1234], [7, 7, large, 1234]], [[99, 99, 0, 0], [0, 0, 0, 0]]], np.int64)
    pred = np.array([[[123, 123, 123, 123], [900, 900, 123, 123]], [[321, 321, 0, 0], [0, 0, 0, 0]]], np.int32)
    gi = [[{"id": 1234, "category_id": 7, "iscrowd": 1}, {"id": large, "category_id": 7, "iscrowd": 1},
           {"id": 7, "category_id": 42}], [{"id": 99, "category_id": 91}]]
    pi = [[{"id": 900, "category_id": 42}, {"id": 123, "category_id": 7}], [{"id": 321, "category_id": 91}]]
    target = pe.PanopticBatch.from_coco(cuda(gt), gi, CATS, compact=False)
    prediction = pe.PanopticBatch.from_coco(cuda(pred), pi, CATS, compact=False)
    assert target.segment_map.dtype == torch.int64
    g_ids, g_slots, gc, g_crowd = opan.from_coco(gt, gi, CATS)
    p_ids, p_slots, pc, _ = opan.from_coco(pred, pi, CATS)
    expected = oracle_state(opan.compact(gt, g_ids, g_slots), opan.compact(pred, p_ids, p_slots), gc, pc,
                            g_crowd, 3).compute()
    for validate in (True, False):
        for ct, cp in ((False, False), (True, False), (False, True)):
            ev = pe.PanopticEvaluator(3, validate=validate)
            ev.update(prediction.compact() if cp else prediction, target.compact() if ct else target)
            compare_panoptic(to_numpy(ev.compute()), expected)
    reused = target.with_maps(target.segment_map.flip(2))
    for name in ("classes", "crowd", "lookup_ids", "lookup_slots"):
        assert getattr(reused, name).data_ptr() == getattr(target, name).data_ptr()
    ev = pe.PanopticEvaluator(3)
    ev.update(prediction, reused)
    flipped = oracle_state(opan.compact(gt[:, :, ::-1].copy(), g_ids, g_slots),
                           opan.compact(pred, p_ids, p_slots), gc, pc, g_crowd, 3).compute()
    assert_exact(ev.compute()["confusion"], flipped["confusion"])


def test_original_ids_across_image_boundaries(pe):  # test_evaluator.py:168-186
    maps = np.array([[[11, 22, 11, 0, 22, 11, 22, 11, 22, 11, 22, 11, 22, 11, 22, 11, 22]]] * 3, np.int64)
    info = [[{"id": 11, "category_id": 7}, {"id": 22, "category_id": 42}],
            [{"id": 22, "category_id": 7}, {"id": 11, "category_id": 42}],
            [{"id": 11, "category_id": 42}, {"id": 22, "category_id": 7}]]
    original = pe.PanopticBatch.from_coco(cuda(maps), info, (7, 42), compact=False)
    ids, slots, classes, crowd = opan.from_coco(maps, info, (7, 42))
    compact = opan.compact(maps, ids, slots)
    assert_exact(original.compact().segment_map, compact)
    expected = oracle_state(compact, compact, classes, classes, crowd, 2).compute()
    for validate in (False, True):
        ev = pe.PanopticEvaluator(2, validate=validate)
        ev.update(original, original)
        compare_panoptic(to_numpy(ev.compute()), expected)


def test_validation_preserves_previous_totals(pe):  # test_evaluator.py:188-210
    maps = torch.tensor([[[1, 1]]], dtype=torch.int32, device="cuda")
    classes = torch.tensor([[-1, 0]], dtype=torch.int32, device="cuda")
    good = pe.PanopticBatch(maps, classes)
    ev = pe.PanopticEvaluator(3)
    ev.update(good, good)
    before = to_numpy(ev.compute())
    unknown = pe.PanopticBatch.from_coco(torch.tensor([[[7, 8]]], dtype=torch.int64, device="cuda"),
                                         [[{"id": 7, "category_id": 7}]], CATS, compact=False)
    cls = lambda v: torch.tensor([v], dtype=torch.int32, device="cuda")  # noqa: E731
    bad_batches = [pe.PanopticBatch(maps + 3, classes), pe.PanopticBatch(maps, cls([-1, 4])),
                   pe.PanopticBatch(maps, cls([-1, -1])), pe.PanopticBatch(maps, cls([-1, 0, 1])),
                   pe.PanopticBatch(maps, classes, torch.tensor([[True, False]], device="cuda")), unknown]
    for index, bad in enumerate(bad_batches):
        with pytest.raises(ValueError):
            ev.update(good, bad)
        if index != 4:  # prediction crowd flags are unused
            with pytest.raises(ValueError):
                ev.update(bad, good)
        after = to_numpy(ev.compute())
        for key in ("tp", "fp", "fn", "iou_sum", "confusion"):
            assert_exact(after[key], before[key], key)


def test_empty_and_noncontiguous(pe):  # test_evaluator.py:212-239
    for shape in ((0, 3, 4), (2, 0, 4)):
        batch = pe.PanopticBatch(torch.empty(shape, dtype=torch.int32, device="cuda"),
                                 torch.full((shape[0], 1), -1, dtype=torch.int32, device="cuda"))
        ev = pe.PanopticEvaluator(3)
        ev.update(batch, batch)
        assert ev.compute()["confusion"].sum().item() == 0
        original = pe.PanopticBatch.from_coco(batch.segment_map.long(), [[] for _ in range(shape[0])], CATS,
                                              compact=False)
        ev.update(original, original)
        assert ev.compute()["confusion"].sum().item() == 0
    empty = pack(pe, [[[0, 0]]], [[]])
    ev = pe.PanopticEvaluator(3)
    ev.update(empty, empty)
    assert ev.compute()["All"]["n"].item() == 0
    prediction = pack(pe, [[[1, 1]]], [[{"id": 1, "category_id": 7}]])
    ev.update(prediction, empty)
    assert ev.compute()["fp"].sum().item() == 0
    ev.update(empty, prediction)
    assert ev.compute()["fn"][0].item() == 1 and ev.compute()["miou"].item() == 0
    maps = torch.ones((2, 4, 6), device="cuda", dtype=torch.int32)[:, ::2, ::2]
    labels = torch.tensor([[-1, -1, 0, -1]], device="cuda", dtype=torch.int32)[:, ::2].expand(2, -1)
    ev.reset()
    ev.update(pe.PanopticBatch(maps, labels), pe.PanopticBatch(maps, labels))
    assert ev.compute()["All"]["pq"].item() == 1


@pytest.mark.parametrize("validate", [False, True])
def test_varied_capacities_and_snapshots(pe, validate):  # test_evaluator.py:283-335
    gen = np.random.default_rng(857)
    ev = pe.PanopticEvaluator(5, validate=False)
    state = opan.PanopticState(5)
    for G, P in ((1, 1), (2, 3), (65, 97), (257, 513), (1025, 769), (2, 3)):
        gt = gen.integers(0, G, (3, 31, 47)).astype(np.int32)
        pred = gen.integers(0, P, gt.shape).astype(np.int32)
        gc = gen.integers(0, 5, (3, G)).astype(np.int32)
        pc = gen.integers(0, 5, (3, P)).astype(np.int32)
        gc[:, 0] = pc[:, 0] = -1
        crowd = gen.random((3, G)) < 0.25
        crowd[:, 0] = False
        if G == 65:
            pred = gt.copy()
            pc[:, :G] = gc
            pc[:, G:] = -1
            pred[:, ::5, :] = 0
        ev.update(pe.PanopticBatch(cuda(pred), cuda(pc)), pe.PanopticBatch(cuda(gt), cuda(gc), cuda(crowd)))
        state.update(gt, pred, gc, pc, crowd, validate=False)
        compare_panoptic(to_numpy(ev.compute()), state.compute())
    snapshot = ev.compute()
    saved = {k: snapshot[k].clone() for k in ("tp", "fp", "fn", "iou_sum", "confusion")}
    pointers = tuple(t.data_ptr() for t in ev._scratch)
    state_pointer = ev.tp.data_ptr()
    ev.reset()
    assert ev.tp.data_ptr() == state_pointer
    ev.update(pe.PanopticBatch(cuda(pred), cuda(pc)), pe.PanopticBatch(cuda(gt), cuda(gc), cuda(crowd)))
    assert tuple(t.data_ptr() for t in ev._scratch) == pointers
    for key in saved:
        assert torch.equal(snapshot[key], saved[key])


def test_cooperative_rows_with_validation(pe):  # test_evaluator.py:283-295
    maps = np.arange(1, 130, dtype=np.int32).reshape(1, 3, 43)
    gc = np.full((1, 137), -1, np.int32)
    pc = np.full((1, 193), -1, np.int32)
    gc[:, 1:130] = pc[:, 1:130] = np.arange(129) % 3
    expected = oracle_state(maps, maps, gc, pc, np.zeros(gc.shape, bool), 3).compute()
    for validate in (False, True):
        ev = pe.PanopticEvaluator(3, validate=validate)
        ev.update(pe.PanopticBatch(cuda(maps), cuda(pc)), pe.PanopticBatch(cuda(maps), cuda(gc)))
        compare_panoptic(to_numpy(ev.compute()), expected)


def test_graph_capture_replays(pe):  # test_evaluator.py:337-357
    maps = torch.tensor([[[1, 1, 2, 2]]], dtype=torch.int32, device="cuda")
    classes = torch.tensor([[-1, 0, 1]], dtype=torch.int32, device="cuda")
    compact = pe.PanopticBatch(maps, classes)
    original = pe.PanopticBatch.from_coco(
        maps.long() * 100003 + 2**40,
        [[{"id": 100003 + 2**40, "category_id": 7}, {"id": 200006 + 2**40, "category_id": 42}]], CATS,
        compact=False)
    for item in (compact, original):
        ev = pe.PanopticEvaluator(3, validate=False)
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            ev.update(item, item)
        torch.cuda.current_stream().wait_stream(side)
        ev.reset()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            ev.update(item, item)
        for replay in range(1, 4):
            graph.replay()
            torch.cuda.synchronize()
            assert ev.compute()["tp"].cpu().tolist() == [replay, replay, 0]
            assert abs(ev.compute()["iou_sum"].sum().item() - 2 * replay) < 1e-12


def test_nondefault_stream_exact(pe):  # test_evaluator.py:359-376
    gen = np.random.default_rng(7)
    gt = gen.integers(0, 17, (4, 128, 129)).astype(np.int32)
    pred = gen.integers(0, 21, gt.shape).astype(np.int32)
    gc = (np.arange(17, dtype=np.int32)[None].repeat(4, 0) % 3).astype(np.int32)
    pc = (np.arange(21, dtype=np.int32)[None].repeat(4, 0) % 3).astype(np.int32)
    gc[:, 0] = pc[:, 0] = -1
    crowd = np.zeros(gc.shape, bool)
    crowd[:, 3] = True
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        ev = pe.PanopticEvaluator(3)
        ev.update(pe.PanopticBatch(cuda(pred), cuda(pc)), pe.PanopticBatch(cuda(gt), cuda(gc), cuda(crowd)))
        result = ev.compute()
    stream.synchronize()
    compare_panoptic(to_numpy(result), oracle_state(gt, pred, gc, pc, crowd, 3).compute())


def test_rewards_known_answers(pe):  # test_rewards.py:12-63
    gt = [[[1, 1, 1, 1]], [[1, 1, 1, 0]], [[1, 0, 0, 0]], [[2, 2, 0, 0]], [[0, 0, 0, 0]], [[1, 1, 1, 1]],
          [[1, 1, 1, 1]], [[1, 1, 2, 2]]]
    pred = [[[1, 1, 0, 0]], [[1, 1, 0, 0]], [[1, 1, 1, 1]], [[2, 2, 2, 2]], [[0, 0, 0, 0]], [[0, 0, 0, 0]],
            [[1, 1, 1, 1]], [[1, 1, 2, 2]]]
    gc = np.array([[-1, 0, -1], [-1, 0, -1], [-1, 0, -1], [-1, -1, 1], [-1, -1, -1], [-1, 0, -1],
                   [-1, 0, -1], [-1, 0, 2]], np.int32)
    pc = np.array([[-1, 0, 2]] * 8, np.int32)
    pc[3, 2] = 1
    crowd = np.zeros(gc.shape, bool)
    crowd[3, 2] = True
    gt, pred = np.array(gt, np.int32), np.array(pred, np.int32)
    expected = opan.panoptic_quality(gt, pred, gc, pc, crowd, 4)
    target, prediction = pe.PanopticBatch(cuda(gt), cuda(gc), cuda(crowd)), pe.PanopticBatch(cuda(pred), cuda(pc))
    for validate in (True, False):
        result = pe.panoptic_quality(prediction, target, 4, validate=validate)
        assert result.shape == (8,) and result.dtype == torch.float64 and result.is_cuda
        assert_close(result, expected)
    assert expected[0] == 0 and abs(expected[1] - 2 / 3) < 1e-15 and expected[2] == 1
    ids = [0, 2**40 + 9, 71]
    gi = [[dict(id=ids[s], category_id=int(gc[b, s]), iscrowd=bool(crowd[b, s])) for s in (1, 2) if gc[b, s] >= 0]
          for b in range(8)]
    pi = [[dict(id=ids[s], category_id=int(pc[b, s])) for s in (1, 2)] for b in range(8)]
    table = torch.tensor(ids, device="cuda")
    prepared_gt = pe.PanopticBatch.from_coco(table[cuda(gt).long()], gi, range(4), compact=False)
    prepared_pred = pe.PanopticBatch.from_coco(table[cuda(pred).long()], pi, range(4), compact=False)
    assert_close(pe.panoptic_quality(prepared_pred, prepared_gt, 4), expected)
    empty = prepared_pred.with_maps(torch.zeros_like(prepared_pred.segment_map))
    assert_close(pe.panoptic_quality(empty, prepared_gt, 4), np.zeros(8))


def test_rewards_invalid_inputs(pe):  # test_rewards.py:88-100
    batch = pe.PanopticBatch(torch.ones((1, 1, 2), device="cuda", dtype=torch.int32),
                             torch.tensor([[-1, 0]], device="cuda", dtype=torch.int32))
    assert pe.panoptic_quality(batch, batch, 1).item() == 1
    with pytest.raises(ValueError):
        pe.panoptic_quality(batch, batch, 1, reduction="mean")
    with pytest.raises(ValueError):
        pe.panoptic_quality(batch.with_maps(torch.full_like(batch.segment_map, 9)), batch, 1)
    with pytest.raises(ValueError):
        pe.PanopticEvaluator(1).update(batch.with_maps(torch.zeros_like(batch.segment_map)), batch)


def test_cpu_device_is_rejected(pe):
    with pytest.raises(ValueError):
        pe.PanopticEvaluator(3, device="cpu")


# ------------------------------------------------------------- full configs

def test_config0_parity(pe):
    import workloads
    pair = workloads.config0(device="cpu")
    arrays = [t.numpy() for t in (pair.gt, pair.pred, pair.gc, pair.pc, pair.crowd)]
    isthing = workloads.isthing_vector(133, 80)
    expected = oracle_state(*arrays, 133, isthing).compute()
    for validate in (True, False):
        ev = pe.PanopticEvaluator(133, isthing=isthing, validate=validate)
        ev.update(pe.PanopticBatch(pair.pred.cuda(), pair.pc.cuda()),
                  pe.PanopticBatch(pair.gt.cuda(), pair.gc.cuda(), pair.crowd.cuda()))
        compare_panoptic(to_numpy(ev.compute()), expected)
    pq = opan.panoptic_quality(*arrays, 133)
    got = pe.panoptic_quality(pe.PanopticBatch(pair.pred.cuda(), pair.pc.cuda()),
                              pe.PanopticBatch(pair.gt.cuda(), pair.gc.cuda(), pair.crowd.cuda()), 133)
    assert_close(got, pq)


def test_config3_histogram_invariants(pe):
    """Full c3 resolution (2 of the 64 images vs the oracle; all 64 through
    size-independent invariants: pixel totals and slot areas)."""
    import workloads
    pair = workloads.config3(device="cuda")
    isthing = workloads.isthing_vector(19, 8)
    ev = pe.PanopticEvaluator(19, isthing=isthing, validate=True)
    areas = ev._update(pe.PanopticBatch(pair.pred, pair.pc), pe.PanopticBatch(pair.gt, pair.gc, pair.crowd),
                       return_areas=True)
    B, H, W = pair.gt.shape
    G, P = pair.gc.shape[1], pair.pc.shape[1]
    # Slot areas of both maps; each image's areas add up to its pixel count.
    assert torch.equal(areas[:, :G].sum(1), torch.full((B,), H * W, device=areas.device))
    assert torch.equal(areas[:, G:].sum(1), torch.full((B,), H * W, device=areas.device))
    want = torch.stack([torch.bincount(pair.gt[b].flatten().long(), minlength=G) for b in range(B)])
    assert torch.equal(areas[:, :G], want)
    want_pred = torch.stack([torch.bincount(pair.pred[b].flatten().long(), minlength=P) for b in range(B)])
    assert torch.equal(areas[:, G:], want_pred)
    # The pair table is handed back zeroed for the next update.
    assert not ev._scratch[0].any().item()
    conf = ev.compute()["confusion"]
    valid = (pair.gc >= 0) & ~pair.crowd
    assert conf.sum().item() == int((want * valid).sum().item())
    sub = [t[:2].cpu().numpy() for t in (pair.gt, pair.pred, pair.gc, pair.pc, pair.crowd)]
    ev2 = pe.PanopticEvaluator(19, isthing=isthing)
    ev2.update(pe.PanopticBatch(pair.pred[:2], pair.pc[:2]), pe.PanopticBatch(pair.gt[:2], pair.gc[:2], pair.crowd[:2]))
    compare_panoptic(to_numpy(ev2.compute()), oracle_state(*sub, 19, isthing).compute())
