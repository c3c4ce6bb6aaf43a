"""Write golden input/output vectors from the reference's CPU path.

Run in the development container only (the reference checkout is not present
on GPU machines):  python tests/golden/make_golden.py /root/reference

It imports the reference package (pure PyTorch CPU code; its CUDA extension is
absent and never needed) and records, on seeded inputs, the outputs of the
functions whose results the known-answer tests and invariants do not fully
pin: PQ matching/accumulation interplay (_update_histogram via
PanopticEvaluator.update/compute), panoptic_quality, original-ID compaction,
_match_ious_cpu, InstanceEvaluator accumulation/summaries,
SegmentationEvaluator, and segmentation_prefixes snapshots.  Only data (inputs
and expected outputs) is written, as compressed .npz files beside this script.
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import torch

OUT = Path(__file__).resolve().parent


def flatten(prefix, value, out):
    if isinstance(value, dict):
        for key, item in value.items():
            flatten(f"{prefix}.{key}" if prefix else str(key), item, out)
    elif isinstance(value, torch.Tensor):
        out[prefix] = value.detach().cpu().numpy()
    else:
        out[prefix] = np.asarray(value)


def panoptic_cases(ref):
    out = {}
    rng = np.random.default_rng(857)
    cases = [(3, 31, 47, 2, 3, 5), (3, 31, 47, 65, 97, 5), (2, 19, 23, 257, 513, 5),
             (4, 40, 56, 30, 36, 7)]
    evaluator = ref.PanopticEvaluator(5, isthing=[True, False, True, True, False],
                                      device="cpu", validate=False)
    for index, (B, H, W, G, P, C) in enumerate(cases):
        gt = rng.integers(0, G, (B, H, W)).astype(np.int32)
        pred = rng.integers(0, P, (B, H, W)).astype(np.int32)
        gc = rng.integers(0, C, (B, G)).astype(np.int32)
        pc = rng.integers(0, C, (B, P)).astype(np.int32)
        gc[:, 0] = pc[:, 0] = -1
        crowd = rng.random((B, G)) < 0.25
        crowd[:, 0] = False
        if index == 1:  # matches, void rows and unused padded slots
            pred = gt.copy()
            pc[:, :G] = gc
            pc[:, G:] = -1
            pred[:, ::5, :] = 0
        if index == 3:  # blocky maps: many >0.5 matches, crowd overlaps
            gt = np.repeat(np.repeat(rng.integers(0, G, (B, H // 8, W // 8)), 8, 1), 8, 2).astype(np.int32)
            pred = np.where(rng.random(gt.shape) < 0.2, rng.integers(0, P, gt.shape), gt).astype(np.int32)
            pc[:, 1:G] = np.where(rng.random((B, G - 1)) < 0.8, gc[:, 1:], pc[:, 1:G])
        C5 = 5
        gc = np.minimum(gc, C5 - 1)
        pc = np.minimum(pc, C5 - 1)
        evaluator.update(ref.PanopticBatch(torch.from_numpy(pred), torch.from_numpy(pc)),
                         ref.PanopticBatch(torch.from_numpy(gt), torch.from_numpy(gc),
                                           torch.from_numpy(crowd)))
        for name, value in (("gt", gt), ("pred", pred), ("gc", gc), ("pc", pc), ("crowd", crowd)):
            out[f"in{index}.{name}"] = value
        flatten(f"out{index}", evaluator.compute(), out)
    out["isthing"] = np.array([True, False, True, True, False])
    out["num_cases"] = np.array(len(cases))
    np.savez_compressed(OUT / "panoptic_accumulate.npz", **out)


def reward_cases(ref):
    out = {}
    rng = np.random.default_rng(19)
    for index in range(6):
        gt = rng.integers(0, 7, (5, 9, 13)).astype(np.int32)
        pred = gt.copy()
        pred[:, ::3] = rng.integers(0, 9, (5, 3, 13))
        gc = rng.integers(0, 4, (5, 7)).astype(np.int32)
        pc = np.concatenate((gc.copy(), np.zeros((5, 2), np.int32)), 1)
        gc[:, 0] = pc[:, 0] = -1
        crowd = rng.random((5, 7)) < 0.25
        crowd[:, 0] = False
        args = (ref.PanopticBatch(torch.from_numpy(pred), torch.from_numpy(pc)),
                ref.PanopticBatch(torch.from_numpy(gt), torch.from_numpy(gc), torch.from_numpy(crowd)), 5)
        out[f"in{index}.gt"], out[f"in{index}.pred"] = gt, pred
        out[f"in{index}.gc"], out[f"in{index}.pc"], out[f"in{index}.crowd"] = gc, pc, crowd
        out[f"out{index}.pq"] = ref.panoptic_quality(*args).numpy()
        out[f"out{index}.stats"] = ref.panoptic_quality(*args, return_stats=True).numpy()
    out["num_cases"] = np.array(6)
    np.savez_compressed(OUT / "panoptic_quality.npz", **out)


def compact_cases(ref):
    out = {}
    large = 2**40 + 17
    gt = np.array([[[0, large, large, 1234], [7, 7, large, 1234]],
                   [[99, 99, 0, 0], [0, 0, 0, 0]]], np.int64)
    gi = [[{"id": 1234, "category_id": 7, "iscrowd": 1}, {"id": large, "category_id": 7, "iscrowd": 1},
           {"id": 7, "category_id": 42}], [{"id": 99, "category_id": 91}]]
    batch = ref.PanopticBatch.from_coco(torch.from_numpy(gt), gi, (7, 42, 91), compact=False)
    out["in.maps"] = gt
    for name in ("classes", "crowd", "lookup_ids", "lookup_slots"):
        out[f"out.{name}"] = getattr(batch, name).numpy()
    out["out.packed"] = batch.compact().segment_map.numpy()
    np.savez_compressed(OUT / "compact.npz", **out)


def matcher_cases(ref):
    from panoptic_evaluator.instance import _match_ious_cpu
    out = {}
    gen = torch.Generator().manual_seed(802)
    cases = ((0, 65, 3), (105, 0, 3), (113, 97, 3), (131, 137, 1), (17, 99, 1), (60, 40, 4))
    for index, (D, G, C) in enumerate(cases):
        pc = torch.randint(C, (D,), generator=gen, dtype=torch.int32)
        gc = torch.randint(C, (G,), generator=gen, dtype=torch.int32)
        crowd = torch.rand(G, generator=gen) < 0.2
        choices = torch.tensor([0.0, 1.0, 1024.0, 1025.0, 9216.0, 9217.0, 1e10, 1e10 + 1], dtype=torch.float64)
        ga = choices[torch.randint(len(choices), (G,), generator=gen)]
        pairs = torch.ones((D, G), dtype=torch.int64)
        pixel_ga = torch.full((G,), 2.0, dtype=torch.float64)
        pixel_pa = torch.ones(D, dtype=torch.float64)
        if D in (17, 60):
            pairs = torch.randint(0, 11, (D, G), generator=gen)
            pixel_ga.fill_(10.0)
            pixel_pa = choices[torch.randint(len(choices), (D,), generator=gen)].clamp_min(10.0)
        union = torch.where(crowd[None], pixel_pa[:, None], pixel_pa[:, None] + pixel_ga[None] - pairs)
        ious = pairs.double() / union.clamp_min(1)
        status, ranks, counts = _match_ious_cpu(ious, pc, gc, pixel_pa, crowd, ga, C)
        for name, value in (("pairs", pairs), ("pc", pc), ("gc", gc), ("crowd", crowd), ("areas", ga),
                            ("pa", pixel_pa), ("ga", pixel_ga), ("C", torch.tensor(C))):
            out[f"in{index}.{name}"] = value.numpy()
        out[f"out{index}.status"], out[f"out{index}.ranks"] = status.numpy(), ranks.numpy()
        out[f"out{index}.counts"] = counts.numpy()
    out["num_cases"] = np.array(len(cases))
    np.savez_compressed(OUT / "coco_matcher.npz", **out)


def instance_cases(ref):
    out = {}
    rng = np.random.default_rng(47)
    evaluator = ref.InstanceEvaluator(4, device="cpu")
    for image in range(8):
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
        evaluator.update(
            ref.InstanceBatch([torch.tensor(pm)], [torch.tensor(pcl)], [torch.tensor(scores)]),
            ref.InstanceBatch([torch.tensor(gm)], [torch.tensor(gcl)], crowd=[torch.tensor(crowd)],
                              areas=[torch.tensor(areas)]))
        for name, value in (("pm", pm), ("gm", gm), ("pc", pcl), ("gc", gcl), ("scores", scores),
                            ("crowd", crowd), ("areas", areas)):
            out[f"in{image}.{name}"] = value
    flatten("out", evaluator.compute(), out)
    out["num_images"] = np.array(8)
    np.savez_compressed(OUT / "instance_eval.npz", **out)


def combined_cases(ref):
    out = {}
    gen = torch.Generator().manual_seed(91)
    maps = torch.randint(0, 6, (2, 13, 35), generator=gen).to(torch.int32)
    classes = torch.tensor([[-1, 0, 1, 0, 2, -1], [-1, 1, 0, 2, 0, 1]], dtype=torch.int32)
    maps[0][maps[0] == 5] = 0
    crowd = torch.tensor([[False, False, True, False, False, False],
                          [False, False, False, False, True, False]])
    pred_maps = maps.clone()
    pred_maps[:, :3] = 0
    masks = [torch.rand((8, 13, 35), generator=gen) > 0.7 for _ in range(2)]
    for b in range(2):
        masks[b][:3] = maps[b][None] == torch.tensor([1, 2, 3])[:, None, None]
    masks[0][3] = True
    labels = [torch.tensor([0, 1, 0, 0, 1, 0, 1, 0]), torch.tensor([1, 0, 0, 1, 0, 1, 0, 1])]
    scores = [torch.tensor([0.9, 0.9, 0.8, 0.5, 0.5, 0.5, 0.1, 0.1], dtype=torch.float64) for _ in range(2)]
    areas = torch.tensor([[0.0, 1024.0, 1025.0, 9216.0, 9217.0, 0.0],
                          [0.0, 100.0, 1024.0, 200.0, 9217.0, 9216.0]], dtype=torch.float64)
    joint = ref.SegmentationEvaluator(3, isthing=[True, True, False], device="cpu")
    for annotation_areas in (None, areas):
        joint.update(ref.PanopticBatch(pred_maps, classes.clone()),
                     ref.InstanceBatch(masks, labels, scores),
                     ref.PanopticBatch(maps, classes, crowd), instance_areas=annotation_areas)
    for name, value in (("maps", maps), ("classes", classes), ("crowd", crowd), ("pred_maps", pred_maps),
                        ("areas", areas)):
        out[f"in.{name}"] = value.numpy()
    for b in range(2):
        out[f"in.masks{b}"] = masks[b].numpy()
        out[f"in.labels{b}"] = labels[b].numpy()
        out[f"in.scores{b}"] = scores[b].numpy()
    flatten("out", joint.compute(), out)
    np.savez_compressed(OUT / "combined.npz", **out)


def prefix_cases(ref):
    out = {}
    gen = torch.Generator().manual_seed(123)
    cases = ((3, 0), (9, 17), (33, 40))
    for index, (g, n) in enumerate(cases):
        maps = torch.randint(g, (1, 19, 23), generator=gen).int()
        maps.flatten()[:g] = torch.arange(g).int()
        labels = (torch.arange(g) % 3).int()[None]
        labels[:, 0] = -1
        crowd = torch.rand((1, g), generator=gen) < 0.3
        crowd[:, 0] = False
        slots = torch.randint(1, g, (n,), generator=gen)
        masks = maps[0][None] == slots[:, None, None]
        if n:
            masks |= torch.rand(masks.shape, generator=gen) < 0.05
        classes = labels[0, slots]
        thresholds = (0.5, 0.75, 0.95)
        result = ref.segmentation_prefixes(
            ref.InstanceBatch([masks], [classes]), ref.PanopticBatch(maps, labels, crowd), 3,
            isthing=[True, False, True], iou_thresholds=thresholds)
        for name, value in (("maps", maps), ("labels", labels), ("crowd", crowd), ("masks", masks),
                            ("classes", classes)):
            out[f"in{index}.{name}"] = value.numpy()
        out[f"in{index}.thresholds"] = np.array(thresholds)
        out[f"out{index}.count"] = np.array(len(result[0]))
        per_snapshot = []
        for snap in result[0]:
            fields = {}
            flatten("", snap, fields)
            per_snapshot.append(fields)
        for key in per_snapshot[0]:
            out[f"out{index}.{key}"] = np.stack([fields[key] for fields in per_snapshot])
    out["num_cases"] = np.array(len(cases))
    np.savez_compressed(OUT / "prefix.npz", **out)


def main():
    root = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    sys.path.insert(0, str(root))
    import panoptic_evaluator as ref  # the reference's CPU path
    assert Path(ref.__file__).resolve().is_relative_to(root.resolve())
    torch.set_num_threads(1)
    panoptic_cases(ref)
    reward_cases(ref)
    compact_cases(ref)
    matcher_cases(ref)
    instance_cases(ref)
    combined_cases(ref)
    prefix_cases(ref)
    for path in sorted(OUT.glob("*.npz")):
        print(path.name, path.stat().st_size)


if __name__ == "__main__":
    main()
