"""Benchmark of the MI355X evaluation path (one JSON line on stdout).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c1]

A step is one reset -> update -> compute pass of the public API over one batch
of synthetic input already resident in HBM (the reference's timing protocol:
DEVELOPMENT.md:47-49).  Two forms are timed over exactly K steps each:

* plain   -- the drop-in API called eagerly: one update() per batch for the
             uniform configs, one update() per image for the uneven batch
             (the reference cannot batch uneven sizes, README.md:57-58);
* batched -- the whole batch (packed when uneven) in one update(): issued
             eagerly, with the update replayed from a HIP graph and compute()
             eager, or with reset, update and compute replayed as one HIP
             graph (graphs.CapturedStep) -- whichever an untimed probe finds
             fastest; the line names the form.

``value`` is the batched form's images/sec.  The dominant kernel's average
duration is measured with HIP events recorded by the library on the kernel's
own launch stream during K eager steps of the batched update (the same
kernel and launch shape as in the graph) and reported against the MI355X HBM
roofline.  ``cpu_baseline`` times a bounded sample of the same workload on
the host: for the SegmentationEvaluator configs, a torch-op restatement of the
reference's CPU path at torch's intra-op threads (with one thread and the
numpy oracle beside it); elsewhere the numpy oracle.

The default run measures c1 (the value) and then, in the same process, the
rest of the metric as sub-objects of the same line: "c2" (the uneven 256-image
batch, batched vs the plain per-image loop) and "c4" (the reward server's
throughput and p50/p99 latency at 200 req/s, then at the ``--load-sweep``
points: higher offered loads, the last under a budget that makes admission
wait).  ``--extras none`` skips them.

Other modes: --copy-calibration runs a plain copy of known size (the reference
point that scales rocprofv3 FETCH_SIZE/WRITE_SIZE into bytes).
"""

from __future__ import annotations

import argparse
import json
import os
import pathlib
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
for _path in (ROOT, ROOT / "panoptic-evaluator_amd"):
    if str(_path) not in sys.path:
        sys.path.insert(0, str(_path))

METRIC = "images/sec PQ+mIoU+mask AP, batched vs plain on 1 GPU; p50/p99 reward latency"
HBM_PEAK_GBS = 8000.0
MFMA_INT8_PEAK_TOPS = 5000.0  # dense int8, MI355X
MFMA_FP4_PEAK_TFLOPS = 10100.0  # dense fp4 (e2m1), MI355X


def log(*args):
    print(*args, file=sys.stderr, flush=True)


# ------------------------------------------------------------------ configs

class PanopticConfig:
    """PanopticEvaluator reset/update/compute over one uniform batch."""

    dominant = "joint_histogram"
    tuned = ("joint_histogram",)
    dtype = "int32"

    def __init__(self, name, maker, classes, things, description):
        self.name, self.maker, self.classes, self.things = name, maker, classes, things
        self.description = description

    def setup(self, device):
        import workloads
        import panoptic_evaluator as pe
        self.pair = self.maker(device=device)
        self.isthing = workloads.isthing_vector(self.classes, self.things)
        self.evaluator = pe.PanopticEvaluator(self.classes, isthing=self.isthing, validate=False)
        self.target = pe.PanopticBatch(self.pair.gt, self.pair.gc, self.pair.crowd)
        self.prediction = pe.PanopticBatch(self.pair.pred, self.pair.pc)
        B, H, W = self.pair.gt.shape
        self.images = B
        self.bytes_per_launch = 8 * B * H * W  # int32 GT + prediction maps
        self.config = {"workload": self.description, "images": B, "height": H, "width": W,
                       "classes": self.classes, "gt_slots": int(self.pair.gc.shape[1]),
                       "pred_slots": int(self.pair.pc.shape[1]), "validate": False}

    def update_args(self):
        return self.prediction, self.target

    def step(self):
        self.evaluator.reset()
        self.evaluator.update(*self.update_args())
        return self.evaluator.compute()

    def cpu_sample(self, budget_s):
        from oracle import panoptic as opan
        arrays = [t.cpu().numpy() for t in (self.pair.gt, self.pair.pred, self.pair.gc, self.pair.pc,
                                            self.pair.crowd)]
        done, start = 0, time.perf_counter()
        state = opan.PanopticState(self.classes, self.isthing)
        while done < self.images:
            sl = slice(done, done + 1)
            state.update(*(a[sl] for a in arrays), validate=False)
            done += 1
            if time.perf_counter() - start > budget_s:
                break
        state.compute()
        elapsed = time.perf_counter() - start
        return done / elapsed, f"{done} of {self.images} images, oracle update+compute"


class RgbConfig(PanopticConfig):
    """PanopticEvaluator over c3's maps stored as decoded COCO panoptic PNGs:
    every slot gets a random 24-bit segment ID, maps are B x H x W x 3 uint8
    (ID = R + 256 G + 65536 B) and PanopticBatch.from_coco(compact=False)
    keeps them that way, so each update decodes and looks up the IDs inside
    the histogram kernel."""

    dtype = "u8 RGB maps (24-bit segment IDs; int64 counts, f64 ratios)"

    def setup(self, device):
        import torch
        import workloads
        import panoptic_evaluator as pe
        pair = self.maker(device=device)
        self.pair = pair
        self.isthing = workloads.isthing_vector(self.classes, self.things)
        gen = torch.Generator().manual_seed(17)
        B, H, W = pair.gt.shape

        def encode(maps, classes, crowd):
            S = classes.shape[1]
            ids = (torch.randperm(2**24 - 1, generator=gen)[:B * S] + 1).view(B, S)
            ids[:, 0] = 0
            table = ids.to(device)
            segment = torch.gather(table, 1, maps.reshape(B, -1).long()).view(B, H, W)
            rgb = torch.stack([(segment >> shift) & 255 for shift in (0, 8, 16)], -1).to(torch.uint8)
            klass, flags = classes.cpu(), None if crowd is None else crowd.cpu()
            info = [[{"id": int(ids[b, s]), "category_id": int(klass[b, s]),
                      "iscrowd": int(flags is not None and bool(flags[b, s]))}
                     for s in range(1, S) if klass[b, s] >= 0] for b in range(B)]
            return rgb.contiguous(), info

        self.rgb_gt, self.gt_info = encode(pair.gt, pair.gc, pair.crowd)
        self.rgb_pred, self.pred_info = encode(pair.pred, pair.pc, None)
        categories = range(self.classes)
        self.target = pe.PanopticBatch.from_coco(self.rgb_gt, self.gt_info, categories, compact=False)
        self.prediction = pe.PanopticBatch.from_coco(self.rgb_pred, self.pred_info, categories, compact=False)
        self.evaluator = pe.PanopticEvaluator(self.classes, isthing=self.isthing, validate=False)
        self.images = B
        self.bytes_per_launch = 6 * B * H * W  # RGB GT + prediction maps
        self.config = {"workload": self.description, "images": B, "height": H, "width": W,
                       "classes": self.classes, "gt_segments": sum(map(len, self.gt_info)),
                       "pred_segments": sum(map(len, self.pred_info)), "validate": False}

    def cpu_sample(self, budget_s):
        from oracle import panoptic as opan
        gt, pred = self.rgb_gt.cpu().numpy(), self.rgb_pred.cpu().numpy()
        categories = list(range(self.classes))
        done, start = 0, time.perf_counter()
        state = opan.PanopticState(self.classes, self.isthing)
        while done < self.images:
            b = done
            g_ids, g_slots, gc, crowd = opan.from_coco(gt[b:b + 1, ..., 0], self.gt_info[b:b + 1], categories)
            p_ids, p_slots, pc, _ = opan.from_coco(pred[b:b + 1, ..., 0], self.pred_info[b:b + 1], categories)
            state.update(opan.compact(opan.rgb_to_id(gt[b:b + 1]), g_ids, g_slots),
                         opan.compact(opan.rgb_to_id(pred[b:b + 1]), p_ids, p_slots), gc, pc, crowd,
                         validate=False)
            done += 1
            if time.perf_counter() - start > budget_s:
                break
        state.compute()
        elapsed = time.perf_counter() - start
        return done / elapsed, f"{done} of {self.images} images, oracle RGB decode + ID compaction + update+compute"


class CombinedConfig:
    """SegmentationEvaluator reset/update/compute over one uniform batch: PQ,
    mIoU and mask AP with the panoptic target's thing segments as AP ground
    truth (the metric's own workload)."""

    dominant = "detection_histogram"
    tuned = ("detection_group", "detection_histogram", "joint_histogram")
    dtype = "u8 masks / int32 maps (int64 counts, f64 ratios)"

    def __init__(self, name, description, **sizes):
        self.name, self.description, self.sizes = name, description, sizes

    def setup(self, device):
        import workloads
        import panoptic_evaluator as pe
        self.pair, self.masks, self.classes, self.scores = workloads.config1(device=device, **self.sizes)
        self.isthing = workloads.isthing_vector(133, 80)
        self.evaluator = pe.SegmentationEvaluator(133, isthing=self.isthing, validate=False)
        self.target = pe.PanopticBatch(self.pair.gt, self.pair.gc, self.pair.crowd)
        self.prediction = pe.PanopticBatch(self.pair.pred, self.pair.pc)
        self.detections = pe.InstanceBatch(self.masks, self.classes, self.scores)
        B, H, W = self.pair.gt.shape
        self.images = B
        dets = sum(m.shape[0] for m in self.masks)
        # detection masks (1 B/px each) + the int32 target map, per image
        self.bytes_per_launch = H * W * (dets + 4 * B)
        self.config = {"workload": self.description, "images": B, "height": H, "width": W, "classes": 133,
                       "detections_per_image": dets // B, "gt_slots": int(self.pair.gc.shape[1]),
                       "pred_slots": int(self.pair.pc.shape[1]), "validate": False,
                       "update_bytes_per_image": H * W * (8 + dets // B)}

    def update_args(self):
        return self.prediction, self.detections, self.target

    def step(self):
        self.evaluator.reset()
        self.evaluator.update(*self.update_args())
        return self.evaluator.compute()

    def cpu_sample(self, budget_s):
        from oracle import combined as ocomb
        state = ocomb.SegmentationState(133, self.isthing, validate=False)
        done, elapsed = 0, 0.0
        while done < self.images and elapsed < budget_s:
            b = done  # the device-to-host copy of the inputs is not timed
            pair = [t[b:b + 1].cpu().numpy() for t in (self.pair.gt, self.pair.pred, self.pair.gc, self.pair.pc,
                                                       self.pair.crowd)]
            det = [dict(masks=self.masks[b].cpu().numpy(), classes=self.classes[b].cpu().numpy(),
                        scores=self.scores[b].cpu().numpy())]
            t0 = time.perf_counter()
            state.update(pair[1], pair[3], det, pair[0], pair[2], pair[4])
            elapsed += time.perf_counter() - t0
            done += 1
        t0 = time.perf_counter()
        state.compute()
        elapsed += time.perf_counter() - t0
        return done / elapsed, f"{done} of {self.images} images, oracle SegmentationEvaluator update+compute"

    def torch_sample(self, budget_s, chunk=4):
        import torch
        from oracle import torch_cpu
        gt, pred, gc, pc, crowd = (t.cpu() for t in (self.pair.gt, self.pair.pred, self.pair.gc, self.pair.pc,
                                                     self.pair.crowd))
        masks = [m.cpu() for m in self.masks]
        classes, scores = [c.cpu() for c in self.classes], [s.cpu() for s in self.scores]
        state = torch_cpu.SegmentationState(133, self.isthing)
        done, elapsed = 0, 0.0
        while done < self.images and elapsed < budget_s:
            part = range(done, min(done + chunk, self.images))
            t0 = time.perf_counter()
            state.update(pred[part.start:part.stop].long(), pc[part.start:part.stop].long(),
                         [(masks[b], classes[b].long(), scores[b].double()) for b in part],
                         gt[part.start:part.stop].long(), gc[part.start:part.stop].long(),
                         crowd[part.start:part.stop])
            elapsed += time.perf_counter() - t0
            done = part.stop
        t0 = time.perf_counter()
        state.compute()
        elapsed += time.perf_counter() - t0
        return done / elapsed, (f"{done} of {self.images} images in updates of {chunk}, torch-op restatement "
                                f"of the reference CPU SegmentationEvaluator update+compute")


class InstanceConfig:
    """InstanceEvaluator reset/update/compute: c1's 100 ellipse detections per
    image against GT instance masks (the target's thing segments as bool
    masks); the mask x mask contraction runs on int8 MFMA."""

    dominant = "mask_overlap_mfma"
    tuned = ("mask_overlap_mfma",)
    dtype = "u8 masks as int8 MFMA operands (int32 accumulate, f64 ratios)"

    def __init__(self, name, description, **sizes):
        self.name, self.description, self.sizes = name, description, sizes

    def setup(self, device):
        import torch
        import workloads
        import panoptic_evaluator as pe
        pair, masks, classes, scores = workloads.config1(device=device, **self.sizes)
        things = torch.tensor(workloads.isthing_vector(133, 80), device=device)
        gt_masks, gt_classes, gt_crowd = [], [], []
        for b in range(pair.gt.shape[0]):
            gc = pair.gc[b]
            slots = [s for s in range(gc.numel()) if gc[s] >= 0 and bool(things[gc[s]])]
            index = torch.tensor(slots, device=device, dtype=torch.int32)
            gt_masks.append((pair.gt[b][None] == index[:, None, None]).contiguous())
            gt_classes.append(gc[index.long()].long())
            gt_crowd.append(pair.crowd[b][index.long()])
        self.masks, self.gt_masks = masks, gt_masks
        self.evaluator = pe.InstanceEvaluator(133, device="cuda", validate=False)
        self.prediction = pe.InstanceBatch(masks, classes, scores)
        self.target = pe.InstanceBatch(gt_masks, gt_classes, crowd=gt_crowd)
        B, H, W = pair.gt.shape
        self.images = B
        dets = [m.shape[0] for m in masks]
        gts = [m.shape[0] for m in gt_masks]
        self.bytes_per_launch = H * W * (sum(dets) + sum(gts))
        self.mfma_ops = sum(2 * d * g * H * W for d, g in zip(dets, gts))
        self.config = {"workload": self.description, "images": B, "height": H, "width": W, "classes": 133,
                       "detections_per_image": sum(dets) // B, "gt_masks": sum(gts), "validate": False}

    def update_args(self):
        return self.prediction, self.target

    def step(self):
        self.evaluator.reset()
        self.evaluator.update(*self.update_args())
        return self.evaluator.compute()

    def cpu_sample(self, budget_s):
        from oracle import instance as oinst
        state = oinst.InstanceState(133, validate=False)
        done, elapsed = 0, 0.0
        while done < self.images and elapsed < budget_s:
            b = done
            pred = [dict(masks=self.masks[b].cpu().numpy(), classes=self.prediction.classes[b].cpu().numpy(),
                         scores=self.prediction.scores[b].cpu().numpy())]
            tgt = [dict(masks=self.gt_masks[b].cpu().numpy(), classes=self.target.classes[b].cpu().numpy(),
                        crowd=self.target.crowd[b].cpu().numpy())]
            t0 = time.perf_counter()
            state.update(pred, tgt)
            elapsed += time.perf_counter() - t0
            done += 1
        t0 = time.perf_counter()
        state.compute()
        elapsed += time.perf_counter() - t0
        return done / elapsed, f"{done} of {self.images} images, oracle InstanceEvaluator update+compute"


class UnevenConfig:
    """SegmentationEvaluator over a batch of uneven images: plain = one
    update() per image; batched = the packed batch in one captured update()."""

    dominant = "detection_histogram"
    tuned = ("detection_group", "detection_histogram", "joint_histogram", "packed_grid")
    dtype = "u8 masks / int32 maps (int64 counts, f64 ratios)"

    def __init__(self, name, description, **sizes):
        self.name, self.description, self.sizes = name, description, sizes

    def setup(self, device):
        import workloads
        import panoptic_evaluator as pe
        u = self.batch = workloads.config2(device=device, **self.sizes)
        self.isthing = workloads.isthing_vector(133, 80)
        self.evaluator = pe.SegmentationEvaluator(133, isthing=self.isthing, validate=False)
        self.target = pe.PanopticBatch.pack(u.gt, u.gc, u.crowd)
        self.prediction = pe.PanopticBatch.pack(u.pred, u.pc)
        self.detections = pe.InstanceBatch(u.masks, u.classes, u.scores)
        self.per_image = [(pe.PanopticBatch(u.pred[i][None], u.pc[i][None]),
                           pe.InstanceBatch([u.masks[i]], [u.classes[i]], [u.scores[i]]),
                           pe.PanopticBatch(u.gt[i][None], u.gc[i][None], u.crowd[i][None]))
                          for i in range(len(u.sizes))]
        self.plain_evaluator = pe.SegmentationEvaluator(133, isthing=self.isthing, validate=False)
        self.images = len(u.sizes)
        pixels = [h * w for h, w in u.sizes]
        dets = [m.shape[0] for m in u.masks]
        self.bytes_per_launch = sum(p * (d + 4) for p, d in zip(pixels, dets))
        self.config = {"workload": self.description, "images": self.images, "pixels": sum(pixels),
                       "detections": sum(dets), "classes": 133,
                       "max_gt_slots": int(self.target.classes.shape[1]),
                       "max_pred_slots": int(self.prediction.classes.shape[1]), "validate": False,
                       "update_bytes": sum(p * (d + 8) for p, d in zip(pixels, dets))}

    def update_args(self):
        return self.prediction, self.detections, self.target

    def step(self):
        # the drop-in API, one image at a time
        ev = self.plain_evaluator
        ev.reset()
        for args in self.per_image:
            ev.update(*args)
        return ev.compute()

    def eager_batched_step(self):
        self.evaluator.reset()
        self.evaluator.update(*self.update_args())
        return self.evaluator.compute()

    def cpu_sample(self, budget_s):
        from oracle import combined as ocomb
        u = self.batch
        state = ocomb.SegmentationState(133, self.isthing, validate=False)
        done, elapsed = 0, 0.0
        while done < self.images and elapsed < budget_s:
            i = done
            arrays = [t.cpu().numpy()[None] for t in (u.gt[i], u.pred[i], u.gc[i], u.pc[i], u.crowd[i])]
            det = [dict(masks=u.masks[i].cpu().numpy(), classes=u.classes[i].cpu().numpy(),
                        scores=u.scores[i].cpu().numpy())]
            t0 = time.perf_counter()
            state.update(arrays[1], arrays[3], det, arrays[0], arrays[2], arrays[4])
            elapsed += time.perf_counter() - t0
            done += 1
        t0 = time.perf_counter()
        state.compute()
        elapsed += time.perf_counter() - t0
        return done / elapsed, f"{done} of {self.images} images, oracle SegmentationEvaluator update+compute"

    def torch_sample(self, budget_s):
        from oracle import torch_cpu
        u = self.batch
        state = torch_cpu.SegmentationState(133, self.isthing)
        done, elapsed = 0, 0.0
        while done < self.images and elapsed < budget_s:
            i = done  # one update per image: the reference cannot batch uneven sizes
            gt, pred, gc, pc, crowd = (t.cpu()[None] for t in (u.gt[i], u.pred[i], u.gc[i], u.pc[i], u.crowd[i]))
            det = [(u.masks[i].cpu(), u.classes[i].cpu().long(), u.scores[i].cpu().double())]
            t0 = time.perf_counter()
            state.update(pred.long(), pc.long(), det, gt.long(), gc.long(), crowd)
            elapsed += time.perf_counter() - t0
            done += 1
        t0 = time.perf_counter()
        state.compute()
        elapsed += time.perf_counter() - t0
        return done / elapsed, (f"{done} of {self.images} images, one update each, torch-op restatement of the "
                                f"reference CPU SegmentationEvaluator update+compute")


class ServingConfig:
    """RewardServer under Poisson arrivals at a fixed offered load: half
    per-image PQ requests, half prefix-reward requests (uneven sizes)."""

    dominant = "prefix_events"
    dtype = "int32 maps / u8 masks (int64 counts, f64 rewards)"

    def __init__(self, name, description, **sizes):
        self.name, self.description, self.sizes = name, description, sizes

    def setup(self, device):
        import workloads
        import panoptic_evaluator as pe
        from panoptic_evaluator.serving import PQRequest, PrefixRequest
        self.isthing = workloads.isthing_vector(133, 80)
        pq, prefix = workloads.config4(device=device, **self.sizes)
        self.raw = (pq, prefix)
        self.pool = []
        for images in pq:
            target = pe.PanopticBatch.pack([i.gt for i in images], [i.gc for i in images], [i.crowd for i in images])
            prediction = pe.PanopticBatch.pack([i.pred for i in images], [i.pc for i in images])
            self.pool.append(PQRequest(prediction, target))
        for images, points in prefix:
            target = pe.PanopticBatch.pack([i.gt for i in images], [i.gc for i in images], [i.crowd for i in images])
            masks = pe.InstanceBatch([i.masks for i in images], [i.classes for i in images])
            self.pool.append(PrefixRequest(masks, target, points))
        # interleave the two kinds
        half = len(self.pool) // 2
        self.pool = [r for pair in zip(self.pool[:half], self.pool[half:]) for r in pair]
        # Algorithmic bytes of the prefix event pass per request: the target map
        # and every mask once, H * W * (4 + K) per image (0 for PQ requests).
        self.prefix_bytes = [sum(m.shape[1] * m.shape[2] * (4 + m.shape[0]) for m in r.prediction.masks)
                             if isinstance(r, PrefixRequest) else 0 for r in self.pool]
        self.config = {"workload": self.description, "pool_requests": len(self.pool),
                       "pool_images": sum(len(images) for images in pq) + sum(len(images) for images, _ in prefix),
                       "classes": 133, "validate": False}

    def run(self, rate, requests, budget, streams, seed=7):
        import numpy as np
        import torch
        from panoptic_evaluator import _native
        from panoptic_evaluator.serving import RewardServer
        rng = np.random.default_rng(seed)
        arrivals = np.cumsum(rng.exponential(1.0 / rate, requests))
        order = [self.pool[i % len(self.pool)] for i in range(requests)]
        server = RewardServer(133, isthing=self.isthing, hbm_budget=budget, streams=streams)
        streamed = sum(self.prefix_bytes[i % len(self.pool)] for i in range(requests))
        lib = _native.library()
        torch.cuda.synchronize()
        lib.pe_profile_reset()
        lib.pe_profile_enable(1)
        tickets, i = [], 0
        start = time.perf_counter()
        while i < requests or not server.idle:
            now = time.perf_counter() - start
            while i < requests and arrivals[i] <= now:
                tickets.append(server.submit(order[i], arrival=start + arrivals[i]))
                i += 1
            # Collect results, waiting at most until the next arrival is due.
            wait = arrivals[i] - (time.perf_counter() - start) if i < requests else 0.002
            server.poll(wait_us=int(max(0.0, min(wait, 0.002)) * 1e6))
        elapsed = max(t.completed for t in tickets) - start
        stats = dict(launches=server.launches, launched=server.launched_requests,
                     blocked=server.admission.blocked)
        server.close()
        torch.cuda.synchronize()
        lib.pe_profile_enable(0)
        kernel_ms, launches = _native.profile_read(self.dominant)
        latencies = np.array([t.latency for t in tickets]) * 1e3
        images = sum(t.images for t in tickets)
        return dict(images=images, elapsed=elapsed, latencies=latencies, kernel_ms=kernel_ms, launches=launches,
                    streamed=streamed, batches=dict(stats["launches"]), tickets=tickets,
                    admission_waits=stats["blocked"], launched_requests=stats["launched"])

    def cpu_sample(self, budget_s):
        from oracle import panoptic as opan
        from oracle import prefix as opre
        pq, prefix = self.raw
        done, elapsed, r = 0, 0.0, 0
        while elapsed < budget_s and r < len(pq):
            images = pq[r]
            for im in images:
                a = [t.cpu().numpy()[None] for t in (im.gt, im.pred, im.gc, im.pc, im.crowd)]
                t0 = time.perf_counter()
                opan.panoptic_quality(a[0], a[1], a[2], a[3], a[4], 133, validate=False)
                elapsed += time.perf_counter() - t0
                done += 1
            images, points = prefix[r]
            for im in images[:2]:
                args = ([im.masks.cpu().numpy()], [im.classes.cpu().numpy()], im.gt.cpu().numpy()[None],
                        im.gc.cpu().numpy()[None], im.crowd.cpu().numpy()[None], 133)
                t0 = time.perf_counter()
                opre.prefixes(*args, isthing=self.isthing, checkpoints=points)
                elapsed += time.perf_counter() - t0
                done += 1
                if elapsed > budget_s:
                    break
            r += 1
        return done / elapsed, f"{done} images of PQ and prefix requests, oracle panoptic_quality/prefixes"


class PackedCombinedConfig(CombinedConfig):
    """c1 with the detection masks arriving bit-packed (PackedMasks, one bit
    per pixel): the detection x slot histogram reads 1/8 of the mask bytes.
    Packing happens once in setup, outside the timed steps."""

    dtype = "bit-packed masks / int32 maps (int64 counts, f64 ratios)"

    def setup(self, device):
        import panoptic_evaluator as pe
        super().setup(device)
        self.detections = pe.InstanceBatch([pe.pack_masks(m) for m in self.masks], self.classes, self.scores)
        self.tuned = ("detection_group", "detection_histogram_bits", "joint_histogram")
        B, H, W = self.pair.gt.shape
        dets = sum(m.shape[0] for m in self.masks)
        # bit rows of the detection masks + the int32 target map
        self.bytes_per_launch = 4 * mask_words(H * W) * dets + 4 * H * W * B
        self.config["mask_format"] = "bits"
        self.config["update_bytes_per_image"] = 8 * H * W + 4 * mask_words(H * W) * dets // B
        self.roofline_extra = extents_note(self.detections.masks)


class RleCombinedConfig(CombinedConfig):
    """c1 with the detection masks arriving as COCO RLE strings (pycocotools'
    compressed form), uploaded to the device once: every step decodes them
    on the GPU into bit rows (rle_runs + rle_raster), then runs c1's
    evaluation on the bit-packed masks.  Encoding happens once in setup."""

    dominant = "rle_raster"
    dtype = "COCO RLE strings -> bit-packed masks / int32 maps (int64 counts, f64 ratios)"

    def setup(self, device):
        import workloads
        import panoptic_evaluator as pe
        super().setup(device)
        log(f"[bench] {self.name}: encoding the detection masks as COCO RLE (untimed)")
        images = [workloads.rle_encode(m) for m in self.masks]
        self.rle_host = images
        self.rle = pe.RleBatch(images, device)
        self.rle.decode(validate=True)
        self.detections = pe.InstanceBatch(self.rle.masks, self.classes, self.scores)
        self.tuned = ("detection_group", "detection_histogram_bits", "joint_histogram")
        B, H, W = self.pair.gt.shape
        dets = sum(m.shape[0] for m in self.masks)
        # The decode's bytes: RLE strings read, bit rows written.
        self.bytes_per_launch = self.rle.input_bytes + 4 * mask_words(H * W) * dets
        self.config["mask_format"] = "COCO RLE (compressed strings), decoded every step"
        self.config["rle_bytes"] = self.rle.input_bytes
        self.config["rle_bytes_per_mask"] = round(self.rle.input_bytes / max(dets, 1), 1)

    def prepare(self):
        self.rle.decode()

    def step(self):
        self.evaluator.reset()
        self.prepare()
        self.evaluator.update(*self.update_args())
        return self.evaluator.compute()

    def cpu_sample(self, budget_s):
        """The oracle's RLE decode (pycocotools' algorithm) and the oracle
        SegmentationEvaluator, one image at a time."""
        from oracle import combined as ocomb
        from oracle import rle as orle
        state = ocomb.SegmentationState(133, self.isthing, validate=False)
        done, elapsed = 0, 0.0
        while done < self.images and elapsed < budget_s:
            b = done  # the device-to-host copy of the inputs is not timed
            pair = [t[b:b + 1].cpu().numpy() for t in (self.pair.gt, self.pair.pred, self.pair.gc, self.pair.pc,
                                                       self.pair.crowd)]
            classes, scores = self.classes[b].cpu().numpy(), self.scores[b].cpu().numpy()
            t0 = time.perf_counter()
            import numpy as np
            masks = np.stack([orle.decode(r) for r in self.rle_host[b]])
            state.update(pair[1], pair[3], [dict(masks=masks, classes=classes, scores=scores)], pair[0], pair[2],
                         pair[4])
            elapsed += time.perf_counter() - t0
            done += 1
        t0 = time.perf_counter()
        state.compute()
        elapsed += time.perf_counter() - t0
        return done / elapsed, (f"{done} of {self.images} images, oracle RLE decode + SegmentationEvaluator "
                                f"update+compute")

    torch_sample = None


class PackedInstanceConfig(InstanceConfig):
    """c1i with bit-packed detection and GT masks: the intersections run on the
    FP4 matrix cores (mask bits expanded to e2m1 operands in registers) or, in
    the other tuner variants, as int8 MFMA over bytes expanded in LDS or as
    popcounts of ANDed words.  Packing happens once in setup, outside the timed
    steps."""

    dominant = "mask_overlap_bits"
    tuned = ("mask_overlap_bits",)
    dtype = "bit-packed masks (e2m1 MFMA operands, exact f32 accumulate of counts below 2^24, f64 ratios)"
    mfma_peak = MFMA_FP4_PEAK_TFLOPS
    mfma_unit = "TFLOP/s (e2m1)"
    mfma_variant = "FP4"

    def setup(self, device):
        import panoptic_evaluator as pe
        super().setup(device)
        self.prediction = pe.InstanceBatch([pe.pack_masks(m) for m in self.masks], self.prediction.classes,
                                           self.prediction.scores)
        self.target = pe.InstanceBatch([pe.pack_masks(m) for m in self.gt_masks], self.target.classes,
                                       crowd=self.target.crowd)
        H, W = self.config["height"], self.config["width"]
        rows = sum(m.shape[0] for m in self.masks) + sum(m.shape[0] for m in self.gt_masks)
        self.bytes_per_launch = 4 * mask_words(H * W) * rows
        self.config["mask_format"] = "bits"
        self.roofline_extra = extents_note(list(self.prediction.masks) + list(self.target.masks))


def mask_words(pixels):
    return (pixels + 127) // 128 * 4


def extents_note(packed):
    """The bytes inside the masks' nonzero-word extents: with extents the
    kernel skips rows outside each work item's range, so it moves fewer bytes
    than the whole bit rows counted as its algorithmic bytes, and `frac` is an
    effective rate (the measured `traffic` shows the bytes moved)."""
    inside = sum(int((p.extents[:, 1] - p.extents[:, 0]).clamp(min=0).sum()) * 4 for p in packed
                 if p.extents is not None)
    return {"bytes_within_extents": inside,
            "note": "bit rows carry nonzero-word extents: rows that are zero over a work item's pixel range are "
                    "skipped, so the kernel reads fewer bytes than the whole rows counted as algorithmic bytes and "
                    "frac is an effective rate"}


def make_configs():
    import workloads
    return {
        "c1": CombinedConfig("c1", "c1: SegmentationEvaluator, 16 x 800x1088, 133 classes, 80 Voronoi seeds, "
                                   "100 ellipse detections per image"),
        "c2": UnevenConfig("c2", "c2: SegmentationEvaluator, 256 uneven images (H, W multiples of 32 in "
                                 "[320, 1344]), 5..120 segments, 0..300 ellipse detections, 133 classes"),
        "c1i": InstanceConfig("c1i", "c1i: InstanceEvaluator, 16 x 800x1088, 100 ellipse detections vs the "
                                     "target's thing segments as GT masks, 133 classes"),
        "c1p": PackedCombinedConfig("c1p", "c1p: c1 (SegmentationEvaluator, 16 x 800x1088, 100 ellipse detections "
                                           "per image) with bit-packed detection masks"),
        "c1ip": PackedInstanceConfig("c1ip", "c1ip: c1i (InstanceEvaluator, 16 x 800x1088, 100 detections vs the "
                                             "target's thing segments) with bit-packed masks on both sides"),
        "c1r": RleCombinedConfig("c1r", "c1r: c1 (SegmentationEvaluator, 16 x 800x1088, 100 ellipse detections "
                                        "per image) with the detections as COCO RLE strings decoded on the GPU "
                                        "every step"),
        "c0": PanopticConfig("c0", workloads.config0, 133, 80,
                             "c0: PanopticEvaluator, 8 x 512x512, 133 classes, 40 Voronoi seeds"),
        "c3": PanopticConfig("c3", workloads.config3, 19, 8,
                             "c3: PanopticEvaluator, 64 x 1024x2048, 19 classes, 150 Voronoi seeds"),
        "c3r": RgbConfig("c3r", workloads.config3, 19, 8,
                         "c3r: PanopticEvaluator, c3's 64 x 1024x2048 maps as decoded COCO panoptic PNGs "
                         "(RGB uint8, 24-bit segment IDs, fused decode + ID lookup), 19 classes"),
        "c4": ServingConfig("c4", "c4: RewardServer, Poisson arrivals; requests of 1..64 images of "
                                  "256^2..1024^2, 10..100 segments; half per-image PQ, half prefix rewards "
                                  "over 8..128 masks (every prefix or every 8th)"),
    }


DEFAULT_CONFIG = "c1"


# ------------------------------------------------------------------ main

def copy_calibration():
    """Plain device copy of a known size: the reference point for rocprofv3
    FETCH_SIZE / WRITE_SIZE scaling (1 GiB read + 1 GiB written per launch)."""
    import torch
    src = torch.ones(1 << 28, dtype=torch.int32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": src.numel() * 4, "copy_bytes_written": dst.numel() * 4,
                      "launches": 3}))


def traffic_for(config_name):
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 counter summary (profiles/traffic_<config>.json), or None."""
    path = ROOT / "profiles" / f"traffic_{config_name}.json"
    if not path.exists():
        return None
    try:
        return json.loads(path.read_text()).get("traffic_bytes_per_launch")
    except (OSError, ValueError):
        return None


READ_CEILING_CAP = 4 << 30


def read_ceiling(nbytes):
    """Bandwidth of a plain non-temporal streaming read (pe_stream_read) of
    the dominant kernel's byte count (capped at 4 GiB), best of three grid
    sizes, in the same process: the measured reference point beside the
    8 TB/s datasheet peak.  Resets the library's profiling counters."""
    import torch
    from panoptic_evaluator import _native
    lib = _native.library()
    nbytes = max(16, min(int(nbytes), READ_CEILING_CAP) // 16 * 16)
    data = torch.ones(nbytes // 4, dtype=torch.int32, device="cuda")
    sink = torch.zeros(1, dtype=torch.int32, device="cuda")
    stream = _native.stream(data.device)
    best = None
    for blocks in (4096, 8192, 16384):
        _native.call("pe_stream_read", data.data_ptr(), nbytes, blocks, sink.data_ptr(), stream)
        torch.cuda.synchronize()
        lib.pe_profile_reset()
        lib.pe_profile_enable(1)
        for _ in range(5):
            _native.call("pe_stream_read", data.data_ptr(), nbytes, blocks, sink.data_ptr(), stream)
        torch.cuda.synchronize()
        lib.pe_profile_enable(0)
        ms, launches = _native.profile_read("stream_read")
        gbs = nbytes / (ms / launches * 1e-3) / 1e9
        if best is None or gbs > best["achieved"]:
            best = {"achieved": round(gbs, 2), "bytes": nbytes, "blocks": blocks}
    del data
    best["frac_of_peak"] = round(best["achieved"] / HBM_PEAK_GBS, 4)
    best["kernel"] = "stream_read (non-temporal 16-byte loads, 4 in flight per lane)"
    return best


def with_ceiling(roofline):
    """Add the measured streaming-read ceiling and the kernel's share of it."""
    if roofline.get("achieved") and roofline.get("algorithmic_bytes_per_launch"):
        ceiling = read_ceiling(roofline["algorithmic_bytes_per_launch"])
        ceiling["kernel_frac_of_ceiling"] = round(roofline["achieved"] / ceiling["achieved"], 4)
        roofline["read_ceiling"] = ceiling
    return roofline


def timed(fn, steps):
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def serve(cfg, args, *, requests=None, rate=None, cpu_budget=None):
    """c4: throughput and p50/p99 request latency of the serving loop as a
    bench-line dictionary."""
    import numpy as np
    requests = requests or args.requests
    rate = rate or args.rate
    cpu_budget = args.cpu_budget if cpu_budget is None else cpu_budget
    log(f"[bench] warm-up: {min(len(cfg.pool), 8)} requests")
    cfg.run(rate=1e6, requests=min(len(cfg.pool), 8), budget=args.hbm_budget, streams=args.streams)
    log(f"[bench] serving {requests} requests at {rate} req/s offered")
    out = cfg.run(rate=rate, requests=requests, budget=args.hbm_budget, streams=args.streams)
    lat = out["latencies"]
    achieved = out["streamed"] / (out["kernel_ms"] * 1e-3) / 1e9 if out["launches"] else None
    cpu = None
    if not args.no_cpu_baseline:
        log("[bench] timing the CPU oracle sample")
        cpu_rate, sample = cfg.cpu_sample(cpu_budget)
        cpu = {"value": round(cpu_rate, 4), "unit": "images/sec", "cores": 1, "kind": "port",
               "sample": sample, "host_threads_available": os.cpu_count()}
    batches = out["batches"]
    launches = sum(batches.values())
    launched = out.get("launched_requests") or requests
    sweep = []
    for point in filter(None, (args.load_sweep or "").split(",")):
        # Higher offered loads, with a budget of their own where one is given
        # (rate:bytes): batch occupancy and admission waits under pressure.
        point_rate, _, point_budget = point.partition(":")
        point_rate, point_budget = float(point_rate), int(point_budget or args.hbm_budget)
        log(f"[bench] serving {requests} requests at {point_rate} req/s offered, budget {point_budget} B")
        run = cfg.run(rate=point_rate, requests=requests, budget=point_budget, streams=args.streams)
        run_launches = sum(run["batches"].values())
        sweep.append({
            "offered_load_rps": point_rate, "hbm_budget_bytes": point_budget,
            "throughput_rps": round(requests / run["elapsed"], 3),
            "images_per_sec": round(run["images"] / run["elapsed"], 3),
            "latency_ms": {"p50": round(float(np.percentile(run["latencies"], 50)), 3),
                           "p99": round(float(np.percentile(run["latencies"], 99)), 3)},
            "batches": run["batches"],
            "requests_per_launch": round((run.get("launched_requests") or requests) / max(run_launches, 1), 3),
            "admission_waits": run.get("admission_waits"),
        })
    return {
        "metric": METRIC,
        "value": round(out["images"] / out["elapsed"], 3),
        "unit": "images/sec",
        "n_gpus": 1,
        "steps": requests,
        "warmup": min(len(cfg.pool), 8),
        "ms_per_step": round(out["elapsed"] * 1e3 / requests, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": cfg.dtype,
        "data": "synthetic",
        "config": dict(cfg.config, offered_load_rps=rate, requests=requests,
                       hbm_budget_bytes=args.hbm_budget, streams=args.streams,
                       form="serving: continuous batching"),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                       "p99": round(float(np.percentile(lat, 99)), 3),
                       "mean": round(float(lat.mean()), 3), "max": round(float(lat.max()), 3)},
        "throughput_rps": round(requests / out["elapsed"], 3),
        "batches": batches,
        "requests_per_launch": round(launched / max(launches, 1), 3),
        "admission_waits": out.get("admission_waits"),
        "load_sweep": sweep,
        "roofline": with_ceiling({
            "bound": "hbm",
            "kernel": cfg.dominant,
            "achieved": None if achieved is None else round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": None if achieved is None else round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic_for(cfg.name),
            "algorithmic_bytes_per_launch": round(out["streamed"] / max(out["launches"], 1)),
            "avg_kernel_ms": round(out["kernel_ms"] / max(out["launches"], 1), 5),
            "launches": out["launches"],
        }),
        "cpu_baseline": cpu,
    }


def cpu_baselines(cfg, budget_s):
    """The CPU baseline of a step config.  Where the config has a torch-op
    restatement of the reference's CPU path (oracle.torch_cpu), that is the
    headline, timed at torch's intra-op threads, with the same at one thread
    and the single-threaded numpy oracle beside it; otherwise the numpy
    oracle alone."""
    import torch
    if getattr(cfg, "torch_sample", None) is None:
        log(f"[bench] {cfg.name}: timing the CPU oracle sample")
        rate, sample = cfg.cpu_sample(budget_s)
        return {"value": round(rate, 4), "unit": "images/sec", "cores": 1, "kind": "port",
                "sample": sample, "host_threads_available": os.cpu_count()}
    threads = torch.get_num_threads()
    timings = {}
    for n in (threads, 1):
        log(f"[bench] {cfg.name}: timing the torch-op CPU restatement at {n} thread(s)")
        torch.set_num_threads(n)
        try:
            timings[n] = cfg.torch_sample(budget_s / 3)
        finally:
            torch.set_num_threads(threads)
    log(f"[bench] {cfg.name}: timing the numpy CPU oracle sample")
    numpy_rate, numpy_sample = cfg.cpu_sample(budget_s / 3)
    rate, sample = timings[threads]
    one_rate, one_sample = timings[1]
    return {"value": round(rate, 4), "unit": "images/sec", "cores": threads, "kind": "port",
            "sample": sample, "restatement": "torch CPU ops (oracle/torch_cpu.py)",
            "host_threads_available": os.cpu_count(),
            "single_thread": {"value": round(one_rate, 4), "cores": 1, "sample": one_sample},
            "numpy_oracle": {"value": round(numpy_rate, 4), "cores": 1, "sample": numpy_sample}}


BATCHED_FORMS = {
    "eager": "batched: reset(), one update() of the whole batch and compute(), eager",
    "graph": "batched: reset(), one update() of the whole batch replayed as a HIP graph, compute()",
    "step_graph": "batched: reset(), one update() of the whole batch and compute(), replayed as one HIP graph",
}


def measure_steps(cfg, args, *, cpu_budget=None):
    """Plain and batched forms of a step config as a bench-line dictionary."""
    import torch
    from panoptic_evaluator import _native
    cpu_budget = args.cpu_budget if cpu_budget is None else cpu_budget
    eager_batched = getattr(cfg, "eager_batched_step", cfg.step)
    prepare = getattr(cfg, "prepare", None)
    for _ in range(args.warmup):
        cfg.step()
        eager_batched()
    torch.cuda.synchronize()
    tuning_report = None
    if not args.no_autotune:
        from panoptic_evaluator import tuning
        log(f"[bench] {cfg.name}: autotuning launch shapes (untimed)")
        report = tuning.autotune(eager_batched, cfg.tuned, repeats=3)
        tuning_report = {kernel: {"variant": report["selection"][kernel],
                                  "describe": report["describe"][kernel][report["selection"][kernel]],
                                  "ms_per_variant": [None if ms is None else round(ms, 5)
                                                     for ms in report["timings"][kernel]]}
                         for kernel in cfg.tuned}
        for _ in range(args.warmup):
            cfg.step()
            eager_batched()
        torch.cuda.synchronize()
    lib = _native.library()

    log(f"[bench] {cfg.name}: timing {args.steps} plain steps")
    plain_elapsed = timed(cfg.step, args.steps)
    # The dominant kernel under the library's HIP events, in eager steps of
    # the batched update (the kernels and launch shapes the graph replays).
    lib.pe_profile_reset()
    lib.pe_profile_enable(1)
    profile_elapsed = timed(eager_batched, args.steps)
    lib.pe_profile_enable(0)
    kernel_ms, launches = _native.profile_read(cfg.dominant)
    avg_ms = kernel_ms / max(launches, 1)
    achieved = cfg.bytes_per_launch / (avg_ms * 1e-3) / 1e9 if launches else None

    batched_elapsed, batched_form, probe = None, None, None
    if not args.plain_only:
        from panoptic_evaluator.graphs import CapturedStep, CapturedUpdate
        log(f"[bench] {cfg.name}: capturing the batched update, and the whole step, as HIP graphs")
        captured = CapturedUpdate(cfg.evaluator, *cfg.update_args())
        # The RLE decode writes only the detection bit rows and extents: the
        # step graph runs it beside the panoptic branch.
        stepped = CapturedStep(cfg.evaluator, *cfg.update_args(), prepare=prepare,
                               prepare_masks_only=prepare is not None)

        def graph_step():
            cfg.evaluator.reset()
            if prepare is not None:
                prepare()
            captured.replay()
            return cfg.evaluator.compute()

        # The batched step is issued eagerly, with its update replayed from a
        # graph, or replayed whole (reset, update and compute in one graph),
        # whichever an untimed probe finds fastest on this workload (graph
        # launch costs more than a few direct launches when the batch is small).
        forms = {"eager": eager_batched, "graph": graph_step, "step_graph": stepped.replay}
        for fn in forms.values():
            for _ in range(args.warmup):
                fn()
        probe_steps = max(3, args.steps // 4)
        probe = {name: round(timed(fn, probe_steps) * 1e3 / probe_steps, 4) for name, fn in forms.items()}
        batched_form = min(probe, key=probe.get)
        log(f"[bench] {cfg.name}: timing {args.steps} batched steps ({batched_form})")
        batched_elapsed = timed(forms[batched_form], args.steps)

    plain_rate = cfg.images * args.steps / plain_elapsed
    value_elapsed = batched_elapsed if batched_elapsed is not None else plain_elapsed
    images_per_s = cfg.images * args.steps / value_elapsed
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baselines(cfg, cpu_budget)
    traffic = traffic_for(cfg.name)
    uneven = isinstance(cfg, UnevenConfig)
    line = {
        "metric": METRIC,
        "value": round(images_per_s, 3),
        "unit": "images/sec",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(value_elapsed / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": cfg.dtype,
        "data": "synthetic",
        "config": dict(cfg.config, form=("plain" if batched_elapsed is None else BATCHED_FORMS[batched_form]),
                       batched_probe_ms_per_step=probe),
        "plain": {"value": round(plain_rate, 3), "ms_per_step": round(plain_elapsed / args.steps * 1e3, 4),
                  "form": "drop-in API, eager, " + ("one update() per image" if uneven else "one update() per batch")},
        "speedup_vs_plain": None if batched_elapsed is None else round(plain_elapsed / batched_elapsed, 3),
        "roofline": with_ceiling({
            "bound": "hbm",
            "kernel": cfg.dominant,
            "achieved": None if achieved is None else round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": None if achieved is None else round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "algorithmic_bytes_per_launch": cfg.bytes_per_launch,
            "avg_kernel_ms": round(avg_ms, 5),
            "launches": launches,
            "measured_over": f"{args.steps} eager steps of the batched update "
                             f"({profile_elapsed * 1e3 / args.steps:.3f} ms/step)",
            **getattr(cfg, "roofline_extra", {}),
        }),
        "cpu_baseline": cpu,
        "tuning": tuning_report,
    }
    variant = (tuning_report or {}).get(cfg.dominant, {}).get("describe", "")
    if getattr(cfg, "mfma_ops", None) and launches and getattr(cfg, "mfma_variant", "") in variant:
        tops = cfg.mfma_ops / (avg_ms * 1e-3) / 1e12
        peak = getattr(cfg, "mfma_peak", MFMA_INT8_PEAK_TOPS)
        line["mfma"] = {"kernel": cfg.dominant, "dense_equivalent_ops_per_launch": cfg.mfma_ops,
                        "achieved": round(tops, 2), "peak": peak, "unit": getattr(cfg, "mfma_unit", "TOP/s"),
                        "frac": round(tops / peak, 4)}
    return line


# Sub-objects of the default line: the rest of the metric, measured in the
# same process after the headline configuration.
EXTRA_KEYS = ("value", "unit", "ms_per_step", "steps", "dtype", "plain", "speedup_vs_plain", "latency_ms",
              "throughput_rps", "batches", "requests_per_launch", "admission_waits", "load_sweep", "roofline",
              "cpu_baseline", "tuning")


def extra(line):
    out = {key: line[key] for key in EXTRA_KEYS if key in line}
    out["workload"] = line["config"]["workload"]
    out["form"] = line["config"]["form"]
    for key in ("offered_load_rps", "hbm_budget_bytes", "streams", "requests"):
        if key in line["config"]:
            out[key] = line["config"][key]
    return out


def run_config(name, args, configs, *, cpu_budget=None, requests=None):
    import gc
    import torch
    cfg = configs[name]
    log(f"[bench] setting up {cfg.name} on cuda:0")
    cfg.setup("cuda")
    torch.cuda.synchronize()
    try:
        if isinstance(cfg, ServingConfig):
            return serve(cfg, args, cpu_budget=cpu_budget, requests=requests)
        return measure_steps(cfg, args, cpu_budget=cpu_budget)
    finally:
        configs[name] = None
        del cfg
        gc.collect()
        torch.cuda.synchronize()
        torch.cuda.empty_cache()


def main(argv=None):
    parser = argparse.ArgumentParser(description=__doc__)
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=20)
    parser.add_argument("--warmup", type=int, default=3)
    parser.add_argument("--config", default=DEFAULT_CONFIG)
    parser.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of CPU oracle sample")
    parser.add_argument("--no-cpu-baseline", action="store_true")
    parser.add_argument("--copy-calibration", action="store_true")
    parser.add_argument("--no-autotune", action="store_true",
                        help="keep the default launch shapes instead of tuning them on this workload")
    parser.add_argument("--plain-only", action="store_true", help="time only the plain eager form")
    parser.add_argument("--rate", type=float, default=200.0, help="c4: offered load, requests/sec")
    parser.add_argument("--requests", type=int, default=400, help="c4: requests served")
    parser.add_argument("--hbm-budget", type=int, default=16 << 30, help="c4: admission budget, bytes")
    parser.add_argument("--streams", type=int, default=3, help="c4: streams for overlapping batches")
    parser.add_argument("--load-sweep", default="800,3200:4294967296",
                        help="c4: further offered loads served after the main one, rate[:budget bytes] "
                             "comma-separated ('' for none)")
    parser.add_argument("--extras", default=None,
                        help="comma-separated configs measured after the main one and reported as "
                             "sub-objects of the same line (default for c1: c2,c4; 'none' for none)")
    args = parser.parse_args(argv)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the evaluation path shards by image with one "
              "exchange at compute(), and no multi-GPU figure is produced here.", file=sys.stderr)
        return 2
    if args.copy_calibration:
        copy_calibration()
        return 0
    import panoptic_evaluator as pe  # noqa: F401

    configs = make_configs()
    extras = args.extras if args.extras is not None else ("c2,c4" if args.config == DEFAULT_CONFIG else "none")
    extras = [] if extras == "none" else [name for name in extras.split(",") if name]
    line = run_config(args.config, args, configs)
    for name in extras:
        # The rest of the metric: batched vs plain on the uneven batch, and the
        # serving loop's latency; half the CPU-sample budget each.
        sub = run_config(name, args, configs, cpu_budget=args.cpu_budget / 2)
        line[name] = extra(sub)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
