"""Continuous-batching server for per-image rewards on one GPU.

Requests are per-image panoptic quality (``PQRequest``: the reference's
``panoptic_quality``, evaluator.py:471-544) or generation-order prefix rewards
(``PrefixRequest``: ``segmentation_prefixes``, prefix.py:11-122).  The server
has no counterpart in the reference; it is the serving form of the same
kernels:

* **admission by HBM budget** -- each request's device footprint (scratch,
  statistics and packed copies) is estimated on submission; requests are
  admitted first-in first-out while the in-use total stays within the budget,
  and their memory returns to the budget when their results are on the host;
* **continuous batching** -- every ``poll()`` merges all newly admitted
  requests of a kind into one packed launch (images of any size, one kernel
  sequence), issued on the next of a few streams, so work admitted while
  earlier batches are still running overlaps with them on the device;
* **device-resident statistics** -- per-image sufficient statistics and the
  FP64 metrics are computed on the GPU; only the final per-image numbers are
  copied (asynchronously, into pinned memory) to the host.

Results do not depend on how requests are batched: per-image PQ folds each
image's IoUs in slot order and prefix snapshots are per image, so a request
gets bit-identical numbers alone or in any mix.  Inputs are not validated
(validation reads a result on the host, which would serialise the loop);
callers validate with the plain API beforehand if needed.
"""

from __future__ import annotations

import collections
import time
from dataclasses import dataclass, field
from typing import Sequence

import torch

from . import _native
from .evaluator import PanopticBatch, _prepare, finalize_arrays, panoptic_quality
from .instance import InstanceBatch
from .prefix import _check_arguments, launch_prefix


@dataclass
class PQRequest:
    """Per-image mean-class PQ of a batch (uniform B x H x W or packed)."""

    prediction: PanopticBatch
    target: PanopticBatch


@dataclass
class PrefixRequest:
    """Prefix rewards of generated masks in order over a target batch;
    ``checkpoints`` None means every prefix 1..N (or [0] when N = 0)."""

    prediction: InstanceBatch
    target: PanopticBatch
    checkpoints: Sequence[int] | None = None


@dataclass
class Ticket:
    """A submitted request and its life cycle (perf_counter seconds)."""

    request: object
    images: int
    footprint: int
    arrival: float
    admitted: float | None = None
    launched: float | None = None
    completed: float | None = None
    result: object = None
    error: BaseException | None = None   # set instead of result when the launch failed

    @property
    def done(self) -> bool:
        return self.completed is not None

    @property
    def latency(self) -> float | None:
        return None if self.completed is None else self.completed - self.arrival


class Admission:
    """First-in first-out admission under a byte budget.  A request larger
    than the whole budget is admitted alone once nothing else is in use."""

    def __init__(self, budget: int):
        if budget <= 0:
            raise ValueError("the HBM budget must be positive")
        self.budget = int(budget)
        self.in_use = 0
        self.running = 0
        self.blocked = 0   # admission passes at which the head request waited for memory
        self.waiting: collections.deque = collections.deque()

    def offer(self, ticket: Ticket):
        self.waiting.append(ticket)

    def admit(self) -> list:
        admitted = []
        while self.waiting:
            ticket = self.waiting[0]
            fits = self.in_use + ticket.footprint <= self.budget
            if not fits and not (self.running == 0 and not admitted):
                self.blocked += 1
                break
            self.waiting.popleft()
            self.in_use += ticket.footprint
            self.running += 1
            admitted.append(ticket)
        return admitted

    def release(self, ticket: Ticket):
        self.in_use -= ticket.footprint
        self.running -= 1


def _images(batch: PanopticBatch, *, crowd: bool):
    """Per-image (H x W map, classes row, crowd row or None, (H, W)) views."""
    batch = batch.compact()
    maps, classes = batch.segment_map, batch.classes
    rows = []
    if batch.packed:
        start = 0
        for b, (h, w) in enumerate(batch.image_sizes):
            rows.append((maps[start:start + h * w].view(h, w), classes[b],
                         batch.crowd[b] if crowd and batch.crowd is not None else None, (h, w)))
            start += h * w
    else:
        for b in range(maps.shape[0]):
            rows.append((maps[b], classes[b], batch.crowd[b] if crowd and batch.crowd is not None else None,
                         tuple(maps.shape[1:])))
    return rows


def _crowd_rows(rows):
    return [c if c is not None else torch.zeros_like(k, dtype=torch.bool) for _, k, c, _ in rows]


class _Launch:
    """One merged launch in flight: its tickets, device outputs kept alive,
    pinned host copies and the completion event."""

    def __init__(self, kind, tickets, stream):
        self.kind, self.tickets, self.stream = kind, tickets, stream
        self.event = torch.cuda.Event()
        self.keep = []
        self.host = None
        self.split = None


class RewardServer:
    """Continuous-batching evaluation of reward requests (see module doc)."""

    def __init__(self, num_classes: int, *, isthing: Sequence[bool] | None = None,
                 iou_thresholds=(0.5, 0.75), hbm_budget: int = 8 << 30, streams: int = 2,
                 device: str | torch.device = "cuda"):
        if not isinstance(num_classes, int) or isinstance(num_classes, bool) or num_classes <= 0:
            raise ValueError("num_classes must be a positive integer")
        if isthing is not None and len(isthing) != num_classes:
            raise ValueError("isthing must contain one flag per category")
        if streams < 1:
            raise ValueError("at least one stream is needed")
        self.num_classes = num_classes
        self.device = _native.require_gpu(device)
        self.things = torch.tensor(list(isthing) if isthing is not None else [True] * num_classes,
                                   dtype=torch.bool, device=self.device)
        self.thresholds, _ = _check_arguments(iou_thresholds, None)
        self.admission = Admission(hbm_budget)
        # Default-priority streams only.
        self.streams = [torch.cuda.Stream(self.device) for _ in range(streams)]
        self._next_stream = 0
        self._pending: list = []
        self._in_flight: collections.deque = collections.deque()
        self.launches = collections.Counter()
        self.prefix_pixels_streamed = 0   # algorithmic bytes of the prefix event passes

    # ------------------------------------------------------------ footprint

    def _footprint(self, request) -> tuple[int, int]:
        """(images, estimated device bytes) of a request."""
        C = self.num_classes
        if isinstance(request, PQRequest):
            tgt, pred = request.target, request.prediction
            B = len(tgt.image_sizes) if tgt.packed else tgt.segment_map.shape[0]
            G, P = tgt.classes.shape[1], pred.classes.shape[1]
            pixels = tgt.segment_map.numel()
            bytes_ = (8 * B * G * P + 8 * B * (G + P) + 4 * B * G + 32 * B * C + 12 * B * P + 8 * pixels
                      + 8 * B * C * 4)
            return B, int(bytes_)
        if isinstance(request, PrefixRequest):
            tgt = request.target
            B = len(request.prediction.masks)
            G, T = tgt.classes.shape[1], len(self.thresholds)
            counts = [m.shape[0] for m in request.prediction.masks]
            rows = sum(len(self._points(request, n)) for n in counts)
            sum_n, sum_tri = sum(counts), sum(n * (n + 1) // 2 for n in counts)
            pixels = sum(m.shape[1] * m.shape[2] for m in request.prediction.masks)
            workspace = _native.library().pe_prefix_workspace_bytes(B, G, C, T, sum_n, sum_tri, pixels)
            per_row = 8 * (4 * C + C * (C + 1)) + 40 * T * C + 8 * (4 * C + 13)
            return B, int(workspace + rows * per_row + 4 * pixels + 16 * B * G)
        raise TypeError("requests are PQRequest or PrefixRequest")

    @staticmethod
    def _points(request, n):
        if request.checkpoints is None:
            return list(range(1, n + 1)) if n else [0]
        return sorted(set(request.checkpoints))

    # ------------------------------------------------------------ interface

    def submit(self, request, arrival: float | None = None) -> Ticket:
        """Queue a request; returns its ticket (``result`` filled on completion)."""
        if isinstance(request, PrefixRequest) and request.checkpoints is not None:
            _, requested = _check_arguments(self.thresholds, request.checkpoints)
            if requested and any(requested[-1] > m.shape[0] for m in request.prediction.masks):
                raise ValueError("checkpoint exceeds prediction count")
        images, footprint = self._footprint(request)
        ticket = Ticket(request, images, footprint, time.perf_counter() if arrival is None else arrival)
        self.admission.offer(ticket)
        return ticket

    @property
    def idle(self) -> bool:
        return not self.admission.waiting and not self._pending and not self._in_flight

    def poll(self) -> int:
        """Admit, launch and complete without blocking; returns how many
        requests completed."""
        now = time.perf_counter()
        for ticket in self.admission.admit():
            ticket.admitted = now
            self._pending.append(ticket)
        if self._pending:
            pq = [t for t in self._pending if isinstance(t.request, PQRequest)]
            prefix = [t for t in self._pending if isinstance(t.request, PrefixRequest)]
            self._pending = []
            done = 0
            for tickets, launch in ((pq, self._launch_pq), (prefix, self._launch_prefix)):
                if not tickets:
                    continue
                try:
                    launch(tickets)
                except Exception as error:  # noqa: BLE001 - reported on the tickets
                    done += self._fail(tickets, error)
            return done + self._complete()
        return self._complete()

    def _fail(self, tickets, error) -> int:
        """A launch that raised: its tickets complete with the error and give
        their memory back, and the server keeps serving."""
        now = time.perf_counter()
        for ticket in tickets:
            ticket.error = error
            ticket.completed = now
            self.admission.release(ticket)
        return len(tickets)

    def drain(self):
        """Poll until every submitted request has completed."""
        while not self.idle:
            if not self.poll() and self._in_flight:
                self._in_flight[0].event.synchronize()

    # ------------------------------------------------------------ launches

    def _stream(self):
        stream = self.streams[self._next_stream]
        self._next_stream = (self._next_stream + 1) % len(self.streams)
        return stream

    @torch.no_grad()
    def _launch_pq(self, tickets):
        stream = self._stream()
        launch = _Launch("pq", tickets, stream)
        stream.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(stream):
            preds, targets, split = [], [], []
            if len(tickets) == 1:
                # A batch of one request is evaluated as submitted (no repacking
                # copy); per-image results do not depend on the batch.
                prediction, target = tickets[0].request.prediction, tickets[0].request.target
                split.append(target.segment_map.shape[0] if not target.packed else len(target.image_sizes))
            else:
                for ticket in tickets:
                    p = _images(ticket.request.prediction, crowd=False)
                    g = _images(ticket.request.target, crowd=True)
                    if [s for *_, s in p] != [s for *_, s in g]:
                        raise ValueError("prediction and target must have matching image sizes")
                    preds += p
                    targets += g
                    split.append(len(g))
                prediction = PanopticBatch.pack([m for m, *_ in preds], [k for _, k, *_ in preds])
                target = PanopticBatch.pack([m for m, *_ in targets], [k for _, k, *_ in targets],
                                            _crowd_rows(targets))
            pq = panoptic_quality(prediction, target, self.num_classes, validate=False)
            launch.host = torch.empty(pq.shape, dtype=pq.dtype, pin_memory=True)
            launch.host.copy_(pq, non_blocking=True)
            launch.event.record(stream)
            launch.keep = [prediction, target, pq]
        launch.split = split
        self._started(launch)

    @torch.no_grad()
    def _launch_prefix(self, tickets):
        stream = self._stream()
        launch = _Launch("prefix", tickets, stream)
        stream.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(stream):
            masks, classes, targets, points, split = [], [], [], [], []
            for ticket in tickets:
                request = ticket.request
                g = _images(request.target, crowd=True)
                if len(g) != len(request.prediction.masks) or any(
                        tuple(m.shape[1:]) != s for m, (*_, s) in zip(request.prediction.masks, g)):
                    raise ValueError("prediction and target must have matching batch and image dimensions")
                masks += [m.contiguous() for m in request.prediction.masks]
                classes += list(request.prediction.classes)
                targets += g
                image_points = [self._points(request, m.shape[0]) for m in request.prediction.masks]
                points += image_points
                split.append(image_points)
            if len(tickets) == 1:
                # One request: its own target maps, without a repacking copy.
                target = tickets[0].request.target.compact()
                if target.crowd is None:
                    target = PanopticBatch(target.segment_map, target.classes,
                                           torch.zeros_like(target.classes, dtype=torch.bool),
                                           image_sizes=target.image_sizes)
            else:
                target = PanopticBatch.pack([m for m, *_ in targets], [k for _, k, *_ in targets],
                                            _crowd_rows(targets))
            maps, gc, crowd = _prepare(target, self.num_classes, False)
            sizes = (list(target.image_sizes) if target.packed
                     else [tuple(target.segment_map.shape[1:])] * target.segment_map.shape[0])
            out = launch_prefix(masks, classes, maps, gc, crowd, sizes, target.packed, points, self.things,
                                self.thresholds, self.num_classes, self.device)
            _, summaries, _, miou, _ = finalize_arrays(out["tp"], out["fp"], out["fn"], out["sums"],
                                                       out["confusion"], self.things, copies=False)
            rows = out["rows"]
            values = torch.cat([summaries[:, 0, :].reshape(rows, 3), miou.reshape(rows, 1)], 1)
            launch.host = torch.empty(values.shape, dtype=values.dtype, pin_memory=True)
            launch.host.copy_(values, non_blocking=True)
            launch.event.record(stream)
            launch.keep = [target, out, summaries, miou, values, masks]
            self.prefix_pixels_streamed += sum(h * w * (4 + m.shape[0]) for (h, w), m in zip(sizes, masks))
        launch.split = split
        self._started(launch)

    def _started(self, launch):
        now = time.perf_counter()
        for ticket in launch.tickets:
            ticket.launched = now
        self.launches[launch.kind] += 1
        self._in_flight.append(launch)

    # ------------------------------------------------------------ completion

    def _complete(self) -> int:
        done = 0
        while self._in_flight:
            # Launches complete in issue order per stream; check each.
            ready = [launch for launch in self._in_flight if launch.event.query()]
            if not ready:
                break
            now = time.perf_counter()
            for launch in ready:
                self._in_flight.remove(launch)
                self._finish(launch, now)
                done += len(launch.tickets)
        return done

    def _finish(self, launch, now):
        host = launch.host.numpy()
        if launch.kind == "pq":
            start = 0
            for ticket, count in zip(launch.tickets, launch.split):
                ticket.result = host[start:start + count].copy()
                start += count
        else:
            row = 0
            for ticket, image_points in zip(launch.tickets, launch.split):
                per_image = []
                for pts in image_points:
                    block = host[row:row + len(pts)]
                    per_image.append({"prefix": list(pts), "pq": block[:, 0].copy(), "sq": block[:, 1].copy(),
                                      "rq": block[:, 2].copy(), "miou": block[:, 3].copy()})
                    row += len(pts)
                ticket.result = per_image
        for ticket in launch.tickets:
            ticket.completed = now
            self.admission.release(ticket)
        launch.keep = None
