"""Continuous-batching server for per-image rewards on one GPU.

Requests are per-image panoptic quality (``PQRequest``: the reference's
``panoptic_quality``, evaluator.py:471-544) or generation-order prefix rewards
(``PrefixRequest``: ``segmentation_prefixes``, prefix.py:11-122).  The server
has no counterpart in the reference; it is the serving form of the same
kernels.  Its loop runs in native code (a worker thread of
libpanoptic_hip.so, ``pe_server_*``):

* **admission by HBM budget** -- the budget is one device arena; a batch is
  admitted when its merged layout fits, requests wait first-in first-out
  while it does not, and a request larger than the whole budget runs alone;
* **continuous batching** -- every pass of the loop merges all waiting
  requests of a kind that fit into one packed launch (images of any size) on
  the next of a few streams, so work admitted while earlier batches run
  overlaps with them on the device;
* **device-resident statistics** -- per-image sufficient statistics and the
  FP64 metrics stay on the GPU; only per-image PQ (or per-checkpoint
  PQ/SQ/RQ/mIoU) is copied to pinned host memory.

This module checks request shapes on the host, hands the server per-image
device addresses, and keeps the request's tensors alive until its result is
back.  Results do not depend on how requests are batched: per-image PQ folds
each image's IoUs in slot order and prefix snapshots are per image, so a
request gets bit-identical numbers alone or in any mix.  Input values are not
validated on the device (validation reads a result on the host, which would
serialise the loop); callers validate with the plain API beforehand if needed.
"""

from __future__ import annotations

import collections
import ctypes
import time
from dataclasses import dataclass, field
from typing import Sequence

import numpy as np
import torch

from . import _native
from .evaluator import PanopticBatch
from .instance import InstanceBatch
from .prefix import _check_arguments

_i64, _int, _ptr = ctypes.c_int64, ctypes.c_int, ctypes.c_void_p
for _name, _sig in {
    "pe_server_create": (_ptr, [_int, _i64, _int, _i64, _ptr, _ptr, _i64, _i64]),
    "pe_server_destroy": (_int, [_ptr]),
    "pe_server_submit": (_i64, [_ptr, _int, _i64, _ptr, _ptr, _ptr, _i64]),
    "pe_server_poll": (_i64, [_ptr, _ptr, _i64, _i64]),
    "pe_server_take": (_int, [_ptr, _i64, _ptr, _i64, _ptr, _ptr, _ptr]),
    "pe_server_stats": (_int, [_ptr, _ptr]),
    "pe_server_footprint": (_i64, [_i64, _i64, _int, _i64, _ptr, _ptr, _ptr]),
}.items():
    _native.register(_name, *_sig)

PQ, PREFIX = 0, 1
PINNED_BYTES = 64 << 20


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
    id: int = -1
    points: list = field(default_factory=list)
    keep: list = field(default_factory=list)

    @property
    def done(self) -> bool:
        return self.completed is not None

    @property
    def latency(self) -> float | None:
        return None if self.completed is None else self.completed - self.arrival


@dataclass
class AdmissionStats:
    """The server's admission state (from the native loop)."""

    budget: int
    in_use: int = 0
    waiting: int = 0
    running: int = 0       # launches in flight
    blocked: int = 0       # loop passes at which the head request waited for memory


def _image_views(batch: PanopticBatch, *, crowd: bool):
    """Per-image (H x W int32 map, classes row, crowd row or None, (H, W))
    views of a compact batch (no copies)."""
    batch = batch.compact()
    maps, classes = batch.segment_map, batch.classes
    if maps.dtype != torch.int32 or classes.dtype != torch.int32 or classes.ndim != 2:
        raise ValueError("served maps must be compact int32 slots with int32 B x S classes")
    if batch.crowd is not None and (batch.crowd.dtype != torch.bool or batch.crowd.shape != classes.shape):
        raise ValueError("crowd must be bool with the same shape as classes")
    maps, classes = maps.contiguous(), classes.contiguous()
    flags = batch.crowd.contiguous() if crowd and batch.crowd is not None else None
    rows = []
    if batch.packed:
        if maps.ndim != 1 or classes.shape[0] != len(batch.image_sizes) or \
                maps.numel() != sum(h * w for h, w in batch.image_sizes):
            raise ValueError("a packed batch needs flat maps of sum(H_i * W_i) pixels and classes B x S")
        start = 0
        for b, (h, w) in enumerate(batch.image_sizes):
            rows.append((maps[start:start + h * w], classes[b], None if flags is None else flags[b], (h, w)))
            start += h * w
    else:
        if maps.ndim != 3 or classes.shape[0] != maps.shape[0]:
            raise ValueError("expected maps B x H x W and classes B x S")
        for b in range(maps.shape[0]):
            rows.append((maps[b], classes[b], None if flags is None else flags[b], tuple(maps.shape[1:])))
    return rows, [maps, classes, flags]


def _points(checkpoints, n):
    if checkpoints is None:
        return list(range(1, n + 1)) if n else [0]
    return sorted(set(checkpoints))


def describe(request, num_classes: int, thresholds):
    """(kind, per-image descriptors int64 [B, 9], checkpoints per image,
    tensors to keep alive) of a request; raises ValueError for requests the
    server cannot evaluate.  Host-side shape checks only."""
    if isinstance(request, PQRequest):
        pred, keep_p = _image_views(request.prediction, crowd=False)
        tgt, keep_t = _image_views(request.target, crowd=True)
        if [s for *_, s in pred] != [s for *_, s in tgt]:
            raise ValueError("prediction and target must have matching image sizes")
        if not tgt:
            raise ValueError("a request needs at least one image")
        device = request.target.segment_map.device
        rows = []
        for (pm, pc, _, (h, w)), (gm, gc, gcrowd, _) in zip(pred, tgt):
            if pm.device != device or gm.device != device:
                raise ValueError("batch tensors must share a device")
            rows.append([gm.data_ptr(), pm.data_ptr(), gc.data_ptr(), pc.data_ptr(),
                         0 if gcrowd is None else gcrowd.data_ptr(), h, w, gc.numel(), pc.numel()])
        return PQ, np.asarray(rows, np.int64), [], keep_p + keep_t
    if isinstance(request, PrefixRequest):
        _, requested = _check_arguments(thresholds, request.checkpoints)
        tgt, keep_t = _image_views(request.target, crowd=True)
        masks, classes = request.prediction.masks, request.prediction.classes
        if len(masks) != len(tgt) or len(classes) != len(masks) or not tgt:
            raise ValueError("prediction and target must have matching batch and image dimensions")
        rows, points, keep = [], [], list(keep_t)
        for m, c, (gm, gc, gcrowd, (h, w)) in zip(masks, classes, tgt):
            if m.ndim != 3 or m.dtype != torch.bool or tuple(m.shape[1:]) != (h, w):
                raise ValueError("masks must be bool N x H x W matching their target image")
            if c.shape != (m.shape[0],) or c.dtype not in (torch.int32, torch.int64):
                raise ValueError("classes must be int32/int64 N tensors")
            if requested and requested[-1] > m.shape[0]:
                raise ValueError("checkpoint exceeds prediction count")
            m, c = m.contiguous(), c.to(torch.int32).contiguous()
            keep += [m, c]
            rows.append([gm.data_ptr(), m.data_ptr(), c.data_ptr(), gc.data_ptr(),
                         0 if gcrowd is None else gcrowd.data_ptr(), h, w, gc.numel(), m.shape[0]])
            points.append(_points(requested, m.shape[0]))
        return PREFIX, np.asarray(rows, np.int64), points, keep
    raise TypeError("requests are PQRequest or PrefixRequest")


def _flat_points(points):
    counts = np.asarray([len(p) for p in points], np.int64)
    flat = np.asarray([q for p in points for q in p] or [0], np.int32)
    return flat, counts


def footprint(request, num_classes: int, iou_thresholds=(0.5, 0.75)) -> int:
    """Device bytes the server reserves for ``request`` when it runs alone."""
    thresholds, _ = _check_arguments(iou_thresholds, None)
    kind, rows, points, _ = describe(request, num_classes, thresholds)
    flat, counts = _flat_points(points)
    size = _native.library().pe_server_footprint(num_classes, len(thresholds), kind, len(rows),
                                                 rows.ctypes.data, flat.ctypes.data, counts.ctypes.data)
    return _native.check(size)


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
        if hbm_budget <= 0:
            raise ValueError("the HBM budget must be positive")
        self.num_classes = num_classes
        self.device = _native.require_gpu(device)
        if self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        self.thresholds, _ = _check_arguments(iou_thresholds, None)
        things = np.asarray(list(isthing) if isthing is not None else [True] * num_classes, np.uint8)
        thresholds = np.asarray(self.thresholds, np.float64)
        lib = _native.library()
        handle = lib.pe_server_create(self.device.index, int(hbm_budget), int(streams), num_classes,
                                      things.ctypes.data, thresholds.ctypes.data, len(thresholds), PINNED_BYTES)
        if not handle:
            raise RuntimeError(f"panoptic HIP kernels: {lib.pe_last_error().decode(errors='replace')}")
        self._handle = handle
        self.budget = int(hbm_budget)
        self._tickets: dict[int, Ticket] = {}
        self._ids = (ctypes.c_int64 * 256)()

    def close(self):
        """Finish every submitted request and release the server."""
        if getattr(self, "_handle", None):
            self.drain()
            _native.library().pe_server_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass

    # ------------------------------------------------------------ interface

    def submit(self, request, arrival: float | None = None) -> Ticket:
        """Queue a request; returns its ticket (``result`` filled on
        completion).  Malformed requests raise ValueError here, before
        anything is queued."""
        kind, rows, points, keep = describe(request, self.num_classes, self.thresholds)
        flat, counts = _flat_points(points)
        arrival = time.perf_counter() if arrival is None else arrival
        ticket = Ticket(request, len(rows), 0, arrival, points=points, keep=keep)
        # Work already queued on the caller's stream (inputs, packing) comes
        # first: the server's streams run independently of it.
        torch.cuda.current_stream(self.device).synchronize()
        ticket.id = _native.check(_native.library().pe_server_submit(
            self._handle, kind, len(rows), rows.ctypes.data, flat.ctypes.data, counts.ctypes.data,
            int(arrival * 1e9)))
        self._tickets[ticket.id] = ticket
        return ticket

    @property
    def idle(self) -> bool:
        return not self._tickets

    def poll(self, wait_us: int = 0) -> int:
        """Collect finished requests without blocking (or up to ``wait_us``
        microseconds); returns how many completed."""
        n = _native.library().pe_server_poll(self._handle, self._ids, len(self._ids), int(wait_us))
        for k in range(n):
            self._finish(self._tickets.pop(self._ids[k]))
        return n

    def drain(self):
        """Wait until every submitted request has completed."""
        while self._tickets:
            self.poll(wait_us=2000)

    @property
    def admission(self) -> AdmissionStats:
        s = self._stats()
        return AdmissionStats(budget=int(s[1]), in_use=int(s[0]), waiting=int(s[2]), running=int(s[3]),
                              blocked=int(s[4]))

    @property
    def launches(self) -> collections.Counter:
        s = self._stats()
        return collections.Counter({"pq": int(s[5]), "prefix": int(s[6])})

    @property
    def launched_requests(self) -> int:
        return int(self._stats()[8])

    def _stats(self):
        out = (ctypes.c_int64 * 9)()
        _native.call("pe_server_stats", self._handle, out)
        return out

    # ------------------------------------------------------------ results

    def _finish(self, ticket: Ticket):
        lib = _native.library()
        length = ctypes.c_int64()
        times = (ctypes.c_int64 * 3)()
        message = ctypes.create_string_buffer(256)
        capacity = ticket.images if isinstance(ticket.request, PQRequest) else 4 * sum(map(len, ticket.points))
        out = np.empty(max(capacity, 1), np.float64)
        code = _native.check(lib.pe_server_take(self._handle, ticket.id, out.ctypes.data, out.size,
                                                ctypes.byref(length), times, message))
        ticket.launched = ticket.admitted = times[1] * 1e-9 if times[1] else None
        ticket.completed = times[2] * 1e-9
        ticket.keep = []
        if code:
            ticket.error = RuntimeError(f"reward launch failed: {message.value.decode(errors='replace')}")
            return
        values = out[:length.value]
        if isinstance(ticket.request, PQRequest):
            ticket.result = values.copy()
            return
        per_image, row = [], 0
        for pts in ticket.points:
            block = values[4 * row:4 * (row + len(pts))].reshape(len(pts), 4)
            per_image.append({"prefix": list(pts), "pq": block[:, 0].copy(), "sq": block[:, 1].copy(),
                              "rq": block[:, 2].copy(), "miou": block[:, 3].copy()})
            row += len(pts)
        ticket.result = per_image
