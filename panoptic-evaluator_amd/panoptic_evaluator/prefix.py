"""Generation-order evaluation (prefix rewards) on AMD GPUs.

Public surface of the reference's panoptic_evaluator/prefix.py
(segmentation_prefixes :11-122).  All images of a call run in two kernel
launches (one pixel pass over every mask; one scoring launch whose
workgroups each take a slice of an image's checkpoints or the instance
matching of a few IoU thresholds) plus the FP64 finaliser over every
checkpoint row.
"""

from __future__ import annotations

import ctypes
import math

import torch

from . import _native
from .evaluator import MSG_ACTIVE, MSG_CLASSES, _prepare, finalize
from .instance import InstanceBatch, _cat, offsets

_i64, _int, _ptr = ctypes.c_int64, ctypes.c_int, ctypes.c_void_p
_native.register("pe_prefix_workspace_bytes", _i64, [_i64] * 7)
_native.register("pe_prefix", _int, [_i64, _ptr, _ptr, _i64, _i64, _i64, _ptr, _ptr, _i64, _i64, _i64, _ptr, _ptr,
                                     _ptr, _ptr,
                                     _ptr, _i64, _ptr, _ptr, _i64, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr,
                                     _ptr, _ptr, _ptr, _ptr, _i64, _int, _ptr])


def _check_arguments(iou_thresholds, checkpoints):
    """prefix.py:28-36."""
    thresholds = tuple(iou_thresholds)
    if not thresholds or any(isinstance(t, bool) or not math.isfinite(t) or not 0 < t <= 1 for t in thresholds):
        raise ValueError("iou_thresholds must be finite values in (0, 1]")
    requested = None if checkpoints is None else tuple(checkpoints)
    if requested is not None and (any(type(k) is not int or k < 0 for k in requested)
                                  or any(a >= b for a, b in zip(requested, requested[1:]))):
        raise ValueError("checkpoints must be strictly increasing nonnegative integers")
    return thresholds, requested


def _raise_prefix_error(code):
    if code & 1:
        raise ValueError("segment map slot outside metadata capacity")
    if code & 2:
        raise ValueError(MSG_CLASSES)
    if code & 4:
        raise ValueError(MSG_ACTIVE)
    if code & 8:
        raise ValueError("classes must be valid category indices")
    if code & 32:
        raise ValueError("checkpoints must be strictly increasing nonnegative integers")
    raise ValueError("unknown target slot or empty prediction mask")


def launch_prefix(masks_list, classes_list, maps, gc, crowd, sizes, packed, points, things, thresholds,
                  num_classes, device):
    """Launch the prefix kernels for images with per-image checkpoints
    (``points[b]``: strictly increasing prefix lengths of image b) on the
    current stream; returns the per-row device outputs (rows in image order)
    and the error word.  ``keep`` holds the launch's scratch alive."""
    B = len(sizes)
    counts = [m.shape[0] for m in masks_list]
    G, C, T = gc.shape[1], num_classes, len(thresholds)
    rows = sum(len(p) for p in points)
    masks_list = [m.contiguous() for m in masks_list]
    n_off = offsets(counts, device)
    k_off = offsets([len(p) for p in points], device)
    flat_points = _native.device_table([k for p in points for k in p], device).int()
    cls = _cat(list(classes_list), torch.int32, device)
    pointers = _native.device_table([m.data_ptr() for m in masks_list], device)
    threshold_tensor = torch.tensor(thresholds, dtype=torch.float64, device=device)
    sum_n, sum_tri = sum(counts), sum(n * (n + 1) // 2 for n in counts)
    longs = dict(dtype=torch.int64, device=device)
    doubles = dict(dtype=torch.float64, device=device)
    tp, fp, fn = (torch.empty((rows, C), **longs) for _ in range(3))
    sums = torch.empty((rows, C), **doubles)
    confusion = torch.empty((rows, C, C + 1), **longs)
    itp, ifp, ifn = (torch.empty((rows, T, C), **longs) for _ in range(3))
    precision, recall = torch.empty((rows, T, C), **doubles), torch.empty((rows, T, C), **doubles)
    error = torch.empty(1, dtype=torch.int32, device=device)
    if packed:
        pix_off, HW = offsets([h * w for h, w in sizes], device), maps.numel()
    else:
        pix_off, HW = None, sizes[0][0] * sizes[0][1] if sizes else 1
    size = max(_native.library().pe_prefix_workspace_bytes(B, G, C, T, sum_n, sum_tri,
                                                           HW if packed else 0), 256)
    workspace = torch.empty(size, dtype=torch.uint8, device=device)
    aligned = int(all((h * w) % 16 == 0 for h, w in sizes) and all(m.data_ptr() % 16 == 0 for m in masks_list))
    stream = _native.stream(device)
    _native.call("pe_prefix", B, maps.data_ptr(), _native.ptr(pix_off), HW, G, C, pointers.data_ptr(),
                 n_off.data_ptr(), sum_n, sum_tri, max(counts, default=0),
                 cls.data_ptr(), gc.data_ptr(), crowd.data_ptr(), things.data_ptr(), threshold_tensor.data_ptr(), T,
                 k_off.data_ptr(), flat_points.data_ptr(), rows, tp.data_ptr(), fp.data_ptr(), fn.data_ptr(),
                 sums.data_ptr(), confusion.data_ptr(), itp.data_ptr(), ifp.data_ptr(), ifn.data_ptr(),
                 precision.data_ptr(), recall.data_ptr(), error.data_ptr(), workspace.data_ptr(), workspace.numel(),
                 aligned, stream)
    return dict(rows=rows, tp=tp, fp=fp, fn=fn, sums=sums, confusion=confusion, itp=itp, ifp=ifp, ifn=ifn,
                precision=precision, recall=recall, error=error, thresholds=threshold_tensor,
                keep=(workspace, pointers, n_off, k_off, flat_points, cls, pix_off))


@torch.no_grad()
def segmentation_prefixes(prediction: InstanceBatch, target, num_classes: int, *, isthing=None, checkpoints=None,
                          iou_thresholds=(0.5, 0.75), validate=True):
    """Evaluate independent masks in generation order, later masks drawn on top.

    Returns one list per image of dictionaries with ``prefix``, the standard
    panoptic/semantic results of the visible canvas, and ``instance``
    counts/precision/recall of shape [threshold, class] from greedy matching in
    generation order (same-class non-crowd GT preferred over reusable crowd GT,
    IoU >= threshold, no detection cap; only thing predictions take part).
    Checkpoints default to 1..N (or [0] for an empty sequence); explicit
    checkpoints must be strictly increasing integers in 0..N for every image.
    """
    if not isinstance(num_classes, int) or isinstance(num_classes, bool) or num_classes <= 0:
        raise ValueError("num_classes must be a positive integer")
    if isthing is not None and len(isthing) != num_classes:
        raise ValueError("isthing must contain one flag per category")
    device = _native.require_gpu(target.segment_map.device)
    thresholds, requested = _check_arguments(iou_thresholds, checkpoints)
    things = torch.tensor(list(isthing) if isthing is not None else [True] * num_classes, dtype=torch.bool,
                          device=device)
    target = target.compact()
    maps, gc, crowd = _prepare(target, num_classes, False)
    masks_list, classes_list = prediction.masks, prediction.classes
    if len(classes_list) != len(masks_list):
        raise ValueError("all fields must contain one tensor per image")
    for masks, classes in zip(masks_list, classes_list):
        if masks.ndim != 3 or masks.dtype != torch.bool or min(masks.shape[1:]) <= 0:
            raise ValueError("masks must be bool N x H x W with positive image dimensions")
        if classes.shape != (masks.shape[0],) or classes.dtype not in (torch.int32, torch.int64):
            raise ValueError("classes must be int32/int64 N tensors")
        if masks.device != device or classes.device != device:
            raise ValueError("inputs must share evaluator device")
    sizes = target.image_sizes if target.packed else [tuple(maps.shape[1:])] * maps.shape[0]
    if len(masks_list) != len(sizes) or any(tuple(m.shape[1:]) != tuple(size) for m, size in zip(masks_list, sizes)):
        raise ValueError("prediction and target must have matching batch and image dimensions")
    counts = [m.shape[0] for m in masks_list]
    if requested and any(requested[-1] > n for n in counts):
        raise ValueError("checkpoint exceeds prediction count")
    points = [sorted(set(requested)) if requested is not None else (list(range(1, n + 1)) if n else [0])
              for n in counts]
    out = launch_prefix(masks_list, classes_list, maps, gc, crowd, sizes, target.packed, points, things,
                        thresholds, num_classes, device)
    if validate:
        code = int(out["error"].item())
        if code:
            _raise_prefix_error(code)
    rows = out["rows"]
    tp, fp, fn, sums, confusion = out["tp"], out["fp"], out["fn"], out["sums"], out["confusion"]
    itp, ifp, ifn, precision, recall = out["itp"], out["ifp"], out["ifn"], out["precision"], out["recall"]
    threshold_tensor = out["thresholds"]
    snapshots = finalize(tp, fp, fn, sums, confusion, things) if rows else []
    # Per-checkpoint views of the instance arrays, one unbind each.
    itp_rows, ifp_rows, ifn_rows, precision_rows, recall_rows = (
        (t.unbind(0) if rows else ()) for t in (itp, ifp, ifn, precision, recall))
    results, row = [], 0
    for image_points in points:
        image = []
        for k in image_points:
            result = snapshots[row]
            result["prefix"] = k
            result["instance"] = dict(iou_thresholds=threshold_tensor, tp=itp_rows[row], fp=ifp_rows[row],
                                      fn=ifn_rows[row], precision=precision_rows[row], recall=recall_rows[row])
            image.append(result)
            row += 1
        results.append(image)
    return results
