"""Drop-in replacement for the reference's pybind11 extension module.

The reference's Python package imports ``panoptic_evaluator._cuda_evaluator``
(evaluator.py:9-12) and calls eight functions on it (bindings.cpp:28-57).
This module provides those eight functions, with the same names, argument
order, dtypes, return conventions and error behaviour, over the C ABI of
libpanoptic_hip.so (include/panoptic_hip.h).  A maintainer who keeps the
reference's Python files copies this file and the built library into that
package; nothing else changes.  It needs only ``torch`` and ``ctypes`` (no
other module of this package), so it can be copied on its own.

Conventions carried over from the reference's native functions:

* argument errors (wrong device, dtype, layout or shape) raise RuntimeError,
  as TORCH_CHECK does;
* every launch goes on the current stream of the input's device; outputs are
  allocated with the torch caching allocator on that stream;
* ``update`` accumulates in place into the caller's state tensors and returns
  the 1/2/4 validation bitmask (0 = success), reporting an error before any
  total changes; with validate=False it performs no host synchronisation;
* ``prefix`` returns its error as a 0-dim int32 tensor (1: target slot
  outside metadata capacity; 2: empty prediction mask, checked only when
  validating), read by the caller.

The library is found through the PANOPTIC_HIP_LIB environment variable, or
next to this file.
"""

from __future__ import annotations

import ctypes
import os
import pathlib
import threading

import torch

_i64, _int, _ptr, _dbl = ctypes.c_int64, ctypes.c_int, ctypes.c_void_p, ctypes.c_double

_SIGNATURES = {
    "pe_last_error": (ctypes.c_char_p, []),
    "pe_update_workspace_bytes": (_i64, [_i64] * 5),
    "pe_update": (_int, [_ptr, _int, _ptr, _int, _i64, _i64, _i64, _ptr, _ptr, _ptr, _ptr, _i64, _i64, _i64,
                         _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr,
                         _ptr, _ptr, _ptr, _i64, _int, _int, _int, _ptr]),
    "pe_compute": (_int, [_ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _i64, _i64, _ptr, _ptr, _ptr, _ptr,
                          _ptr, _ptr, _ptr, _ptr, _ptr, _ptr]),
    "pe_pack": (_int, [_ptr, _int, _i64, _i64, _ptr, _ptr, _i64, _ptr, _ptr, _ptr]),
    "pe_mask_overlap": (_int, [_ptr, _i64, _ptr, _i64, _i64, _ptr, _ptr, _ptr, _ptr]),
    "pe_coco_match": (_int, [_i64, _i64, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _int,
                             _i64, _i64, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _i64, _int, _ptr]),
    "pe_ap_workspace_bytes": (_i64, [_i64, _i64]),
    "pe_ap_compute": (_int, [_i64, _ptr, _ptr, _ptr, _ptr, _ptr, _i64, _ptr, _ptr, _ptr, _ptr, _i64, _ptr]),
    "pe_detection_histogram": (_int, [_i64, _ptr, _ptr, _ptr, _ptr, _ptr, _i64, _i64, _i64, _i64, _ptr, _int,
                                      _int, _ptr, _i64, _ptr]),
    "pe_prefix_workspace_bytes": (_i64, [_i64] * 7),
    "pe_prefix": (_int, [_i64, _ptr, _ptr, _i64, _i64, _i64, _ptr, _ptr, _i64, _i64, _i64, _ptr, _ptr, _ptr, _ptr,
                         _ptr, _i64, _ptr, _ptr, _i64, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr,
                         _ptr, _ptr, _i64, _int, _ptr]),
}

_lock = threading.Lock()
_lib = None
INT_MAX = 2 ** 31 - 1


def _library():
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                path = os.environ.get("PANOPTIC_HIP_LIB") or str(
                    pathlib.Path(__file__).with_name("libpanoptic_hip.so"))
                lib = ctypes.CDLL(path)
                for name, (restype, argtypes) in _SIGNATURES.items():
                    fn = getattr(lib, name)
                    fn.restype, fn.argtypes = restype, argtypes
                _lib = lib
    return _lib


def _require(condition, message):
    if not condition:
        raise RuntimeError(message)


def _call(name, *args):
    lib = _library()
    code = getattr(lib, name)(*args)
    if code < 0:
        raise RuntimeError(lib.pe_last_error().decode(errors="replace"))
    return code


def _p(tensor):
    return None if tensor is None else tensor.data_ptr()


def _stream(device):
    return torch.cuda.current_stream(device).cuda_stream


def _check(tensor, reference, dtype, message="incorrect tensor device, layout or dtype"):
    _require(tensor.device == reference.device and tensor.is_contiguous() and tensor.dtype == dtype, message)


def _table(values, device):
    """Small int64 device array (offsets, pointer tables) on the current stream."""
    return torch.tensor(list(values), dtype=torch.int64).to(device, non_blocking=False)


# ------------------------------------------------------------------ panoptic

def update(gt, pred, gc, pc, crowd, pairs, areas, matched, tp, fp, fn, iou_sum, confusion,
           gt_ids=None, gt_slots=None, pred_ids=None, pred_slots=None, error=None, validate=True, reward=False):
    """Joint slot histogram, validation and PQ accumulation (bindings.cpp:4-9,
    :29-49; update_cuda, histogram_kernel.cu:257-341).  Returns the 1/2/4
    validation bitmask; totals change only when it is 0."""
    _require(gt.is_cuda and gt.dim() == 3, "gt must be CUDA B x H x W")
    for t in (gt, pred):
        _require(t.device == gt.device and t.is_contiguous() and t.dtype in (torch.int32, torch.int64),
                 "maps must be contiguous int32 or int64 on the same CUDA device")
    _require(error is not None, "error must be a CUDA int32 tensor")
    for t in (gc, pc, matched, error):
        _check(t, gt, torch.int32)
    for t in (pairs, areas, tp, fp, fn):
        _check(t, gt, torch.int64)
    if not reward:
        _check(confusion, gt, torch.int64)
    _check(crowd, gt, torch.bool)
    _check(iou_sum, gt, torch.float64)
    _require(pred.shape == gt.shape and gc.dim() == 2 and pc.dim() == 2 and gc.shape[0] == gt.shape[0]
             and pc.shape[0] == gt.shape[0] and crowd.shape == gc.shape, "incorrect input shapes")
    B, H, W = gt.shape
    G, P = gc.shape[1], pc.shape[1]
    C = tp.shape[1] if reward and tp.dim() == 2 else tp.numel()
    _require(0 < C < INT_MAX and 0 < G < INT_MAX and 0 < P < INT_MAX, "invalid capacities")
    _require(pairs.shape == (B, G, P) and areas.shape == (B, G + P) and matched.shape == gc.shape
             and error.numel() == 1, "incorrect scratch shapes")
    _require((tp.dim() == 2 and tp.shape[0] == B if reward else tp.dim() == 1) and fp.shape == tp.shape
             and fn.shape == tp.shape and iou_sum.shape == tp.shape
             and (reward or (confusion.dim() == 2 and confusion.shape == (C, C + 1))), "incorrect state shapes")
    lookups = []
    for ids, slots, classes in ((gt_ids, gt_slots, gc), (pred_ids, pred_slots, pc)):
        _require((ids is None) == (slots is None), "lookup IDs and slots must be supplied together")
        if ids is not None:
            _check(ids, gt, torch.int64)
            _check(slots, gt, torch.int32)
            _require(ids.shape == classes.shape and slots.shape == classes.shape, "incorrect lookup shape")
        lookups += [ids, slots]
    size = max(_library().pe_update_workspace_bytes(B, G, P, C, 0), 256)
    workspace = torch.empty(size, dtype=torch.uint8, device=gt.device)
    # The reference's state keeps no exact IoU words (iou_exact = NULL): each
    # update adds its correctly rounded per-class IoU sum to iou_sum, so a
    # caller that zeroes its tensors on reset (evaluator.py:274-278) starts
    # from zero.  pairs is caller scratch of unknown content: zeroed first.
    return _call("pe_update", gt.data_ptr(), gt.element_size(), pred.data_ptr(), pred.element_size(), B, H, W,
                 None, gc.data_ptr(), pc.data_ptr(), crowd.data_ptr(), G, P, C, pairs.data_ptr(),
                 areas.data_ptr(), matched.data_ptr(), tp.data_ptr(), fp.data_ptr(), fn.data_ptr(),
                 iou_sum.data_ptr(), None if reward else confusion.data_ptr(), *(_p(t) for t in lookups),
                 error.data_ptr(), None, workspace.data_ptr(), workspace.numel(), int(bool(validate)),
                 int(bool(reward)), 0, _stream(gt.device))


def _finalize(tp, fp, fn, sums, confusion, things):
    """pe_compute over K rows: [per_class K x 4 x C, summaries K x 3 x 3,
    counts K x 3, miou K, copies of tp, fp, fn, sums, confusion]."""
    K, C = tp.shape
    doubles = dict(dtype=torch.float64, device=tp.device)
    per_class = torch.empty((K, 4, C), **doubles)
    summaries = torch.empty((K, 3, 3), **doubles)
    counts = torch.empty((K, 3), dtype=torch.int64, device=tp.device)
    miou = torch.empty(K, **doubles)
    copies = [torch.empty_like(t) for t in (tp, fp, fn, sums, confusion)]
    if K:
        _call("pe_compute", tp.data_ptr(), fp.data_ptr(), fn.data_ptr(), sums.data_ptr(), confusion.data_ptr(),
              things.data_ptr(), K, C, per_class.data_ptr(), summaries.data_ptr(), counts.data_ptr(),
              miou.data_ptr(), *(t.data_ptr() for t in copies), _stream(tp.device))
    return [per_class, summaries, counts, miou, *copies]


def compute(tp, fp, fn, sums, confusion, things):
    """FP64 finaliser and state snapshots (bindings.cpp:10-11; compute_cuda,
    compute_kernel.cu:66-95): [per_class 4 x C, summaries 3 x 3, counts 3,
    miou (0-dim), tp, fp, fn, sums, confusion copies]."""
    _require(tp.is_cuda and tp.dim() == 1 and 0 < tp.numel() < INT_MAX, "tp must be nonempty CUDA category counts")
    for t in (tp, fp, fn, sums, confusion, things):
        _require(t.device == tp.device and t.is_contiguous(), "incorrect device or layout")
    for t in (tp, fp, fn, confusion):
        _require(t.dtype == torch.int64, "counts must be int64")
    C = tp.numel()
    _require(fp.shape == tp.shape and fn.shape == tp.shape and sums.shape == tp.shape and things.shape == tp.shape
             and sums.dtype == torch.float64 and things.dtype == torch.bool and confusion.dim() == 2
             and confusion.shape == (C, C + 1), "incorrect state shapes or dtypes")
    out = _finalize(tp[None], fp[None], fn[None], sums[None], confusion[None], things)
    per_class, summaries, counts, miou = out[0][0], out[1][0], out[2][0], out[3][0]
    return [per_class, summaries, counts, miou] + [t[0] for t in out[4:]]


def pack(maps, ids, slots):
    """Original IDs -> compact int32 slots (bindings.cpp:12; pack_cuda,
    histogram_kernel.cu:343-366): [packed B x H x W int32, error int32[1]]."""
    _require(maps.is_cuda and maps.dim() == 3 and maps.is_contiguous()
             and maps.dtype in (torch.int32, torch.int64), "maps must be contiguous CUDA int32 or int64 B x H x W")
    _check(ids, maps, torch.int64)
    _check(slots, maps, torch.int32)
    _require(ids.dim() == 2 and slots.shape == ids.shape and ids.shape[0] == maps.shape[0]
             and 0 < ids.shape[1] < INT_MAX, "incorrect lookup shapes")
    B, H, W = maps.shape
    output = torch.empty((B, H, W), dtype=torch.int32, device=maps.device)
    error = torch.zeros(1, dtype=torch.int32, device=maps.device)
    _call("pe_pack", maps.data_ptr(), maps.element_size(), B, H * W, ids.data_ptr(), slots.data_ptr(), ids.shape[1],
          output.data_ptr(), error.data_ptr(), _stream(maps.device))
    return [output, error]


# ------------------------------------------------------------------ instance

def _check_metadata(reference, pc, gc, pa, ga, crowd, areas, D, G):
    """instance_kernel.cu:271-289."""
    for t in (pc, gc):
        _check(t, reference, torch.int32, "invalid instance tensor device, layout or dtype")
    for t in (pa, ga, areas):
        _check(t, reference, torch.float64, "invalid instance tensor device, layout or dtype")
    _check(crowd, reference, torch.bool, "invalid instance tensor device, layout or dtype")
    _require(D <= INT_MAX and G <= INT_MAX and D * G <= INT_MAX, "too many instance pairs")
    for t in (pc, pa):
        _require(t.dim() == 1 and t.numel() == D, "prediction metadata shape mismatch")
    for t in (gc, ga, crowd, areas):
        _require(t.dim() == 1 and t.numel() == G, "target metadata shape mismatch")


def _match(inter, inter_bytes, pc, gc, pa, ga, crowd, areas, C):
    """COCO greedy matching of one image (match_instances,
    instance_kernel.cu:153-174): [flags 40 x D int8, ranks D int32, counts
    C x 4 int64]."""
    D, G = pc.numel(), gc.numel()
    device = pc.device
    status = torch.empty((D, 40), dtype=torch.int8, device=device)
    ranks = torch.empty(D, dtype=torch.int32, device=device)
    counts = torch.zeros((C, 4), dtype=torch.int64, device=device)
    det_off, gt_off, inter_off = _table([0, D], device), _table([0, G], device), _table([0, D * G], device)
    _call("pe_coco_match", 1, C, det_off.data_ptr(), gt_off.data_ptr(), inter_off.data_ptr(), pc.data_ptr(),
          pa.data_ptr(), None, gc.data_ptr(), ga.data_ptr(), crowd.data_ptr(), areas.data_ptr(), inter.data_ptr(),
          inter_bytes, D, G, status.data_ptr(), ranks.data_ptr(), counts.data_ptr(), None, None, None, 0, 0,
          _stream(device))
    # The library writes one 40-flag row per detection; the reference's
    # layout is 40 x D (area * 10 + threshold, detection).
    return [status.t().contiguous(), ranks, counts]


def instance_update(pm, gm, pc, gc, pa, ga, crowd, areas, C):
    """Mask intersections and COCO matching of one image (bindings.cpp:13-15;
    instance_update_cuda, instance_kernel.cu:282-309).  pm: D x H x W bool in
    score order, gm: G x H x W bool; pa/ga: f64 pixel areas; areas: f64 GT
    annotation areas (area ranges)."""
    _require(pm.is_cuda and pm.dim() == 3 and gm.dim() == 3 and 0 < C <= INT_MAX // 120,
             "invalid instance masks or category count")
    _check(pm, pm, torch.bool, "invalid instance tensor device, layout or dtype")
    _check(gm, pm, torch.bool, "invalid instance tensor device, layout or dtype")
    _require(pm.shape[1:] == gm.shape[1:] and pm.shape[1] > 0 and pm.shape[2] > 0,
             "instance mask dimensions must match")
    D, G = pm.shape[0], gm.shape[0]
    _check_metadata(pm, pc, gc, pa, ga, crowd, areas, D, G)
    inter = torch.empty((D, G), dtype=torch.int32, device=pm.device)
    if D and G:
        _call("pe_mask_overlap", pm.data_ptr(), D, gm.data_ptr(), G, pm.shape[1] * pm.shape[2], inter.data_ptr(),
              None, None, _stream(pm.device))
    return _match(inter, 4, pc, gc, pa, ga, crowd, areas, C)


def instance_compute(classes, ranks, flags, counts):
    """COCO AP/AR accumulation over records already in global score order
    (bindings.cpp:16-17; instance_compute_cuda, instance_kernel.cu:311-333):
    [precision 10 x 101 x C x 4 x 3, recall 10 x C x 4 x 3], -1 where a
    (class, area) has no positives."""
    _require(classes.is_cuda, "instance accumulation requires CUDA")
    _check(classes, classes, torch.int32, "invalid instance tensor device, layout or dtype")
    _check(ranks, classes, torch.int32, "invalid instance tensor device, layout or dtype")
    _check(flags, classes, torch.int8, "invalid instance tensor device, layout or dtype")
    _check(counts, classes, torch.int64, "invalid instance tensor device, layout or dtype")
    _require(classes.dim() == 1 and ranks.shape == classes.shape and flags.dim() == 2 and flags.shape[0] == 40
             and flags.shape[1] == classes.numel() and counts.dim() == 2 and counts.shape[1] == 4
             and 0 < counts.shape[0] <= INT_MAX // 120, "invalid instance state shapes")
    device, N, C = classes.device, classes.numel(), counts.shape[0]
    # The library sorts records by descending score (stable); strictly
    # decreasing stand-in scores keep the caller's global order.
    scores = torch.arange(N, 0, -1, dtype=torch.float64, device=device)
    status = flags.t().contiguous()
    precision = torch.empty((10, 101, C, 4, 3), dtype=torch.float64, device=device)
    recall = torch.empty((10, C, 4, 3), dtype=torch.float64, device=device)
    summary = torch.empty(12, dtype=torch.float64, device=device)
    size = max(_library().pe_ap_workspace_bytes(N, C), 256)
    workspace = torch.empty(size, dtype=torch.uint8, device=device)
    _call("pe_ap_compute", N, scores.data_ptr(), classes.data_ptr(), ranks.data_ptr(), status.data_ptr(),
          counts.data_ptr(), C, precision.data_ptr(), recall.data_ptr(), summary.data_ptr(), workspace.data_ptr(),
          workspace.numel(), _stream(device))
    return [precision, recall]


def instance_histogram(masks, target, order, S):
    """Detection x shared-target-slot intersections of one image
    (bindings.cpp:18-19; instance_histogram_cuda, instance_kernel.cu:335-363):
    pairs D x S int64, row d = masks[order[d]]."""
    _require(masks.is_cuda and masks.dim() == 3 and target.dim() == 2 and 0 < S < INT_MAX,
             "invalid shared ground-truth inputs")
    _check(masks, masks, torch.bool, "invalid instance tensor device, layout or dtype")
    _check(target, masks, torch.int32, "invalid instance tensor device, layout or dtype")
    _check(order, masks, torch.int64, "invalid instance tensor device, layout or dtype")
    _require(target.shape == masks.shape[1:] and target.numel() > 0 and order.dim() == 1
             and order.numel() == masks.shape[0], "shared target and detection dimensions must match")
    D, HW = masks.shape[0], target.numel()
    _require(D <= INT_MAX and D * S <= INT_MAX, "too many shared detection histogram bins or blocks")
    pairs = torch.empty((D, S), dtype=torch.int64, device=masks.device)
    if D:
        pointers = _table([masks.data_ptr()], masks.device)
        det_off = _table([0, D], masks.device)
        aligned = int(HW % 16 == 0 and masks.data_ptr() % 16 == 0)
        _call("pe_detection_histogram", 1, pointers.data_ptr(), det_off.data_ptr(), order.data_ptr(),
              target.data_ptr(), None, HW, S, D, D, pairs.data_ptr(), 0, aligned, None, 0, _stream(masks.device))
    return pairs


def instance_match_histogram(pairs, pc, gc, pa, ga, crowd, areas, C):
    """COCO matching from shared-target intersections (bindings.cpp:20-22;
    instance_match_histogram_cuda, instance_kernel.cu:365-381): pairs D x G
    int64; GT with class -1 are not candidates."""
    _require(pairs.is_cuda and pairs.dim() == 2 and 0 < C <= INT_MAX // 120,
             "invalid shared instance intersections or category count")
    _check(pairs, pairs, torch.int64, "invalid instance tensor device, layout or dtype")
    D, G = pairs.shape
    _check_metadata(pairs, pc, gc, pa, ga, crowd, areas, D, G)
    return _match(pairs, 8, pc, gc, pa, ga, crowd, areas, C)


# -------------------------------------------------------------------- prefix

def prefix(gt, masks, pc, gc, crowd, things, thresholds, points, validate):
    """Generation-order prefix evaluation of one image (bindings.cpp:24-26;
    prefix_cuda, prefix_kernel.cu:142-199): the nine outputs of the FP64
    finaliser over the K checkpoints, then instance tp, fp, fn (K x T x C
    int64), precision, recall (K x T x C f64), then the error word."""
    _require(gt.is_cuda, "prefix inputs must be CUDA tensors")
    for t in (gt, masks, pc, gc, crowd, things, thresholds):
        _require(t.device == gt.device and t.is_contiguous(), "incorrect device or layout")
    _require(gt.dtype == torch.int32 and masks.dtype == torch.bool and pc.dtype == torch.int32
             and gc.dtype == torch.int32 and crowd.dtype == torch.bool and things.dtype == torch.bool
             and thresholds.dtype == torch.float64, "incorrect prefix input dtypes")
    _require(gt.dim() == 2 and masks.dim() == 3 and pc.dim() == 1 and gc.dim() == 1 and crowd.shape == gc.shape
             and things.dim() == 1 and thresholds.dim() == 1 and masks.shape[0] == pc.numel()
             and masks.shape[1:] == gt.shape, "incorrect prefix input shapes")
    N, G, C, T, K = pc.numel(), gc.numel(), things.numel(), thresholds.numel(), len(points)
    _require(N < INT_MAX and 0 < G < INT_MAX and 0 < C < INT_MAX and T > 0 and T * C < INT_MAX and gt.numel() > 0
             and K < INT_MAX, "prefix dimensions outside supported range")
    last = -1
    for q in points:
        _require(int(q) > last and 0 <= int(q) <= N, "invalid checkpoints")
        last = int(q)
    device, HW = gt.device, gt.numel()
    longs, doubles = dict(dtype=torch.int64, device=device), dict(dtype=torch.float64, device=device)
    tp, fp, fn = (torch.empty((K, C), **longs) for _ in range(3))
    sums = torch.empty((K, C), **doubles)
    confusion = torch.empty((K, C, C + 1), **longs)
    itp, ifp, ifn = (torch.empty((K, T, C), **longs) for _ in range(3))
    precision, recall = torch.empty((K, T, C), **doubles), torch.empty((K, T, C), **doubles)
    error = torch.empty(1, dtype=torch.int32, device=device)
    pointers = _table([masks.data_ptr()], device)
    n_off, k_off = _table([0, N], device), _table([0, K], device)
    flat_points = torch.tensor([int(q) for q in points] or [0], dtype=torch.int32).to(device)
    sum_tri = N * (N + 1) // 2
    size = max(_library().pe_prefix_workspace_bytes(1, G, C, T, N, sum_tri, 0), 256)
    workspace = torch.empty(size, dtype=torch.uint8, device=device)
    aligned = int(HW % 16 == 0 and masks.data_ptr() % 16 == 0)
    _call("pe_prefix", 1, gt.data_ptr(), None, HW, G, C, pointers.data_ptr(), n_off.data_ptr(), N, sum_tri, N,
          pc.data_ptr(), gc.data_ptr(), crowd.data_ptr(), things.data_ptr(), thresholds.data_ptr(), T,
          k_off.data_ptr(), flat_points.data_ptr(), K, tp.data_ptr(), fp.data_ptr(), fn.data_ptr(), sums.data_ptr(),
          confusion.data_ptr(), itp.data_ptr(), ifp.data_ptr(), ifn.data_ptr(), precision.data_ptr(),
          recall.data_ptr(), error.data_ptr(), workspace.data_ptr(), workspace.numel(), aligned, _stream(device))
    result = _finalize(tp, fp, fn, sums, confusion, things)
    # The reference's error word: 1 for a target slot outside 0..G-1 (always
    # checked), 2 for an empty prediction mask (checked when validating).
    code = error[0] & 1
    if validate:
        code = code | ((error[0] >> 3) & 2)
    result += [itp, ifp, ifn, precision, recall, code.to(torch.int32)]
    return result
