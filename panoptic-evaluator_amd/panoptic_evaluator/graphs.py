"""HIP-graph capture of an evaluator's update(), or of a whole step.

``CapturedUpdate(evaluator, *update_args)`` records one ``evaluator.update``
(PanopticEvaluator, InstanceEvaluator or SegmentationEvaluator, built with
``validate=False``) as a HIP graph.  ``replay()`` re-runs every kernel of that
update on the same input buffers -- refill them in place between replays --
and accumulates into the evaluator's state exactly as an eager update would;
``compute()`` and ``reset()`` work as usual.  The batch layout (sizes, slot
capacities, detection counts, buffer addresses) is fixed at capture.

Capture needs no host work inside the graph: the update path allocates its
scratch on the first call, keeps small metadata tables in a device cache, and
clears memory with fill kernels (no memset nodes).  One untimed eager update
runs first to populate those; the evaluator's state is restored afterwards.

``CapturedStep(evaluator, *update_args)`` records ``reset(); update(...);
compute()`` as one graph, for loops that evaluate batches of one layout one
at a time (reward computation, per-batch metrics): a replay runs the three
with no launch gaps between them and returns compute()'s result.
"""

from __future__ import annotations

import torch

from . import _native


def _parts(evaluator):
    """(panoptic evaluator or None, instance evaluator or None)."""
    from .combined import SegmentationEvaluator
    from .evaluator import PanopticEvaluator
    from .instance import InstanceEvaluator
    if isinstance(evaluator, SegmentationEvaluator):
        return evaluator.panoptic, evaluator.instance
    if isinstance(evaluator, PanopticEvaluator):
        return evaluator, None
    if isinstance(evaluator, InstanceEvaluator):
        return None, evaluator
    raise TypeError("expected a PanopticEvaluator, InstanceEvaluator or SegmentationEvaluator")


def _scratch_of(evaluator):
    """The per-shape scratch tensors an evaluator keeps between updates."""
    from .combined import SegmentationEvaluator
    keep = []
    panoptic, _ = _parts(evaluator)
    if panoptic is not None:
        keep += [panoptic._scratch, panoptic._workspace, panoptic._default_crowd]
    if isinstance(evaluator, SegmentationEvaluator):
        keep.append(evaluator._pairs)
    return keep


class CapturedUpdate:
    """One update() of ``evaluator`` recorded as a HIP graph."""

    def __init__(self, evaluator, *args, **kwargs):
        panoptic, instance = _parts(evaluator)
        if any(part is not None and part.validate for part in (panoptic, instance)):
            raise ValueError("capture needs validate=False: validation reads its result on the host")
        self.evaluator = evaluator
        self._instance = instance
        # A combined evaluator defers the matching of captured updates to its
        # next compute() (or flushes it before its next update or reset).
        self._flush = getattr(evaluator, "_flush", None)
        if self._flush is not None:
            self._flush()
        device = (panoptic or instance).device
        saved = self._save(panoptic, instance)
        # Warm-up on a side stream (default priority): scratch allocation and
        # the device metadata cache are filled outside the capture.
        side = torch.cuda.Stream(device)
        side.wait_stream(torch.cuda.current_stream(device))
        with torch.cuda.stream(side):
            evaluator.update(*args, **kwargs)
        torch.cuda.current_stream(device).wait_stream(side)
        torch.cuda.synchronize(device)
        self._restore(panoptic, instance, saved)
        self.graph = torch.cuda.CUDAGraph()
        self._record = None
        self._pending = None
        if instance is not None:
            # The captured update's detection records live in graph memory;
            # every replay appends them, and they are copied out only if the
            # evaluator still holds them when the next replay would overwrite
            # them (a reset() in between makes the copy unnecessary).
            instance._commit = self._keep
        try:
            # The graph's kernels read the device metadata tables by address:
            # keep every table the captured update used for the graph's life.
            with _native.pinned_tables() as tables, torch.cuda.graph(self.graph):
                evaluator.update(*args, **kwargs)
            self._tables = tables
            # The deferred matching of the captured update (in graph memory)
            # is registered again after every replay.
            self._pending = getattr(evaluator, "_pending", None)
            if self._pending is not None:
                evaluator._pending = None
            # Scratch the evaluator allocated outside the capture (per-shape
            # buffers, kept between updates) is read by the graph's kernels by
            # address: keep it alive even if the evaluator later replaces it.
            self._scratch = _scratch_of(evaluator)
        finally:
            if instance is not None:
                del instance._commit  # back to the class's method
        torch.cuda.synchronize(device)

    def _keep(self, record):
        self._record = record

    @staticmethod
    def _save(panoptic, instance):
        state = {}
        if panoptic is not None:
            state["panoptic"] = panoptic._state.clone()
        if instance is not None:
            state["counts"] = instance._counts.clone()
            state["records"] = list(instance._records)
        return state

    @staticmethod
    def _restore(panoptic, instance, state):
        if panoptic is not None:
            panoptic._state.copy_(state["panoptic"])
        if instance is not None:
            instance._counts.copy_(state["counts"])
            instance._records = state["records"]

    def replay(self):
        """Run the recorded update once more (on the current stream's device)."""
        if self._flush is not None:
            # The previous update's deferred matching reads the table this
            # replay refills, and completes the flags of its record.
            self._flush()
        if self._record is not None:
            records = self._instance._records
            for i, record in enumerate(records):
                if record is self._record:
                    records[i] = tuple(t.clone() for t in record)
        self.graph.replay()
        if self._record is not None:
            self._instance._records.append(self._record)
        if self._pending is not None:
            self.evaluator._pending = self._pending
        return self.evaluator


class CapturedStep:
    """``reset(); prepare(); update(*args); compute()`` of ``evaluator``
    recorded as one HIP graph (``prepare``, optional: device work that makes
    the update's inputs, such as decoding RLE masks).  ``replay()`` re-runs the three on the same input buffers
    (refill them in place between replays) and returns compute()'s result,
    whose tensors live in the graph's memory: the next replay overwrites them,
    so clone what must outlive it.  After a replay the evaluator holds that
    update's state, as after the eager calls (a later eager update() adds to
    it, compute() recomputes it).  The batch layout is fixed at capture."""

    def __init__(self, evaluator, *args, prepare=None, prepare_masks_only=False, **kwargs):
        panoptic, instance = _parts(evaluator)
        if any(part is not None and part.validate for part in (panoptic, instance)):
            raise ValueError("capture needs validate=False: validation reads its result on the host")
        self.evaluator = evaluator
        self._instance = instance
        self._flush = getattr(evaluator, "_flush", None)
        # prepare(): device work that makes the update's inputs (such as
        # RleBatch.decode), recorded between reset and update.  With
        # prepare_masks_only (prepare writes only PackedMasks bit rows and
        # extents, as RleBatch.decode does) a combined evaluator's step runs it
        # on a branch of its own: the panoptic branch, score order and grouping
        # do not wait for it, only the detection histogram does.
        early = prepare is not None and prepare_masks_only and hasattr(evaluator, "_masks_ready")
        prepare = prepare or (lambda: None)
        device = (panoptic or instance).device
        # Warm-up step on a side stream (default priority): scratch,
        # workspaces and the device metadata cache are filled outside the
        # capture.
        side = torch.cuda.Stream(device)
        side.wait_stream(torch.cuda.current_stream(device))
        with torch.cuda.stream(side):
            evaluator.reset()
            prepare()
            evaluator.update(*args, **kwargs)
            evaluator.compute()
        torch.cuda.current_stream(device).wait_stream(side)
        torch.cuda.synchronize(device)
        self.graph = torch.cuda.CUDAGraph()
        # A combined evaluator clears its accumulators beside the detection
        # histogram and matches inside compute()'s launch when the three are
        # in one graph.
        stepwise = hasattr(evaluator, "_capturing_step") and not evaluator.validate
        branch = torch.cuda.Stream(device) if early and stepwise else None
        if stepwise:
            evaluator._capturing_step = True
        try:
            with _native.pinned_tables() as tables, torch.cuda.graph(self.graph):
                if branch is not None:
                    # Recorded first, so that the graph starts it first: it
                    # leads to the detection histogram, the longest path.
                    main = torch.cuda.current_stream(device)
                    branch.wait_stream(main)
                    with torch.cuda.stream(branch):
                        prepare()
                    evaluator._masks_ready = lambda: main.wait_stream(branch)
                evaluator.reset()
                if branch is None:
                    prepare()
                evaluator.update(*args, **kwargs)
                join = getattr(evaluator, "_masks_ready", None)
                if join is not None:  # the branch joins before compute() in any case
                    evaluator._masks_ready = None
                    join()
                self._result = evaluator.compute()
        finally:
            if stepwise:
                evaluator._capturing_step = False
                evaluator._masks_ready = None
        self._tables = tables
        self._scratch = _scratch_of(evaluator)
        self._records = list(instance._records) if instance is not None else None
        torch.cuda.synchronize(device)

    def replay(self):
        """Run reset, update and compute once more; returns the result."""
        if self._flush is not None:
            # An eager update's deferred matching still owns rows of the
            # detection x slot table that the replay refills.
            self._flush()
        self.graph.replay()
        if self._instance is not None:
            self._instance._records = list(self._records)
        return self._result
