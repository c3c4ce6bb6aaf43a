"""Launch-shape autotuner for the streaming kernels.

Each tuning knob of libpanoptic_hip.so ("joint_histogram",
"detection_histogram", "detection_histogram_bits", "mask_overlap_mfma",
"packed_grid", "detection_group", "mask_overlap_bits") has a fixed list of variants -- pixel tile per
workgroup, masks in flight per lane, mask groups and pixel range per
contraction workgroup, persistent grid size over packed batches, detections
per workgroup (LDS bins) -- that compute identical results.  ``autotune`` times every
variant on the caller's own workload with HIP events recorded around the
kernel's launches and keeps the fastest; the selection is process-wide.
Choices can be saved as JSON and applied again with ``apply``.
"""

from __future__ import annotations

import json
import pathlib
from typing import Callable, Iterable

import torch

from . import _native

KERNELS = ("joint_histogram", "detection_histogram", "mask_overlap_mfma", "packed_grid", "detection_group",
           "mask_overlap_bits", "detection_histogram_bits")

# Kernels whose time decides each knob ("packed_grid" shapes the persistent
# grids of both histograms over packed batches of uneven images).
MEASURED = {"packed_grid": ("joint_histogram", "detection_histogram"),
            "detection_group": ("detection_histogram",),
            "detection_histogram_bits": ("detection_histogram",)}


def variants(kernel: str) -> list[str]:
    """Descriptions of a kernel's variants, indexed by variant number."""
    lib = _native.library()
    count = lib.pe_tune_count(kernel.encode())
    if count <= 0:
        raise ValueError(f"unknown tunable kernel {kernel!r}")
    return [lib.pe_tune_describe(kernel.encode(), v).decode() for v in range(count)]


def selected(kernel: str) -> int:
    variants(kernel)
    return _native.library().pe_tune_get(kernel.encode())


def select(kernel: str, variant: int) -> None:
    """Use ``variant`` for every later launch of ``kernel`` (-1 = default)."""
    _native.call("pe_tune_set", kernel.encode(), int(variant))


def selection() -> dict[str, int]:
    return {kernel: selected(kernel) for kernel in KERNELS}


def apply(choices: dict[str, int] | str | pathlib.Path) -> None:
    """Apply a selection dict, or one saved by ``save``."""
    if not isinstance(choices, dict):
        choices = json.loads(pathlib.Path(choices).read_text())["selection"]
    for kernel, variant in choices.items():
        select(kernel, variant)


def save(path, report: dict) -> None:
    pathlib.Path(path).write_text(json.dumps(report, indent=1, sort_keys=True))


def _kernel_ms(step: Callable[[], object], knob: str, repeats: int):
    """Milliseconds per step spent in the kernels a knob shapes, and their
    launch count (0: the step launches none of them)."""
    lib = _native.library()
    torch.cuda.synchronize()
    lib.pe_profile_reset()
    lib.pe_profile_enable(1)
    try:
        for _ in range(repeats):
            step()
        torch.cuda.synchronize()
    finally:
        lib.pe_profile_enable(0)
    total, launches = 0.0, 0
    for kernel in MEASURED.get(knob, (knob,)):
        ms, count = _native.profile_read(kernel)
        total, launches = total + ms, launches + count
    return (total / repeats if launches else None), launches


def autotune(step: Callable[[], object], kernels: Iterable[str] = KERNELS, *, repeats: int = 5,
             warmup: int = 1) -> dict:
    """Time each variant of each knob on ``step`` and keep the fastest.
    Knobs whose kernels ``step`` never launches keep their selection.

    Returns {"selection": {knob: variant}, "timings": {knob: [ms per step in
    the knob's kernels, or None, ...]}, "describe": {knob: [...]}}."""
    report = {"selection": {}, "timings": {}, "describe": {}}
    for kernel in kernels:
        names = variants(kernel)
        before = selected(kernel)
        timings = []
        for variant in range(len(names)):
            select(kernel, variant)
            for _ in range(warmup):
                step()
            ms, launches = _kernel_ms(step, kernel, repeats)
            timings.append(ms)
            if not launches:
                break  # the workload does not launch this kernel
        measured = [(ms, v) for v, ms in enumerate(timings) if ms is not None]
        best = min(measured)[1] if measured else before
        select(kernel, best)
        report["selection"][kernel] = best
        report["timings"][kernel] = timings
        report["describe"][kernel] = names
    return report
