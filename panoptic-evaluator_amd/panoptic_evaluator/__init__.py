"""panoptic_evaluator for AMD Instinct MI355X (gfx950).

Drop-in for the reference package's public API (its __init__.py:1-14): PQ/SQ/RQ,
semantic mIoU, COCO mask AP/AR and generation-prefix rewards, computed by
hand-written HIP kernels behind a C ABI (libpanoptic_hip.so).
"""

from .evaluator import PanopticBatch, PanopticEvaluator, panoptic_quality
from .instance import InstanceBatch, InstanceEvaluator, PackedMasks, pack_masks
from .combined import SegmentationEvaluator
from .prefix import segmentation_prefixes
from .rle import RleBatch, decode_rle, decode_rle_batch

__all__ = [
    "PanopticBatch",
    "PanopticEvaluator",
    "panoptic_quality",
    "InstanceBatch",
    "InstanceEvaluator",
    "SegmentationEvaluator",
    "segmentation_prefixes",
    "PackedMasks",
    "pack_masks",
    "RleBatch",
    "decode_rle",
    "decode_rle_batch",
]
