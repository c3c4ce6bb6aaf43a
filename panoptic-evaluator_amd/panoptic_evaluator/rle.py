"""COCO run-length encoded masks as an input format of the instance evaluators.

The reference takes bool mask tensors and meets RLE only where its tests hand
masks to pycocotools (tests/test_instance.py:110: ``mask_utils.encode(
np.asfortranarray(pixels, dtype=np.uint8))``).  COCO annotation files and
detector outputs carry masks in that form, so here they can stay compressed
until they reach the GPU: ``decode_rle`` uploads the strings (or count lists)
as they are, in one copy, and two kernels decode them into bit-packed masks
(``PackedMasks``), which ``InstanceBatch`` takes wherever it takes a bool
tensor.

An RLE is a dict with ``size`` = [h, w] and ``counts``: the compressed
string (bytes or str) that ``mask_utils.encode`` writes, or a list of run
lengths (column-major runs of 0s, then 1s, ..., summing to h * w).
"""

from __future__ import annotations

import ctypes
from typing import Sequence

import numpy as np
import torch

from . import _native
from .instance import PackedMasks, mask_words

_i64, _int, _ptr = ctypes.c_int64, ctypes.c_int, ctypes.c_void_p
_native.register("pe_rle_workspace_bytes", _i64, [_i64, _i64, _i64])
_native.register("pe_rle_decode", _int, [_i64, _ptr, _ptr, _i64, _i64, _i64, _ptr, _ptr, _ptr, _ptr, _i64, _ptr])

_ERRORS = ((1, "a character outside the RLE alphabet"), (2, "a count of more than seven 5-bit groups or an "
           "unterminated count"), (4, "a negative run length"), (8, "run lengths that do not sum to h * w"))


def _size(rle) -> tuple[int, int]:
    try:
        h, w = (int(v) for v in rle["size"])
    except (KeyError, TypeError, ValueError):
        raise ValueError("an RLE needs a 'size' of [h, w]") from None
    if h <= 0 or w <= 0 or h * w >= 2 ** 31:
        raise ValueError("RLE masks need positive image dimensions of fewer than 2^31 pixels")
    return h, w


def _payload(rle) -> tuple[int, bytes, int]:
    """(format, bytes, units) of one RLE's counts."""
    counts = rle.get("counts") if isinstance(rle, dict) else None
    if isinstance(counts, str):
        counts = counts.encode("ascii")
    if isinstance(counts, (bytes, bytearray)):
        return 0, bytes(counts), len(counts)
    if counts is None:
        raise ValueError("an RLE needs 'counts'")
    values = np.asarray(counts, dtype=np.int64).reshape(-1)
    if values.size and (values.min() < 0 or values.max() > np.iinfo(np.int32).max):
        raise ValueError("RLE run lengths must be non-negative int32 values")
    return 1, values.astype("<i4").tobytes(), int(values.size)


class RleBatch:
    """The COCO RLE masks of a batch of images (one list per image, all masks
    of an image of one size), uploaded to the device once, as they are, and
    decoded on demand into bit-packed masks.  The decoded rows live in one
    buffer owned by the batch (``masks`` are PackedMasks views of it), so
    every decode writes to the same addresses: a HIP graph may capture
    ``decode()``.  ``sizes`` (one (h, w) per image) is needed only for images
    without masks."""

    def __init__(self, images: Sequence[Sequence[dict]], device="cuda", *,
                 sizes: Sequence[tuple[int, int]] | None = None):
        device = _native.require_gpu(device)
        if sizes is not None and len(sizes) != len(images):
            raise ValueError("sizes must give one (h, w) per image")
        rows, chunks = [], []
        data_bytes = units_total = columns = out_words = 0
        max_width = 0
        shapes = []
        for i, masks in enumerate(images):
            shape = None if sizes is None else (int(sizes[i][0]), int(sizes[i][1]))
            for rle in masks:
                h, w = _size(rle)
                if shape is not None and (h, w) != shape:
                    raise ValueError("all RLE masks of one image must share its size")
                shape = (h, w)
            if shape is None:
                raise ValueError("an image without masks needs its size")
            h, w = shape
            if h <= 0 or w <= 0:
                raise ValueError("RLE masks need positive image dimensions")
            words = mask_words(h * w)
            shapes.append((len(masks), h, w, out_words))
            for rle in masks:
                form, payload, units = _payload(rle)
                data_bytes = (data_bytes + 3) // 4 * 4  # int32 counts stay aligned
                rows.append((data_bytes, units, out_words, columns, h, w, form, units_total))
                chunks.append((data_bytes, payload))
                data_bytes += len(payload)
                units_total += units
                columns += w
                out_words += words
                max_width = max(max_width, w)
        self.device = device
        self.count = len(rows)
        self.units, self.columns, self.max_width = units_total, columns, max_width
        self.input_bytes = data_bytes
        self.bits = torch.zeros(max(out_words, 4), dtype=torch.int32, device=device)
        self.error = torch.zeros(1, dtype=torch.int32, device=device)
        self.table_bytes = 64 * self.count
        # One host buffer, one copy: the table (N x 8 int64), then the data.
        host = np.zeros(self.table_bytes + max(data_bytes, 1), dtype=np.uint8)
        if rows:
            host[:self.table_bytes] = np.frombuffer(np.asarray(rows, dtype="<i8").tobytes(), dtype=np.uint8)
        for offset, payload in chunks:
            start = self.table_bytes + offset
            host[start:start + len(payload)] = np.frombuffer(payload, dtype=np.uint8)
        self.upload = torch.from_numpy(host).to(device)
        size = _native.library().pe_rle_workspace_bytes(self.count, units_total, columns)
        self.workspace = torch.empty(max(size, 256), dtype=torch.uint8, device=device)
        # Each mask's nonzero-word range, written by the decode itself.
        self.extents = torch.zeros((max(self.count, 1), 2), dtype=torch.int32, device=device)
        self.masks, first = [], 0
        for n, h, w, start in shapes:
            self.masks.append(PackedMasks(self.bits[start:start + n * mask_words(h * w)].view(n, mask_words(h * w)),
                                          h, w, self.extents[first:first + n]))
            first += n

    @torch.no_grad()
    def decode(self, *, validate: bool = False) -> list[PackedMasks]:
        """Decode every mask into ``masks`` (pe_rle_decode's two launches,
        which rewrite every bit row whole).  ``validate`` reads back the
        decoder's error word (a host synchronisation) and raises ValueError for
        a malformed RLE; without it a malformed mask decodes as all-zero."""
        if self.count:
            self.error.zero_()
            _native.call("pe_rle_decode", self.count, self.upload.data_ptr(),
                         self.upload.data_ptr() + self.table_bytes, self.units, self.columns, self.max_width,
                         self.bits.data_ptr(), self.error.data_ptr(), self.extents.data_ptr(), self.workspace.data_ptr(),
                         self.workspace.numel(), _native.stream(self.device))
            if validate:
                code = int(self.error.item())
                if code:
                    reasons = ", ".join(text for bit, text in _ERRORS if code & bit)
                    raise ValueError(f"malformed RLE mask: {reasons}")
        return self.masks


def decode_rle_batch(images: Sequence[Sequence[dict]], device="cuda", *,
                     sizes: Sequence[tuple[int, int]] | None = None, validate: bool = True) -> list[PackedMasks]:
    """Decode the RLE masks of several images (one list per image, all masks
    of an image of one size) into one PackedMasks per image, with one upload
    and two kernel launches for the whole batch.  ``sizes`` (one (h, w) per
    image) is needed only for images without masks.

    ``validate`` reads back the decoder's error word and raises ValueError
    for a malformed RLE; without it a malformed mask decodes as all-zero."""
    return RleBatch(images, device, sizes=sizes).decode(validate=validate)


def decode_rle(rles: Sequence[dict], device="cuda", *, size: tuple[int, int] | None = None,
               validate: bool = True) -> PackedMasks:
    """Decode one image's RLE masks into a PackedMasks (N x h x w).  ``size``
    gives (h, w) for an image without masks."""
    return decode_rle_batch([rles], device, sizes=None if size is None else [size], validate=validate)[0]
