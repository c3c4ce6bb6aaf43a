"""Oracle for COCO run-length encoded (RLE) instance masks.

Test infrastructure only (see oracle/__init__.py).

The reference feeds its masks to pycocotools as RLE in its parity tests
(tests/test_instance.py:110: ``mask_utils.encode(np.asfortranarray(pixels,
dtype=np.uint8))``; examples/benchmark_instance.py:42 likewise) and never
decodes RLE itself.  pycocotools is not part of the reference checkout and is
not installed here, so this module restates the published format of
pycocotools 2.0 (cocodataset/cocoapi, common/maskApi.c: rleEncode,
rleDecode, rleToString, rleFrString), which is what ``mask_utils.encode``
writes and what COCO annotation files hold:

* A mask of size [h, w] is scanned in column-major (Fortran) order, pixel
  (y, x) at position x * h + y.  ``counts`` are the lengths of alternating
  runs of 0s and 1s, starting with a (possibly empty) run of 0s; they sum to
  h * w.
* The compressed string form stores count i as the difference from count
  i - 2 when i > 2 (counts 0, 1 and 2 are stored as they are), each value
  as little-endian 5-bit groups of its two's complement.  A character is
  48 + group, plus 32 when more groups follow.  The last group's bit 4 is
  the sign: a decoder sign-extends from it, and an encoder continues until
  the remaining value is 0 (bit 4 clear) or -1 (bit 4 set).

Parity pins (tests/test_oracle.py): the format's defining identities on
tiny inputs (hand-derived strings for multi-group counts, delta coding and
negative deltas), decode(encode(m)) == m on random masks of many shapes,
counts summing to h * w, and the area as the sum of the odd counts
(pycocotools' ``area``).
"""

from __future__ import annotations

import numpy as np


def runs_of(mask):
    """Run lengths of a bool h x w mask in column-major order (rleEncode)."""
    flat = np.asarray(mask, dtype=bool).reshape(-1, order="F")
    # Run boundaries: every position whose value differs from the previous
    # one, with a virtual 0 before the first pixel (so a mask starting with 1
    # opens with an empty run of 0s).
    changes = np.flatnonzero(np.diff(np.concatenate([[False], flat]).astype(np.int8))) if flat.size else []
    edges = np.concatenate([[0], changes, [flat.size]]).astype(np.int64)
    return [int(v) for v in np.diff(edges)]


def counts_to_string(counts):
    """The compressed string form of a counts list (rleToString)."""
    out = []
    for i, count in enumerate(counts):
        x = int(count)
        if i > 2:
            x -= int(counts[i - 2])
        more = True
        while more:
            c = x & 0x1F
            x >>= 5  # arithmetic shift, as for C's long
            more = (x != -1) if (c & 0x10) else (x != 0)
            if more:
                c |= 0x20
            out.append(chr(c + 48))
    return "".join(out).encode("ascii")


def counts_from_string(text):
    """The counts list of a compressed string (rleFrString)."""
    data = text.encode("ascii") if isinstance(text, str) else bytes(text)
    counts = []
    p = 0
    while p < len(data):
        x, k, more = 0, 0, True
        while more:
            c = data[p] - 48
            x |= (c & 0x1F) << (5 * k)
            more = bool(c & 0x20)
            p += 1
            k += 1
            if not more and (c & 0x10):
                x |= -1 << (5 * k)
        if len(counts) > 2:
            x += counts[-2]
        counts.append(x)
    return counts


def encode(mask):
    """COCO RLE dict of a bool h x w mask (mask_utils.encode on a Fortran
    uint8 array gives the same ``size`` and ``counts``)."""
    mask = np.asarray(mask, dtype=bool)
    h, w = mask.shape
    return {"size": [int(h), int(w)], "counts": counts_to_string(runs_of(mask))}


def decode(rle):
    """The bool h x w mask of an RLE dict with compressed or list counts
    (rleDecode)."""
    h, w = (int(v) for v in rle["size"])
    counts = rle["counts"]
    if isinstance(counts, (bytes, str)):
        counts = counts_from_string(counts)
    counts = [int(c) for c in counts]
    if any(c < 0 for c in counts) or sum(counts) != h * w:
        raise ValueError("RLE counts must be non-negative and sum to h * w")
    flat = np.zeros(h * w, dtype=bool)
    position, value = 0, False
    for count in counts:
        if value:
            flat[position:position + count] = True
        position += count
        value = not value
    return flat.reshape((h, w), order="F")


def area(rle):
    """Foreground pixels: the sum of the odd-indexed counts (rleArea)."""
    counts = rle["counts"]
    if isinstance(counts, (bytes, str)):
        counts = counts_from_string(counts)
    return int(sum(counts[1::2]))
