"""Host-side logic that needs no GPU: input-format checks for bit-packed masks
and decoded COCO panoptic PNG maps, and the metadata from_coco builds for
them (the pixel work itself is covered by the GPU tests)."""

import numpy as np
import pytest
import torch

import panoptic_evaluator as pe
from oracle import panoptic as opan
from panoptic_evaluator.evaluator import id_map_ok, is_rgb, map_bytes
from panoptic_evaluator.instance import mask_words


def test_mask_words_rows_are_whole_16_byte_units():
    for pixels, words in ((1, 4), (128, 4), (129, 8), (800 * 1088, 27200), (0, 0)):
        assert mask_words(pixels) == words
        assert mask_words(pixels) * 32 >= pixels and mask_words(pixels) % 4 == 0


def test_packed_masks_validate_their_layout():
    H, W = 5, 7
    bits = torch.zeros((3, mask_words(H * W)), dtype=torch.int32)
    packed = pe.PackedMasks(bits, H, W)
    assert packed.shape == (3, H, W) and packed.ndim == 3 and packed.contiguous() is packed
    assert packed.data_ptr() == bits.data_ptr()
    for bad in (bits[:, :2].clone(), bits.to(torch.int64), bits.t().contiguous().t()):
        with pytest.raises(ValueError):
            pe.PackedMasks(bad, H, W)
    with pytest.raises(ValueError):
        pe.PackedMasks(bits, 0, W)
    offset = torch.zeros(3 * mask_words(H * W) + 1, dtype=torch.int32)[1:].view(3, mask_words(H * W))
    with pytest.raises(ValueError):
        pe.PackedMasks(offset, H, W)  # rows must be 16-byte aligned


def test_packed_masks_unpack_follows_the_bit_order():
    H, W = 3, 50
    bits = torch.zeros((1, mask_words(H * W)), dtype=torch.int32)
    bits[0, 0] = 1            # pixel 0
    bits[0, 4] = 1 << 21      # pixel 4 * 32 + 21 = 149 = (2, 49)
    masks = pe.PackedMasks(bits, H, W).to_bool()
    assert masks.shape == (1, H, W) and masks.sum() == 2
    assert masks[0, 0, 0] and masks[0, 2, 49]


def test_rgb_maps_are_recognised():
    rgb = torch.zeros((2, 4, 5, 3), dtype=torch.uint8)
    assert is_rgb(rgb) and id_map_ok(rgb) and map_bytes(rgb) == 3
    ids = torch.zeros((2, 4, 5), dtype=torch.int64)
    assert not is_rgb(ids) and id_map_ok(ids) and map_bytes(ids) == 8
    assert not id_map_ok(torch.zeros((2, 4, 5, 4), dtype=torch.uint8))
    assert not id_map_ok(torch.zeros((2, 4, 5), dtype=torch.float32))


def test_from_coco_keeps_rgb_maps_and_builds_the_id_tables():
    ids = np.array([[[0, 0x010203, 0x010203], [77, 77, 0xABCDEF]]], np.int64)
    info = [[{"id": 0xABCDEF, "category_id": 42, "iscrowd": 1}, {"id": 77, "category_id": 7},
             {"id": 0x010203, "category_id": 7}]]
    rgb = torch.from_numpy(opan.id_to_rgb(ids))
    batch = pe.PanopticBatch.from_coco(rgb, info, (7, 42), compact=False)
    assert batch.segment_map is rgb
    want_ids, want_slots, want_classes, want_crowd = opan.from_coco(ids, info, (7, 42))
    assert np.array_equal(batch.lookup_ids.numpy(), want_ids)
    assert np.array_equal(batch.lookup_slots.numpy(), want_slots)
    assert np.array_equal(batch.classes.numpy(), want_classes)
    assert np.array_equal(batch.crowd.numpy(), want_crowd)
    # an RGB map needs its ID tables: as compact slots it is refused
    from panoptic_evaluator.evaluator import _prepare
    maps, classes, crowd = _prepare(batch, 2, False)
    assert maps.dtype == torch.uint8 and maps.shape == (1, 2, 3, 3)
    with pytest.raises(ValueError, match="compact maps must be int32"):
        _prepare(pe.PanopticBatch(rgb, batch.classes), 2, False)


def test_pack_pads_metadata_rows():
    """PanopticBatch.pack (CPU tensors): maps one after another, classes and
    crowd padded to the largest S with -1 / False, sizes kept."""
    maps = [torch.arange(6, dtype=torch.int32).reshape(2, 3), torch.zeros((1, 4), dtype=torch.int32),
            torch.ones((3, 1), dtype=torch.int32)]
    classes = [torch.tensor([-1, 4], dtype=torch.int32), torch.tensor([-1, 1, 2, 3], dtype=torch.int32),
               torch.tensor([-1], dtype=torch.int32)]
    crowd = [torch.tensor([False, True]), torch.tensor([False, False, True, False]), torch.tensor([False])]
    batch = pe.PanopticBatch.pack(maps, classes, crowd)
    assert batch.image_sizes == ((2, 3), (1, 4), (3, 1))
    assert batch.segment_map.tolist() == [0, 1, 2, 3, 4, 5, 0, 0, 0, 0, 1, 1, 1]
    assert batch.classes.tolist() == [[-1, 4, -1, -1], [-1, 1, 2, 3], [-1, -1, -1, -1]]
    assert batch.crowd.tolist() == [[False, True, False, False], [False, False, True, False],
                                    [False, False, False, False]]
    with pytest.raises(ValueError, match="crowd"):
        pe.PanopticBatch.pack(maps, classes, [torch.tensor([1, 0], dtype=torch.int32)] * 3)


def test_evaluator_signatures_follow_the_reference():
    """Default devices of the reference (evaluator.py:254, instance.py:57,
    combined.py:30); the CPU default of InstanceEvaluator is refused with the
    build's GPU-only error."""
    import inspect
    assert inspect.signature(pe.PanopticEvaluator).parameters["device"].default == "cuda"
    assert inspect.signature(pe.SegmentationEvaluator).parameters["device"].default == "cuda"
    assert inspect.signature(pe.InstanceEvaluator).parameters["device"].default == "cpu"
    with pytest.raises(ValueError, match="AMD GPUs"):
        pe.InstanceEvaluator(3)


def test_workload_rle_encoder_matches_oracle():
    """The synthetic-input RLE encoder (workloads.rle_encode) writes exactly
    the oracle's pycocotools strings."""
    import numpy as np
    import torch
    import workloads
    from oracle import rle as orle
    rng = np.random.default_rng(3)
    for n, h, w in ((6, 7, 9), (3, 1, 5), (4, 33, 17), (2, 16, 64)):
        m = torch.from_numpy(rng.random((n, h, w)) < 0.4)
        m[0] = True
        m[1] = False
        for got, mask in zip(workloads.rle_encode(m), m.numpy()):
            assert got == orle.encode(mask)
