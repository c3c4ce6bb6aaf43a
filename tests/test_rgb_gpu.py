"""Decoded COCO panoptic PNGs (B x H x W x 3 uint8 on the device) as segment
maps: PanopticBatch.from_coco decodes R + 256 * G + 65536 * B inside the
compaction and histogram kernels (README.md:67-68).  Every result must equal
the same batch given as integer ID maps, checked against the oracle."""

import numpy as np
import pytest
import torch

from helpers import assert_close, assert_exact, compare_panoptic, to_numpy
from oracle import panoptic as opan

pytestmark = pytest.mark.gpu

CATS = (3, 17, 40, 41)


@pytest.fixture(scope="module")
def pe():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import panoptic_evaluator
    return panoptic_evaluator


def id_batch(seed, B, H, W, segments):
    """Voronoi-like ID maps with COCO-style 24-bit IDs (every byte used)."""
    rng = np.random.default_rng(seed)
    maps, infos = [], []
    for _ in range(B):
        ids = rng.choice(np.arange(1, 2**24), segments, replace=False)
        ids[0] = 0x010203  # all three bytes nonzero
        cy, cx = rng.integers(0, H, segments), rng.integers(0, W, segments)
        yy, xx = np.mgrid[0:H, 0:W]
        nearest = np.argmin((yy[..., None] - cy) ** 2 + (xx[..., None] - cx) ** 2, axis=-1)
        image = ids[nearest]
        image[rng.random((H, W)) < 0.1] = 0  # void
        maps.append(image)
        infos.append(segments_of(image, rng))
    return np.stack(maps).astype(np.int64), infos


def segments_of(image, rng, known=()):
    """segments_info for exactly the IDs present in an image (metadata and
    pixels correspond, as validation requires); known IDs keep their entry."""
    known = {e["id"]: e for e in known}
    return [known.get(int(i), {"id": int(i), "category_id": int(rng.choice(CATS)),
                               "iscrowd": int(rng.random() < 0.1)})
            for i in np.unique(image) if i != 0]


def rgb_tensor(ids, misaligned=False):
    rgb = torch.from_numpy(opan.id_to_rgb(ids)).cuda()
    if not misaligned:
        return rgb
    flat = torch.zeros(rgb.numel() + 1, dtype=torch.uint8, device="cuda")
    flat[1:] = rgb.reshape(-1)
    view = flat[1:].view(rgb.shape)
    assert view.data_ptr() % 4 != 0 and view.is_contiguous()
    return view


def expected_state(gt, pred, gi, pi):
    g_ids, g_slots, gc, crowd = opan.from_coco(gt, gi, CATS)
    p_ids, p_slots, pc, _ = opan.from_coco(pred, pi, CATS)
    state = opan.PanopticState(len(CATS), [True, True, False, False])
    state.update(opan.compact(gt, g_ids, g_slots), opan.compact(pred, p_ids, p_slots), gc, pc, crowd,
                 validate=False)
    return state.compute()


@pytest.mark.parametrize("H,W,misaligned", [(37, 53, False), (64, 96, False), (29, 31, True)])
def test_rgb_compact_matches_id_maps(pe, H, W, misaligned):
    gt, gi = id_batch(1, 3, H, W, 12)
    ids, slots, _, _ = opan.from_coco(gt, gi, CATS)
    batch = pe.PanopticBatch.from_coco(rgb_tensor(gt, misaligned), gi, CATS)
    assert batch.segment_map.dtype == torch.int32 and batch.segment_map.shape == (3, H, W)
    assert_exact(batch.segment_map, opan.compact(gt, ids, slots))


@pytest.mark.parametrize("H,W,misaligned", [(37, 53, False), (128, 160, False), (29, 31, True)])
def test_rgb_fused_lookup_evaluator(pe, H, W, misaligned):
    gt, gi = id_batch(2, 4, H, W, 15)
    pred, _ = id_batch(3, 4, H, W, 18)
    # Predictions reuse the GT IDs of the top half so that segments match.
    pred[:, : H // 2] = gt[:, : H // 2]
    rng = np.random.default_rng(4)
    pi = [[dict(e, iscrowd=0) for e in segments_of(p, rng, g)] for p, g in zip(pred, gi)]
    expected = expected_state(gt, pred, gi, pi)
    rgb_gt = pe.PanopticBatch.from_coco(rgb_tensor(gt, misaligned), gi, CATS, compact=False)
    rgb_pred = pe.PanopticBatch.from_coco(rgb_tensor(pred), pi, CATS, compact=False)
    assert rgb_gt.segment_map.dtype == torch.uint8
    id_pred = pe.PanopticBatch.from_coco(torch.from_numpy(pred).int().cuda(), pi, CATS, compact=False)
    for validate in (False, True):
        for target, prediction in ((rgb_gt, rgb_pred), (rgb_gt, rgb_pred.compact()), (rgb_gt.compact(), rgb_pred),
                                   (rgb_gt, id_pred)):
            ev = pe.PanopticEvaluator(len(CATS), isthing=[True, True, False, False], validate=validate)
            ev.update(prediction, target)
            compare_panoptic(to_numpy(ev.compute()), expected)


def test_rgb_panoptic_quality_and_unknown_ids(pe):
    gt, gi = id_batch(5, 2, 40, 44, 10)
    pred = gt.copy()
    pred[:, :, :7] = gt[:, :, 30:37]
    target = pe.PanopticBatch.from_coco(rgb_tensor(gt), gi, CATS, compact=False)
    prediction = pe.PanopticBatch.from_coco(rgb_tensor(pred), gi, CATS, compact=False)
    g_ids, g_slots, gc, crowd = opan.from_coco(gt, gi, CATS)
    want = opan.panoptic_quality(opan.compact(gt, g_ids, g_slots), opan.compact(pred, g_ids, g_slots), gc, gc,
                                 crowd, len(CATS))
    got = pe.panoptic_quality(prediction, target, len(CATS))
    assert_close(got, want)
    # An ID absent from segments_info: compaction and validated updates refuse it.
    bad = gt.copy()
    bad[0, 0, 0] = 0xABCDEF
    with pytest.raises(ValueError):
        pe.PanopticBatch.from_coco(rgb_tensor(bad), gi, CATS)
    unknown = pe.PanopticBatch.from_coco(rgb_tensor(bad), gi, CATS, compact=False)
    ev = pe.PanopticEvaluator(len(CATS), validate=True)
    with pytest.raises(ValueError, match="unknown segment ID"):
        ev.update(prediction, unknown)


def test_rgb_lookup_with_global_pair_bins(pe):
    """~300 segments per side: the G x P pair table exceeds the LDS bins, so the
    histogram counts into global memory while the ID hashes stay in LDS."""
    gt, gi = id_batch(7, 2, 96, 96, 300)
    pred, _ = id_batch(8, 2, 96, 96, 320)
    pred[:, :48] = gt[:, :48]
    rng = np.random.default_rng(9)
    pi = [[dict(e, iscrowd=0) for e in segments_of(p, rng, g)] for p, g in zip(pred, gi)]
    G, P = 1 + max(map(len, gi)), 1 + max(map(len, pi))
    assert G * P * 2 > 128 * 1024  # 16-bit bins beyond the LDS budget
    expected = expected_state(gt, pred, gi, pi)
    target = pe.PanopticBatch.from_coco(rgb_tensor(gt), gi, CATS, compact=False)
    prediction = pe.PanopticBatch.from_coco(rgb_tensor(pred), pi, CATS, compact=False)
    for validate in (False, True):
        ev = pe.PanopticEvaluator(len(CATS), isthing=[True, True, False, False], validate=validate)
        ev.update(prediction, target)
        compare_panoptic(to_numpy(ev.compute()), expected)


def same_home_ids(count, bits, bucket):
    """24-bit IDs whose home bucket in a table of 2**bits buckets is `bucket`
    under the histogram's multiplicative hash (x * 0x9E3779B1 mod 2^32, top
    bits), so that buckets overflow and lookups probe on."""
    out, x = [], 1
    while len(out) < count:
        if ((x * 0x9E3779B1) & 0xFFFFFFFF) >> (32 - bits) == bucket:
            out.append(x)
        x += 1
    return out


def test_lookup_with_overflowing_buckets(pe):
    """20 segments per image (21 slots: 16 buckets of four): twelve IDs share
    one home bucket, so eight of them live in the buckets after it.  RGB, int32
    and int64 maps of the same IDs give the oracle's result; an unknown ID
    homed in the full bucket is refused by validation."""
    H, W, S = 48, 64, 20
    rng = np.random.default_rng(31)
    crowded = same_home_ids(12, 4, 5)
    maps, infos = [], []
    yy, xx = np.mgrid[0:H, 0:W]
    for b in range(2):
        others = rng.choice(np.arange(1 << 16, 1 << 24), S - len(crowded), replace=False)
        ids = np.concatenate([crowded, others])
        rng.shuffle(ids)
        cy, cx = rng.integers(0, H, S), rng.integers(0, W, S)
        nearest = np.argmin((yy[..., None] - cy) ** 2 + (xx[..., None] - cx) ** 2, axis=-1)
        image = ids[nearest]
        image[rng.random((H, W)) < 0.05] = 0
        maps.append(image)
        infos.append(segments_of(image, rng))
    gt = np.stack(maps).astype(np.int64)
    pred = np.roll(gt, 3, axis=2)
    pi = [[dict(e, iscrowd=0) for e in segments_of(p, rng, g)] for p, g in zip(pred, infos)]
    assert 1 + max(map(len, infos)) <= 21 and sum(int(i) in crowded for i in np.unique(gt)) > 8
    expected = expected_state(gt, pred, infos, pi)
    forms = {
        "rgb": (rgb_tensor(gt), rgb_tensor(pred)),
        "int32": (torch.from_numpy(gt).int().cuda(), torch.from_numpy(pred).int().cuda()),
        "int64": (torch.from_numpy(gt).cuda(), torch.from_numpy(pred).cuda()),
    }
    for name, (g, p) in forms.items():
        target = pe.PanopticBatch.from_coco(g, infos, CATS, compact=False)
        prediction = pe.PanopticBatch.from_coco(p, pi, CATS, compact=False)
        for validate in (False, True):
            ev = pe.PanopticEvaluator(len(CATS), isthing=[True, True, False, False], validate=validate)
            ev.update(prediction, target)
            compare_panoptic(to_numpy(ev.compute()), expected)
    unknown_id = same_home_ids(13, 4, 5)[-1]
    bad = gt.copy()
    bad[1, 5, 7] = unknown_id
    for g in (rgb_tensor(bad), torch.from_numpy(bad).int().cuda()):
        unknown = pe.PanopticBatch.from_coco(g, infos, CATS, compact=False)
        ev = pe.PanopticEvaluator(len(CATS), validate=True)
        with pytest.raises(ValueError, match="unknown segment ID"):
            ev.update(pe.PanopticBatch.from_coco(forms["int32"][1], pi, CATS, compact=False), unknown)
