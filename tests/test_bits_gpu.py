"""Bit-packed instance masks (PackedMasks: one bit per pixel) on the GPU.

The packing kernel is checked against numpy.packbits (little bit order); the
popcount contraction, the bit form of the detection x slot histogram and
every evaluator taking packed masks must give exactly what the bool-mask path
and the CPU oracle give (integer outputs bit-exact, AP/AR arrays exact)."""

import numpy as np
import pytest
import torch

from helpers import assert_close, assert_exact, compare_panoptic, to_numpy
from oracle import combined as ocomb
from oracle import instance as oinst

pytestmark = pytest.mark.gpu

SUMMARY = ("ap", "ap50", "ap75", "ap_small", "ap_medium", "ap_large", "ar1", "ar10", "ar100", "ar_small",
           "ar_medium", "ar_large")


@pytest.fixture(scope="module")
def pe():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import panoptic_evaluator
    return panoptic_evaluator


@pytest.fixture
def tuning(pe):
    from panoptic_evaluator import tuning
    yield tuning
    for kernel in tuning.KERNELS:
        tuning.select(kernel, -1)


def cuda(x):
    return torch.as_tensor(np.asarray(x)).cuda()


def compare_instance(actual, expected):
    assert_exact(actual["precision"], expected["precision"], "precision")
    assert_exact(actual["recall"], expected["recall"], "recall")
    for key in SUMMARY:
        assert_close(actual[key], expected[key], what=key)


def packbits_words(masks):
    """numpy.packbits reference: little-endian bit order, rows zero-padded to a
    multiple of 128 pixels, read as uint32 words."""
    N, H, W = masks.shape
    flat = masks.reshape(N, H * W)
    words = (flat.shape[1] + 127) // 128 * 4
    padded = np.zeros((N, words * 32), bool)
    padded[:, :flat.shape[1]] = flat
    return np.packbits(padded, axis=1, bitorder="little").view("<u4").reshape(N, words)


@pytest.mark.parametrize("N,H,W,offset", [(3, 37, 53, 0), (2, 64, 64, 0), (4, 31, 33, 1), (0, 5, 5, 0),
                                          (1, 800, 1088, 0)])
def test_pack_matches_packbits(pe, N, H, W, offset):
    rng = np.random.default_rng(N * H + W)
    masks = rng.random((N, H, W)) < 0.37
    dev = cuda(masks)
    if offset:
        flat = torch.zeros(dev.numel() + offset, dtype=torch.bool, device="cuda")
        flat[offset:] = dev.reshape(-1)
        dev = flat[offset:].view(N, H, W)
        assert dev.data_ptr() % 16 != 0
    packed = pe.pack_masks(dev)
    assert packed.shape == (N, H, W) and packed.bits.dtype == torch.int32
    assert_exact(packed.bits.cpu().numpy().view(np.uint32), packbits_words(masks))
    assert_exact(packed.to_bool(), masks)


@pytest.mark.parametrize("D,G,H,W", [(17, 140, 24, 37), (33, 20, 160, 300), (150, 70, 40, 52), (0, 9, 30, 31),
                                     (9, 0, 8, 8), (1, 1, 4, 4)])
def test_overlap_bits_variants(pe, tuning, D, G, H, W):
    from panoptic_evaluator import _native
    rng = np.random.default_rng(D + 3 * G)
    pm = rng.random((D, H, W)) < 0.3
    gm = rng.random((G, H, W)) < 0.4
    pflat = pm.reshape(D, H * W).astype(np.int64)
    gflat = gm.reshape(G, H * W).astype(np.int64)
    pbits, gbits = pe.pack_masks(cuda(pm)), pe.pack_masks(cuda(gm))
    table = lambda xs: _native.device_table(xs, "cuda")  # noqa: E731
    for variant in range(len(tuning.variants("mask_overlap_bits"))):
        tuning.select("mask_overlap_bits", variant)
        inter = torch.empty(D * G, dtype=torch.int32, device="cuda")
        pa = torch.empty(D, dtype=torch.int32, device="cuda")
        ga = torch.empty(G, dtype=torch.int32, device="cuda")
        _native.call("pe_mask_overlap_bits_batch", 1, table([pbits.data_ptr()]).data_ptr(),
                     table([gbits.data_ptr()]).data_ptr(), table([0, D]).data_ptr(), table([0, G]).data_ptr(),
                     table([0, D * G]).data_ptr(), None, H * W, D, G, H * W, D, G, D * G, inter.data_ptr(),
                     pa.data_ptr(), ga.data_ptr(), None, None, _native.stream())
        assert_exact(inter.view(D, G).long(), pflat @ gflat.T, f"variant {variant}")
        assert_exact(pa.long(), pflat.sum(1))
        assert_exact(ga.long(), gflat.sum(1))


def ellipse_masks(rng, n, H, W):
    yy, xx = np.mgrid[:H, :W]
    cy, cx = rng.uniform(0, H, n), rng.uniform(0, W, n)
    ry, rx = rng.uniform(4, H / 3, n), rng.uniform(4, W / 3, n)
    return ((yy[None] - cy[:, None, None]) / ry[:, None, None]) ** 2 + \
        ((xx[None] - cx[:, None, None]) / rx[:, None, None]) ** 2 <= 1


def test_overlap_bits_variants_full_size_uneven_batch(pe, tuning):
    """Every launch variant on two uneven full-size images in one batch (the
    FP4 forms size their pixel ranges from the batch; ranges end inside a
    block, groups are partly filled) against an exact float32 product, with
    the masks' exact extents, and the FP4 forms also without extents and with
    extents covering whole rows."""
    from panoptic_evaluator import _native
    rng = np.random.default_rng(11)
    shapes = [(100, 48, 800, 1088), (37, 70, 333, 501)]
    pms = [ellipse_masks(rng, d, h, w) for d, g, h, w in shapes]
    gms = [ellipse_masks(rng, g, h, w) for d, g, h, w in shapes]
    expected = [(p.reshape(len(p), -1).astype(np.float32) @ g.reshape(len(g), -1).astype(np.float32).T)
                .astype(np.int64) for p, g in zip(pms, gms)]
    pbits = [pe.pack_masks(cuda(p)) for p in pms]
    gbits = [pe.pack_masks(cuda(g)) for g in gms]
    dets = [d for d, _, _, _ in shapes]
    gts = [g for _, g, _, _ in shapes]
    det_off, gt_off = np.cumsum([0] + dets), np.cumsum([0] + gts)
    inter_off = np.cumsum([0] + [d * g for d, g in zip(dets, gts)])
    pixels = [h * w for _, _, h, w in shapes]
    table = lambda xs: _native.device_table(list(map(int, xs)), "cuda")  # noqa: E731
    whole = lambda b: torch.tensor([[0, b.bits.shape[1]]] * b.bits.shape[0], dtype=torch.int32,  # noqa: E731
                                   device="cuda")
    extents = {"none": (None, None),
               "exact": ([b.extents for b in pbits], [b.extents for b in gbits]),
               "whole rows": ([whole(b) for b in pbits], [whole(b) for b in gbits])}
    cases = [(v, "exact") for v in range(len(tuning.variants("mask_overlap_bits")))]
    cases += [(v, e) for v in (0, 11, 18) for e in ("none", "whole rows")]
    for variant, form in cases:
        tuning.select("mask_overlap_bits", variant)
        pm_ext, gm_ext = (None if e is None else table([x.data_ptr() for x in e]).data_ptr()
                          for e in extents[form])
        inter = torch.empty(int(inter_off[-1]), dtype=torch.int32, device="cuda")
        pa = torch.empty(int(det_off[-1]), dtype=torch.int32, device="cuda")
        ga = torch.empty(int(gt_off[-1]), dtype=torch.int32, device="cuda")
        _native.call("pe_mask_overlap_bits_batch", 2, table([b.bits.data_ptr() for b in pbits]).data_ptr(),
                     table([b.bits.data_ptr() for b in gbits]).data_ptr(), table(det_off).data_ptr(),
                     table(gt_off).data_ptr(), table(inter_off).data_ptr(), table(pixels).data_ptr(), 0,
                     max(dets), max(gts), max(pixels), int(det_off[-1]), int(gt_off[-1]), int(inter_off[-1]),
                     inter.data_ptr(), pa.data_ptr(), ga.data_ptr(), pm_ext, gm_ext, _native.stream())
        got = inter.cpu().numpy().astype(np.int64)
        for b in range(2):
            block = got[inter_off[b]:inter_off[b + 1]].reshape(dets[b], gts[b])
            assert_exact(block, expected[b], f"variant {variant}, extents {form}, image {b}")
        assert_exact(pa.cpu().numpy(), np.concatenate([p.reshape(len(p), -1).sum(1) for p in pms]))
        assert_exact(ga.cpu().numpy(), np.concatenate([g.reshape(len(g), -1).sum(1) for g in gms]))


def uneven_instance_inputs(seed):
    rng = np.random.default_rng(seed)
    sizes = [(24, 37), (40, 64), (17, 29), (33, 50)]
    dets, gts = [9, 0, 21, 6], [6, 4, 0, 11]
    pm, gm, pcl, gcl, sc, cr = [], [], [], [], [], []
    for (h, w), d, g in zip(sizes, dets, gts):
        gmask = rng.random((g, h, w)) > 0.6
        pmask = rng.random((d, h, w)) > 0.6
        if d and g:
            pmask[0] = gmask[0]
        gm.append(gmask)
        pm.append(pmask)
        gcl.append(rng.integers(0, 3, g))
        pcl.append(rng.integers(0, 3, d))
        sc.append(rng.choice([0.2, 0.6, 0.9], d))
        cr.append(rng.random(g) < 0.2)
    return pm, gm, pcl, gcl, sc, cr


@pytest.mark.parametrize("packed_side", ["both", "prediction", "target"])
def test_instance_evaluator_packed_masks(pe, packed_side):
    pm, gm, pcl, gcl, sc, cr = uneven_instance_inputs(5)
    pred_masks = [cuda(m) for m in pm]
    gt_masks = [cuda(m) for m in gm]
    if packed_side in ("both", "prediction"):
        pred_masks = [pe.pack_masks(m) for m in pred_masks]
    if packed_side in ("both", "target"):
        gt_masks = [pe.pack_masks(m) for m in gt_masks]
    ev = pe.InstanceEvaluator(3, device="cuda")
    ev.update(pe.InstanceBatch(pred_masks, [cuda(c) for c in pcl], [cuda(s) for s in sc]),
              pe.InstanceBatch(gt_masks, [cuda(c) for c in gcl], crowd=[cuda(c) for c in cr]))
    state = oinst.InstanceState(3)
    state.update([dict(masks=m, classes=c, scores=s) for m, c, s in zip(pm, pcl, sc)],
                 [dict(masks=m, classes=c, crowd=k) for m, c, k in zip(gm, gcl, cr)])
    compare_instance(to_numpy(ev.compute()), state.compute())


def test_combined_packed_detections_all_variants(pe, tuning):
    import workloads
    pair, masks, classes, scores = workloads.config1(device="cpu", seed=5, B=2, H=200, W=333, detections=37)
    isthing = workloads.isthing_vector(133, 80)
    state = ocomb.SegmentationState(133, isthing)
    state.update(pair.pred.numpy(), pair.pc.numpy(),
                 [dict(masks=m.numpy(), classes=c.numpy(), scores=s.numpy()) for m, c, s in zip(masks, classes, scores)],
                 pair.gt.numpy(), pair.gc.numpy(), pair.crowd.numpy())
    expected = state.compute()
    packed = [pe.pack_masks(m.cuda()) for m in masks]
    # With the exact extents pack_masks computes, without extents, and with
    # extents covering whole rows: the same counts.
    forms = {
        "extents": packed,
        "none": [pe.PackedMasks(p.bits, p.height, p.width) for p in packed],
        "whole rows": [pe.PackedMasks(p.bits, p.height, p.width,
                                      torch.tensor([[0, p.bits.shape[1]]] * p.bits.shape[0], dtype=torch.int32,
                                                   device="cuda")) for p in packed],
    }
    for name, form in forms.items():
        dets = pe.InstanceBatch(form, [c.cuda() for c in classes], [s.cuda() for s in scores])
        for variant in range(len(tuning.variants("detection_histogram_bits"))):
            tuning.select("detection_histogram_bits", variant)
            joint = pe.SegmentationEvaluator(133, isthing=isthing)
            joint.update(pe.PanopticBatch(pair.pred.cuda(), pair.pc.cuda()), dets,
                         pe.PanopticBatch(pair.gt.cuda(), pair.gc.cuda(), pair.crowd.cuda()))
            result = to_numpy(joint.compute())
            compare_panoptic(result, expected)
            compare_instance(result["instance"], expected["instance"])


@pytest.mark.parametrize("N,H,W", [(6, 37, 53), (5, 200, 333), (3, 1, 4096)])
def test_mask_extents(pe, N, H, W):
    """pack_masks' extents are the first nonzero word and one past the last
    (0, 0 for an empty mask), against numpy on ellipse, empty, full and
    one-pixel masks."""
    rng = np.random.default_rng(N + H)
    masks = ellipse_masks(rng, N, H, W) if H > 1 else rng.random((N, H, W)) < 0.01
    masks[0] = False
    masks[1] = True
    masks[2] = False
    masks[2, H - 1, W - 1] = True
    packed = pe.pack_masks(cuda(masks))
    words = packbits_words(masks)
    expected = np.zeros((N, 2), np.int64)
    for n in range(N):
        nz = np.flatnonzero(words[n])
        if nz.size:
            expected[n] = nz[0], nz[-1] + 1
    assert_exact(packed.extents.cpu().numpy(), expected)


def test_combined_packed_uneven_batch(pe, tuning):
    """Packed detection masks over a packed batch of uneven images, flat and
    persistent grids."""
    import workloads
    batch = workloads.config2(seed=9, B=5, sizes=(2, 6), segments=(3, 25), detections=(0, 40))
    isthing = workloads.isthing_vector(133, 80)
    state = ocomb.SegmentationState(133, isthing)
    for i in range(len(batch.sizes)):
        state.update(batch.pred[i][None].numpy(), batch.pc[i][None].numpy(),
                     [dict(masks=batch.masks[i].numpy(), classes=batch.classes[i].numpy(),
                           scores=batch.scores[i].numpy())],
                     batch.gt[i][None].numpy(), batch.gc[i][None].numpy(), batch.crowd[i][None].numpy())
    expected = state.compute()
    dev = lambda xs: [x.cuda() for x in xs]  # noqa: E731
    target = pe.PanopticBatch.pack(dev(batch.gt), dev(batch.gc), dev(batch.crowd))
    prediction = pe.PanopticBatch.pack(dev(batch.pred), dev(batch.pc))
    dets = pe.InstanceBatch([pe.pack_masks(m.cuda()) for m in batch.masks], dev(batch.classes), dev(batch.scores))
    for grid in range(len(tuning.variants("packed_grid"))):
        for variant in (0, 1, len(tuning.variants("detection_histogram_bits")) - 1):
            tuning.select("packed_grid", grid)
            tuning.select("detection_histogram_bits", variant)
            joint = pe.SegmentationEvaluator(133, isthing=isthing)
            joint.update(prediction, dets, target)
            result = to_numpy(joint.compute())
            compare_panoptic(result, expected)
            compare_instance(result["instance"], expected["instance"])


def test_packed_masks_rejected_where_unsupported(pe):
    masks = pe.pack_masks(torch.ones((2, 8, 8), dtype=torch.bool, device="cuda"))
    with pytest.raises(ValueError):
        pe.PackedMasks(masks.bits[:, :3], 8, 8)  # wrong word count
    # Prefix rewards paint bool masks onto a canvas; they do not take bits.
    target = pe.PanopticBatch(torch.zeros((1, 8, 8), dtype=torch.int32, device="cuda"),
                              torch.tensor([[-1]], dtype=torch.int32, device="cuda"))
    prediction = pe.InstanceBatch([masks], [torch.zeros(2, dtype=torch.int64, device="cuda")])
    with pytest.raises(ValueError, match="bool"):
        pe.segmentation_prefixes(prediction, target, 3)
