"""COCO RLE masks decoded on the GPU into bit-packed masks.

Every decoded mask must equal the oracle's decode of the same RLE (and the
mask it was encoded from) bit for bit, for compressed strings and count
lists, widths with and without 32-pixel alignment, single rows and columns,
empty and full masks, and batches of images of different sizes.  Evaluators
fed decoded masks must give exactly the bool-mask results and the oracle's.
The RLE strings are written by the oracle's encoder (pycocotools' format)."""

import numpy as np
import pytest
import torch

from helpers import assert_close, assert_exact, to_numpy
from oracle import instance as oinst
from oracle import rle as orle

pytestmark = pytest.mark.gpu

SUMMARY = ("ap", "ap50", "ap75", "ap_small", "ap_medium", "ap_large", "ar1", "ar10", "ar100", "ar_small",
           "ar_medium", "ar_large")


@pytest.fixture(scope="module")
def pe():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import panoptic_evaluator
    return panoptic_evaluator


def cuda(x):
    return torch.as_tensor(np.asarray(x)).cuda()


def random_masks(rng, N, H, W):
    """Masks of mixed character: empty, full, sparse noise, dense noise,
    column stripes, a rectangle, and ellipses."""
    masks = np.zeros((N, H, W), bool)
    ys, xs = np.mgrid[0:H, 0:W]
    for n in range(N):
        kind = n % 7
        if kind == 1:
            masks[n] = True
        elif kind == 2:
            masks[n] = rng.random((H, W)) < 0.03
        elif kind == 3:
            masks[n] = rng.random((H, W)) < 0.8
        elif kind == 4:
            masks[n, :, rng.integers(0, W, max(1, W // 3))] = True
        elif kind == 5:
            y0, x0 = rng.integers(0, H), rng.integers(0, W)
            masks[n, y0:y0 + rng.integers(1, H + 1), x0:x0 + rng.integers(1, W + 1)] = True
        elif kind == 6:
            cy, cx = rng.random() * H, rng.random() * W
            a, b = 1 + rng.random() * H / 2, 1 + rng.random() * W / 2
            masks[n] = ((ys - cy) / a) ** 2 + ((xs - cx) / b) ** 2 <= 1
    return masks


def counts_form(code):
    return {"size": code["size"], "counts": orle.counts_from_string(code["counts"])}


@pytest.mark.parametrize("N,H,W", [(7, 37, 53), (14, 64, 64), (7, 1, 70), (7, 45, 1), (9, 31, 33), (7, 100, 96),
                                   (3, 129, 257)])
@pytest.mark.parametrize("form", ["string", "counts"])
def test_decode_matches_oracle(pe, N, H, W, form):
    rng = np.random.default_rng(N * 1000 + H * 7 + W)
    masks = random_masks(rng, N, H, W)
    codes = [orle.encode(m) for m in masks]
    if form == "counts":
        codes = [counts_form(c) for c in codes]
    packed = pe.decode_rle(codes, "cuda")
    assert packed.shape == (N, H, W)
    expected = np.stack([orle.decode(c) for c in codes])
    assert_exact(expected, masks)
    assert_exact(packed.to_bool(), expected)
    # The same bits as packing the bool masks on the device, and the same
    # nonzero-word extents (an empty mask's range is empty in both).
    bool_packed = pe.pack_masks(cuda(masks))
    assert_exact(packed.bits, bool_packed.bits.cpu())
    assert_exact(normal_extents(packed.extents), normal_extents(bool_packed.extents))


def normal_extents(ext):
    """Extents with every empty range as (0, 0)."""
    ext = ext.cpu().numpy().astype(np.int64)
    empty = ext[:, 1] <= ext[:, 0]
    ext[empty] = 0
    return ext


def test_decode_batch_of_uneven_images(pe):
    rng = np.random.default_rng(5)
    shapes = [(40, 33), (8, 8), (1, 1), (96, 128), (17, 200), (50, 7)]
    images, originals = [], []
    for i, (h, w) in enumerate(shapes):
        masks = random_masks(rng, 3 + i, h, w)
        codes = [orle.encode(m) for m in masks]
        if i % 2:
            codes = [counts_form(c) if k % 2 else c for k, c in enumerate(codes)]
        images.append(codes)
        originals.append(masks)
    images.insert(2, [])
    originals.insert(2, np.zeros((0, 12, 9), bool))
    decoded = pe.decode_rle_batch(images, "cuda", sizes=[m.shape[1:] for m in originals])
    assert len(decoded) == len(images)
    for packed, masks in zip(decoded, originals):
        assert packed.shape == masks.shape
        assert_exact(packed.to_bool(), masks)
        if len(masks):
            assert_exact(normal_extents(packed.extents), normal_extents(pe.pack_masks(cuda(masks)).extents))


def block_aligned_masks(H, W):
    """Runs that start and end on the decoder's 32-row blocks, or one pixel
    either side of them, in every column of a 64-column band, and bands with
    no set pixel between bands with some."""
    masks = np.zeros((6, H, W), bool)
    for x in range(W):
        band = x // 64
        if band % 2:
            continue  # an empty band
        for k, (y0, y1) in enumerate(((32, 64), (31, 65), (33, 63), (0, 32), (H - 32, H), (64, 96))):
            masks[k, max(y0, 0):min(y1, H), x] = (x + k) % 3 != 0
    return masks


def with_empty_runs(counts, rng):
    """The same mask with pairs of zero-length runs inserted (a 0-run and a
    1-run, or the other way round, of length 0 leave every later run's kind
    unchanged), at the start, between runs and at the end."""
    out = list(counts)
    for _ in range(1 + len(out) // 4):
        at = int(rng.integers(0, len(out) + 1))
        out[at:at] = [0, 0]
    return out + [0, 0]


@pytest.mark.parametrize("form", ["string", "counts"])
def test_decode_runs_on_block_edges_and_empty_runs(pe, form):
    """The rasteriser skips the 32-row blocks that runs of 0s cover: runs that
    end on a block edge, one row before or after it, empty bands and
    zero-length runs anywhere in the list must all decode exactly."""
    rng = np.random.default_rng(41)
    for H, W in ((160, 200), (97, 256), (129, 70)):
        masks = np.concatenate([block_aligned_masks(H, W), random_masks(rng, 7, H, W)])
        codes = []
        for m in masks:
            counts = with_empty_runs(orle.counts_from_string(orle.encode(m)["counts"]), rng)
            code = {"size": [H, W], "counts": counts}
            if form == "string":
                code = {"size": [H, W], "counts": orle.counts_to_string(counts)}
            assert_exact(orle.decode(code), m)
            codes.append(code)
        packed = pe.decode_rle(codes, "cuda")
        assert_exact(packed.to_bool(), masks)
        bool_packed = pe.pack_masks(cuda(masks))
        assert_exact(packed.bits, bool_packed.bits.cpu())
        assert_exact(normal_extents(packed.extents), normal_extents(bool_packed.extents))


def test_decode_large_ellipses(pe):
    """c1-sized images: 100 ellipse masks of 800 x 1088, and the same at an
    unaligned width."""
    import workloads
    gen = torch.Generator().manual_seed(3)
    for H, W in ((800, 1088), (333, 501)):
        masks = workloads.ellipse_masks(gen, 100, H, W).numpy()
        codes = [orle.encode(m) for m in masks]
        packed = pe.decode_rle(codes, "cuda")
        assert_exact(packed.bits, pe.pack_masks(cuda(masks)).bits.cpu())


def test_malformed_rle(pe):
    good = orle.encode(np.eye(4, dtype=bool))
    cases = [
        {"size": [4, 4], "counts": good["counts"] + b"1"},    # one pixel too many
        {"size": [4, 4], "counts": good["counts"][:-1]},      # too few
        {"size": [4, 4], "counts": b"0\x7f"},                 # outside the alphabet
        {"size": [4, 4], "counts": b"0h"},                    # the last count carries on
        {"size": [4, 4], "counts": b"hhhhhhhh0"},             # more than seven groups
        {"size": [4, 4], "counts": [3, 14]},                  # list total
    ]
    for bad in cases:
        with pytest.raises(ValueError, match="malformed RLE"):
            pe.decode_rle([good, bad], "cuda")
    # Without validation a malformed mask decodes as all-zero, the rest as usual.
    packed = pe.decode_rle([good, cases[0], good], "cuda", validate=False)
    got = packed.to_bool().cpu().numpy()
    assert_exact(got[0], np.eye(4, dtype=bool))
    assert not got[1].any()
    assert_exact(got[2], np.eye(4, dtype=bool))
    with pytest.raises(ValueError):
        pe.decode_rle([{"size": [4, 4], "counts": [-1, 17]}], "cuda")
    with pytest.raises(ValueError):
        pe.decode_rle([good, orle.encode(np.ones((3, 4), bool))], "cuda")
    with pytest.raises(ValueError):
        pe.decode_rle([], "cuda")


def test_instance_evaluator_on_rle(pe):
    """The reference's adversarial instance case shape (test_instance.py:
    85-160: duplicated masks, exact ties, crowd, area edges) with both sides
    given as RLE, against the bool-mask path and the oracle."""
    rng = np.random.default_rng(802)
    preds, targets, pm_all, gm_all, pcl, gcl, sc, cr, ar = [], [], [], [], [], [], [], [], []
    for _ in range(4):
        h, w = rng.integers(20, 60, 2)
        gm = rng.random((6, h, w)) > 0.55
        pm = np.concatenate([gm[:4].copy(), rng.random((5, h, w)) > 0.65])
        pm[1] &= rng.random((h, w)) > 0.2
        pm[2] = gm[0]
        gc = rng.integers(0, 3, 6)
        pc = np.concatenate([gc[:4], rng.integers(0, 3, 5)])
        pc[2] = gc[0]
        pm_all.append(pm)
        gm_all.append(gm)
        pcl.append(pc)
        gcl.append(gc)
        sc.append(rng.choice([0.1, 0.5, 0.9], 9))
        cr.append(rng.random(6) < 0.3)
        ar.append(rng.choice([100.0, 1024.0, 1025.0, 9216.0, 9217.0], 6))
        preds.append([orle.encode(m) for m in pm])
        targets.append([orle.encode(m) for m in gm])
    pred_bits = pe.decode_rle_batch(preds, "cuda")
    gt_bits = pe.decode_rle_batch(targets, "cuda")
    results = []
    for pred_masks, gt_masks in ((pred_bits, gt_bits), ([cuda(m) for m in pm_all], [cuda(m) for m in gm_all])):
        ev = pe.InstanceEvaluator(3, device="cuda")
        ev.update(pe.InstanceBatch(pred_masks, [cuda(c) for c in pcl], [cuda(s) for s in sc]),
                  pe.InstanceBatch(gt_masks, [cuda(c) for c in gcl], crowd=[cuda(c) for c in cr],
                                   areas=[cuda(a) for a in ar]))
        results.append(to_numpy(ev.compute()))
    state = oinst.InstanceState(3)
    state.update([dict(masks=m, classes=c, scores=s) for m, c, s in zip(pm_all, pcl, sc)],
                 [dict(masks=m, classes=c, crowd=k, areas=a) for m, c, k, a in zip(gm_all, gcl, cr, ar)])
    expected = state.compute()
    for key in ("precision", "recall", *SUMMARY):
        assert_exact(results[0][key], results[1][key], key)
    assert_exact(results[0]["precision"], expected["precision"], "precision")
    assert_exact(results[0]["recall"], expected["recall"], "recall")
    for key in SUMMARY:
        assert_close(results[0][key], expected[key], what=key)


@pytest.mark.parametrize("masks_only", [False, True])
def test_rle_batch_decodes_in_a_step_graph(pe, masks_only):
    """RleBatch keeps its strings and decoded rows on the device: decode()
    may run inside a whole-step graph (CapturedStep's prepare, in line or on
    a branch of its own beside the panoptic branch), and every replay equals
    the eager step on bool masks."""
    import workloads
    from panoptic_evaluator.graphs import CapturedStep
    pair, masks, classes, scores = workloads.config1(device="cuda", seed=12, B=2, H=120, W=176, detections=30)
    isthing = workloads.isthing_vector(133, 80)
    rle = pe.RleBatch([workloads.rle_encode(m) for m in masks], "cuda")
    decoded = rle.decode(validate=True)
    for packed, m in zip(decoded, masks):
        assert_exact(packed.bits, pe.pack_masks(m).bits.cpu())
    target = pe.PanopticBatch(pair.gt, pair.gc, pair.crowd)
    prediction = pe.PanopticBatch(pair.pred, pair.pc)
    eager = pe.SegmentationEvaluator(133, isthing=isthing, validate=False)
    eager.update(prediction, pe.InstanceBatch(masks, classes, scores), target)
    want = to_numpy(eager.compute())
    ev = pe.SegmentationEvaluator(133, isthing=isthing, validate=False)
    step = CapturedStep(ev, prediction, pe.InstanceBatch(rle.masks, classes, scores), target, prepare=rle.decode,
                        prepare_masks_only=masks_only)
    for _ in range(3):
        rle.bits.fill_(-1)  # every replay decodes again from the strings
        got = to_numpy(step.replay())
        assert_exact(got["tp"], want["tp"], "tp")
        assert_exact(got["instance"]["precision"], want["instance"]["precision"], "precision")
        assert_exact(got["instance"]["recall"], want["instance"]["recall"], "recall")
