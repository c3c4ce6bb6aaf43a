"""pe_update_noalloc (workspace in a static device array) against pe_update
(caller workspace): identical results through the C ABI on two calls in a
row, the second on a smaller batch, in dataset and reward mode, for B x H x W
maps and a packed batch of uneven images; oversized batches are rejected on
the host before anything runs."""

import ctypes

import numpy as np
import pytest
import torch

from oracle import panoptic as opan

pytestmark = pytest.mark.gpu

C = 9
ISTHING = [c % 3 != 2 for c in range(C)]
_i64, _int, _ptr = ctypes.c_int64, ctypes.c_int, ctypes.c_void_p


@pytest.fixture(scope="module")
def native():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from panoptic_evaluator import _native
    _native.register("pe_update_noalloc_max_workspace_bytes", _i64, [])
    _native.register("pe_update_noalloc", _int, [_ptr, _int, _ptr, _int, _i64, _i64, _i64, _ptr, _ptr, _ptr, _ptr,
                                                 _i64, _i64, _i64, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr,
                                                 _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _int, _int, _int, _ptr])
    return _native


def batch(seed, B, H, W, G, P):
    rng = np.random.default_rng(seed)
    gt = rng.integers(0, G, (B, H, W)).astype(np.int32)
    pred = np.where(rng.random((B, H, W)) < 0.7, gt, rng.integers(0, P, (B, H, W))).astype(np.int32)
    gc = rng.integers(0, C, (B, G)).astype(np.int32)
    pc = np.concatenate([gc, rng.integers(0, C, (B, P - G))], 1).astype(np.int32)
    gc[:, 0] = pc[:, 0] = -1
    for maps, classes in ((gt, gc), (pred, pc)):
        for b in range(B):  # slots without pixels are inactive metadata
            present = np.bincount(maps[b].ravel(), minlength=classes.shape[1]) > 0
            classes[b, 1:] = np.where(present[1:], classes[b, 1:], -1)
    crowd = (rng.random((B, G)) < 0.15) & (gc >= 0)
    return gt, pred, gc, pc, crowd


class Call:
    """Device buffers of one pe_update call; ``run(name)`` issues it through
    the C ABI with fresh accumulators and returns them on the host."""

    def __init__(self, native, arrays, *, reward, exact, packed_sizes=None):
        self.native = native
        gt, pred, gc, pc, crowd = arrays
        self.B, self.G, self.P = gc.shape[0], gc.shape[1], pc.shape[1]
        if packed_sizes is not None:
            flat = lambda maps: np.concatenate([m[:h, :w].ravel() for m, (h, w) in zip(maps, packed_sizes)])  # noqa: E731
            offsets = np.concatenate([[0], np.cumsum([h * w for h, w in packed_sizes])]).astype(np.int64)
            self.gt, self.pred = torch.from_numpy(flat(gt)).cuda(), torch.from_numpy(flat(pred)).cuda()
            self.pix_off = torch.from_numpy(offsets).cuda()
            self.H, self.W = 1, int(offsets[-1])
        else:
            self.gt, self.pred = torch.from_numpy(gt).cuda(), torch.from_numpy(pred).cuda()
            self.pix_off = None
            self.H, self.W = gt.shape[1], gt.shape[2]
        self.gc, self.pc = torch.from_numpy(gc).cuda(), torch.from_numpy(pc).cuda()
        self.crowd = torch.from_numpy(crowd).cuda()
        self.reward, self.exact = reward, exact
        self.validate = int(packed_sizes is None)  # cropped packed images may leave slots without pixels

    def run(self, name):
        B, G, P = self.B, self.G, self.P
        dev = dict(device="cuda")
        shape = (B, C) if self.reward else (C,)
        tp = torch.zeros(shape, dtype=torch.int64, **dev)
        fp, fn = torch.zeros_like(tp), torch.zeros_like(tp)
        sums = torch.zeros(shape, dtype=torch.float64, **dev)
        confusion = None if self.reward else torch.zeros((C, C + 1), dtype=torch.int64, **dev)
        exact = torch.zeros(2 * C, dtype=torch.int64, **dev) if self.exact and not self.reward else None
        pairs = torch.full((B, G, P), 3, dtype=torch.int64, **dev)
        areas = torch.empty((B, G + P), dtype=torch.int64, **dev)
        matched = torch.empty((B, G), dtype=torch.int32, **dev)
        error = torch.empty(1, dtype=torch.int32, **dev)
        p = self.native.ptr
        head = (self.gt.data_ptr(), 4, self.pred.data_ptr(), 4, B, self.H, self.W, p(self.pix_off),
                self.gc.data_ptr(), self.pc.data_ptr(), self.crowd.data_ptr(), G, P, C, pairs.data_ptr(),
                areas.data_ptr(), matched.data_ptr(), tp.data_ptr(), fp.data_ptr(), fn.data_ptr(), sums.data_ptr(),
                p(confusion), None, None, None, None, error.data_ptr(), p(exact))
        tail = (self.validate, int(self.reward), 0, self.native.stream("cuda"))
        if name == "pe_update":
            size = self.native.library().pe_update_workspace_bytes(B, G, P, C, self.W if self.pix_off is not None
                                                                   else 0)
            workspace = torch.empty(max(size, 256), dtype=torch.uint8, **dev)
            code = self.native.call("pe_update", *head, workspace.data_ptr(), workspace.numel(), *tail)
        else:
            code = self.native.call("pe_update_noalloc", *head, *tail)
        assert code == 0
        out = {"tp": tp, "fp": fp, "fn": fn, "iou_sum": sums, "areas": areas, "pairs": pairs}
        if confusion is not None:
            out["confusion"] = confusion
        return {k: v.cpu().numpy() for k, v in out.items()}


def assert_same(a, b):
    assert a.keys() == b.keys()
    for key in a:
        np.testing.assert_array_equal(a[key], b[key], err_msg=key)


@pytest.mark.parametrize("reward,exact", [(False, False), (False, True), (True, False)])
def test_noalloc_equals_update_twice_second_smaller(native, reward, exact):
    large = batch(1, 4, 61, 83, 17, 23)
    small = batch(2, 2, 29, 35, 6, 9)
    for arrays in (large, small):
        call = Call(native, arrays, reward=reward, exact=exact)
        want, got = call.run("pe_update"), call.run("pe_update_noalloc")
        assert_same(got, want)
        assert not got["pairs"].any()  # the pair scratch is left zeroed
        if not reward:
            oracle = opan.PanopticState(C, ISTHING)
            oracle.update(*arrays, validate=True)
            expected = oracle.compute()
            for key in ("tp", "fp", "fn", "confusion"):
                np.testing.assert_array_equal(got[key], expected[key], err_msg=key)
            np.testing.assert_allclose(got["iou_sum"], expected["iou_sum"], rtol=1e-12, atol=1e-15)


def test_noalloc_packed_batch_twice_second_smaller(native):
    for seed, sizes in ((3, [(40, 52), (17, 61), (33, 33)]), (4, [(21, 19), (12, 40)])):
        H = max(h for h, _ in sizes)
        W = max(w for _, w in sizes)
        arrays = batch(seed, len(sizes), H, W, 11, 14)
        call = Call(native, arrays, reward=False, exact=False, packed_sizes=sizes)
        assert_same(call.run("pe_update_noalloc"), call.run("pe_update"))


def test_noalloc_rejects_batches_beyond_its_scratch(native):
    limit = native.library().pe_update_noalloc_max_workspace_bytes()
    assert limit == 8 << 20
    # B * P prediction records of 12 bytes each exceed 8 MiB: refused before
    # any launch (the pointers are never dereferenced).
    B, G, P = 4096, 2, 256
    assert native.library().pe_update_workspace_bytes(B, G, P, C, 0) > limit
    code = native.library().pe_update_noalloc(None, 4, None, 4, B, 8, 8, None, None, None, None, G, P, C, None, None,
                                              None, None, None, None, None, None, None, None, None, None, None, None,
                                              0, 0, 0, None)
    assert code < 0
    assert b"static scratch maximum" in native.library().pe_last_error()
