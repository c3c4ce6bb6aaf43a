"""Host-side parts of the reward server (no GPU needed): the native loop's
first-fit arena, which holds the HBM budget, and the request checks that run
on submission before anything is queued."""

import ctypes
import pathlib

import pytest
import torch

ROOT = pathlib.Path(__file__).resolve().parents[1]
LIB = ROOT / "panoptic-evaluator_amd" / "panoptic_evaluator" / "libpanoptic_hip.so"


@pytest.fixture(scope="module")
def arena_api():
    if not LIB.exists():
        pytest.skip("library not built")
    lib = ctypes.CDLL(str(LIB))
    i64, p = ctypes.c_int64, ctypes.c_void_p
    lib.pe_arena_create.restype, lib.pe_arena_create.argtypes = p, [i64]
    lib.pe_arena_alloc.restype, lib.pe_arena_alloc.argtypes = i64, [p, i64]
    lib.pe_arena_release.argtypes = [p, i64, i64]
    lib.pe_arena_in_use.restype, lib.pe_arena_in_use.argtypes = i64, [p]
    lib.pe_arena_largest_free.restype, lib.pe_arena_largest_free.argtypes = i64, [p]
    lib.pe_arena_destroy.argtypes = [p]
    return lib


def test_arena_first_fit_and_budget(arena_api):
    lib = arena_api
    a = lib.pe_arena_create(1024)
    x = lib.pe_arena_alloc(a, 300)          # rounded to 512
    y = lib.pe_arena_alloc(a, 256)
    assert (x, y) == (0, 512) and lib.pe_arena_in_use(a) == 768
    assert lib.pe_arena_alloc(a, 512) == -1  # only 256 left: the request waits
    z = lib.pe_arena_alloc(a, 200)
    assert z == 768 and lib.pe_arena_in_use(a) == 1024
    lib.pe_arena_release(a, x, 300)
    assert lib.pe_arena_largest_free(a) == 512 and lib.pe_arena_alloc(a, 512) == 0
    lib.pe_arena_destroy(a)


def test_arena_coalesces_freed_neighbours(arena_api):
    lib = arena_api
    a = lib.pe_arena_create(4096)
    blocks = [lib.pe_arena_alloc(a, 1024) for _ in range(4)]
    assert blocks == [0, 1024, 2048, 3072] and lib.pe_arena_alloc(a, 1) == -1
    for off in (1024, 3072, 2048):          # middle, end, then the gap between
        lib.pe_arena_release(a, off, 1024)
    assert lib.pe_arena_largest_free(a) == 3072
    lib.pe_arena_release(a, 0, 1024)
    assert lib.pe_arena_in_use(a) == 0 and lib.pe_arena_largest_free(a) == 4096
    assert lib.pe_arena_alloc(a, 4096) == 0
    lib.pe_arena_destroy(a)


def pq_request(h=8, w=6, g=3, p=4):
    import panoptic_evaluator as pe
    from panoptic_evaluator.serving import PQRequest
    target = pe.PanopticBatch(torch.zeros((2, h, w), dtype=torch.int32), torch.full((2, g), -1, dtype=torch.int32))
    prediction = pe.PanopticBatch(torch.zeros((2, h, w), dtype=torch.int32), torch.full((2, p), -1, dtype=torch.int32))
    return PQRequest(prediction, target)


def test_malformed_requests_rejected_before_queueing():
    import panoptic_evaluator as pe
    from panoptic_evaluator.serving import PQRequest, PrefixRequest, describe
    good = pq_request()
    kind, rows, points, keep = describe(good, 5, (0.5,))
    assert kind == 0 and rows.shape == (2, 9) and list(rows[0, 5:]) == [8, 6, 3, 4]
    mismatched = PQRequest(good.prediction, pe.PanopticBatch(torch.zeros((2, 8, 7), dtype=torch.int32),
                                                             good.target.classes))
    with pytest.raises(ValueError, match="image sizes"):
        describe(mismatched, 5, (0.5,))
    int64_maps = PQRequest(pe.PanopticBatch(good.prediction.segment_map.long(), good.prediction.classes),
                           good.target)
    with pytest.raises(ValueError, match="int32"):
        describe(int64_maps, 5, (0.5,))
    masks = torch.zeros((3, 8, 6), dtype=torch.bool)
    prefix = PrefixRequest(pe.InstanceBatch([masks, masks], [torch.zeros(3, dtype=torch.int64)] * 2), good.target,
                           [0, 4])
    with pytest.raises(ValueError, match="checkpoint"):
        describe(prefix, 5, (0.5,))
    with pytest.raises(TypeError):
        describe(object(), 5, (0.5,))


def test_footprint_grows_with_the_request():
    from panoptic_evaluator.serving import footprint
    small, large = footprint(pq_request(), 5), footprint(pq_request(h=64, w=64, g=9, p=11), 5)
    assert 0 < small < large
