"""CPU checks of the C ABI: the built library loads and exports every entry
point declared in include/panoptic_hip.h (no compute calls without a GPU)."""

import ctypes
import pathlib
import re
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "panoptic_hip.h"
LIB = ROOT / "panoptic-evaluator_amd" / "panoptic_evaluator" / "libpanoptic_hip.so"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(pe_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "pe_update" in names and "pe_compute" in names and len(names) >= 6


@pytest.mark.skipif(not LIB.exists(), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(LIB))
    missing = [name for name in declared_functions() if not hasattr(lib, name)]
    assert not missing, missing
    lib.pe_abi_version.restype = ctypes.c_int
    assert lib.pe_abi_version() == 3


@pytest.mark.skipif(not LIB.exists(), reason="library not built")
def test_library_has_no_torch_dependency():
    dynamic = subprocess.run(["readelf", "-d", str(LIB)], capture_output=True, text=True).stdout
    needed = " ".join(line for line in dynamic.splitlines() if "(NEEDED)" in line)
    assert "torch" not in needed and "c10" not in needed
    assert "libamdhip64" in needed


@pytest.mark.skipif(not LIB.exists(), reason="library not built")
def test_workspace_queries_are_host_only():
    lib = ctypes.CDLL(str(LIB))
    fn = lib.pe_update_workspace_bytes
    fn.restype = ctypes.c_int64
    fn.argtypes = [ctypes.c_int64] * 5
    small, large = fn(1, 2, 3, 4, 0), fn(8, 41, 47, 133, 0)
    assert 0 < small <= large and large % 256 == 0
    # A packed batch adds its tile list (one int per 4096 pixels plus two per image).
    packed = fn(8, 41, 47, 133, 8 * 512 * 512)
    assert packed >= large + 4 * (8 * 512 * 512 // 4096)
    det = lib.pe_detection_histogram_workspace_bytes
    det.restype = ctypes.c_int64
    det.argtypes = [ctypes.c_int64] * 2
    assert det(16, 0) == 0 and det(16, 16 * 800 * 1088) >= 4 * (16 * 800 * 1088 // 4096)


REFERENCE_BINDINGS = ("update", "prefix", "compute", "pack", "instance_update", "instance_compute",
                      "instance_histogram", "instance_match_histogram")  # bindings.cpp:29-56


def test_shim_module_has_the_reference_bindings():
    import inspect

    from panoptic_evaluator import _cuda_evaluator
    for name in REFERENCE_BINDINGS:
        assert callable(getattr(_cuda_evaluator, name)), name
    # Argument order of the reference's update binding (bindings.cpp:30-49).
    params = list(inspect.signature(_cuda_evaluator.update).parameters)
    assert params == ["gt", "pred", "gc", "pc", "crowd", "pairs", "areas", "matched", "tp", "fp", "fn", "iou_sum",
                      "confusion", "gt_ids", "gt_slots", "pred_ids", "pred_slots", "error", "validate", "reward"]
    assert inspect.signature(_cuda_evaluator.update).parameters["reward"].default is False


@pytest.mark.skipif(not LIB.exists(), reason="library not built")
def test_shim_rejects_host_tensors_like_torch_check():
    import torch

    from panoptic_evaluator import _cuda_evaluator
    maps = torch.zeros((1, 4, 4), dtype=torch.int32)
    ids = torch.zeros((1, 2), dtype=torch.int64)
    with pytest.raises(RuntimeError):
        _cuda_evaluator.pack(maps, ids, ids.int())
    with pytest.raises(RuntimeError):
        _cuda_evaluator.compute(*(torch.zeros(3, dtype=torch.int64) for _ in range(3)),
                                torch.zeros(3, dtype=torch.float64), torch.zeros((3, 4), dtype=torch.int64),
                                torch.zeros(3, dtype=torch.bool))
