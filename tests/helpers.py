"""Comparison helpers shared by the GPU parity tests."""

import numpy as np
import torch

INT_KEYS = ("tp", "fp", "fn", "confusion")


def to_numpy(value):
    if isinstance(value, dict):
        return {k: to_numpy(v) for k, v in value.items()}
    if isinstance(value, list):
        return [to_numpy(v) for v in value]
    if isinstance(value, torch.Tensor):
        return value.detach().cpu().numpy()
    return value


def assert_exact(actual, expected, what=""):
    actual, expected = np.asarray(to_numpy(actual)), np.asarray(expected)
    assert actual.shape == expected.shape, (what, actual.shape, expected.shape)
    np.testing.assert_array_equal(actual, expected, err_msg=what)


def assert_close(actual, expected, rtol=1e-12, what=""):
    actual = np.asarray(to_numpy(actual), np.float64)
    expected = np.asarray(expected, np.float64)
    assert actual.shape == expected.shape, (what, actual.shape, expected.shape)
    np.testing.assert_allclose(actual, expected, rtol=rtol, atol=1e-15, equal_nan=True, err_msg=what)


def compare_panoptic(actual, expected, *, exact_iou_sum=False):
    """GPU compute() dict vs oracle compute() dict: integers bit-exact; iou_sum
    and the ratios <= 1e-12 relative (the GPU's dataset iou_sum is the exact
    sum rounded once, the oracle's an in-order FP64 fold)."""
    for key in INT_KEYS:
        assert_exact(actual[key], expected[key], key)
    if exact_iou_sum:
        assert_exact(actual["iou_sum"], expected["iou_sum"], "iou_sum")
    else:
        assert_close(actual["iou_sum"], expected["iou_sum"], what="iou_sum")
    for name in ("All", "Things", "Stuff"):
        for key in ("pq", "sq", "rq"):
            assert_close(actual[name][key], expected[name][key], what=f"{name}.{key}")
        assert_exact(actual[name]["n"], expected[name]["n"], f"{name}.n")
    for key in ("pq", "sq", "rq", "iou"):
        assert_close(actual["per_class"][key], expected["per_class"][key], what=f"per_class.{key}")
    assert_close(actual["miou"], expected["miou"], what="miou")
