"""Loader for the golden .npz fixtures written by tests/golden/make_golden.py."""

import pathlib

import numpy as np

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"


def load(name):
    with np.load(GOLDEN / f"{name}.npz", allow_pickle=False) as data:
        return {key: data[key] for key in data.files}


def group(data, prefix):
    """Sub-dictionary of keys starting with ``prefix.`` (prefix stripped)."""
    start = prefix + "."
    return {key[len(start):]: value for key, value in data.items() if key.startswith(start)}
