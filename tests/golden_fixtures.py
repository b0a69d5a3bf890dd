"""Loading helpers for the PyElastica golden vectors extracted from the
reference's FullRodTest sources (tests/golden/*.json, written by
tests/golden/extract_golden.py)."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(GOLDEN, name)) as fh:
        return json.load(fh)


def as_field(entry, tag, shrink_diagonal=False):
    """Fixture array -> the (ncomp, slots) layout used by the block fields.

    Tensors (3,3,n) become 9 lines (page-major, the blaze DynamicTensor order);
    diagonal-only tensors (rigidities, inertia and its inverse) keep their
    diagonal as 3 lines, like shrink_using_diagonal (FullRodTest/Helpers.cpp:19-30).
    """
    a = np.array(entry["data"], dtype=np.float64).reshape(entry["shape"])
    if entry["kind"] == "vector":
        return a.reshape(1, -1)
    if entry["kind"] == "matrix":
        return a
    if tag == "Director":
        return a.reshape(9, a.shape[2])
    # diagonal storage for S, B, J, J^-1
    off = a.copy()
    for i in range(3):
        off[i, i, :] = 0.0
    assert not off.any(), "non-diagonal %s fixture" % tag
    return np.stack([a[i, i, :] for i in range(3)])


def rel_err_ok(actual, expected, eps):
    """Approx::custom() semantics (elastica/Utilities/Approx.hpp:36-40, scale 0,
    margin 0): |a - b| <= eps * |b| elementwise; an expected 0 must be hit exactly."""
    actual = np.asarray(actual, dtype=np.float64)
    expected = np.asarray(expected, dtype=np.float64)
    return np.abs(actual - expected) <= eps * np.abs(expected)
