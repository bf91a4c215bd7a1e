"""Shared pytest configuration: the ``gpu`` marker and golden helpers."""

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


def load_golden(op, method, variant):
    """Load a reference golden ``.npz`` (keys ``out`` or ``out0``, ``out1``)."""
    path = os.path.join(GOLDEN, f"{op}__{method}__{variant}.npz")
    with np.load(path, allow_pickle=False) as d:
        return {k: np.asarray(d[k]) for k in d.files}


def canonical_inputs():
    """The reference's canonical 18x18 inputs (``tests/fixtures/inputs.py:141-275``)."""
    n = 18
    x = np.linspace(0.0, 1.0, n)
    y = np.linspace(0.0, 1.0, n)
    Y, X = np.meshgrid(y, x, indexing="ij")
    h = np.sin(2.0 * np.pi * X) * np.cos(np.pi * Y) + 0.5 * (X + Y)
    u = np.cos(np.pi * X) * np.sin(2.0 * np.pi * Y)
    v = -np.sin(np.pi * X) * np.cos(2.0 * np.pi * Y)
    q = np.sin(np.pi * X) * np.sin(np.pi * Y)
    hm = np.zeros((n, n), dtype=bool)
    hm[1:-1, 1:-1] = True
    hm[1, :] = hm[-2, :] = False
    hm[:, 1] = hm[:, -2] = False
    hm[6:9, 6:11] = False
    hm[5:10, 7:10] = False
    dx = dy = 1.0 / 16.0
    return dict(h=h, u=u, v=v, q=q, hmask=hm, dx=dx, dy=dy)


@pytest.fixture
def canon():
    return canonical_inputs()
