"""Shared test set-up.

Markers: `gpu` tests need an MI355X and call the HIP path through the C-ABI;
everything else runs on the CPU (oracle, host logic, library loading).
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (HIP path)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import load_oracle
    return load_oracle()


@pytest.fixture(scope="session")
def ref():
    from oracle import load_ref, ref_available
    if not ref_available():
        pytest.skip("compiled reference (oracle/_ref) not built")
    return load_ref()


@pytest.fixture(scope="session")
def gpu():
    """Skips unless a GPU is visible; returns the torch module."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def adversarial_angles():
    """Angles that stress the traversal: axis-aligned and diagonal in the
    float conversions the reference uses, their float neighbours, and
    angles outside [0, pi) (quadrant rule vs slope sign)."""
    kpi = np.float32(3.14159265358979)
    base = [0.0, kpi / 4, kpi / 2, 3 * kpi / 4, kpi, 5 * kpi / 4, 3 * kpi / 2,
            7 * kpi / 4, 2 * kpi, -kpi / 4, -kpi / 2, 2.5 * kpi, 1e-7, -1e-7]
    out = []
    for b in base:
        b = np.float32(b)
        out += [b, np.nextafter(b, np.float32(10)), np.nextafter(b, np.float32(-10))]
    return np.array(out, np.float32)
