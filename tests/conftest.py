"""Shared pytest setup: the ``gpu`` marker, repo import path, golden loader."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built libmetalmom.so")


def load_golden(name):
    return np.load(os.path.join(GOLDEN_DIR, name if name.endswith(".npy") else name + ".npy"))


@pytest.fixture(scope="session")
def signal_440():
    return load_golden("signal_440hz_22050sr")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
