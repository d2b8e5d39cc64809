import importlib
import os
import sys

import numpy as np
import pytest
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (ROCm device) and the built library")


def pkg():
    return importlib.import_module("mlff-distiller_amd")


def pkg_mod(name):
    return importlib.import_module(f"mlff-distiller_amd.{name}")


def golden(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def tiny_state(fx):
    return {k[len("param/"):]: torch.tensor(fx[k]) for k in fx.files if k.startswith("param/")}


@pytest.fixture(scope="session")
def gpu():
    if not torch.cuda.is_available():
        pytest.skip("no ROCm device")
    return torch.device("cuda:0")


@pytest.fixture(autouse=True)
def _seed():
    torch.manual_seed(42)
    np.random.seed(42)
