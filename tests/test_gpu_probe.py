"""The matrix-core probe behind bench.py's measured peaks runs and reports a
throughput between a tenth of and the published dense peak."""

import importlib
import sys

import pytest
import torch

from conftest import ROOT

pytestmark = pytest.mark.gpu


def test_measured_peaks_are_plausible(gpu):
    print("[case] matrix-core and copy probe", flush=True)
    sys.path.insert(0, ROOT)
    bench = importlib.import_module("bench")
    peaks = bench.measured_peaks(torch.device(gpu), iters=500, copy_mib=256)
    assert 15.0 < peaks["fp32_mfma_tflops"] <= 1.05 * 157.3
    assert 250.0 < peaks["bf16_mfma_tflops"] <= 1.05 * 2500.0
    assert 800.0 < peaks["hbm_copy_gbps"] <= 1.05 * 8000.0
