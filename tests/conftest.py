import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def nw():
    import neural_wavelet_prt_demo_amd as nw
    nw.lib()  # fail loudly if the HIP library was not built
    return nw


@pytest.fixture(scope="session")
def orc():
    from oracle import binding
    binding.lib()
    return binding


@pytest.fixture(scope="session")
def ctx(nw):
    c = nw.Context(0)
    yield c
    c.close()


def psnr(ref: np.ndarray, test: np.ndarray) -> float:
    """PSNR on linear radiance with peak = max of the reference image."""
    mse = float(np.mean((np.asarray(ref, np.float64) - np.asarray(test, np.float64)) ** 2))
    peak = float(np.max(ref))
    return float("inf") if mse == 0 else 10.0 * np.log10(peak * peak / mse)


def max_rel_err(ref: np.ndarray, test: np.ndarray, floor: float = 1e-3) -> float:
    """max |test - ref| / max(|ref|, floor * max|ref|)."""
    ref = np.asarray(ref, np.float64)
    den = np.maximum(np.abs(ref), floor * np.max(np.abs(ref)))
    return float(np.max(np.abs(np.asarray(test, np.float64) - ref) / den))
