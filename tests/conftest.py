import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
for p in (str(ROOT / "mars-md_amd"), str(ROOT)):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long-running")


@pytest.fixture(scope="session")
def cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test collected without a GPU")
    return torch


def sorted_pairs(p: np.ndarray) -> np.ndarray:
    p = np.asarray(p).reshape(-1, 2)
    lo = np.minimum(p[:, 0], p[:, 1])
    hi = np.maximum(p[:, 0], p[:, 1])
    order = np.lexsort((hi, lo))
    return np.stack([lo[order], hi[order]], axis=1).astype(np.int64)


def rel_force_error(f, ref) -> float:
    """max_i |F_i - F_ref,i| / max_i |F_ref,i| (normwise relative error)."""
    f = np.asarray(f, np.float64)[:, :3]
    ref = np.asarray(ref, np.float64)[:, :3]
    scale = np.abs(ref).max()
    return float(np.abs(f - ref).max() / max(scale, 1e-30))
