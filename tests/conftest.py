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


def seam_pair_positions(n=20000, L=160.01, seed=21, planted=40):
    """Uniform particles in a periodic cube of 16 cells of L/16 (>= 10 A) per axis, with
    `planted` pairs per axis placed across the periodic seam: one bead in the
    second-to-last cell (x in [150.05, 150.13)), one at x = 0.02, equal y and z, so the
    pair is 9.9-9.98 A apart but two cells apart after wrapping."""
    rs = np.random.default_rng(seed)
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rs.uniform(0, L, size=(n, 3))
    k = 0
    for a in range(3):
        for j in range(planted):
            other = rs.uniform(20.0, L - 20.0, size=3)
            pos[k, :3] = other
            pos[k + 1, :3] = other
            pos[k, a] = 150.05 + 0.002 * j
            pos[k + 1, a] = 0.02
            k += 2
    return pos
