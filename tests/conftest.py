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


def canonical_topology(spec, ids):
    """Bonded/exclusion tuples of the spec (written in global ids) in the patch's
    current canonical indices: canonical c holds global id ids[c]."""
    canon_of_id = np.empty(len(ids), np.int64)
    canon_of_id[ids] = np.arange(len(ids))
    out = {}
    for name in ("bonds", "angles", "exclusions", "dihedrals"):
        arr = getattr(spec, name)
        out[name] = None if arr is None else canon_of_id[np.asarray(arr)].astype(np.int32)
    r = spec.restraint_ids
    if r is None:
        out["restraint_ids"] = None
    else:
        r = np.asarray(r)
        out["restraint_ids"] = np.where(r >= 0, canon_of_id[np.maximum(r, 0)], -1).astype(np.int32)
    return out


def min_image(d, L):
    """|d| folded to the minimum image of a periodic length L (per component)."""
    d = np.abs(np.asarray(d, np.float64))
    return np.minimum(d, L - d)


def host_threads(cap: int = 16) -> int:
    """CPU threads this process may use (its affinity mask, at most `cap`)."""
    import os

    return max(1, min(cap, len(os.sched_getaffinity(0))))
