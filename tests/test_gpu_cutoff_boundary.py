"""The pair term tests the cutoff as d > t, t precomputed on the host as the largest
float whose FP32 square rounds to at most cutoff^2; the reference tests d * d > cutoff2
(PairNonbondedComputer, Pairwise.h:184-305). Since fl(d * d) never decreases as d grows
the two agree for every d. Pinned here at the boundary itself: isolated bead pairs are
placed so that the distance the kernel computes (the FP32 difference, its square, the
correctly rounded square root) runs over the floats around t, for cutoffs whose square
is and is not a float, and each pair must feel a force exactly when the reference's
FP32 predicate keeps it.
"""
import numpy as np
import pytest

import marsmd
from marsmd import configs

pytestmark = pytest.mark.gpu


def _threshold(c2):
    t = np.sqrt(c2, dtype=np.float32)
    while np.float32(t * t) > c2:
        t = np.nextafter(t, np.float32(0))
    while np.float32(np.nextafter(t, np.float32(np.inf)) ** 2) <= c2:
        t = np.nextafter(t, np.float32(np.inf))
    return np.float32(t)


def _pairs(cutoff):
    """One pair per float x1 from t - 4 ulp to t + 4 ulp: beads at (0, y, 10) and
    (x1, y, 10), 50 A apart in y from the next pair, so the kernel's difference is x1
    itself and its distance sqrt(fl(x1 * x1))."""
    c2 = np.float32(np.float32(cutoff) * np.float32(cutoff))
    t = _threshold(c2)
    x1 = t
    for _ in range(4):
        x1 = np.nextafter(x1, np.float32(0))
    pos, dist = [], []
    for k in range(9):
        y = np.float32(10.0 + 50.0 * k)
        pos.append([(0.0, y, 10.0), (x1, y, 10.0)])
        dist.append(np.sqrt(np.float32(x1 * x1), dtype=np.float32))
        x1 = np.nextafter(x1, np.float32(np.inf))
    keep = [not (np.float32(d * d) > c2) for d in dist]
    return np.array(pos, np.float32).reshape(-1, 3), np.array(dist, np.float32), np.array(keep), t


@pytest.mark.parametrize("cutoff", [12.3, 21.0, 7.77])  # inside the 30 A table
def test_pair_cutoff_boundary(cuda, cutoff):
    L = 500.0
    pos, dist, keep, t = _pairs(cutoff)
    assert keep.any() and not keep.all(), (dist, t)  # both sides of the boundary
    n = len(pos)
    spec = configs.config0(n=n, L=L, seed=3)
    spec.positions = pos
    spec.params.cutoff = float(cutoff)
    spec.params.pairlist_cutoff = float(cutoff) + 2.0
    patch = marsmd.Patch(spec)
    P = spec.params
    patch.calculate_nonbonded_forces(P.pairlist_cutoff, P.cutoff, 1, P.rebuild_period)
    cuda.cuda.synchronize()
    f = patch.copy_particles_to_host()["force"][:, :3]
    fx = f[:, 0].reshape(-1, 2)
    for k in range(len(dist)):
        felt = fx[k, 0] != 0.0 and fx[k, 1] != 0.0
        assert felt or (fx[k] == 0.0).all()
        assert felt == keep[k], f"pair {k}: d {dist[k]!r} t {t!r} kept {keep[k]} force {fx[k]}"
    assert np.abs(f[:, 1:]).max() == 0.0
