"""Oracle pins for the dihedral and harmonic-restraint terms (BondComputer.h:84-131,
281-346; DihedralGeometry BondGeometry.h:56-87).

What fixes them beyond restating the functors:
* the shared FP32 arctangent against libm's double atan2 (<= 3 ulp, signed zeros);
* the dihedral angle's magnitude against the angle between the two plane normals;
* forces against -grad E by central differences of the FP64 energy, with a table
  linear in phi (so the difference quotient cannot straddle two table bins);
* the +-1000 magnitude clamp and the degenerate (near-collinear) cut-off;
* the restraint's closed form F = -k wrap(x - x0), U = k/2 |wrap(x - x0)|^2, across
  the periodic seam.
"""
import numpy as np

from marsmd import PeriodicBox
from marsmd.core import TabulatedPotential
from oracle import pyoracle as orc

L = 60.0
BOX = PeriodicBox((L, L, L))


def _ulp32(x):
    x = np.abs(np.float32(x))
    return float(np.spacing(x)) if x > 0 else float(np.spacing(np.float32(0)))


def test_atan2_series_accuracy_and_special_values():
    rng = np.random.default_rng(5)
    worst = 0.0
    for scale_y, scale_x in [(1.0, 1.0), (1e-3, 1.0), (1.0, 1e-3), (1e4, 1e-2), (1e-20, 1e-20)]:
        y = (rng.standard_normal(4000) * scale_y).astype(np.float32)
        x = (rng.standard_normal(4000) * scale_x).astype(np.float32)
        for yy, xx in zip(y, x):
            got = orc.atan2_32(float(yy), float(xx))
            ref = np.arctan2(np.float64(yy), np.float64(xx))
            worst = max(worst, abs(got - ref) / _ulp32(ref))
    assert worst <= 3.0, f"atan2_series error {worst} ulp"
    pi32 = float(np.float32(np.pi))
    assert orc.atan2_32(0.0, 1.0) == 0.0
    assert orc.atan2_32(0.0, -1.0) == pi32
    assert orc.atan2_32(-0.0, -1.0) == -pi32
    assert np.signbit(orc.atan2_32(-0.0, 1.0))
    assert orc.atan2_32(0.0, -0.0) == pi32
    assert orc.atan2_32(1.0, 0.0) == float(np.float32(np.pi / 2))
    assert orc.atan2_32(-1.0, 0.0) == -float(np.float32(np.pi / 2))
    assert abs(orc.atan2_32(1.0, 1.0) - np.pi / 4) <= _ulp32(np.pi / 4)
    assert np.isnan(orc.atan2_32(float("nan"), 1.0))


def _linear_dihedral_table(slope=1.0, knots=721):
    """U(phi) = slope * phi on [-pi, pi] (a sawtooth when read periodically)."""
    step = np.float32(2 * np.pi / (knots - 1))
    start = np.float32(-np.pi)
    phi = start + np.arange(knots) * np.float64(step)
    return TabulatedPotential((slope * phi).astype(np.float32), float(start), float(step), is_periodic=True)


def _cos_dihedral_table(k=1.5, n=3, delta=0.4, knots=361):
    step = np.float32(2 * np.pi / (knots - 1))
    start = np.float32(-np.pi)
    phi = start + np.arange(knots) * np.float64(step)
    return TabulatedPotential((k * (1 + np.cos(n * phi - delta))).astype(np.float32), float(start), float(step),
                              is_periodic=True)


def _four(phi, bend=(0.3, 0.5)):
    """i, j, k, l with the j-k axis along x, the i arm in the xy plane and the l arm
    rotated by phi about x (all inside the box, away from its faces)."""
    c = np.array([20.0, 21.0, 22.0])
    p1 = c
    p2 = c + np.array([1.5, 0.0, 0.0])
    p0 = p1 + np.array([-bend[0], 1.2, 0.0])
    p3 = p2 + np.array([bend[1], 1.1 * np.cos(phi), 1.1 * np.sin(phi)])
    pos = np.zeros((4, 4), np.float32)
    pos[:, :3] = np.stack([p0, p1, p2, p3])
    return pos


def _plane_angle(pos):
    p = pos[:, :3].astype(np.float64)
    n1 = np.cross(p[1] - p[0], p[2] - p[1])
    n2 = np.cross(p[2] - p[1], p[3] - p[2])
    c = np.dot(n1, n2) / (np.linalg.norm(n1) * np.linalg.norm(n2))
    return np.arccos(np.clip(c, -1, 1))


def test_dihedral_angle_magnitude_and_one_sign_convention():
    tab = _cos_dihedral_table()
    signs = set()
    for phi in np.deg2rad([-170, -120, -90, -45, -10, 10, 45, 90, 120, 170]):
        pos = _four(phi)
        _, ph64 = orc.dihedral_restraint_forces(pos, BOX, [[0, 1, 2, 3]], None, [tab], mode=2)
        _, ph32 = orc.dihedral_restraint_forces(pos, BOX, [[0, 1, 2, 3]], None, [tab], mode=1)
        mag = _plane_angle(pos)
        assert abs(abs(ph64[0]) - mag) < 1e-9
        assert abs(ph32[0] - ph64[0]) < 1e-5
        signs.add(np.sign(ph64[0]) * np.sign(phi))
    # one global sign relation between the construction angle and the reference's phi
    assert len(signs) == 1


def test_dihedral_forces_are_minus_energy_gradient():
    rng = np.random.default_rng(9)
    tab = _linear_dihedral_table(slope=0.8)
    for trial in range(12):
        phi = rng.uniform(-2.8, 2.8)
        pos = _four(phi, bend=tuple(rng.uniform(-0.6, 0.6, 2)))
        pos[:, :3] += rng.normal(0, 0.05, (4, 3)).astype(np.float32)
        f, _ = orc.dihedral_restraint_forces(pos, BOX, [[0, 1, 2, 3]], None, [tab], mode=2)
        assert abs(f[:, :3].sum(axis=0)).max() < 1e-10  # Newton's third law
        h = 2e-3
        for i in range(4):
            for a in range(3):
                pp, pm = pos.copy(), pos.copy()
                pp[i, a] += h
                pm[i, a] -= h
                dx = float(pp[i, a]) - float(pm[i, a])
                ep = orc.dihedral_restraint_forces(pp, BOX, [[0, 1, 2, 3]], None, [tab], mode=2)[0][:, 3].sum()
                em = orc.dihedral_restraint_forces(pm, BOX, [[0, 1, 2, 3]], None, [tab], mode=2)[0][:, 3].sum()
                fd = -(ep - em) / dx
                assert abs(f[i, a] - fd) <= 1e-4 * max(1.0, abs(fd)), (trial, i, a, f[i, a], fd)


def test_dihedral_clamp_and_degenerate():
    pos = _four(0.7)
    steep = _linear_dihedral_table(slope=-5e4)
    at_limit = _linear_dihedral_table(slope=-1000.0)
    fs, _ = orc.dihedral_restraint_forces(pos, BOX, [[0, 1, 2, 3]], None, [steep], mode=2)
    fl, _ = orc.dihedral_restraint_forces(pos, BOX, [[0, 1, 2, 3]], None, [at_limit], mode=2)
    # the table force is -slope = 5e4 > 1000: clamped to the forces of slope -1000 (whose
    # FP32 knots give a table force within a few 1e-6 of 1000)
    assert np.abs(fs[:, :3] - fl[:, :3]).max() <= 1e-5 * np.abs(fl[:, :3]).max()
    # i almost on the j-k axis: |ab|^2 |bc|^2 / |ab x bc|^2 > 100 -> no force
    deg = pos.copy()
    deg[0, :3] = deg[1, :3] - np.array([1.0, 0.05, 0.0], np.float32)
    fd, _ = orc.dihedral_restraint_forces(deg, BOX, [[0, 1, 2, 3]], None, [_cos_dihedral_table()], mode=2)
    assert np.all(fd[:, :3] == 0.0)


def test_restraint_closed_form_across_the_seam():
    rng = np.random.default_rng(3)
    n = 50
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform(0, L, (n, 3))
    anchors = (pos[:, :3] + rng.normal(0, 3.0, (n, 3))).astype(np.float32)
    anchors[:5, 0] = pos[:5, 0] + L - 1.0  # across the seam: the wrapped offset is +1
    anchors = np.mod(anchors, L).astype(np.float32)
    ids = np.arange(n, dtype=np.int32)
    ids[7] = -1  # skipped
    k = rng.uniform(0.1, 2.0, n).astype(np.float32)
    f, _ = orc.dihedral_restraint_forces(pos, BOX, restraint_ids=ids, restraint_anchors=anchors, restraint_k=k, mode=2)
    d = pos[:, :3].astype(np.float64) - anchors
    d -= L * np.round(d / L)
    ref = -k[:, None] * d
    ref[7] = 0
    assert np.abs(f[:, :3] - ref).max() < 1e-4 * np.abs(ref).max()
    e = 0.5 * k * (d ** 2).sum(axis=1)
    e[7] = 0
    assert np.abs(f[:, 3] - e).max() < 1e-4 * e.max()
    f32, _ = orc.dihedral_restraint_forces(pos, BOX, restraint_ids=ids, restraint_anchors=anchors, restraint_k=k,
                                           mode=1)
    assert np.abs(f32[:, :3] - ref).max() < 1e-5 * np.abs(ref).max()
