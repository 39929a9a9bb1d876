"""Pins the oracle's analytical pair terms (Coulomb, Debye-Hueckel, ONC, Gaussian,
softcore; Pairwise.h:184-305, Columb.h:11-101).

The cases of Unit_test/Interactions/AnalyticalNB.cpp:133-266 (closed forms in double
precision, sign conventions, Newton's third law, additivity, empty mask, cutoff,
softcore ordering), with the dielectric of ONC in its exponential form
eps = Sz (1 - (r/z)^2 e^(r/z) / (e^(r/z) - 1)^2) rather than the sinh form the
functor uses; then -dU/dr of every term against central differences of the oracle's
own FP64 energy on both sides of each branch (ONC small-v series, v >= 1, the 1e-2
floor; softcore inside and outside the core), and FP32 against the FP64 twin.
"""
import numpy as np
import pytest

from marsmd import PairParams, ParticleType, PeriodicBox
from marsmd import _capi as api
from oracle import pyoracle as orc

C = 332.0636
BOX = PeriodicBox((1000.0, 1000.0, 1000.0))
PP = PairParams(debye_length=10.0, dielectric=80.0, onck_kappa=0.1, onck_sz=80.0, onck_z=6.86, gaussian_amp=2.5,
                gaussian_sigma=3.0)
COUL, DH, ONCK, GAUSS, SOFT = (api.PAIR_TERM_COULOMB, api.PAIR_TERM_DEBYE_HUCKEL, api.PAIR_TERM_ONCK,
                               api.PAIR_TERM_GAUSSIAN, api.PAIR_TERM_SOFTCORE)
ALL = COUL | DH | ONCK | GAUSS | SOFT


def run_pair(terms, r, q1=1.0, cutoff2=0.0, mode=1, params=PP, types=None, direction=(1.0, 0.0, 0.0)):
    """Particle 0 (charge +1) at the origin, particle 1 (charge q1) at r along
    `direction`; both of unit radius and eps. Returns (N, 4) float64 force/energy."""
    pos = np.zeros((2, 4), np.float32)
    pos[1, :3] = np.asarray(direction, np.float64) * r
    types = types or [ParticleType(charge=1.0, radius=1.0, eps=1.0), ParticleType(charge=q1, radius=1.0, eps=1.0)]
    return orc.pair_forces(np.array([[0, 1]]), pos, np.array([0, 1]), None, [], BOX, cutoff2, mode=mode,
                           enabled_terms=terms, particle_types=types, pair_params=params)


def fd_force(u, r):
    h = 1e-5 * max(1.0, r)
    return -(u(r + h) - u(r - h)) / (2.0 * h)


def dh_energy(r, qi=1.0, qj=1.0):
    return qi * qj * C / 80.0 * np.exp(-r / 10.0) / r


def onc_energy(r, qi=1.0, qj=1.0, kappa=0.1, sz=80.0, z=6.86):
    e = np.exp(r / z)
    eps = sz * (1.0 - (r * r) / (z * z) * e / (e - 1.0) ** 2)
    return qi * qj * C * np.exp(-kappa * r) / (eps * r)


def gauss_energy(r):
    return 2.5 * np.exp(-(r * r) / 9.0)


def test_coulomb_closed_form():
    r = 2.0
    out = run_pair(COUL, r)
    u = lambda d: C / d
    assert out[0, 0] == pytest.approx(-fd_force(u, r), rel=1e-5)
    assert out[1, 0] == pytest.approx(fd_force(u, r), rel=1e-5)
    assert out[:, 3].sum() == pytest.approx(u(r), rel=1e-6)


def test_debye_huckel_closed_form():
    r = 4.0
    out = run_pair(DH, r)
    assert out[0, 0] == pytest.approx(-fd_force(dh_energy, r), rel=1e-5)
    assert out[:, 3].sum() == pytest.approx(dh_energy(r), rel=1e-6)


@pytest.mark.parametrize("r", [5.0, 20.0])  # v = r / 2z below and above 1
def test_onc_closed_form_exponential_dielectric(r):
    out = run_pair(ONCK, r)
    assert out[0, 0] == pytest.approx(-fd_force(onc_energy, r), rel=1e-4)
    assert out[:, 3].sum() == pytest.approx(onc_energy(r), rel=1e-5)


def test_gaussian_closed_form_ignores_charge():
    r = 2.0
    out = run_pair(GAUSS, r)
    assert out[0, 0] == pytest.approx(-fd_force(gauss_energy, r), rel=1e-5)
    assert out[:, 3].sum() == pytest.approx(gauss_energy(r), rel=1e-6)
    flipped = run_pair(GAUSS, r, q1=-1.0)
    assert flipped[0, 0] == out[0, 0]


def test_gaussian_own_cutoff():
    p = PairParams(gaussian_amp=2.5, gaussian_sigma=3.0, gaussian_cutoff_squared=4.0)
    assert run_pair(GAUSS, 1.9, params=p)[0, 0] != 0.0
    assert np.all(run_pair(GAUSS, 2.1, params=p) == 0.0)


def test_opposite_charges_attract():
    like = run_pair(COUL, 2.0, 1.0)
    opp = run_pair(COUL, 2.0, -1.0)
    assert like[0, 0] < 0.0 < opp[0, 0]
    assert opp[0, 0] == pytest.approx(-like[0, 0], rel=1e-6)


def test_newton_third_law_all_terms():
    out = run_pair(ALL, 2.0, direction=(0.6, 0.0, 0.8))
    assert np.abs(out[0, :3] + out[1, :3]).max() <= 1e-6 * np.abs(out[0, :3]).max()


def test_terms_add():
    a = run_pair(COUL, 2.0)[0, 0]
    b = run_pair(GAUSS, 2.0)[0, 0]
    assert run_pair(COUL | GAUSS, 2.0)[0, 0] == pytest.approx(a + b, rel=1e-6)


def test_empty_mask_and_cutoff():
    assert np.all(run_pair(0, 2.0) == 0.0)
    assert run_pair(COUL, 2.0, cutoff2=9.0)[0, 0] != 0.0
    assert np.all(run_pair(COUL, 5.0, cutoff2=9.0) == 0.0)


def test_softcore_repels_harder_when_closer():
    near = run_pair(SOFT, 0.7)[0, 0]
    far = run_pair(SOFT, 0.9)[0, 0]
    assert near < far < 0.0


def _energy64(terms, r, params=PP, types=None):
    return run_pair(terms, r, mode=2, params=params, types=types)[:, 3].sum()


def fd_force_at_f32(u, r, rel_h=1e-4):
    """-dU/dr at float32 positions: the two points are rounded to float32 (where the
    oracle places the particle) and their true spacing is the denominator."""
    h = rel_h * r
    rp, rm = float(np.float32(r + h)), float(np.float32(r - h))
    return -(u(rp) - u(rm)) / (rp - rm)


@pytest.mark.parametrize("terms,radii", [
    (COUL, [1.5, 3.0, 25.0]),
    (DH, [1.5, 4.0, 25.0]),
    (ONCK, [0.5, 5.0, 13.0, 14.5, 30.0]),  # v = r / 2z on both sides of 1
    (GAUSS, [0.5, 2.0, 4.0]),
    (SOFT, [0.6, 0.8, 0.95]),  # inside the core (r < rad): U is the shifted 12-6 well
])
def test_force_is_minus_energy_gradient_fp64(terms, radii):
    for r in radii:
        f = run_pair(terms, r, mode=2)[1, 0]
        assert f == pytest.approx(fd_force_at_f32(lambda d: _energy64(terms, d), r), rel=2e-5, abs=1e-9), f"r={r}"


def test_onc_small_v_series_and_floor():
    # a wide sigmoid (z = 1e3) puts v = r / 2z below 1e-5 at r < 0.02: the sz v^2 / 3
    # branch; below r = 0.01 the distance is floored, so the energy is flat there
    p = PairParams(onck_z=1000.0)
    for r in [0.012, 0.015, 0.03]:
        f = run_pair(ONCK, r, mode=2, params=p)[1, 0]
        assert f == pytest.approx(fd_force_at_f32(lambda d: _energy64(ONCK, d, p), r), rel=2e-5), f"r={r}"
    e1 = _energy64(ONCK, 0.004, p)
    e2 = _energy64(ONCK, 0.008, p)
    assert e1 == e2


def test_softcore_outside_core_keeps_force_without_energy():
    # softcoreForce (Pairwise.h:80-97) returns the 12-6 force with zero energy when
    # d^6 >= rad^6
    out = run_pair(SOFT, 1.3)
    assert out[:, 3].sum() == 0.0
    d2 = 1.3 ** 2
    d6 = d2 ** 3
    expect = -(-12.0 * (1.0 / (d6 * d6 * d2) - 1.0 / (d6 * d2))) * 1.3
    assert out[1, 0] == pytest.approx(expect, rel=1e-5)


def test_softcore_mixing_rules():
    # rad = (r_i + r_j) / 2, eps = sqrt(eps_i eps_j): a (0.5, 1.5) pair with eps (4, 1)
    # equals a unit-radius pair with eps 2
    mixed = [ParticleType(charge=1.0, radius=0.5, eps=4.0), ParticleType(charge=1.0, radius=1.5, eps=1.0)]
    unit = [ParticleType(charge=1.0, radius=1.0, eps=2.0), ParticleType(charge=1.0, radius=1.0, eps=2.0)]
    a = run_pair(SOFT, 0.8, types=mixed)
    b = run_pair(SOFT, 0.8, types=unit)
    assert np.allclose(a, b, rtol=1e-6)


def test_fp32_terms_match_fp64_twin():
    rs = np.random.default_rng(3)
    for terms in (COUL, DH, ONCK, GAUSS, SOFT, ALL):
        for r in rs.uniform(0.5, 25.0, 20):
            f32 = run_pair(terms, r, mode=1, direction=(0.48, 0.6, 0.64))
            f64 = run_pair(terms, r, mode=2, direction=(0.48, 0.6, 0.64))
            scale = np.abs(f64[:, :3]).max()
            assert np.abs(f32[:, :3] - f64[:, :3]).max() <= 1e-5 * scale + 1e-12, (terms, r)
