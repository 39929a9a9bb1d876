"""Pins the oracle's tables, pair/bond/angle forces, grid samplers and integrators
against closed forms, finite differences, the FP64 twin, invariants and statistics
(NonbondedEnergy.cpp:149-265, TabulatedBonded.cpp:189-403, BDKernels.cpp:92-320,
Pmf.cpp:107-292, grid_boundary_tests.cpp:37-237, StochasticIntegrators.cpp:157-172)."""
from pathlib import Path

import numpy as np
import pytest

import marsmd
from marsmd import PeriodicBox, configs
from marsmd import _capi as api
from oracle import pyoracle as orc

BOX = PeriodicBox((1000.0, 1000.0, 1000.0))


def _table(fn, n=2048, rmax=10.0):
    step = rmax / n
    r = np.arange(n) * np.float32(step)
    return marsmd.TabulatedPotential(fn(r).astype(np.float32), 0.0, step)


def test_table_lookup_closed_form():
    t = _table(lambda r: (r - 3.0) ** 2 - 1.0)
    # U(3) = -1 exactly at a knot neighbourhood; linear interpolation between knots
    assert orc.table_compute(t, 3.0)[1] == pytest.approx(-1.0, abs=1e-5)
    for x in (0.3, 1.7, 3.5, 9.2):
        w = np.float32(np.float32(x) * np.float32(t.step_inv))
        h = int(np.floor(w))
        e = (t.pot[h + 1] - t.pot[h]) * (w - h) + t.pot[h]
        f = -(t.pot[h + 1] - t.pot[h]) * np.float32(t.step_inv)
        got = orc.table_compute(t, x)
        assert got[1] == pytest.approx(float(e), rel=1e-6)
        assert got[0] == pytest.approx(float(f), rel=1e-6)
    # beyond the last knot: zero force, last value (TabulatedPotential.h:32-35)
    assert list(orc.table_compute(t, 11.0)) == [0.0, t.pot[-1]]
    # below start: home clamps to 0, w is not clamped
    got = orc.table_compute(t, -0.01)
    assert got[0] == pytest.approx(-(t.pot[1] - t.pot[0]) * t.step_inv, rel=1e-6)


def test_periodic_table_wraps():
    t = configs.cosine_angle_table()
    f_lo = orc.table_compute(t, 0.1)
    f_wrapped = orc.table_compute(t, 0.1 + len(t.pot) / t.step_inv)
    assert np.allclose(f_lo, f_wrapped, rtol=1e-4, atol=1e-5)


def test_pair_energy_known_answers():
    t = _table(lambda r: (r - 3.0) ** 2 - 1.0)
    tag = np.array([[marsmd.make_pair_tag(api.PAIR_TERM_TABULATED, 0)]], np.uint32)

    def energy(pos, pairs, cut2=0.0):
        p = np.zeros((len(pos), 4), np.float32)
        p[:, :3] = pos
        out = orc.pair_forces(np.array(pairs), p, np.zeros(len(pos)), tag, [t], BOX, cut2, mode=1)
        return out

    o = energy([[0, 0, 0], [3.5, 0, 0]], [[0, 1]])
    assert o[:, 3].sum() == pytest.approx(0.25 - 1.0, abs=1e-4)
    assert o[0, 3] == pytest.approx(o[1, 3])
    assert energy([[0, 0, 0], [3.0, 0, 0]], [[0, 1]])[:, 3].sum() == pytest.approx(-1.0, abs=1e-4)
    # cutoff drop: (0,2) at 8 A dropped by cutoff 5
    o = energy([[0, 0, 0], [3.5, 0, 0], [8, 0, 0]], [[0, 1], [0, 2], [1, 2]], 25.0)
    assert o[:, 3].sum() == pytest.approx((0.25 - 1.0) + (2.25 - 1.0), abs=1e-3)
    # force sign: repulsive wall pushes particle 1 along +x (F = -dU/dr)
    o = energy([[0, 0, 0], [1.0, 0, 0]], [[0, 1]])
    assert o[1, 0] > 0 and o[0, 0] < 0


def test_pair_force_is_minus_gradient_of_energy_and_fp64_twin():
    spec = configs.config0(n=400, L=60.0, seed=1)
    pos = spec.pos4()
    pairs = orc.build_pairlist(pos, spec.box, 35.0)["pairs"]
    tags = spec.tag_matrix()
    f1 = orc.pair_forces(pairs, pos, spec.type_ids, tags, spec.nb_tables, spec.box, 900.0, mode=1)
    f0 = orc.pair_forces(pairs, pos, spec.type_ids, tags, spec.nb_tables, spec.box, 900.0, mode=0)
    f2 = orc.pair_forces(pairs, pos, spec.type_ids, tags, spec.nb_tables, spec.box, 900.0, mode=2)
    scale = np.abs(f1[:, :3]).max()
    assert np.abs(f0[:, :3] - f1[:, :3]).max() < 1e-5 * scale  # FP32 vs FP64 summation
    # the FP64 twin differs only where a distance sits within FP32 rounding of a
    # table bin edge (piecewise-constant force); the large majority agrees to 1e-5
    close = np.abs(f2[:, :3] - f1[:, :3]).max(axis=1) < 1e-5 * scale
    assert close.mean() > 0.95
    # momentum conservation (Newton's third law on every pair)
    assert np.abs(f1[:, :3].sum(axis=0)).max() < 1e-6 * scale * len(pos)
    # -dE/dx for one particle, central differences of the linear interpolant within a bin
    i = int(np.argmax(np.abs(f1[:, 0])))
    h = 1e-4
    def total_energy(x):
        p = pos.copy()
        p[i, 0] = x
        return orc.pair_forces(pairs, p, spec.type_ids, tags, spec.nb_tables, spec.box, 900.0, mode=2)[:, 3].sum()
    fd = -(total_energy(pos[i, 0] + h) - total_energy(pos[i, 0] - h)) / (2 * h)
    assert fd == pytest.approx(f2[i, 0], rel=2e-2, abs=1e-3)


def test_bond_chain_fixture():
    g = np.load(Path(__file__).parent / "golden" / "bond-19.190-3.800.npz")
    X, Y = g["X"].astype(np.float32), g["Y"].astype(np.float32)
    tab = marsmd.TabulatedPotential(Y, float(X[0]), float(np.float32(X[1] - X[0])))
    n = 129
    pos = np.zeros((n, 4), np.float32)
    pos[:, 2] = 1.885 * np.arange(n, dtype=np.float32)
    bonds = np.stack([np.arange(n - 1), np.arange(1, n)], axis=1)
    f = orc.bonded_forces(pos, PeriodicBox((500.0,) * 3), bonds=bonds, bond_tables=[tab])
    w = (np.float32(1.885) - X[0]) / (X[1] - X[0])
    h = int(np.floor(w))
    fref = -(Y[h + 1] - Y[h]) / (X[1] - X[0])
    assert f[0, 2] == pytest.approx(-fref, abs=0.05)
    assert f[-1, 2] == pytest.approx(fref, abs=0.05)
    assert np.abs(f[1:-1, :3]).max() < 0.05
    # the fixture is a force-capped harmonic: linear far from r0 = 3.8, quadratic near it
    assert fref == pytest.approx(19.19 * (3.8 - 1.885), rel=0.02) or abs(fref) > 10


def test_shared_acos_accuracy():
    # the series arccosine used by kernels and oracle: within 2 ulp of acos over [-1, 1]
    xs = np.concatenate([np.linspace(-1, 1, 20001), [-1.0, -0.5, 0.0, 0.5, 1.0, 0.49999997, 0.50000006]])
    for x in np.float32(xs):
        got = np.float32(orc.acos32(float(x)))
        ref = np.arccos(np.float64(x))
        ulp = np.spacing(np.float32(max(ref, 1e-30)))
        assert abs(float(got) - ref) <= 2.5 * ulp, (x, got, ref)


def test_angle_forces_finite_difference_and_invariants():
    # ramp table U(theta) = 2 theta (degrees abscissa, periodic), TabulatedBonded.cpp:189-260
    knots = 361
    deg = np.arange(knots, dtype=np.float64)
    ramp = marsmd.TabulatedPotential((2.0 * np.deg2rad(deg)).astype(np.float32), 0.0,
                                     float(np.float32(1.0) * configs.DEG_TO_RAD), is_periodic=True)
    rs = np.random.default_rng(3)
    for _ in range(20):
        p = np.zeros((3, 4), np.float32)
        p[:, :3] = rs.normal(size=(3, 3)) * 3.0 + 50.0
        ang = np.array([[0, 1, 2]])
        f = orc.bonded_forces(p, BOX, angles=ang, angle_tables=[ramp], mode=2)
        # momentum and torque conservation
        assert np.abs(f[:, :3].sum(axis=0)).max() < 1e-9
        tq = np.cross(p[:, :3].astype(np.float64) - p[1, :3], f[:, :3]).sum(axis=0)
        assert np.abs(tq).max() < 1e-8
        # in-plane
        nrm = np.cross(p[0, :3] - p[1, :3], p[2, :3] - p[1, :3]).astype(np.float64)
        assert abs(np.dot(f[0, :3], nrm)) < 1e-6 * np.linalg.norm(nrm) * (np.abs(f).max() + 1)
        # finite differences of the energy on particle 0
        h = 1e-5
        for a in range(3):
            q1 = p.astype(np.float64).copy(); q1[0, a] += h
            q2 = p.astype(np.float64).copy(); q2[0, a] -= h
            th = lambda q: np.arccos(np.dot(q[0, :3] - q[1, :3], q[2, :3] - q[1, :3]) /
                                     (np.linalg.norm(q[0, :3] - q[1, :3]) * np.linalg.norm(q[2, :3] - q[1, :3])))
            fd = -(2.0 * th(q1) - 2.0 * th(q2)) / (2 * h)
            assert f[0, a] == pytest.approx(fd, rel=1e-3, abs=1e-4)


@pytest.mark.parametrize("scheme", [api.INTERP_LINEAR, api.INTERP_CUBIC])
def test_grid_samplers_exact_on_polynomials(scheme):
    # trilinear reproduces linear functions; Catmull-Rom reproduces quadratics
    n = 12
    c = np.arange(n, dtype=np.float64)
    x, y, z = np.meshgrid(c, c, c, indexing="ij")
    if scheme == api.INTERP_LINEAR:
        v = 0.5 * x - 0.25 * y + 0.75 * z + 1.0
        grad = lambda p: np.array([0.5, -0.25, 0.75]) / 2.0
        val = lambda p: 0.5 * p[0] / 2 - 0.25 * p[1] / 2 + 0.75 * p[2] / 2 + 1.0
    else:
        v = 0.1 * x * x - 0.2 * y * z + 0.3 * z + 0.05 * x * y
        val = lambda p: 0.1 * (p[0] / 2) ** 2 - 0.2 * (p[1] / 2) * (p[2] / 2) + 0.3 * p[2] / 2 + 0.05 * (p[0] / 2) * (p[1] / 2)
        grad = lambda p: np.array([0.2 * p[0] / 2 + 0.05 * p[1] / 2, -0.2 * p[2] / 2 + 0.05 * p[0] / 2,
                                   -0.2 * p[1] / 2 + 0.3]) / 2.0
    g = marsmd.Grid(v.astype(np.float32), origin=(0.0, 0.0, 0.0), spacing=(2.0, 2.0, 2.0),
                    boundary_condition=api.BC_NEUMANN)
    rs = np.random.default_rng(4)
    pts = np.zeros((50, 4), np.float32)
    pts[:, :3] = rs.uniform(4.0, 2.0 * (n - 3), size=(50, 3))
    out = orc.grid_sample(g, pts, scheme, fp64=True)
    for i in range(50):
        assert out[i, 0] == pytest.approx(val(pts[i, :3].astype(np.float64)), abs=1e-5)
        assert np.allclose(out[i, 1:], grad(pts[i, :3].astype(np.float64)), atol=1e-5)


@pytest.mark.parametrize("bc", [api.BC_PERIODIC, api.BC_NEUMANN, api.BC_DIRICHLET])
@pytest.mark.parametrize("scheme", [api.INTERP_LINEAR, api.INTERP_CUBIC])
def test_grid_gradient_matches_finite_difference(bc, scheme):
    rs = np.random.default_rng(bc * 7 + scheme)
    g = marsmd.Grid(rs.normal(size=(10, 9, 8)).astype(np.float32), origin=(1.0, -2.0, 0.5),
                    spacing=(1.5, 2.0, 1.25), boundary_condition=bc)
    pts = np.zeros((40, 4), np.float32)
    pts[:, 0] = rs.uniform(-3, 18, 40)
    pts[:, 1] = rs.uniform(-6, 18, 40)
    pts[:, 2] = rs.uniform(-3, 12, 40)
    base = orc.grid_sample(g, pts, scheme, fp64=True)
    f32 = orc.grid_sample(g, pts, scheme, fp64=False)
    assert np.allclose(f32, base, rtol=1e-4, atol=1e-4)
    h = 1e-4
    for a in range(3):
        pp = pts.copy(); pm = pts.copy()
        pp[:, a] += h; pm[:, a] -= h
        # positions are float32: divide by the perturbation actually applied
        dx = pp[:, a].astype(np.float64) - pm[:, a].astype(np.float64)
        fd = (orc.grid_sample(g, pp, scheme, fp64=True)[:, 0] - orc.grid_sample(g, pm, scheme, fp64=True)[:, 0]) / dx
        # skip points within h of a grid line (the linear interpolant's kinks)
        gp = (pts[:, a] - g.origin[a]) / g.spacing[a]
        ok = np.abs(gp - np.round(gp)) > 1e-2
        assert np.allclose(fd[ok], base[ok, 1 + a], rtol=1e-3, atol=2e-3)


def test_bd_single_step_closed_form():
    spec = configs.config0(n=50, L=500.0, seed=5)
    spec.nb_tables = [marsmd.TabulatedPotential(np.zeros(4, np.float32), 0.0, 1.0)]
    sim = orc.Sim(spec)
    P = spec.params
    sim.calculate_nonbonded_forces(P.pairlist_cutoff, P.cutoff, 1, P.rebuild_period)
    f = np.zeros((spec.n, 4), np.float32)
    f[:, :3] = np.random.default_rng(1).normal(size=(spec.n, 3))
    sim.set_state(force=f)
    p0 = spec.pos4()
    sim.integrate_motion(P.timestep, P.kT, P.integrator, 1)
    out = sim.copy_particles_to_host()["pos"]
    seed = orc.mix_seed(spec.seed, 0, 1)
    D = np.float32(150.0)
    for i in (0, 7, 49):
        raw = orc.philox_draw(seed, i, 1, 0x5324120)[0]
        rn = orc.box_muller([orc.u01(int(x)) for x in raw])
        exp = p0[i, :3] + (D / np.float32(P.kT)) * f[i, :3] * np.float32(P.timestep) + \
            np.sqrt(np.float32(2.0) * D * np.float32(P.timestep)) * rn
        assert np.allclose(out[i, :3], np.mod(exp, 500.0), atol=2e-5)


def test_bd_free_diffusion_msd():
    # BDKernels.cpp:92-220: <|dr|^2> = 6 D t within 10%
    n, steps = 4000, 200
    spec = configs.config0(n=n, L=3000.0, seed=2)
    spec.nb_tables = [marsmd.TabulatedPotential(np.zeros(4, np.float32), 0.0, 1.0)]
    spec.params.cutoff, spec.params.pairlist_cutoff = 2.0, 3.0
    sim = orc.Sim(spec)
    x0 = spec.pos4()[:, :3].astype(np.float64)
    sim.run(steps)
    out = sim.copy_particles_to_host()  # canonical order changes at each reorder
    x1 = out["pos"][:, :3].astype(np.float64)
    d = x1 - x0[out["id"]]
    d -= 3000.0 * np.round(d / 3000.0)
    msd = (d ** 2).sum(axis=1).mean()
    assert msd == pytest.approx(6 * 150.0 * steps * 1e-4, rel=0.1)


def test_baoab_equipartition():
    # StochasticIntegrators.cpp:157-172: <p_i^2> = kT m SQRT_CAL_TO_JOULE^2 within 8%
    n, steps = 3000, 3000
    mass, gamma, kT = 100.0, 5000.0, 0.593
    spec = configs.config0(n=n, L=5000.0, seed=3)
    spec.types = [marsmd.ParticleType(mass=mass, trans_damping=(gamma,) * 3)]
    spec.nb_tables = [marsmd.TabulatedPotential(np.zeros(4, np.float32), 0.0, 1.0)]
    spec.params.cutoff, spec.params.pairlist_cutoff = 2.0, 3.0
    spec.params.integrator = api.INTEGRATOR_LANGEVIN
    sim = orc.Sim(spec)
    sim.run(steps)
    m = sim.copy_particles_to_host()["mom"][:, :3].astype(np.float64)
    want = kT * mass * 2.046167337 ** 2
    assert np.allclose((m ** 2).mean(axis=0), want, rtol=0.08)


def test_threaded_timing_mode_matches_the_serial_oracle():
    # the OpenMP mode used for the multi-core CPU baseline: same pairs, forces summed per
    # thread (FP32 rounding only), per-particle integrator loops unchanged
    spec = configs.config0(n=3000, L=170.0, seed=3)
    serial, threaded = orc.Sim(spec), orc.Sim(spec)
    threaded.set_threads(4)
    serial.run(5)
    threaded.run(5)
    a, b = serial.copy_particles_to_host(), threaded.copy_particles_to_host()
    assert np.array_equal(serial.get_pairs(), threaded.get_pairs())
    assert np.abs(a["pos"][:, :3] - b["pos"][:, :3]).max() < 1e-3
    fa, fb = a["force"][:, :3].astype(np.float64), b["force"][:, :3].astype(np.float64)
    assert np.abs(fa - fb).max() <= 1e-5 * np.abs(fa).max()
