"""GPU parity of the force and integrator kernels through the C ABI.

Forces: max_i |F_i - F_i^ref| <= 1e-5 max_i |F_i^ref| where F^ref sums the oracle's
FP32 per-pair terms in FP64 (the per-pair term is the reference functor's FP32
expression, identical bit for bit on both sides; only summation order differs).
Energies likewise. Integrators: same Philox counters; positions agree to FP32
rounding of the device transcendental functions (tolerance 1e-4 A).
"""
import numpy as np
import pytest

import marsmd
from marsmd import PeriodicBox, configs
from marsmd import _capi as api
from oracle import pyoracle as orc
from tests.conftest import canonical_topology as _canonical_topology, rel_force_error, sorted_pairs

pytestmark = pytest.mark.gpu
FORCE_TOL = 1e-5


def _dev(cuda, a):
    return cuda.from_numpy(np.ascontiguousarray(a)).cuda()


def _min_image(d, L):
    d = np.abs(d)
    return np.minimum(d, L - d)


def test_pairwise_nonbonded_launcher(cuda):
    spec = configs.config0(n=4000, L=160.0, seed=5)
    pos = spec.pos4()
    pairs = orc.build_pairlist(pos, spec.box, 35.0)["pairs"]
    tags = spec.tag_matrix()
    types = np.zeros(len(pos), np.int32)
    tag_d = cuda.zeros(len(pairs), dtype=cuda.int32, device="cuda")
    pd, td = _dev(cuda, pairs), _dev(cuda, types)
    marsmd.launch_resolve_pair_tables(pd, td, _dev(cuda, tags.view(np.int32)), 1, tag_d, len(pairs))
    fe = cuda.zeros((len(pos), 4), dtype=cuda.float32, device="cuda")
    tabs = marsmd.DeviceTables(spec.nb_tables)
    view = marsmd.particle_view(pos=_dev(cuda, pos), force_energy=fe, type_id=td)
    marsmd.launch_pairwise_nonbonded(pd, view, tag_d, tabs, spec.box, True, len(pairs), 30.0 * 30.0)
    cuda.cuda.synchronize()
    ref = orc.pair_forces(pairs, pos, types, tags, spec.nb_tables, spec.box, 900.0, mode=1)
    g = fe.cpu().numpy()
    assert rel_force_error(g, ref) <= FORCE_TOL
    assert abs(g[:, 3].sum() - ref[:, 3].sum()) <= 1e-5 * np.abs(ref[:, 3]).sum()


def test_nonbonded_energy_known_answers(cuda):
    # NonbondedEnergy.cpp:149-265: U = (r-3)^2 - 1 tabulated on [0, 10) with 2048 knots
    step = 10.0 / 2048
    r = np.arange(2048) * np.float32(step)
    tab = marsmd.TabulatedPotential(((r - 3.0) ** 2 - 1.0).astype(np.float32), 0.0, step)
    tabs = marsmd.DeviceTables([tab])
    box = PeriodicBox((1000.0, 1000.0, 1000.0))

    def run(positions, pairs, cutoff):
        pos = np.zeros((len(positions), 4), np.float32)
        pos[:, :3] = positions
        fe = cuda.zeros((len(pos), 4), dtype=cuda.float32, device="cuda")
        tag = _dev(cuda, np.full(len(pairs), marsmd.make_pair_tag(api.PAIR_TERM_TABULATED, 0), np.int32))
        view = marsmd.particle_view(pos=_dev(cuda, pos), force_energy=fe)
        marsmd.launch_pairwise_nonbonded(_dev(cuda, np.array(pairs, np.int32)), view, tag, tabs, box, True,
                                         len(pairs), cutoff * cutoff if cutoff > 0 else 0.0)
        cuda.cuda.synchronize()
        return fe.cpu().numpy()

    def interp(x):
        w = x / np.float32(step)
        h = int(np.floor(w))
        return (tab.pot[h + 1] - tab.pot[h]) * (w - h) + tab.pot[h]

    out = run([[0, 0, 0], [3.5, 0, 0]], [[0, 1]], 0)
    assert out[:, 3].sum() == pytest.approx(interp(3.5), rel=1e-4)
    assert out[0, 3] == pytest.approx(interp(3.5) / 2, rel=1e-4)
    assert run([[0, 0, 0], [3.0, 0, 0]], [[0, 1]], 0)[:, 3].sum() == pytest.approx(-1.0, rel=1e-3)
    assert run([[0, 0, 0], [1.0, 0, 0]], [[0, 1]], 0)[:, 3].sum() > 0
    pos3 = [[0, 0, 0], [3.5, 0, 0], [8.0, 0, 0]]
    e = run(pos3, [[0, 1], [0, 2], [1, 2]], 5.0)[:, 3].sum()
    assert e == pytest.approx(interp(3.5) + interp(4.5), rel=1e-3)


def test_bonded_launchers_and_chain_fixture(cuda):
    from pathlib import Path

    g = np.load(Path(__file__).parent / "golden" / "bond-19.190-3.800.npz")
    X, Y = g["X"].astype(np.float32), g["Y"].astype(np.float32)
    tab = marsmd.TabulatedPotential(Y, float(X[0]), float(np.float32(X[1] - X[0])))
    n = 129
    pos = np.zeros((n, 4), np.float32)
    pos[:, 2] = 1.885 * np.arange(n, dtype=np.float32)
    bonds = np.stack([np.arange(n - 1), np.arange(1, n)], axis=1).astype(np.int32)
    fe = cuda.zeros((n, 4), dtype=cuda.float32, device="cuda")
    box = PeriodicBox((500.0, 500.0, 500.0))
    marsmd.launch_tabulated_bonds(_dev(cuda, bonds), _dev(cuda, pos), fe, marsmd.DeviceTables([tab]),
                                  _dev(cuda, np.zeros(n - 1, np.int32)), None, box, True, n - 1)
    cuda.cuda.synchronize()
    f = fe.cpu().numpy()
    w = (np.float32(1.885) - X[0]) / (X[1] - X[0])
    h = int(np.floor(w))
    fref = -(Y[h + 1] - Y[h]) / (X[1] - X[0])
    # BDKernels.cpp:295-313: ends carry -/+ F_ref, the interior nets to zero
    assert f[0, 2] == pytest.approx(-fref, abs=0.05)
    assert f[-1, 2] == pytest.approx(fref, abs=0.05)
    assert np.abs(f[1:-1, :3]).max() < 0.05
    ref = orc.bonded_forces(pos, box, bonds=bonds, bond_tables=[tab], mode=1)
    assert rel_force_error(f, ref) <= FORCE_TOL


def test_bonds_angles_config1(cuda):
    spec = configs.config1(chains=20, beads=200, L=300.0)
    pos = spec.pos4()
    fe = cuda.zeros((spec.n, 4), dtype=cuda.float32, device="cuda")
    pd = _dev(cuda, pos)
    a4 = np.zeros((len(spec.angles), 4), np.int32)
    a4[:, :3] = spec.angles
    marsmd.launch_tabulated_bonds(_dev(cuda, spec.bonds), pd, fe, marsmd.DeviceTables(spec.bond_tables),
                                  _dev(cuda, spec.bond_table), None, spec.box, True, len(spec.bonds))
    marsmd.launch_tabulated_angles(_dev(cuda, a4), pd, fe, marsmd.DeviceTables(spec.angle_tables),
                                   _dev(cuda, spec.angle_table), None, spec.box, True, len(spec.angles))
    cuda.cuda.synchronize()
    ref = orc.bonded_forces(pos, spec.box, spec.bonds, spec.bond_table, spec.bond_tables, spec.angles,
                            spec.angle_table, spec.angle_tables, mode=1)
    g = fe.cpu().numpy()
    assert rel_force_error(g, ref) <= FORCE_TOL
    assert abs(g[:, 3].sum() - ref[:, 3].sum()) <= 1e-5 * np.abs(ref[:, 3]).sum()
    assert np.abs(g[:, :3].sum(axis=0)).max() < 1e-2  # momentum conservation


@pytest.mark.parametrize("scheme", [api.INTERP_LINEAR, api.INTERP_CUBIC])
@pytest.mark.parametrize("bc", [api.BC_PERIODIC, api.BC_NEUMANN, api.BC_DIRICHLET])
def test_pmf_launcher(cuda, scheme, bc):
    rs = np.random.default_rng(scheme * 10 + bc)
    grid = marsmd.Grid(rs.normal(size=(24, 20, 16)).astype(np.float32), origin=(-5.0, 2.0, 1.0),
                       spacing=(2.0, 2.5, 3.0), boundary_condition=bc)
    n = 5000
    pos = np.zeros((n, 4), np.float32)
    pos[:, 0] = rs.uniform(-12, 50, n)
    pos[:, 1] = rs.uniform(-5, 60, n)
    pos[:, 2] = rs.uniform(-5, 55, n)
    types = marsmd.make_types([marsmd.ParticleType(charge=0.5, pmf_terms=[(0, 1.5, -1)])])
    gv, keep = marsmd.device_grid_views([grid])
    fe = cuda.zeros((n, 4), dtype=cuda.float32, device="cuda")
    view = marsmd.particle_view(pos=_dev(cuda, pos), force_energy=fe, type_id=_dev(cuda, np.zeros(n, np.int32)))
    marsmd.launch_PMF(view, types, n, gv, electric_field=(0.1, -0.2, 0.3), interpolation=scheme)
    cuda.cuda.synchronize()
    s = orc.grid_sample(grid, pos, scheme)
    ref = np.zeros((n, 4))
    ref[:, :3] = -1.5 * s[:, 1:4] + 0.5 * np.array([0.1, -0.2, 0.3])
    ref[:, 3] = 1.5 * s[:, 0]
    g = fe.cpu().numpy()
    assert rel_force_error(g, ref) <= FORCE_TOL
    assert np.abs(g[:, 3] - ref[:, 3]).max() <= 1e-5 * np.abs(ref[:, 3]).max()


def test_bd_and_baoab_launchers_single_step(cuda):
    spec = configs.config0(n=3000, L=150.0, seed=9)
    sim = orc.Sim(spec)
    P = spec.params
    sim.calculate_nonbonded_forces(P.pairlist_cutoff, P.cutoff, 1, P.rebuild_period)
    o = sim.copy_particles_to_host()
    fe = _dev(cuda, o["force"])
    pos = _dev(cuda, o["pos"])
    types = marsmd.make_types(spec.types)
    view = marsmd.particle_view(pos=pos, force_energy=fe, type_id=_dev(cuda, spec.type_ids))
    seed = orc.mix_seed(spec.seed, 0, 1)
    marsmd.launch_BD(view, types, P.timestep, 1, P.kT, spec.n, spec.box, seed)
    sim.integrate_motion(P.timestep, P.kT, P.integrator, 1)
    cuda.cuda.synchronize()
    o2 = sim.copy_particles_to_host()
    assert _min_image(pos.cpu().numpy()[:, :3] - o2["pos"][:, :3], 150.0).max() < 1e-4
    # BAOAB with momenta
    t = [marsmd.ParticleType(mass=100.0, trans_damping=(1000.0,) * 3)]
    mom = np.zeros((spec.n, 4), np.float32)
    mom[:, :3] = np.random.default_rng(1).normal(size=(spec.n, 3)) * 10
    p0 = spec.pos4()
    f0 = np.zeros((spec.n, 4), np.float32)
    f0[:, :3] = np.random.default_rng(2).normal(size=(spec.n, 3))
    pd, md, fd = _dev(cuda, p0), _dev(cuda, mom), _dev(cuda, f0)
    view = marsmd.particle_view(pos=pd, force_energy=fd, type_id=_dev(cuda, spec.type_ids), mom=md)
    marsmd.launch_BAOAB(view, marsmd.make_types(t), spec.box, 1e-4, 3, 0.593, spec.n, seed)
    cuda.cuda.synchronize()
    # closed form with the oracle's Philox + Box-Muller
    hdt = np.float32(0.5) * np.float32(1e-4)
    FCF = np.float32(4.18679994e4)
    for i in (0, 1, 777, spec.n - 1):
        m = mom[i, :3].copy()
        p = p0[i, :3].astype(np.float32).copy()
        m = m + hdt * f0[i, :3] * FCF
        p = p + hdt * m * np.float32(1 / np.float32(100.0)) * np.float32(1e4)
        c = np.float32(np.exp(np.float32(-1000.0) * np.float32(1e-4)))
        ns = np.float32(np.sqrt(np.float32(0.593) * np.float32(100.0) * (np.float32(1) - c * c))) * np.float32(2.046167337)
        raw = orc.philox_draw(seed, i, 3, 0x1356914)[0]
        rn = orc.box_muller([orc.u01(int(x)) for x in raw])
        m = c * m + ns * rn
        p = p + hdt * m * np.float32(1 / np.float32(100.0)) * np.float32(1e4)
        p = np.mod(p, 150.0)
        assert np.abs(md.cpu().numpy()[i, :3] - m).max() < 1e-3
        assert _min_image(pd.cpu().numpy()[i, :3] - p, 150.0).max() < 1e-4


def _check_step(cuda, patch, spec, step, check_pairs=True):
    """One SimManager step on the GPU patch, each stage checked against the oracle on
    the same input: pair set (bit for bit), nonbonded + bonded forces (1e-5), then
    the integrator replayed by the oracle from the GPU's pre-integration state."""
    P = spec.params
    L = spec.box.size[0]
    patch.maybe_reorder(step)
    energy = True
    patch.calculate_nonbonded_forces(P.pairlist_cutoff, P.cutoff, step, P.rebuild_period, P.electric_field,
                                     P.interpolation, energy)
    patch._prepared = True
    patch.calculate_bonded_forces(energy)
    cuda.cuda.synchronize()
    g = patch.copy_particles_to_host()
    topo = _canonical_topology(spec, g["id"])
    gp = sorted_pairs(patch.get_pairlist().get_neighbor_pairs().cpu().numpy())
    if check_pairs:
        # the pairlist was built from the positions of its build step; those are the
        # current ones on a rebuild step
        st = patch.get_pairlist().get_statistics()
        if st.num_particles and (step - 1) % P.rebuild_period == 0:
            off = exc = None
            ne = 0
            if topo["exclusions"] is not None:
                off, exc, ne = orc.exclusion_csr(topo["exclusions"])
            ref = orc.build_pairlist(g["pos"], spec.box, P.pairlist_cutoff, off, exc, ne)
            assert np.array_equal(gp, ref["pairs"].astype(np.int64)), f"pair set differs at step {step}"
    fref = orc.pair_forces(gp, g["pos"], g["type"], spec.tag_matrix(), spec.nb_tables, spec.box,
                           P.cutoff * P.cutoff, mode=1, enabled_terms=spec.enabled_terms, particle_types=spec.types,
                           pair_params=spec.pair_params)
    if topo["bonds"] is not None:
        fref += orc.bonded_forces(g["pos"], spec.box, topo["bonds"], spec.bond_table, spec.bond_tables,
                                  topo["angles"], spec.angle_table, spec.angle_tables, mode=1)
    if topo["dihedrals"] is not None or topo["restraint_ids"] is not None:
        fref += orc.dihedral_restraint_forces(g["pos"], spec.box, topo["dihedrals"], spec.dihedral_table,
                                              spec.dihedral_tables, topo["restraint_ids"], spec.restraint_anchors,
                                              spec.restraint_k, mode=1)[0]
    assert rel_force_error(g["force"], fref) <= FORCE_TOL, f"force error at step {step}"
    assert abs(g["force"][:, 3].sum() - fref[:, 3].sum()) <= 1e-5 * np.abs(fref[:, 3]).sum() + 1e-6
    # integrator replay
    sim = orc.Sim(spec)
    sim.set_state(g["pos"], g["mom"], g["force"], g["type"], g["id"], deferred_kick=(step > 1 and
                  P.integrator == api.INTEGRATOR_LANGEVIN))
    sim.set_field(P.electric_field, P.interpolation)
    patch.integrate_motion(P.timestep, P.kT, P.integrator, step)
    sim.integrate_motion(P.timestep, P.kT, P.integrator, step)
    cuda.cuda.synchronize()
    g2 = patch.copy_particles_to_host()
    o2 = sim.copy_particles_to_host()
    assert np.array_equal(g2["id"], o2["id"])
    assert _min_image(g2["pos"][:, :3] - o2["pos"][:, :3], L).max() < 1e-4, f"positions differ at step {step}"
    if P.integrator == api.INTEGRATOR_LANGEVIN:
        assert np.abs(g2["mom"][:, :3] - o2["mom"][:, :3]).max() <= 1e-4 * np.abs(o2["mom"][:, :3]).max()
    return g2


def test_patch_steps_config0(cuda):
    spec = configs.config0(n=5000, L=190.0, seed=11)
    patch = marsmd.Patch(spec)
    for step in range(1, 44):  # rebuilds at 1, 21, 41; reorders at 21, 41
        _check_step(cuda, patch, spec, step, check_pairs=True)


def test_patch_steps_config1_polymers(cuda):
    spec = configs.config1(chains=30, beads=200, L=350.0)
    patch = marsmd.Patch(spec)
    for step in range(1, 24):
        _check_step(cuda, patch, spec, step)


def test_patch_steps_config2_langevin_grid(cuda):
    spec = configs.config2(n=20_000, L=400.0, grid_n=64)
    patch = marsmd.Patch(spec)
    for step in range(1, 23):
        _check_step(cuda, patch, spec, step)


def _open_y_spec():
    """The config-0 system in a box open along y (periodic in x and z)."""
    spec = configs.config0(n=4000, L=170.0, seed=13)
    spec.box = PeriodicBox(spec.box.size, periodic=(True, False, True))
    return spec


def test_patch_steps_open_axis(cuda):
    # the pair term's wrap runs the same operations on an open axis (a 3e38 length with
    # a zero inverse): stage by stage against the oracle, rebuilds at 1 and 21
    spec = _open_y_spec()
    patch = marsmd.Patch(spec)
    for step in range(1, 24):
        _check_step(cuda, patch, spec, step)


def test_fused_run_open_axis(cuda):
    spec = _open_y_spec()
    fused = marsmd.Patch(spec)
    unfused = marsmd.Patch(spec)
    steps = 25
    fused.run(steps)
    for s in range(1, steps + 1):
        unfused.step_unfused(s)
    cuda.cuda.synchronize()
    a = fused.copy_particles_to_host()
    b = unfused.copy_particles_to_host()
    assert np.array_equal(a["id"], b["id"])
    assert np.abs(a["pos"][:, :3] - b["pos"][:, :3]).max() < 1e-3


@pytest.mark.parametrize("which", ["config0", "config1", "config2"])
def test_fused_run_matches_unfused_and_oracle(cuda, which):
    if which == "config0":
        spec = configs.config0(n=6000, L=200.0, seed=12)
    elif which == "config1":
        spec = configs.config1(chains=30, beads=200, L=350.0)
    else:
        spec = configs.config2(n=20_000, L=400.0, grid_n=64)
    L = spec.box.size[0]
    fused = marsmd.Patch(spec)
    unfused = marsmd.Patch(spec)
    sim = orc.Sim(spec)
    # stiff WCA contacts amplify FP32 summation-order differences chaotically in the
    # dense polymer melt; its whole-run comparison stays within a few steps
    steps = 3 if which == "config1" else 25
    fused.run(steps)
    for s in range(1, steps + 1):
        unfused.step_unfused(s)
    sim.run(steps)
    cuda.cuda.synchronize()
    a = fused.copy_particles_to_host()
    b = unfused.copy_particles_to_host()
    o = sim.copy_particles_to_host()
    assert np.array_equal(a["id"], b["id"])
    assert _min_image(a["pos"][:, :3] - b["pos"][:, :3], L).max() < 1e-3
    # FP32 noise grows slowly over a short run; the oracle's trajectory stays close
    close = _min_image(a["pos"][:, :3] - o["pos"][:, :3], L).max(axis=1) < 1e-2
    assert close.mean() > 0.999


@pytest.mark.parametrize("moves", ["below_half_skin", "beyond_half_skin"])
def test_period_step_rebuilds_only_beyond_half_the_skin(cuda, moves):
    # Patch::calculate_nonbonded_forces (Patch.cpp:85-121): a rebuild-period step
    # rebuilds only when a particle has moved more than skin / 2 since the last build;
    # reordering is off, so only that check decides (BD time steps of 1e-6 and 5e-4 ns)
    spec = configs.config0(n=4000, L=170.0, seed=3)
    dt = 1e-6 if moves == "below_half_skin" else 5e-4
    P = configs.RunParams(cutoff=30.0, pairlist_cutoff=35.0, rebuild_period=5, reorder_period=10**6, timestep=dt,
                          kT=configs.KT, integrator=api.INTEGRATOR_BROWNIAN)
    skin = P.pairlist_cutoff - P.cutoff
    L = spec.box.size[0]
    patch = marsmd.Patch(spec)
    patch.run(5, params=P)  # step 1 builds
    cuda.cuda.synchronize()
    g = patch.copy_particles_to_host()  # the positions step 6 starts from
    assert np.array_equal(g["id"], np.arange(spec.n))
    moved = np.sqrt((_min_image(g["pos"][:, :3] - spec.pos4()[:, :3], L) ** 2).sum(axis=1)).max()
    patch.run(1, first_step=5, params=P)  # step 6: a rebuild-period step
    cuda.cuda.synchronize()
    gp = sorted_pairs(patch.get_pairlist().get_neighbor_pairs().cpu().numpy())
    first = orc.build_pairlist(spec.pos4(), spec.box, P.pairlist_cutoff)["pairs"].astype(np.int64)
    fresh = orc.build_pairlist(g["pos"], spec.box, P.pairlist_cutoff)["pairs"].astype(np.int64)
    print(f"case: {moves}, largest displacement {moved:.3f} A", flush=True)
    if moves == "below_half_skin":
        assert moved < 0.5 * skin
        assert not np.array_equal(first, fresh)  # the positions did change the set
        assert np.array_equal(gp, first)         # but the list is step 1's
    else:
        assert moved > 0.5 * skin
        assert np.array_equal(gp, fresh)
