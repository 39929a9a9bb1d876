"""GPU parity of the analytical pair terms (Coulomb, Debye-Hueckel, ONC, Gaussian,
softcore; PairNonbondedComputer Pairwise.h:184-305, AnalyticalPairKernel
AnalyticalPairKernels.h:21-118) against the oracle.

A three-type system whose pair tags mix every analytical term with the tabulated one:
* launch_pairwise_nonbonded and launch_analytical_pair_nonbonded over an explicit pair
  list: forces within 1e-5 of the oracle's FP32 terms summed in FP64, energies summed;
* the Patch: each of 23 steps (rebuilds and reorders included) checked stage by stage
  as the tabulated-only systems are;
* the replica batch: each replica against its own GPU Patch over 25 steps.
"""
import numpy as np
import pytest

import marsmd
from marsmd import PairParams, ParticleType, configs
from marsmd import _capi as api
from marsmd.batch import Batch
from oracle import pyoracle as orc
from tests.conftest import rel_force_error
from tests.test_gpu_forces_integrators import _check_step, _dev, _min_image

pytestmark = pytest.mark.gpu
FORCE_TOL = 1e-5
TAB = api.PAIR_TERM_TABULATED
COUL, DH, ONCK, GAUSS, SOFT = (api.PAIR_TERM_COULOMB, api.PAIR_TERM_DEBYE_HUCKEL, api.PAIR_TERM_ONCK,
                               api.PAIR_TERM_GAUSSIAN, api.PAIR_TERM_SOFTCORE)
ANALYTIC = COUL | DH | ONCK | GAUSS | SOFT


def analytic_spec(n=4000, L=170.0, seed=7):
    """The config-0 system with three types (charges 0.3, -0.3, 0.15; softcore radii
    2, 3, 2.5 and eps 0.5, 1, 0.2) and a tag per type pair mixing the terms."""
    spec = configs.config0(n=n, L=L, seed=seed)
    D = (150.0, 150.0, 150.0)
    spec.types = [ParticleType(mass=1.0, charge=0.3, diffusion=D, radius=2.0, eps=0.5),
                  ParticleType(mass=1.0, charge=-0.3, diffusion=D, radius=3.0, eps=1.0),
                  ParticleType(mass=1.0, charge=0.15, diffusion=D, radius=2.5, eps=0.2)]
    spec.type_ids = (np.arange(n) % 3).astype(np.int32)
    tag = marsmd.make_pair_tag
    terms = {(0, 0): TAB | DH | SOFT, (0, 1): COUL | ONCK | TAB, (1, 1): GAUSS | SOFT,
             (0, 2): DH, (1, 2): TAB | SOFT, (2, 2): ANALYTIC | TAB}
    m = np.zeros((3, 3), np.uint32)
    for (i, j), t in terms.items():
        m[i, j] = m[j, i] = tag(t, 0)
    spec.term_matrix = m
    spec.enabled_terms = ANALYTIC | TAB
    spec.pair_params = PairParams(debye_length=12.0, dielectric=78.5, onck_kappa=0.12, onck_sz=78.0, onck_z=7.0,
                                  gaussian_amp=2.5, gaussian_sigma=3.0, gaussian_cutoff_squared=400.0)
    spec.name = f"analytic-{n}-beads-L{L:g}"
    return spec


def _launch_inputs(cuda, spec):
    pos = spec.pos4()
    pairs = orc.build_pairlist(pos, spec.box, spec.params.pairlist_cutoff)["pairs"]
    tags = spec.tag_matrix()
    pd, td = _dev(cuda, pairs), _dev(cuda, spec.type_ids)
    tag_d = cuda.zeros(len(pairs), dtype=cuda.int32, device="cuda")
    marsmd.launch_resolve_pair_tables(pd, td, _dev(cuda, tags.view(np.int32)), 3, tag_d, len(pairs))
    return pos, pairs, tags, pd, td, tag_d


def test_pairwise_launcher_all_terms(cuda):
    spec = analytic_spec()
    pos, pairs, tags, pd, td, tag_d = _launch_inputs(cuda, spec)
    fe = cuda.zeros((len(pos), 4), dtype=cuda.float32, device="cuda")
    view = marsmd.particle_view(pos=_dev(cuda, pos), force_energy=fe, type_id=td)
    cut2 = spec.params.cutoff ** 2
    marsmd.launch_pairwise_nonbonded(pd, view, tag_d, marsmd.DeviceTables(spec.nb_tables), spec.box, True, len(pairs),
                                     cut2, spec.enabled_terms, types=spec.types, params=spec.pair_params)
    cuda.cuda.synchronize()
    ref = orc.pair_forces(pairs, pos, spec.type_ids, tags, spec.nb_tables, spec.box, cut2, mode=1,
                          enabled_terms=spec.enabled_terms, particle_types=spec.types, pair_params=spec.pair_params)
    g = fe.cpu().numpy()
    assert rel_force_error(g, ref) <= FORCE_TOL
    assert abs(g[:, 3].sum() - ref[:, 3].sum()) <= 1e-5 * np.abs(ref[:, 3]).sum()


@pytest.mark.parametrize("terms", [COUL, DH, ONCK, GAUSS, SOFT, ANALYTIC])
def test_analytical_launcher(cuda, terms):
    spec = analytic_spec(n=3000, L=150.0, seed=9)
    pos, pairs, tags, pd, td, tag_d = _launch_inputs(cuda, spec)
    fe = cuda.zeros((len(pos), 4), dtype=cuda.float32, device="cuda")
    cut2 = spec.params.cutoff ** 2
    marsmd.launch_analytical_pair_nonbonded(pd, _dev(cuda, pos), fe, td, tag_d, spec.types, spec.box, terms, True,
                                            len(pairs), cut2, params=spec.pair_params)
    cuda.cuda.synchronize()
    ref = orc.pair_forces(pairs, pos, spec.type_ids, tags, [], spec.box, cut2, mode=1, enabled_terms=terms,
                          particle_types=spec.types, pair_params=spec.pair_params)
    g = fe.cpu().numpy()
    if np.abs(ref[:, :3]).max() == 0.0:  # no pair carries this term
        assert np.abs(g).max() == 0.0
        return
    assert rel_force_error(g, ref) <= FORCE_TOL, f"terms {terms}"
    assert abs(g[:, 3].sum() - ref[:, 3].sum()) <= 1e-5 * np.abs(ref[:, 3]).sum() + 1e-6


def test_patch_steps_with_analytical_terms(cuda):
    spec = analytic_spec()
    patch = marsmd.Patch(spec)
    for step in range(1, 24):  # rebuilds at 1 and 21
        _check_step(cuda, patch, spec, step)


def test_fused_run_and_batch_with_analytical_terms(cuda):
    specs = [analytic_spec(n=3000, L=155.0, seed=21), analytic_spec(n=2200, L=140.0, seed=22)]
    for r, s in enumerate(specs):
        s.patch_id = r
        s.seed = 42  # the batch's base seed is shared; the patch id tells the streams apart
    steps = 25
    b = Batch(specs)
    b.run(steps)
    cuda.cuda.synchronize()
    got = b.copy_particles_to_host()
    for r, s in enumerate(specs):
        L = s.box.size[0]
        p = marsmd.Patch(s)
        p.run(steps)
        sim = orc.Sim(s)
        sim.run(steps)
        cuda.cuda.synchronize()
        g = p.copy_particles_to_host()
        o = sim.copy_particles_to_host()
        assert np.array_equal(got[r]["id"], g["id"])
        assert _min_image(got[r]["pos"][:, :3] - g["pos"][:, :3], L).max() < 1e-3, f"replica {r}"
        close = _min_image(g["pos"][:, :3] - o["pos"][:, :3], L).max(axis=1) < 1e-2
        assert close.mean() > 0.999, f"replica {r}"
