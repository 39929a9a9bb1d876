"""The Patch's launch configurations give the same bits.

marsmd_patch_run's step kernel may run with fewer persistent workgroups per CU and,
with grids, take the position-dependent force from a kernel of its own; the
autotuner picks among these on real steps. Each configuration is checked against the
default launch, bit for bit on positions, momenta, forces and ids, after each of
several runs (rebuilds, reorders and energy steps inside them), for Langevin with a
tricubic grid, BD with a linear one and the config-1 polymers (bonded forces summed
per particle in a fixed order); the autotuned run equals the default run.
"""
import numpy as np
import pytest

import marsmd
from marsmd import configs
from marsmd import _capi as api

pytestmark = pytest.mark.gpu


def _state(p):
    g = p.copy_particles_to_host()
    return np.concatenate([g["pos"].view(np.uint32).ravel(), g["mom"].view(np.uint32).ravel(),
                           g["force"].view(np.uint32).ravel(), g["id"].view(np.uint32)])


def _spec(integrator):
    spec = configs.config2(n=20_000, L=400.0, grid_n=64)
    if integrator == api.INTEGRATOR_BROWNIAN:
        spec.params.integrator = api.INTEGRATOR_BROWNIAN
        spec.params.interpolation = api.INTERP_LINEAR
    spec.params.energy_period = 7
    spec.params.reorder_period = 20
    return spec


def _polymers():
    spec = configs.config1(chains=20, beads=200, L=300.0)
    spec.params.energy_period = 5
    return spec


@pytest.mark.parametrize("integrator", [api.INTEGRATOR_LANGEVIN, api.INTEGRATOR_BROWNIAN, "polymers"])
def test_patch_launch_configurations_are_bit_identical(cuda, integrator):
    # the polymers' bonded shares are summed per particle in a fixed order, so their
    # runs repeat bit for bit too
    spec = _polymers() if integrator == "polymers" else _spec(integrator)
    chunks = [13, 9, 25]
    ref = marsmd.Patch(spec)
    want, step = [], 0
    for c in chunks:
        ref.run(c, first_step=step)
        step += c
        cuda.cuda.synchronize()
        want.append(_state(ref))
    for bpc, split in [(0, True), (1, False), (2, True)]:
        print(f"case: integrator {integrator} blocks_per_cu {bpc} split_pmf {split}", flush=True)
        p = marsmd.Patch(spec)
        p.set_launch(bpc, split)
        assert p.get_launch() == {"blocks_per_cu": bpc, "split_pmf": split}
        step = 0
        for k, c in enumerate(chunks):
            p.run(c, first_step=step)
            step += c
            cuda.cuda.synchronize()
            assert np.array_equal(_state(p), want[k]), f"after run {k}"


def test_patch_autotune_keeps_the_trajectory(cuda):
    spec = _spec(api.INTEGRATOR_LANGEVIN)
    ref = marsmd.Patch(spec)
    ref.run(60)
    cuda.cuda.synchronize()
    p = marsmd.Patch(spec)
    choice = p.autotune(60)
    cuda.cuda.synchronize()
    assert choice["ms_per_step"] > 0.0
    assert p.get_launch() == {"blocks_per_cu": choice["blocks_per_cu"], "split_pmf": choice["split_pmf"]}
    assert np.array_equal(_state(p), _state(ref))


def test_patch_cell_configuration_keeps_the_pair_set(cuda):
    # the walk's cells change the order of neighbours in a row (so the FP32 summation
    # order of the forces), not the pair set
    from oracle import pyoracle as orc
    from tests.conftest import rel_force_error, sorted_pairs

    spec = _spec(api.INTEGRATOR_LANGEVIN)
    ref = marsmd.Patch(spec)
    ref.run(1)
    p = marsmd.Patch(spec)
    assert p.get_cells() == (24, 2)
    p.set_cells(8, 3)
    assert p.get_cells() == (8, 3)
    p.run(1)
    cuda.cuda.synchronize()
    a = sorted_pairs(p.get_pairlist().get_neighbor_pairs().cpu().numpy())
    b = sorted_pairs(ref.get_pairlist().get_neighbor_pairs().cpu().numpy())
    assert np.array_equal(a, b)
    fa, fb = p.copy_particles_to_host()["force"], ref.copy_particles_to_host()["force"]
    assert rel_force_error(fa, fb.astype(np.float64)) <= 1e-5
    with pytest.raises(marsmd.MarsError):
        p.set_cells(0, 2)
    # the tuner keeps the pair set of the build it re-walks and steps on
    q = marsmd.Patch(spec)
    choice = q.autotune_cells(20)
    cuda.cuda.synchronize()
    print(f"case: tuned cells {choice}", flush=True)
    assert choice["walk_ms"] > 0.0 and q.get_cells() == (choice["home_target"], choice["fine_offset"])
    q2 = marsmd.Patch(spec)
    q2.set_cells(choice["home_target"], choice["fine_offset"])
    q2.run(1)
    cuda.cuda.synchronize()
    ref_pairs = orc.build_pairlist(spec.pos4(), spec.box, spec.params.pairlist_cutoff)["pairs"].astype(np.int64)
    assert np.array_equal(sorted_pairs(q2.get_pairlist().get_neighbor_pairs().cpu().numpy()), ref_pairs)
