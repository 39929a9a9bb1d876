"""The Patch's launch configurations give the same bits.

marsmd_patch_run's step kernel may run with fewer persistent workgroups per CU and,
with grids, take the position-dependent force from a kernel of its own; the
autotuner picks among these on real steps. Each configuration is checked against the
default launch, bit for bit on positions, momenta, forces and ids, after each of
several runs (rebuilds, reorders and energy steps inside them), for Langevin with a
tricubic grid and BD with a linear one; the autotuned run equals the default run.
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


@pytest.mark.parametrize("integrator", [api.INTEGRATOR_LANGEVIN, api.INTEGRATOR_BROWNIAN])
def test_patch_launch_configurations_are_bit_identical(cuda, integrator):
    spec = _spec(integrator)
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
