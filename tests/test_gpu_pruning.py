"""The fused step's pruned rows give the full rows' bits.

With pruning on, after each build marsmd_patch_run walks a copy of the neighbour rows
pruned to cutoff + 2 margin (placed by length within each block of 256 slots, or at
their own slots); a step falls back to the full rows, on the device, once a particle
has moved a margin from its position at the prune. An entry the copy drops adds
exactly +-0 to a particle's sums while no particle has moved that far, and the copy
keeps the list order, so every mode (off, sorted only, pruned sorted or in place with
the adaptive margin, with margins the particles outgrow at once or mid-life, a margin
past skin / 2 that keeps every entry) must give
the same bits on positions, momenta, forces and ids after each of several runs:
Langevin with a tricubic grid (also with the grid force in its own kernel), BD with a
linear grid, the config-1 polymers (bonds, angles, exclusions; their bonded shares are
summed per particle in a fixed order) and the analytical pair terms. Unfused steps
between runs (positions moved outside the displacement record) must leave it so too.
"""
import numpy as np
import pytest

import marsmd
from marsmd import configs
from marsmd import _capi as api
from tests.test_gpu_analytical_terms import analytic_spec

pytestmark = pytest.mark.gpu

MODES = [(1, -1.0), (2, -1.0), (2, 1e-4), (2, 0.05), (2, 100.0), (3, -1.0), (3, 0.05)]


def _state(p):
    g = p.copy_particles_to_host()
    return np.concatenate([g["pos"].view(np.uint32).ravel(), g["mom"].view(np.uint32).ravel(),
                           g["force"].view(np.uint32).ravel(), g["id"].view(np.uint32)])


def _langevin():
    spec = configs.config2(n=20_000, L=400.0, grid_n=64)
    spec.params.energy_period = 7
    spec.params.reorder_period = 20
    return spec


def _langevin_slow():
    # a fiftieth of the time step: displacements well inside the skin, so the adaptive
    # margin prunes
    spec = _langevin()
    spec.params.timestep = spec.params.timestep / 50.0
    return spec


def _bd():
    spec = _langevin()
    spec.params.integrator = api.INTEGRATOR_BROWNIAN
    spec.params.interpolation = api.INTERP_LINEAR
    return spec


def _polymers():
    spec = configs.config1(chains=20, beads=200, L=300.0)
    spec.params.energy_period = 5
    return spec


def _analytic():
    spec = analytic_spec(n=6000, L=190.0)
    spec.params.energy_period = 4
    return spec


CASES = {"langevin_grid": (_langevin, False), "langevin_grid_split": (_langevin, True),
         "langevin_slow": (_langevin_slow, False), "bd_grid": (_bd, False), "polymers": (_polymers, False),
         "analytic": (_analytic, False)}


def _runs(spec, mode, margin, split, chunks, unfused=0):
    p = marsmd.Patch(spec)
    p.set_pruning(mode, margin)
    if split:
        p.set_launch(0, True)
    out, step = [], 0
    for c in chunks:
        p.run(c, first_step=step)
        step += c
        for _ in range(unfused):
            step += 1
            p.step_unfused(step)
        marsmd_sync()
        out.append(_state(p))
    return p, out


def marsmd_sync():
    import torch
    torch.cuda.synchronize()


@pytest.mark.parametrize("case", list(CASES))
def test_pruned_rows_are_bit_identical(cuda, case):
    make, split = CASES[case]
    spec = make()
    chunks = [13, 9, 23]
    _, want = _runs(spec, 0, -1.0, split, chunks)
    for mode, margin in MODES:
        print(f"case: {case} pruning mode {mode} margin {margin}", flush=True)
        p, got = _runs(spec, mode, margin, split, chunks)
        for k in range(len(chunks)):
            assert np.array_equal(got[k], want[k]), f"{case}: mode {mode} margin {margin}, after run {k}"
        live = p.get_pruning()
        assert live["mode"] == mode
        if mode == 2 and margin == 100.0:
            assert live["keep_all"]
        if mode in (2, 3) and margin < 0.0 and case == "langevin_slow":
            # the adaptive margin follows the displacement of the previous life
            skin = spec.params.pairlist_cutoff - spec.params.cutoff
            assert 0.0 < live["margin"] < 0.5 * skin and not live["keep_all"], live
            assert 0.0 < live["displacement"] < live["margin"], live


def test_unfused_steps_between_runs(cuda):
    spec = _langevin()
    chunks = [11, 12, 10]
    _, want = _runs(spec, 0, -1.0, False, chunks, unfused=2)
    _, got = _runs(spec, 2, -1.0, False, chunks, unfused=2)
    for k in range(len(chunks)):
        assert np.array_equal(got[k], want[k]), f"after run {k}"


def test_pruning_arguments(cuda):
    p = marsmd.Patch(_langevin())
    assert p.get_pruning()["mode"] == 0  # off unless asked for
    with pytest.raises(marsmd.MarsError):
        p.set_pruning(4)
    with pytest.raises(marsmd.MarsError):
        p.set_pruning(2, float("inf"))
    p.set_pruning(1)
    assert p.get_pruning() == {"mode": 1, "margin": 0.0, "keep_all": False, "displacement": 0.0}
