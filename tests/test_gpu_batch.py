"""GPU replica batch (marsmd_batch): every replica stepped as the reference's Patch
would step it, checked per replica against the oracle's Patch + SimManager loop with
that replica's patch id.

* pair set of each replica's first build: bit for bit (oracle build_pairlist);
* 25 steps (step 1 build, reorders and forced rebuilds at 21): identical canonical
  relabelling (exported ids) and positions within the whole-run trajectory tolerance
  of the single-patch test;
* the batch equals the same replicas stepped one by one as GPU Patches, also when one
  replica is above the one-workgroup sort's 32768 particles (the rebuild then takes the
  global two-key radix sort);
* a replica whose reference cell grid misses pairs across the periodic seam (after the
  pair-buffer growth of its first build) keeps the reference's pair set;
* every launch configuration (workgroups per CU, hipGraph segment replay) gives the
  same bits as the default eager launch, checked after each of several runs with host
  work and synchronisation between the replays, for BD and Langevin; the autotuner
  steps the same trajectory and its cached choice is reused;
* particles outside their box are rejected; a coordinate that the wrap's rounding
  leaves an ulp below the box face is kept, with the reference's pair set;
* the walk's cell configuration per replica size bucket keeps every pair set (forces
  within 1e-5: only their FP32 summation order changes), and its autotuner picks one
  per bucket present, caches it, and a first build with it gives the oracle's pair sets.
"""
import numpy as np
import pytest

import marsmd
from marsmd import configs
from marsmd.batch import Batch
from oracle import pyoracle as orc
from tests.conftest import seam_pair_positions, sorted_pairs

pytestmark = pytest.mark.gpu


def _min_image(d, L):
    return np.abs(d - L * np.round(d / L))


def _specs():
    return configs.config3(replicas=6, n_min=300, n_max=4000, density=4e-4, seed=17)


def test_first_build_pair_sets(cuda):
    specs = _specs()
    b = Batch(specs)
    b.run(1)
    cuda.cuda.synchronize()
    for r, s in enumerate(specs):
        print(f"case: replica {r} n={s.n} L={s.box.size[0]:.1f}", flush=True)
        gp = sorted_pairs(b.get_neighbor_pairs(r))
        ref = orc.build_pairlist(s.pos4(), s.box, s.params.pairlist_cutoff)
        assert np.array_equal(gp, ref["pairs"].astype(np.int64)), f"replica {r}"


def test_steps_match_oracle_and_single_patches(cuda):
    specs = _specs()
    b = Batch(specs)
    steps = 25
    b.run(steps)
    cuda.cuda.synchronize()
    got = b.copy_particles_to_host()
    for r, s in enumerate(specs):
        L = s.box.size[0]
        sim = orc.Sim(s)
        sim.run(steps)
        o = sim.copy_particles_to_host()
        assert np.array_equal(got[r]["id"], o["id"]), f"replica {r}: canonical order differs"
        close = _min_image(got[r]["pos"][:, :3] - o["pos"][:, :3], L).max(axis=1) < 1e-2
        assert close.mean() > 0.999, f"replica {r}"
        p = marsmd.Patch(s)
        p.run(steps)
        cuda.cuda.synchronize()
        g = p.copy_particles_to_host()
        assert np.array_equal(got[r]["id"], g["id"])
        assert _min_image(got[r]["pos"][:, :3] - g["pos"][:, :3], L).max() < 1e-3, f"replica {r}"


def test_replica_above_the_workgroup_sort_matches_single_patches(cuda):
    specs = configs.config3(replicas=3, n_min=300, n_max=3000, density=4e-4, seed=17)
    big = configs.config3(replicas=1, n_min=34_000, n_max=36_000, density=4e-4, seed=17)[0]
    big.patch_id = len(specs)
    specs.append(big)
    assert big.n > 32768
    b = Batch(specs)
    steps = 22  # the reorder at step 21 sorts again
    b.run(steps)
    cuda.cuda.synchronize()
    got = b.copy_particles_to_host()
    for r, s in enumerate(specs):
        print(f"case: replica {r} n={s.n}", flush=True)
        p = marsmd.Patch(s)
        p.run(steps)
        cuda.cuda.synchronize()
        g = p.copy_particles_to_host()
        assert np.array_equal(got[r]["id"], g["id"]), f"replica {r}: canonical order differs"
        assert _min_image(got[r]["pos"][:, :3] - g["pos"][:, :3], s.box.size[0]).max() < 1e-3, f"replica {r}"


def test_incomplete_reference_grid_after_pair_buffer_growth(cuda):
    # first build: pair buffer = n = 20000 selects a single cell; the buffer then grows to
    # 1.25 x the pair count, and the rebuild of step 21 selects 16 cells of 10.0006 A with
    # r = 1, whose stencil misses the planted seam pairs. dt = 0 keeps the positions fixed.
    n, L, rc = 20000, 160.01, 10.0
    seam = configs.config0(n=n, L=L, seed=5)
    seam.positions = seam_pair_positions(n, L)[:, :3].copy()
    small = _specs()[0]
    seam.seed = small.seed  # one base seed per batch
    params = configs.RunParams(cutoff=rc, pairlist_cutoff=rc, rebuild_period=20, timestep=0.0, kT=configs.KT)
    b = Batch([seam, small])
    b.run(1, params=params)
    cuda.cuda.synchronize()
    first = orc.build_pairlist(seam.pos4(), seam.box, rc)
    assert first["grid"][0] == 0
    assert np.array_equal(sorted_pairs(b.get_neighbor_pairs(0)), first["pairs"].astype(np.int64))
    grown = len(first["pairs"]) + len(first["pairs"]) // 4
    b.run(24, first_step=1, params=params)
    cuda.cuda.synchronize()
    got = b.copy_particles_to_host()
    ref = orc.build_pairlist(got[0]["pos"], seam.box, rc, max_pairs=grown)
    assert ref["grid"][:2] == (4, (1, 1, 1))
    assert len(ref["pairs"]) + 120 <= len(first["pairs"])
    assert np.array_equal(sorted_pairs(b.get_neighbor_pairs(0)), ref["pairs"].astype(np.int64))


def test_wrap_rounding_at_the_box_face_is_kept(cuda):
    # L = (1474 / 1e-4)^(1/3) in FP32: PeriodicBox::wrap_scalar sends the float just below
    # L to a small negative coordinate (floor((x - o) inv_L) rounds to 1). The next build
    # keeps the particle (its Morton code clamped, as the reference's) and the pair set;
    # dt = 0 keeps the positions fixed.
    n = 1474
    L = float((n / 1e-4) ** (1.0 / 3.0))
    spec = configs.config0(n=n, L=L, seed=8)
    spec.positions = np.array(spec.positions, np.float32)
    spec.positions[0, 0] = np.nextafter(np.float32(L), np.float32(0.0))
    small = _specs()[0]
    spec.seed = small.seed
    rc = 30.0
    params = configs.RunParams(cutoff=rc, pairlist_cutoff=rc, rebuild_period=5, timestep=0.0, kT=configs.KT)
    b = Batch([spec, small])
    b.run(7, params=params)
    cuda.cuda.synchronize()
    got = b.copy_particles_to_host()
    x = got[0]["pos"][:, 0]
    assert x.min() < 0.0 and x.min() > -1e-3, x.min()
    ref = orc.build_pairlist(got[0]["pos"], spec.box, rc)
    assert np.array_equal(sorted_pairs(b.get_neighbor_pairs(0)), ref["pairs"].astype(np.int64))


def _state(b):
    """Every replica's positions, momenta, forces and ids as raw 32-bit words."""
    return [np.concatenate([r["pos"].view(np.uint32).ravel(), r["mom"].view(np.uint32).ravel(),
                            r["force"].view(np.uint32).ravel(), r["id"].view(np.uint32)])
            for r in b.copy_particles_to_host()]


@pytest.mark.parametrize("integrator", [configs.api.INTEGRATOR_BROWNIAN, configs.api.INTEGRATOR_LANGEVIN])
def test_launch_configurations_are_bit_identical(cuda, integrator):
    specs = _specs()[:4]
    params = configs.RunParams(cutoff=30.0, pairlist_cutoff=35.0, rebuild_period=10, reorder_period=20,
                               timestep=1e-4 if integrator == configs.api.INTEGRATOR_BROWNIAN else 2e-5,
                               kT=configs.KT, integrator=integrator, energy_period=7)
    chunks = [13, 9, 25]  # runs end inside segments; replays resume after host work
    ref = Batch(specs)
    ref_states, step = [], 0
    for c in chunks:
        ref.run(c, first_step=step, params=params)
        step += c
        cuda.cuda.synchronize()
        ref_states.append(_state(ref))
    for bpc, graph in [(1, False), (0, True), (2, True)]:
        print(f"case: integrator {integrator} blocks_per_cu {bpc} graph {graph}", flush=True)
        b = Batch(specs)
        b.set_launch(bpc, graph)
        assert b.get_launch() == {"blocks_per_cu": bpc, "use_graph": graph}
        step = 0
        for k, c in enumerate(chunks):
            b.run(c, first_step=step, params=params)
            step += c
            cuda.cuda.synchronize()
            got = _state(b)
            for r in range(len(specs)):
                assert np.array_equal(got[r], ref_states[k][r]), f"replica {r} after run {k}"


def test_autotune_keeps_the_trajectory_and_caches_its_choice(cuda, tmp_path):
    specs = _specs()[:3]
    params = specs[0].params
    ref = Batch(specs)
    ref.run(120, params=params)
    cuda.cuda.synchronize()
    want = _state(ref)
    cache = tmp_path / "autotune.json"
    b = Batch(specs)
    choice = b.autotune(120, params=params, cache=cache)
    cuda.cuda.synchronize()
    assert not choice["cached"] and choice["ms_per_step"] > 0.0
    assert b.get_launch() == {"blocks_per_cu": choice["blocks_per_cu"], "use_graph": choice["use_graph"]}
    got = _state(b)
    for r in range(len(specs)):
        assert np.array_equal(got[r], want[r]), f"replica {r}"
    b2 = Batch(specs)
    again = b2.autotune(120, params=params, cache=cache)
    cuda.cuda.synchronize()
    assert again["cached"] and again["blocks_per_cu"] == choice["blocks_per_cu"]
    got2 = _state(b2)
    for r in range(len(specs)):
        assert np.array_equal(got2[r], want[r]), f"replica {r}"


def test_particles_outside_their_box_are_rejected(cuda):
    specs = _specs()[:2]
    specs[1].positions = np.array(specs[1].positions, np.float32)
    specs[1].positions[3, 0] = specs[1].box.size[0] + 5.0
    with pytest.raises(marsmd.MarsError):
        Batch(specs)


def test_cell_configuration_per_size_bucket(cuda, tmp_path):
    # the walk's cells per size bucket change the order of neighbours in a row (so the
    # FP32 summation order) but not the pair set
    specs = _specs()
    ref = Batch(specs)
    ref.run(1)
    b = Batch(specs)
    buckets = b.get_cells()
    assert set(buckets) == {Batch.size_bucket(s.n) for s in specs}
    assert all(v == (24, 2) for v in buckets.values())
    b.set_cells({bk: (8, 3) for bk in buckets})
    assert all(v == (8, 3) for v in b.get_cells().values())
    b.run(1)
    cuda.cuda.synchronize()
    got, want = b.copy_particles_to_host(), ref.copy_particles_to_host()
    for r, s in enumerate(specs):
        print(f"case: cells (8, 3), replica {r} n={s.n}", flush=True)
        assert np.array_equal(sorted_pairs(b.get_neighbor_pairs(r)), sorted_pairs(ref.get_neighbor_pairs(r)))
        f, fr = got[r]["force"][:, :3].astype(np.float64), want[r]["force"][:, :3].astype(np.float64)
        assert np.abs(f - fr).max() <= 1e-5 * np.abs(fr).max(), f"replica {r}"
    with pytest.raises(marsmd.MarsError):
        b.set_cells({min(buckets): (0, 2)})
    # the autotuner: a choice per bucket present, cached per workload
    cache = tmp_path / "cells.json"
    c = Batch(specs)
    choice = c.autotune_cells(45, cache=cache)
    cuda.cuda.synchronize()
    print(f"case: tuned cells {choice}", flush=True)
    assert not choice["cached"] and choice["walk_ms"] > 0.0
    assert set(choice["cells"]) == set(buckets)
    assert c.get_cells() == choice["cells"]
    again = Batch(specs).autotune_cells(45, cache=cache)
    assert again["cached"] and again["cells"] == choice["cells"]
    # a first build with the chosen cells gives the reference's pair sets
    d = Batch(specs)
    d.set_cells(choice["cells"])
    d.run(1)
    cuda.cuda.synchronize()
    for r, s in enumerate(specs):
        ref_pairs = orc.build_pairlist(s.pos4(), s.box, s.params.pairlist_cutoff)["pairs"].astype(np.int64)
        assert np.array_equal(sorted_pairs(d.get_neighbor_pairs(r)), ref_pairs), f"replica {r}"
