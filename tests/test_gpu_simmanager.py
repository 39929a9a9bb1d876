"""The SimManager loop (marsmd.simmanager) on the GPU Patch: energy output with
momenta settled first, and the native restart format with firstStep continuation.

* Restart continuation (Unit_test/System/RestartContinuation.cpp:125-175): 64 free BD
  beads, a run of 2 x 50 steps against a 50-step run, its .restart file, and a run
  resumed from that file with firstStep 50. The file round-trips positions bit for
  bit; the resumed run equals the continuous one bit for bit; resuming with firstStep
  0 replays segment 1's noise instead. The same for Langevin with the momentum restart
  file (the closing half-kick settled at the output step in both runs).
* Energy output (SimManager.cpp:1000-1075): energy.dat lines in the reference's text
  format at every energy step; the potential energy equals the oracle's from the same
  state (FP32 pair terms summed in FP64) and, for Langevin, the kinetic energy in kT
  equals its closed form from the exported momenta; a run's trajectory does not depend
  on its output periods.
"""
import re

import numpy as np
import pytest

import marsmd
from marsmd import configs
from marsmd import _capi as api
from marsmd.simmanager import SimManager, continued_run, float_to_chars, read_restart
from oracle import pyoracle as orc

pytestmark = pytest.mark.gpu

BOLTZMANN = 0.001987191


def _free_spec(integrator):
    """RestartContinuation.cpp:27-96: 64 beads on a 4x4x4 lattice of 10 A near the
    centre of a 400 A box, 300 K, dt = 20 fs, cutoff = pairlist cutoff = 10 A, no pair
    interactions, Ar mass 39.948 and D = 149 (BD) or gamma = 50/ns (Langevin)."""
    i = np.arange(64)
    pos = np.stack([180.0 + 10.0 * (i % 4), 180.0 + 10.0 * ((i // 4) % 4), 180.0 + 10.0 * (i // 16)], axis=1)
    kT = np.float32(BOLTZMANN) * np.float32(300.0)
    if integrator == api.INTEGRATOR_BROWNIAN:
        t = marsmd.ParticleType(mass=39.948, diffusion=(149.0,) * 3)
    else:
        g = 50.0
        t = marsmd.ParticleType(mass=39.948, diffusion=(float(kT) / (g * 39.948),) * 3, trans_damping=(g,) * 3)
    zero = marsmd.TabulatedPotential(np.zeros(64, np.float32), 0.0, 10.0 / 63)
    spec = marsmd.SystemSpec(
        box=marsmd.PeriodicBox((400.0, 400.0, 400.0)), types=[t], positions=pos.astype(np.float32),
        type_ids=np.zeros(64, np.int32),
        momenta=np.zeros((64, 3), np.float32) if integrator == api.INTEGRATOR_LANGEVIN else None,
        params=marsmd.RunParams(cutoff=10.0, pairlist_cutoff=10.0, rebuild_period=1000, timestep=2e-5,
                                kT=float(kT), integrator=integrator),
        nb_tables=[zero], seed=42, name="restart-continuation")
    return spec


def _by_id(sm):
    g = sm.gather()
    return g["pos"][:, :3].copy(), g["mom"][:, :3].copy()


@pytest.mark.parametrize("integrator", [api.INTEGRATOR_BROWNIAN, api.INTEGRATOR_LANGEVIN])
def test_restart_with_first_step_continues_the_noise_stream(cuda, tmp_path, integrator):
    print(f"case: restart continuation, integrator {integrator}", flush=True)
    spec = _free_spec(integrator)
    K = 50
    cont = SimManager(spec, tmp_path / "continuous", 2 * K, output_period=K, energy_output_period=K).run()
    pos_c, mom_c = _by_id(cont)
    seg1 = SimManager(spec, tmp_path / "seg1", K, output_period=K, energy_output_period=K).run()
    pos_1, mom_1 = _by_id(seg1)
    types, restart = read_restart(tmp_path / "seg1.restart")
    assert len(types) == 64
    assert np.array_equal(restart.view(np.uint32), pos_1.view(np.uint32)), "restart file does not round-trip"
    if integrator == api.INTEGRATOR_LANGEVIN:
        _, mr = read_restart(tmp_path / "seg1.0.momentum.restart")
        assert np.array_equal(mr.view(np.uint32), mom_1.view(np.uint32))
    seg2 = continued_run(spec, tmp_path / "seg2", tmp_path / "seg1", K, K, output_period=K,
                         energy_output_period=K).run()
    pos_2, mom_2 = _by_id(seg2)
    assert np.array_equal(pos_2.view(np.uint32), pos_c.view(np.uint32)), \
        f"resumed run differs from the continuous one by {np.abs(pos_2 - pos_c).max()}"
    if integrator == api.INTEGRATOR_LANGEVIN:
        assert np.array_equal(mom_2.view(np.uint32), mom_c.view(np.uint32))
    # firstStep 0 replays segment 1's kicks (RestartContinuation.cpp:160-175)
    replay = continued_run(spec, tmp_path / "replay", tmp_path / "seg1", 0, K, output_period=K,
                           energy_output_period=K).run()
    pos_r, _ = _by_id(replay)
    if integrator == api.INTEGRATOR_BROWNIAN:
        d1 = pos_1 - spec.positions
        d2 = pos_r - restart
        assert np.abs(d2 - d1).max() < 1e-3
    assert (np.linalg.norm(pos_r - pos_c, axis=1) < 1e-2).sum() == 0


def _energy_lines(path):
    text = path.read_text()
    pat = re.compile(r"Kinetic Energy: (\S+) \(kT\) \nPotential Energy: (\S+) \(kcal/mol\) \n")
    rows = pat.findall(text)
    assert "".join(f"Kinetic Energy: {a} (kT) \nPotential Energy: {b} (kcal/mol) \n" for a, b in rows) == text
    return [(float(a), float(b)) for a, b in rows]


def test_energy_output_bd_matches_the_oracle(cuda, tmp_path):
    spec = configs.config0(n=3000, L=160.0, seed=23)
    P = spec.params
    steps, period = 40, 20
    print(f"case: energy output, BD n={spec.n}, steps {steps}, period {period}", flush=True)
    sm = SimManager(spec, tmp_path / "bd", steps, energy_output_period=period).run()
    rows = _energy_lines(tmp_path / "bd.energy.dat")
    assert len(rows) == steps // period
    for k, (step, ke, pot) in enumerate(sm.energies):
        # BD has no pending kick: the energies are those of the step's own force pass,
        # from the positions before the step (a twin Patch steps to them bit for bit)
        twin = marsmd.Patch(spec)
        twin.run(step - 1, params=sm.params)
        cuda.cuda.synchronize()
        g = twin.copy_particles_to_host()
        sim = orc.Sim(spec)
        sim.set_state(g["pos"], g["mom"], g["force"], g["type"], g["id"])
        sim.calculate_nonbonded_forces(P.pairlist_cutoff, P.cutoff, step, P.rebuild_period, compute_energy=True)
        ref = float(np.sum(sim.copy_particles_to_host()["force64"][:, 3]))
        assert abs(pot - ref) <= 1e-5 * abs(ref) + 1e-6, (step, pot, ref)
        assert rows[k][1] == pytest.approx(pot, rel=1e-5, abs=1e-5)
        assert ke == 0.0 and rows[k][0] == 0.0
    # the output periods do not change the trajectory
    plain = marsmd.Patch(spec)
    plain.run(steps)
    cuda.cuda.synchronize()
    a = plain.copy_particles_to_host()
    b = sm.patch.copy_particles_to_host()
    assert np.array_equal(a["pos"].view(np.uint32), b["pos"].view(np.uint32))


def test_energy_output_langevin_grid_settles_momenta(cuda, tmp_path):
    spec = configs.config2(n=20_000, L=400.0, grid_n=64)
    steps, period = 24, 12
    print(f"case: energy output, Langevin n={spec.n}, grid 64^3, steps {steps}, period {period}", flush=True)
    sm = SimManager(spec, tmp_path / "lg", steps, energy_output_period=period).run()
    rows = _energy_lines(tmp_path / "lg.energy.dat")
    assert len(rows) == steps // period
    # the last output step's state: closing half-kick applied (nothing pending)
    assert not sm.patch.has_deferred_kick()
    g = sm.gather(need_energy=True)
    ke, pot = sm.energies_of(g)
    step, ke_out, pot_out = sm.energies[-1]
    assert step == steps and ke_out == ke and pot_out == pot
    # kinetic energy in kT from the momenta, in FP64
    m = np.array([t.mass for t in spec.types])[g["type"]]
    ke_ref = float(np.sum(0.5 * (g["mom"][:, :3].astype(np.float64) ** 2).sum(axis=1) / m)) / (
        2.046167337 ** 2 * float(np.float32(spec.params.kT)))
    assert ke == pytest.approx(ke_ref, rel=1e-6)
    # equipartition over 3N degrees of freedom, loosely (the momenta start Maxwellian)
    assert 0.9 * 1.5 * spec.n < ke < 1.1 * 1.5 * spec.n
    # the potential energy at the output step is the settled force pass's (pair terms at
    # the step's final positions)
    P = spec.params
    sim = orc.Sim(spec)
    h = sm.patch.copy_particles_to_host()
    sim.set_state(h["pos"], h["mom"], h["force"], h["type"], h["id"])
    sim.calculate_nonbonded_forces(P.pairlist_cutoff, P.cutoff, steps, P.rebuild_period, compute_energy=True)
    ref = float(np.sum(sim.copy_particles_to_host()["force64"][:, 3]))
    assert abs(pot - ref) <= 1e-5 * abs(ref) + 1e-6, (pot, ref)
    assert rows[-1][1] == pytest.approx(pot, rel=1e-5, abs=1e-5)
    # a momentum restart file was written next to the position one
    t, mr = read_restart(tmp_path / "lg.0.momentum.restart")
    assert len(t) == spec.n and np.array_equal(mr.view(np.uint32), g["mom"][:, :3].copy().view(np.uint32))
