"""The C++ interface (MARS::Patch, MARS::Pairlist, MARS::Exception in
mars-md_amd/host/MARS.hpp) driven from a compiled program the way the reference's
SimManager drives Patch (SimManager.cpp:403-425, 485), checked against the oracle:
pair set bit-exact (Patch pairlist and a stand-alone Pairlist), step-1 forces within
1e-5 of the FP64 sum, positions after 5 BD steps within 1e-3 A, and the reference's
exception categories for an unsupported builder and an out-of-range type."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from conftest import rel_force_error, sorted_pairs

DRIVER = Path(__file__).resolve().parent / "cpp" / "mirror_driver"
STEPS = 5


def _write_input(path: Path, spec):
    P = spec.params
    t = spec.nb_tables[0]
    n = spec.n
    with open(path, "wb") as f:
        np.array([n, len(t.pot), STEPS], np.int64).tofile(f)
        np.array([spec.seed], np.uint64).tofile(f)
        np.array([spec.box.size[0], P.cutoff, P.pairlist_cutoff, P.timestep, P.kT, spec.types[0].diffusion[0],
                  t.start, t.step_inv], np.float32).tofile(f)
        spec.pos4().astype(np.float32).tofile(f)
        np.asarray(spec.type_ids, np.int32).tofile(f)
        t.pot.astype(np.float32).tofile(f)


@pytest.mark.gpu
def test_cpp_interface_against_oracle(cuda, tmp_path):
    from marsmd import configs
    from oracle import pyoracle

    if not DRIVER.exists():
        pytest.fail(f"{DRIVER} is not built (the build step compiles it)")
    spec = configs.config0(n=3000, L=160.0, seed=11)
    P = spec.params
    inp = tmp_path / "in.bin"
    out = tmp_path / "out"
    _write_input(inp, spec)
    print(f"case: C++ driver, config0 n={spec.n} L=160 seed=11", flush=True)
    r = subprocess.run([str(DRIVER), str(inp), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    summary = r.stdout.split()
    assert "exceptions" in summary and summary[summary.index("exceptions") + 1] == "2", r.stdout

    sim = pyoracle.Sim(spec)
    sim.calculate_nonbonded_forces(P.pairlist_cutoff, P.cutoff, 1, P.rebuild_period)
    o = sim.copy_particles_to_host()
    op = sim.get_pairs()
    n = spec.n
    pairs = np.fromfile(str(out) + ".pairs", np.int32).reshape(-1, 2)
    assert np.array_equal(sorted_pairs(pairs), sorted_pairs(op)), "Patch pair set differs from the oracle"
    pairs_sa = np.fromfile(str(out) + ".pairs_standalone", np.int32).reshape(-1, 2)
    assert np.array_equal(sorted_pairs(pairs_sa), sorted_pairs(op)), "stand-alone pair set differs"
    force = np.fromfile(str(out) + ".force", np.float32).reshape(n, 4)
    err = rel_force_error(force, o["force64"])
    assert err <= 1e-5, f"force error {err}"

    sim.integrate_motion(P.timestep, P.kT, P.integrator, 1)
    for s in range(2, STEPS + 1):
        sim.calculate_nonbonded_forces(P.pairlist_cutoff, P.cutoff, s, P.rebuild_period)
        sim.integrate_motion(P.timestep, P.kT, P.integrator, s)
    o = sim.copy_particles_to_host()
    pos = np.fromfile(str(out) + ".pos", np.float32).reshape(n, 4)
    d = np.abs(pos[:, :3].astype(np.float64) - o["pos"][:, :3])
    d = np.minimum(d, 160.0 - d)
    assert d.max() < 1e-3, f"position difference {d.max()}"
