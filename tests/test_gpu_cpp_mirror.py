"""The C++ interface (mars-md_amd/host/MARS.hpp) driven from a compiled program with the
reference's Patch constructor, containers and SimManager call sequence unchanged
(SimManager.cpp:343-349, 401-425, 485; tests/cpp/mirror_driver.cpp), on bead-spring
polymers (2 types, 3 pair tables, bond and angle tables with the angle abscissa in
degrees, 1-2 and 1-3 exclusions, reorder period 20), checked against the oracle and
the Python front end:

* step-1 pair set bit for bit (the Patch's pairlist and a stand-alone Pairlist);
* step-1 forces (pair + bonded) within 1e-5 of the oracle's FP32 terms summed in FP64;
* positions after 25 steps (a reorder and rebuild at step 21) bit for bit equal to the
  Python Patch stepped through the same calls, compared by global id;
* ZOrderSort displacement tracking; the reference's exception categories for an
  unsupported builder, an out-of-range type, a host range past the data, and containers
  swapped after the first force call."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from conftest import rel_force_error, sorted_pairs

DRIVER = Path(__file__).resolve().parent / "cpp" / "mirror_driver"
STEPS = 25


def _write_input(path: Path, spec, temperature_K):
    P = spec.params
    nb = spec.nb_tables
    tm = spec.tag_matrix()
    T = len(spec.types)
    entries = []
    for a in range(T):
        for b in range(a, T):
            entries.append((a, b, int(tm[a, b]) >> 8))
    bt, at = spec.bond_tables[0], spec.angle_tables[0]
    bx = (np.arange(len(bt.pot)) * np.float32(bt.step)).astype(np.float32) + np.float32(bt.start)
    # the angle table's abscissa in degrees, as a table file holds it
    deg = np.float32(180.0 / (len(at.pot) - 1))
    ax = (np.arange(len(at.pot)) * deg).astype(np.float32)
    with open(path, "wb") as f:
        np.array([spec.n, T, len(entries), len(nb[0].pot), len(nb), len(spec.bonds), len(spec.angles),
                  len(spec.exclusions), len(bt.pot), len(at.pot), STEPS, P.rebuild_period, P.reorder_period, 10],
                 np.int64).tofile(f)
        np.array([spec.seed], np.uint64).tofile(f)
        np.array([spec.box.size[0], P.cutoff, P.pairlist_cutoff, P.timestep, temperature_K, nb[0].step],
                 np.float32).tofile(f)
        np.array([[t.mass, t.charge, t.diffusion[0]] for t in spec.types], np.float32).tofile(f)
        np.array(entries, np.int32).tofile(f)
        np.concatenate([t.pot for t in nb]).astype(np.float32).tofile(f)
        bx.tofile(f)
        bt.pot.astype(np.float32).tofile(f)
        ax.tofile(f)
        at.pot.astype(np.float32).tofile(f)
        spec.pos4().astype(np.float32).tofile(f)
        np.asarray(spec.type_ids, np.int32).tofile(f)
        np.asarray(spec.bonds, np.int32).tofile(f)
        np.asarray(spec.angles, np.int32).tofile(f)
        np.asarray(spec.exclusions, np.int32).tofile(f)


@pytest.mark.gpu
def test_cpp_interface_replays_simmanager_against_oracle(cuda, tmp_path):
    import marsmd
    from marsmd import configs
    from oracle import pyoracle

    if not DRIVER.exists():
        pytest.fail(f"{DRIVER} is not built (the build step compiles it)")
    spec = configs.config1(chains=20, beads=150, L=300.0)
    P = spec.params
    # the temperature crosses as Kelvin (Temperature(T).kT = T * BOLTZMANN in FP32)
    temperature_K = np.float32(P.kT / 0.001987191)
    spec.params.kT = float(np.float32(temperature_K * np.float32(0.001987191)))
    n = spec.n
    inp = tmp_path / "in.bin"
    out = tmp_path / "out"
    _write_input(inp, spec, temperature_K)
    print(f"case: C++ driver, polymers n={n} L=300, {STEPS} steps", flush=True)
    r = subprocess.run([str(DRIVER), str(inp), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    words = r.stdout.split()
    assert words[words.index("exceptions") + 1] == "4", r.stdout
    assert words[words.index("energy_steps") + 1] == str(STEPS // 10), r.stdout
    assert words[words.index("needs_update") + 1] == "0", r.stdout
    assert float(words[words.index("max_displacement") + 1]) == 0.0, r.stdout

    # step 1 against the oracle
    sim = pyoracle.Sim(spec)
    sim.calculate_nonbonded_forces(P.pairlist_cutoff, P.cutoff, 1, P.rebuild_period)
    op = sim.get_pairs()
    pairs = np.fromfile(str(out) + ".pairs", np.int32).reshape(-1, 2)
    assert np.array_equal(sorted_pairs(pairs), sorted_pairs(op)), "Patch pair set differs from the oracle"
    standalone = pyoracle.build_pairlist(spec.pos4(), spec.box, P.pairlist_cutoff)["pairs"]
    pairs_sa = np.fromfile(str(out) + ".pairs_standalone", np.int32).reshape(-1, 2)
    assert np.array_equal(sorted_pairs(pairs_sa), sorted_pairs(standalone)), "stand-alone pair set differs"
    force = np.fromfile(str(out) + ".force", np.float32).reshape(n, 4)
    fref = pyoracle.pair_forces(op, spec.pos4(), spec.type_ids, spec.tag_matrix(), spec.nb_tables, spec.box,
                                P.cutoff * P.cutoff, mode=1)
    fref += pyoracle.bonded_forces(spec.pos4(), spec.box, spec.bonds, spec.bond_table, spec.bond_tables, spec.angles,
                                   spec.angle_table, spec.angle_tables, mode=1)
    err = rel_force_error(force, fref)
    print(f"  step-1 force error {err:.2e}", flush=True)
    assert err <= 1e-5, f"force error {err}"

    # the whole run against the Python front end issuing the same calls
    patch = marsmd.Patch(spec)
    run = marsmd.RunParams(**{**P.__dict__, "energy_period": 10})
    for s in range(1, STEPS + 1):
        patch.step_unfused(s, run)
    cuda.cuda.synchronize()
    h = patch.copy_particles_to_host()
    want = np.zeros((n, 4), np.float32)
    want[h["id"]] = h["pos"]
    pos = np.fromfile(str(out) + ".pos", np.float32).reshape(n, 4)
    assert np.array_equal(pos[:, :3].view(np.uint32), want[:, :3].view(np.uint32)), \
        f"positions differ by up to {np.abs(pos[:, :3] - want[:, :3]).max()}"
