"""GPU parity of the dihedral and harmonic-restraint kernels (launch_tabulated_dihedrals,
launch_harmonic_restraints; BondComputer.h:84-131, 281-346) through the C ABI, and of
a Patch whose bonded pass includes them (Patch.cpp:208-236).

Dihedrals share particles along chains, so the device sums terms with atomics in no
fixed order: forces within 1e-5 of the oracle's FP32 terms summed in FP64. Each
particle carries at most one restraint, so restraint forces are the oracle's FP32
values bit for bit (0 + f is exact).
"""
import numpy as np
import pytest

import marsmd
from marsmd import PeriodicBox, configs
from oracle import pyoracle as orc
from tests.conftest import rel_force_error

pytestmark = pytest.mark.gpu


def _chains(n_chains=300, beads=30, L=120.0, seed=4):
    rng = np.random.default_rng(seed)
    pos = np.zeros((n_chains, beads, 3))
    pos[:, 0] = rng.uniform(0, L, (n_chains, 3))
    for b in range(1, beads):
        d = rng.standard_normal((n_chains, 3))
        pos[:, b] = pos[:, b - 1] + 3.8 * d / np.linalg.norm(d, axis=1, keepdims=True)
    p4 = np.zeros((n_chains * beads, 4), np.float32)
    p4[:, :3] = np.mod(pos.reshape(-1, 3), L)
    idx = np.arange(n_chains * beads).reshape(n_chains, beads)
    dih = np.stack([idx[:, :-3].ravel(), idx[:, 1:-2].ravel(), idx[:, 2:-1].ravel(), idx[:, 3:].ravel()], axis=1)
    return p4, dih.astype(np.int32)


def test_dihedral_launcher_matches_oracle(cuda):
    L = 120.0
    box = PeriodicBox((L, L, L))
    pos, dih = _chains(L=L)
    tabs = [configs.cosine_dihedral_table(k=1.2, n=3, delta=0.3), configs.cosine_dihedral_table(k=0.7, n=1)]
    tidx = (np.arange(len(dih)) % 2).astype(np.int32)
    dt = marsmd.DeviceTables(tabs)
    d_pos = cuda.tensor(pos, device="cuda")
    d_dih = cuda.tensor(dih, device="cuda")
    d_tidx = cuda.tensor(tidx, device="cuda")
    d_fe = cuda.zeros((len(pos), 4), dtype=cuda.float32, device="cuda")
    print(f"case: dihedral launcher, {len(dih)} dihedrals", flush=True)
    marsmd.launch_tabulated_dihedrals(d_dih, d_pos, d_fe, dt, d_tidx, None, box, True, len(dih))
    cuda.cuda.synchronize()
    ref, _ = orc.dihedral_restraint_forces(pos, box, dih, tidx, tabs, mode=1)
    got = d_fe.cpu().numpy()
    assert rel_force_error(got, ref) <= 1e-5
    assert np.abs(got[:, 3] - ref[:, 3]).max() <= 1e-5 * np.abs(ref[:, 3]).max()
    # one isolated dihedral per launch record: the four FP32 terms bit for bit
    lone = dih[::27][:40]
    d_lone = cuda.tensor(lone, device="cuda")
    d_fe.zero_()
    marsmd.launch_tabulated_dihedrals(d_lone, d_pos, d_fe, dt, None, None, box, False, len(lone))
    cuda.cuda.synchronize()
    ref1, _ = orc.dihedral_restraint_forces(pos, box, lone, None, tabs, mode=1)
    assert np.array_equal(d_fe.cpu().numpy()[:, :3].astype(np.float64), ref1[:, :3])


def test_restraint_launcher_bit_exact(cuda):
    L = 80.0
    box = PeriodicBox((L, L, L))
    rng = np.random.default_rng(8)
    n = 20000
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform(0, L, (n, 3))
    ids = rng.permutation(n)[: n // 2].astype(np.int32)
    ids[::97] = -1
    anchors = np.zeros((len(ids), 4), np.float32)
    anchors[:, :3] = np.mod(pos[np.maximum(ids, 0), :3] + rng.normal(0, 4.0, (len(ids), 3)), L)
    k = rng.uniform(0.1, 3.0, len(ids)).astype(np.float32)
    d_fe = cuda.zeros((n, 4), dtype=cuda.float32, device="cuda")
    print(f"case: restraint launcher, {len(ids)} restraints", flush=True)
    marsmd.launch_harmonic_restraints(cuda.tensor(ids, device="cuda"), cuda.tensor(pos, device="cuda"), d_fe,
                                      cuda.tensor(anchors, device="cuda"), cuda.tensor(k, device="cuda"), box, True,
                                      len(ids))
    cuda.cuda.synchronize()
    ref, _ = orc.dihedral_restraint_forces(pos, box, restraint_ids=ids, restraint_anchors=anchors[:, :3],
                                           restraint_k=k, mode=1)
    assert np.array_equal(d_fe.cpu().numpy().astype(np.float64), ref)


def test_patch_steps_with_dihedrals_and_restraints(cuda):
    from tests.test_gpu_forces_integrators import _check_step

    spec = configs.config1(chains=20, beads=200, L=300.0, dihedrals=True, restraints=True)
    patch = marsmd.Patch(spec)
    print(f"case: Patch steps, {len(spec.dihedrals)} dihedrals, {len(spec.restraint_ids)} restraints", flush=True)
    for step in range(1, 24):  # reorder at 21 remaps dihedrals and restraint ids
        _check_step(cuda, patch, spec, step)
