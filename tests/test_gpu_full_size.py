"""GPU parity at the workloads' stated sizes, through the fused step kernel the
throughput runs use.

* 1M beads, 3 types, 2000 A box, 256^3 tricubic membrane grid, Langevin: step 1 (first
  build) at the default launch and step 21 (reorder, rebuild) at the launch the
  autotuner keeps after timing its candidates on steps 2..20. Each checked step starts
  the oracle from the GPU's state before it: pair set bit for bit, forces (pair + grid)
  within 1e-5 of the oracle's FP32 per-pair terms summed in FP64, positions within
  1e-4 A (2 ulp of the coordinate where that is coarser: FP32 coordinates above 1024 A
  are 1.2e-4 A apart) and momenta within 1e-4 of their largest after the integrator.
* 200 chains x 500 beads (bonds, angles, 1-2 and 1-3 exclusions), BD, reorder period 20:
  the same checks at steps 1 and 21.
* 512 replicas (N log-uniform in 500..20000, density 1e-4): every replica's first pair
  set bit for bit; the 8 smallest and 8 largest replicas checked at step 21 as above
  and over a 25-step run against the oracle's trajectory.
* Force matching on 64 BD snapshots x 100k beads (10 tables x 512 knots): forces within
  1e-5, the FP64 adjoint within 1e-4 of the oracle's FP64 gradient, central differences
  with h = 1e-3 on sampled knots, 200 Adam steps.

The oracle runs on the host's cores where its calls are independent (ctypes releases
the GIL), so each case takes seconds of CPU time.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import marsmd
from marsmd import configs
from marsmd import _capi as api
from marsmd.batch import Batch
from oracle import pyoracle as orc
from tests.conftest import canonical_topology, host_threads, min_image, rel_force_error, sorted_pairs

pytestmark = pytest.mark.gpu
FORCE_TOL = 1e-5


def _say(msg):
    print(msg, flush=True)


def _grid_force(spec, pos4, types):
    """Position-dependent force and energy (Pmf.h:68-124) of every particle, FP64 sums of
    the oracle's FP32 samples."""
    out = np.zeros((len(pos4), 4))
    P = spec.params
    for g_idx, grid in enumerate(spec.grids):
        s = orc.grid_sample(grid, pos4, P.interpolation)
        for t, pt in enumerate(spec.types):
            sel = types == t
            for gi, scale, _ in pt.pmf_terms:
                if gi != g_idx:
                    continue
                out[sel, :3] += -scale * s[sel, 1:4]
                out[sel, 3] += scale * s[sel, 0]
    return out


def _reference_force(spec, mid, pairs, with_grid):
    """Forces at the oracle's force-time state `mid`: pair terms (FP32 per pair, FP64
    sums) over `pairs`, bonded terms in the current canonical indices, and the grid
    force when the fused kernel folds it into the stored force (Langevin)."""
    P = spec.params
    f = orc.pair_forces(pairs, mid["pos"], mid["type"], spec.tag_matrix(), spec.nb_tables, spec.box,
                        P.cutoff * P.cutoff, mode=1, enabled_terms=spec.enabled_terms, particle_types=spec.types,
                        pair_params=spec.pair_params)
    topo = canonical_topology(spec, mid["id"])
    if topo["bonds"] is not None:
        f += orc.bonded_forces(mid["pos"], spec.box, topo["bonds"], spec.bond_table, spec.bond_tables,
                               topo["angles"], spec.angle_table, spec.angle_tables, mode=1)
    if with_grid and spec.grids:
        f += _grid_force(spec, mid["pos"], mid["type"])
    return f


def _oracle_step(spec, prev, step, threads):
    """The oracle's step `step` started from the GPU's state `prev` after step - 1 (None:
    the initial state): the maybe_reorder of SimManager, forces, integrator. Returns the
    force-time state, its pair set and the state after the integrator."""
    P = spec.params
    sim = orc.Sim(spec)
    sim.set_threads(threads)
    rp = P.reorder_period if P.reorder_period > 0 else P.rebuild_period
    if prev is not None:
        sim.set_state(prev["pos"], prev["mom"], prev["force"], prev["type"], prev["id"],
                      deferred_kick=P.integrator == api.INTEGRATOR_LANGEVIN)
        if rp > 0 and (step - 1) % rp == 0:
            sim.reorder_particles()
    sim.set_field(P.electric_field, P.interpolation)
    sim.calculate_nonbonded_forces(P.pairlist_cutoff, P.cutoff, step, P.rebuild_period, P.electric_field,
                                   P.interpolation, False)
    mid = sim.copy_particles_to_host()
    pairs = sim.get_pairs()
    sim.calculate_bonded_forces(False)
    sim.integrate_motion(P.timestep, P.kT, P.integrator, step)
    return mid, pairs, sim.copy_particles_to_host()


def _check_fused_step(spec, prev, got, got_pairs, step, threads, label):
    """The GPU's fused step `step` (state `got`, stored forces, pair set `got_pairs`)
    against the oracle's step from the same state `prev`."""
    P = spec.params
    L = spec.box.size[0]
    langevin = P.integrator == api.INTEGRATOR_LANGEVIN
    mid, pairs, out = _oracle_step(spec, prev, step, threads)
    assert np.array_equal(got["id"], out["id"]), f"{label}: canonical order differs at step {step}"
    assert np.array_equal(got_pairs, pairs.astype(np.int64)), f"{label}: pair set differs at step {step}"
    fref = _reference_force(spec, mid, pairs, with_grid=langevin)
    err = rel_force_error(got["force"], fref)
    _say(f"  {label} step {step}: {len(pairs)} pairs, force error {err:.2e}")
    assert err <= FORCE_TOL, f"{label}: force error {err} at step {step}"
    # 1e-4 A, or 2 ulp of the coordinate where that is coarser (FP32 positions above
    # 1024 A are spaced 1.2e-4 A apart)
    dx = min_image(got["pos"][:, :3] - out["pos"][:, :3], L)
    tol = np.maximum(1e-4, 2.0 * np.spacing(np.abs(out["pos"][:, :3]).astype(np.float32)).astype(np.float64))
    worst = (dx / tol).max()
    assert worst <= 1.0, f"{label}: positions differ by {dx.max()} ({worst:.2f} of the bar) at step {step}"
    if langevin:
        dp = np.abs(got["mom"][:, :3] - out["mom"][:, :3]).max() / np.abs(out["mom"][:, :3]).max()
        assert dp <= 1e-4, f"{label}: momenta differ by {dp} (relative) at step {step}"


def _patch_pairs(patch):
    return sorted_pairs(patch.get_pairlist().get_neighbor_pairs().cpu().numpy())


def _run_and_check_patch(cuda, spec, label, tune):
    threads = host_threads()
    patch = marsmd.Patch(spec)
    # step 1: the first build, at the default launch (as many workgroups as fit)
    patch.run(1)
    cuda.cuda.synchronize()
    _check_fused_step(spec, None, patch.copy_particles_to_host(), _patch_pairs(patch), 1, threads, label)
    # steps 2..20: the autotuner's real steps (or plain steps), then the launch it keeps
    if tune:
        choice = patch.autotune(19, first_step=1)
        _say(f"  {label} launch kept: {choice}")
    else:
        patch.run(19, first_step=1)
    cuda.cuda.synchronize()
    prev = patch.copy_particles_to_host()
    # step 21: reorder, forced rebuild
    patch.run(1, first_step=20)
    cuda.cuda.synchronize()
    _check_fused_step(spec, prev, patch.copy_particles_to_host(), _patch_pairs(patch), 21, threads, label)


def test_langevin_1M_beads_grid256_fused(cuda):
    _say("case: 1,000,000 beads, 3 types, L 2000, 256^3 tricubic grid, Langevin")
    spec = configs.config2()
    _run_and_check_patch(cuda, spec, "1M", tune=True)


def test_polymers_200x500_fused(cuda):
    _say("case: 200 chains x 500 beads, bonds + angles + exclusions, BD, reorder period 20")
    spec = configs.config1()
    assert spec.n == 100_000 and spec.params.reorder_period == 20
    _run_and_check_patch(cuda, spec, "polymers", tune=False)


def test_batch_512_replicas(cuda):
    specs = configs.config3()
    ns = np.array([s.n for s in specs])
    _say(f"case: {len(specs)} replicas, {ns.sum()} beads, {ns.min()}..{ns.max()} per replica")
    threads = host_threads()
    b = Batch(specs)
    b.run(1)
    cuda.cuda.synchronize()
    got_pairs = [sorted_pairs(b.get_neighbor_pairs(r)) for r in range(len(specs))]

    def build(r):
        s = specs[r]
        return orc.build_pairlist(s.pos4(), s.box, s.params.pairlist_cutoff)["pairs"].astype(np.int64)

    with ThreadPoolExecutor(threads) as ex:
        refs = list(ex.map(build, range(len(specs))))
    for r in range(len(specs)):
        assert np.array_equal(got_pairs[r], refs[r]), f"replica {r} (n={specs[r].n}): first pair set differs"
    _say(f"  first builds of all {len(specs)} replicas equal")
    order = np.argsort(ns, kind="stable")
    picked = [int(r) for r in list(order[:8]) + list(order[-8:])]
    b.run(19, first_step=1)
    cuda.cuda.synchronize()
    prev = b.copy_particles_to_host()
    b.run(1, first_step=20)
    cuda.cuda.synchronize()
    got = b.copy_particles_to_host()
    for r in picked:
        _check_fused_step(specs[r], prev[r], got[r], sorted_pairs(b.get_neighbor_pairs(r)), 21, threads,
                          f"replica {r} (n={specs[r].n})")
    b.run(4, first_step=21)
    cuda.cuda.synchronize()
    got = b.copy_particles_to_host()

    def trajectory(r):
        sim = orc.Sim(specs[r])
        sim.run(25)
        return sim.copy_particles_to_host()

    with ThreadPoolExecutor(threads) as ex:
        outs = list(ex.map(trajectory, picked))
    for r, o in zip(picked, outs):
        L = specs[r].box.size[0]
        assert np.array_equal(got[r]["id"], o["id"]), f"replica {r}: canonical order differs after 25 steps"
        close = min_image(got[r]["pos"][:, :3] - o["pos"][:, :3], L).max(axis=1) < 1e-2
        assert close.mean() > 0.999, f"replica {r}: trajectory left the oracle's"


def test_force_matching_64x100k(cuda):
    torch = cuda
    from marsmd.fm import ForceMatching, force_matching_loss

    n, L, S = 100_000, 1000.0, 64
    _say(f"case: force matching, {S} BD snapshots x {n} beads, 10 tables x 512 knots")
    frames, type_ids = configs.bd_snapshots(n, L, S)
    w = configs.config4(n=n, L=L, snapshots=S, frames=frames)
    assert np.array_equal(w.type_ids, type_ids)
    fm = ForceMatching(w.positions, w.type_ids, w.box, w.tables, w.term_matrix, w.cutoff)
    k_start = np.concatenate([t.pot.astype(np.float64) for t in w.tables])
    k_target = np.concatenate([t.pot.astype(np.float64) for t in w.target_tables])
    threads = host_threads()

    def snap(s, knots, target=None):
        return orc.fm_loss_grad(w.positions[s:s + 1], w.type_ids, w.term_matrix, w.tables, knots, w.box, w.cutoff,
                                target=None if target is None else target[s:s + 1])

    with ThreadPoolExecutor(threads) as ex:
        tgt = np.concatenate([r[0] for r in ex.map(lambda s: snap(s, k_target), range(S))])
    with ThreadPoolExecutor(threads) as ex:
        res = list(ex.map(lambda s: snap(s, k_start, tgt), range(S)))
    f_ref = np.concatenate([r[0] for r in res])
    g_ref = np.mean([r[2] for r in res], axis=0)  # each snapshot's gradient is normalised by n x 3
    knots = fm.knots()
    f = fm.forces(knots).cpu().numpy()[..., :3]
    ferr = np.abs(f - f_ref).max() / np.abs(f_ref).max()
    _say(f"  forces: error {ferr:.2e} over {fm.num_pairs()} pairs")
    assert ferr <= FORCE_TOL
    G = 2.0 * (f.astype(np.float64) - tgt) / f.size
    g = fm.adjoint(knots, torch.from_numpy(G.astype(np.float32)).cuda()).cpu().numpy()
    gerr = np.abs(g - g_ref).max() / np.abs(g_ref).max()
    _say(f"  adjoint: error {gerr:.2e} against the FP64 gradient")
    assert gerr <= 1e-4
    # central differences of the loss with h = 1e-3 on sampled knots; the loss difference
    # is formed from the force difference, sum (f+ - f-)(f+ + f- - 2 t) / count, so the
    # two losses' common part does not cancel in FP64
    target = torch.from_numpy(tgt.astype(np.float32)).cuda()
    k = knots.clone().requires_grad_(True)
    loss = force_matching_loss(k, fm, target)
    loss.backward()
    grad = k.grad.double()
    used = torch.nonzero(grad).flatten().cpu().numpy()
    rs = np.random.default_rng(3)
    h = 1e-3
    gmax = grad.abs().max().item()
    worst = 0.0
    for kk in rs.choice(used, 16, replace=False):
        kp = knots.clone()
        km = knots.clone()
        kp[kk] += h
        km[kk] -= h
        fp = fm.forces(kp)[..., :3].double()
        fmi = fm.forces(km)[..., :3].double()
        t64 = target.double()
        fd = ((fp - fmi) * (fp + fmi - 2.0 * t64)).sum().item() / (f.size * 2.0 * h)
        worst = max(worst, abs(fd - grad[kk].item()) / gmax)
        assert abs(fd - grad[kk].item()) <= 1e-3 * gmax + 1e-3 * abs(grad[kk].item()), (kk, fd, grad[kk].item())
    _say(f"  central differences (h = 1e-3, 16 knots): worst error {worst:.2e} of the largest gradient")
    opt = torch.optim.Adam([k], lr=0.02)
    first = None
    for _ in range(200):
        opt.zero_grad()
        loss = force_matching_loss(k, fm, target)
        loss.backward()
        opt.step()
        first = loss.item() if first is None else first
    _say(f"  Adam, 200 steps: loss {first:.4e} -> {loss.item():.4e}")
    assert np.isfinite(loss.item()) and loss.item() < 0.5 * first
