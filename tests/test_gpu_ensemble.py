"""Ensemble statistics of the HIP path, each asserted within its computed statistical
error.

* Free Brownian diffusion, MSD = 6 D t: the reference's own case (BDKernels.cpp:92-220:
  100 beads x 5 seeds, D = 149, dt = 20 fs, 10,000 steps, 10% bar) through launch_BD,
  and a 40,960-sample case whose bar is 4 standard errors of a chi-square(3) mean.
* BAOAB equipartition <p_i^2> = kT m SQRT_CAL_TO_JOULE^2 (StochasticIntegrators.cpp:
  157-172: 4,096 walkers from rest, gamma = 50/ns, 8,000 steps, 8% bar) and its
  decorrelation across adjacent seeds (|r| < 0.15), through launch_BAOAB.
* The same two laws through the Patch's fused step kernel (rebuilds and reorders every
  20 steps, displacements tracked by global id) on non-interacting beads.
* The 10,000-bead WCA+DH system: its radial distribution function and diffusion
  coefficient from GPU Patches (32 seeds) and from the oracle (4 other seeds) agree
  within 5 standard errors, the per-seed spread taken from the GPU's seeds.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import marsmd
from marsmd import PeriodicBox, configs
from marsmd import _capi as api
from oracle import pyoracle as orc
from tests.conftest import host_threads

pytestmark = pytest.mark.gpu

BOLTZMANN = 0.001987191  # kcal/mol/K (Constants.h:25)
SQRT_CAL_TO_JOULE = 2.046167337


def _say(msg):
    print(msg, flush=True)


def _dev(cuda, a):
    return cuda.from_numpy(np.ascontiguousarray(a)).cuda()


def _wrap_disp(d, L):
    return d - L * np.round(d / L)


def _bd_walk(cuda, n, seed, steps, D, dt, kT, L, mass=39.948):
    types = marsmd.make_types([marsmd.ParticleType(mass=mass, diffusion=(D, D, D))])
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = L / 2
    pd = _dev(cuda, pos)
    fe = cuda.zeros((n, 4), dtype=cuda.float32, device="cuda")
    view = marsmd.particle_view(pos=pd, force_energy=fe, type_id=_dev(cuda, np.zeros(n, np.int32)))
    box = PeriodicBox((L, L, L))
    for step in range(steps):
        # BDKernels.cpp:150-164: current step and base counter both the step
        marsmd.launch_BD(view, types, dt, step, kT, n, box, seed, step, interpolation=api.INTERP_CUBIC)
    cuda.cuda.synchronize()
    return _wrap_disp(pd.cpu().numpy()[:, :3].astype(np.float64) - L / 2, L)


def test_bd_free_diffusion_msd_reference_case(cuda):
    D, dt, steps, kT = 149.0, 2e-5, 10000, BOLTZMANN * 300.0
    disp = np.concatenate([_bd_walk(cuda, 100, s, steps, D, dt, kT, 100.0) for s in (203, 123414, 1240, 342, 603)])
    msd = (disp ** 2).sum(axis=1).mean()
    want = 6.0 * D * dt * steps
    _say(f"  MSD {msd:.3f} A^2, 6Dt {want:.3f} (ratio {msd / want:.4f}), 500 samples")
    assert msd == pytest.approx(want, rel=0.10)


def test_bd_free_diffusion_msd_statistical(cuda):
    D, dt, steps, kT, n = 149.0, 2e-5, 2000, BOLTZMANN * 300.0, 8192
    disp = np.concatenate([_bd_walk(cuda, n, 7000 + s, steps, D, dt, kT, 1000.0) for s in range(5)])
    t = dt * steps
    r2 = (disp ** 2).sum(axis=1)
    want = 6.0 * D * t
    # |disp|^2 / (2 D t) ~ chi-square(3): relative standard error sqrt(2/3) / sqrt(N)
    se = want * np.sqrt(2.0 / 3.0) / np.sqrt(len(r2))
    _say(f"  MSD {r2.mean():.4f}, 6Dt {want:.4f}, standard error {se:.4f} ({len(r2)} samples)")
    assert abs(r2.mean() - want) <= 4.0 * se
    # each axis: variance 2 D t, no drift
    for a in range(3):
        var_se = 2.0 * D * t * np.sqrt(2.0 / len(r2))
        assert abs((disp[:, a] ** 2).mean() - 2.0 * D * t) <= 4.0 * var_se, a
        assert abs(disp[:, a].mean()) <= 4.0 * np.sqrt(2.0 * D * t / len(r2)), a


def _baoab_walk(cuda, n, seed, steps, gamma=50.0, mass=39.948, dt=2e-5, kT=BOLTZMANN * 300.0):
    D = kT / (gamma * mass)
    types = marsmd.make_types([marsmd.ParticleType(mass=mass, diffusion=(D, D, D),
                                                   trans_damping=(gamma, gamma, gamma))])
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = 50.0
    pd, md = _dev(cuda, pos), cuda.zeros((n, 4), dtype=cuda.float32, device="cuda")
    fe = cuda.zeros((n, 4), dtype=cuda.float32, device="cuda")
    view = marsmd.particle_view(pos=pd, force_energy=fe, type_id=_dev(cuda, np.zeros(n, np.int32)), mom=md)
    box = PeriodicBox((100.0, 100.0, 100.0))
    for s in range(steps):
        marsmd.launch_BAOAB(view, types, box, dt, s, kT, n, seed, 0)
    cuda.cuda.synchronize()
    return md.cpu().numpy()[:, :3].astype(np.float64)


def test_baoab_equipartition_reference_case(cuda):
    mom = _baoab_walk(cuda, 4096, 20260816, 8000)
    want = BOLTZMANN * 300.0 * 39.948 * SQRT_CAL_TO_JOULE ** 2
    for a in range(3):
        got = (mom[:, a] ** 2).mean()
        _say(f"  axis {a}: <p^2> {got:.4f}, kT m U^2 {want:.4f} (ratio {got / want:.4f})")
        assert got == pytest.approx(want, rel=0.08)


def test_baoab_equipartition_statistical_and_decorrelated(cuda):
    n = 65536
    a = _baoab_walk(cuda, n, 20260816, 8000)
    b = _baoab_walk(cuda, n, 20260817, 2000)
    want = BOLTZMANN * 300.0 * 39.948 * SQRT_CAL_TO_JOULE ** 2
    p2 = (a ** 2).mean()
    se = want * np.sqrt(2.0 / a.size)  # a Gaussian component's squared mean
    _say(f"  <p^2> {p2:.4f}, want {want:.4f}, standard error {se:.4f} ({a.size} samples)")
    assert abs(p2 - want) <= 4.0 * se
    for ax in range(3):
        r = np.corrcoef(a[:, ax], b[:, ax])[0, 1]
        assert abs(r) < 0.15, (ax, r)


def _free_spec(n, L, seed, integrator):
    """Non-interacting beads (an all-zero pair table) in a periodic cube."""
    zero = marsmd.TabulatedPotential(np.zeros(1024, np.float32), 0.0, 30.0 / 1023)
    if integrator == api.INTEGRATOR_BROWNIAN:
        t = marsmd.ParticleType(mass=1.0, diffusion=(150.0, 150.0, 150.0))
    else:
        t = marsmd.ParticleType(mass=100.0, diffusion=(configs.KT / (100.0 * 1000.0),) * 3,
                                trans_damping=(1000.0,) * 3)
    spec = configs.config0(n=n, L=L, seed=seed)
    spec.types = [t]
    spec.nb_tables = [zero]
    spec.params.integrator = integrator
    if integrator == api.INTEGRATOR_LANGEVIN:
        spec.momenta = np.zeros((n, 3), np.float32)
    return spec


def _track(patch, steps, every, L):
    """Run `steps` steps, unwrapping each bead's displacement (by global id) from
    positions every `every` steps."""
    h = patch.copy_particles_to_host()
    n = len(h["id"])
    last = np.empty((n, 3))
    last[h["id"]] = h["pos"][:, :3]
    total = np.zeros((n, 3))
    for s in range(0, steps, every):
        patch.run(every, first_step=s)
        h = patch.copy_particles_to_host()
        cur = np.empty((n, 3))
        cur[h["id"]] = h["pos"][:, :3]
        total += _wrap_disp(cur - last, L)
        last = cur
    return total, h


def test_patch_fused_free_diffusion_and_equipartition(cuda):
    n, L, steps = 20000, 600.0, 200
    spec = _free_spec(n, L, 31, api.INTEGRATOR_BROWNIAN)
    disp, _ = _track(marsmd.Patch(spec), steps, 20, L)
    t = steps * spec.params.timestep
    r2 = (disp ** 2).sum(axis=1)
    want = 6.0 * 150.0 * t
    se = want * np.sqrt(2.0 / 3.0) / np.sqrt(n)
    _say(f"  Patch BD: MSD {r2.mean():.4f}, 6Dt {want:.4f}, standard error {se:.4f}")
    assert abs(r2.mean() - want) <= 4.0 * se
    spec = _free_spec(n, L, 32, api.INTEGRATOR_LANGEVIN)
    _, h = _track(marsmd.Patch(spec), steps, 20, L)
    mom = h["mom"][:, :3].astype(np.float64)
    want = configs.KT * 100.0 * SQRT_CAL_TO_JOULE ** 2
    se = want * np.sqrt(2.0 / mom.size)
    _say(f"  Patch Langevin: <p^2> {(mom ** 2).mean():.4f}, want {want:.4f}, standard error {se:.4f}")
    assert abs((mom ** 2).mean() - want) <= 4.0 * se


# --------------------------------------------------------------- interacting system
R_MAX, BIN = 20.0, 0.5
EDGES = np.arange(2.0, R_MAX + 1e-9, BIN)


def _rdf(pos, L):
    from scipy.spatial import cKDTree

    p = np.mod(np.asarray(pos, np.float64)[:, :3], L)
    tree = cKDTree(p, boxsize=L)
    d = tree.query_pairs(R_MAX, output_type="ndarray")
    r = p[d[:, 0]] - p[d[:, 1]]
    r = np.sqrt((_wrap_disp(r, L) ** 2).sum(axis=1))
    h, _ = np.histogram(r, EDGES)
    n = len(p)
    shell = 4.0 / 3.0 * np.pi * (EDGES[1:] ** 3 - EDGES[:-1] ** 3)
    return h / (0.5 * n * (n - 1) / L ** 3 * shell)


FRAMES = (200, 250, 300, 350, 400)


def _observables(frames, L):
    """(g(r) averaged over the frames, D from the MSD between the first and last frame)."""
    g = np.mean([_rdf(f, L) for f in frames.values()], axis=0)
    keys = sorted(frames)
    disp = np.zeros((len(frames[keys[0]]), 3))
    for a, b in zip(keys[:-1], keys[1:]):
        disp += _wrap_disp(frames[b] - frames[a], L)
    t = (keys[-1] - keys[0]) * 1e-4
    return g, (disp ** 2).sum(axis=1).mean() / (6.0 * t)


def _by_id(h):
    out = np.empty((len(h["id"]), 3))
    out[h["id"]] = h["pos"][:, :3]
    return out


def _gpu_seed(seed):
    spec = configs.config0(seed=seed)
    patch = marsmd.Patch(spec)
    frames, step = {}, 0
    for f in FRAMES:
        patch.run(f - step, first_step=step)
        step = f
        frames[f] = _by_id(patch.copy_particles_to_host())
    return _observables(frames, spec.box.size[0])


def _oracle_seed(seed, threads):
    spec = configs.config0(seed=seed)
    sim = orc.Sim(spec)
    sim.set_threads(threads)
    frames, step = {}, 0
    for f in FRAMES:
        sim.run(f - step, first_step=step)
        step = f
        frames[f] = _by_id(sim.copy_particles_to_host())
    return _observables(frames, spec.box.size[0])


def test_config0_rdf_and_diffusion_match_the_oracle(cuda):
    _say("case: 10,000 beads, L 250, WCA+DH, BD; 32 GPU seeds against 4 oracle seeds, 400 steps")
    gpu = [_gpu_seed(1000 + s) for s in range(32)]
    threads = host_threads()
    per = max(1, threads // 4)
    with ThreadPoolExecutor(4) as ex:
        orc_res = list(ex.map(lambda s: _oracle_seed(2000 + s, per), range(4)))
    g_gpu = np.array([g for g, _ in gpu])
    g_orc = np.array([g for g, _ in orc_res])
    d_gpu = np.array([d for _, d in gpu])
    d_orc = np.array([d for _, d in orc_res])
    # the spread of one seed's estimate, from the 32 GPU seeds
    sd = g_gpu.std(axis=0, ddof=1)
    se = sd * np.sqrt(1.0 / len(g_orc) + 1.0 / len(g_gpu))
    ok = sd > 0
    z = (g_orc.mean(axis=0) - g_gpu.mean(axis=0))[ok] / se[ok]
    _say(f"  g(r): max |z| {np.abs(z).max():.2f}, mean z^2 {np.mean(z ** 2):.2f} over {ok.sum()} bins; "
         f"peak g {g_gpu.mean(axis=0).max():.3f}")
    assert np.abs(z).max() <= 5.0
    assert np.mean(z ** 2) <= 2.5
    d_se = d_gpu.std(ddof=1) * np.sqrt(1.0 / len(d_orc) + 1.0 / len(d_gpu))
    _say(f"  D: GPU {d_gpu.mean():.3f} +- {d_gpu.std(ddof=1):.3f}, oracle {d_orc.mean():.3f}, "
         f"standard error {d_se:.3f} A^2/ns")
    assert abs(d_orc.mean() - d_gpu.mean()) <= 5.0 * d_se
    assert 0.5 * 150.0 < d_gpu.mean() < 150.0  # interactions slow free diffusion
