"""Seeded synthetic systems for the five workloads of the hot path.

Every constant the workload descriptions leave open is pinned here once:
  * pair tables: WCA(sigma, eps) + Debye-Hueckel 332.0636 q_i q_j / eps_r e^{-r/lambda} / r
    with eps_r = 80 (the DebyeHuckelPotential default, Columb.h:30-43), lambda = 10 A;
    below r_cap = 0.8 sigma the potential continues linearly with its slope at r_cap
    (finite knots near r = 0); shifted so U(r_max) = 0; knots on [0, r_max] with step
    r_max / (K - 1), so the last knot sits on the cutoff;
  * BD: D = 150 A^2/ns, kT = 0.593 kcal/mol, dt = 1e-4 ns; Langevin: gamma = 1/ps
    (trans_damping 1000 ns^-1), mass 100 amu, Maxwell-Boltzmann momenta
    p = sqrt(kT m) * SQRT_CAL_TO_JOULE * N(0,1) (SimManager.cpp:969-991);
  * positions: Philox stream STREAM_POSITIONS, uniform in the box, candidates closer
    than 3 A (minimum image) to an earlier accepted bead are redrawn;
  * rebuild period 20 and skin 5 A (pairlist cutoff = cutoff + 5).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from . import _capi as api
from . import rng
from .core import Grid, ParticleType, PeriodicBox, TabulatedPotential, make_pair_tag
from .system import RunParams, SystemSpec

COULOMB = 332.0636
KT = 0.593
SQRT_CAL_TO_JOULE = 2.046167337
DEG_TO_RAD = np.float32(math.pi / 180.0)


def wca_dh(r: np.ndarray, sigma: float, eps: float, qq: float, eps_r: float = 80.0, debye: float = 10.0):
    r = np.asarray(r, np.float64)

    def raw(x):
        sr6 = (sigma / x) ** 6
        wca = np.where(x < 2.0 ** (1.0 / 6.0) * sigma, 4.0 * eps * (sr6 * sr6 - sr6) + eps, 0.0)
        dh = COULOMB * qq / eps_r * np.exp(-x / debye) / x
        return wca + dh

    rc = 0.8 * sigma
    h = 1e-6
    slope = (raw(np.array(rc + h)) - raw(np.array(rc - h))) / (2 * h)
    return np.where(r < rc, raw(np.array(rc)) + slope * (r - rc), raw(np.maximum(r, rc)))


def pair_table(sigma=5.0, eps=0.5, qq=1.0, rmax=30.0, knots=1024, eps_r=80.0, debye=10.0) -> TabulatedPotential:
    step = rmax / (knots - 1)
    r = np.arange(knots) * step
    u = wca_dh(r, sigma, eps, qq, eps_r, debye) - wca_dh(np.array(rmax), sigma, eps, qq, eps_r, debye)
    return TabulatedPotential(u.astype(np.float32), 0.0, step)


def harmonic_bond_table(k=10.0, r0=5.0, rmax=20.0, knots=512) -> TabulatedPotential:
    step = rmax / (knots - 1)
    r = np.arange(knots) * step
    return TabulatedPotential((0.5 * k * (r - r0) ** 2).astype(np.float32), 0.0, step)


def cosine_angle_table(k=2.0, knots=256) -> TabulatedPotential:
    """U(theta) = k (1 + cos theta) on [0, 180] degrees, abscissa converted to radians
    as Table::convert_angular_abscissa_to_radians (Tables.h:109-119); periodic."""
    step_deg = np.float32(180.0 / (knots - 1))
    th = np.arange(knots) * float(step_deg)
    u = k * (1.0 + np.cos(np.deg2rad(th)))
    start = np.float32(0.0) * DEG_TO_RAD
    step = np.float32(step_deg * DEG_TO_RAD)
    return TabulatedPotential(u.astype(np.float32), float(start), float(step), is_periodic=True)


def cosine_dihedral_table(k=1.0, n=3, delta=0.0, knots=361) -> TabulatedPotential:
    """U(phi) = k (1 + cos(n phi - delta)) on [-180, 180] degrees, the abscissa converted
    to radians as Table::convert_angular_abscissa_to_radians (Tables.h:109-119);
    periodic, as dihedral tables are (DeviceBondedInteraction.h:245-255)."""
    step_deg = np.float32(360.0 / (knots - 1))
    ph = -180.0 + np.arange(knots) * float(step_deg)
    u = k * (1.0 + np.cos(n * np.deg2rad(ph) - delta))
    start = np.float32(-180.0) * DEG_TO_RAD
    step = np.float32(step_deg * DEG_TO_RAD)
    return TabulatedPotential(u.astype(np.float32), float(start), float(step), is_periodic=True)


def uniform_positions(n: int, L, seed: int, min_sep: float = 3.0, origin=(0.0, 0.0, 0.0)) -> np.ndarray:
    """Uniform positions with rejection of candidates within min_sep of an earlier
    accepted bead (minimum image). Deterministic for a seed."""
    from scipy.spatial import cKDTree

    L = np.asarray(L, np.float64) * np.ones(3)
    pos = np.zeros((0, 3))
    next_ctr = 0
    need = n
    accepted = np.zeros((0, 3))
    rounds = 0
    while need > 0:
        rounds += 1
        if rounds > 64:
            raise RuntimeError("could not place particles at this density")
        u = rng.uniform(seed, rng.STREAM_POSITIONS, need, offset=next_ctr)[:, :3].astype(np.float64)
        next_ctr += need
        cand = np.minimum(u * L, np.nextafter(L, 0))
        allp = np.vstack([accepted, cand])
        tree = cKDTree(allp, boxsize=L)
        pairs = tree.query_pairs(min_sep, output_type="ndarray")
        ok = np.ones(len(allp), bool)
        na = len(accepted)
        if len(pairs):
            # a candidate is rejected when an accepted bead, or an earlier surviving
            # candidate, lies within min_sep: process candidates in order
            pairs = np.sort(pairs, axis=1)
            nbrs = {}
            for a, b in pairs:
                nbrs.setdefault(int(b), []).append(int(a))
            for b in sorted(nbrs):
                if b < na:
                    continue
                if any(ok[a] for a in nbrs[b]):
                    ok[b] = False
        accepted = allp[ok]
        need = n - len(accepted)
    return (accepted[:n] + np.asarray(origin)).astype(np.float32)


def maxwell_momenta(n: int, masses: np.ndarray, kT: float, seed: int) -> np.ndarray:
    z = rng.normal(seed, rng.STREAM_MOMENTA, n)
    return (np.sqrt(kT * masses)[:, None] * SQRT_CAL_TO_JOULE * z).astype(np.float32)


# ------------------------------------------------------------------------ configs
def config0(n: int = 10_000, L: float = 250.0, seed: int = 42, steps: int = 1000) -> SystemSpec:
    """Plain BD: point beads of one type in a periodic cube, WCA+DH table (1024 knots),
    cutoff 30, skin 5, D = 150, kT = 0.593, dt = 100 fs, Philox seed 42."""
    pos = uniform_positions(n, L, seed)
    t = ParticleType(mass=1.0, diffusion=(150.0, 150.0, 150.0))
    spec = SystemSpec(
        box=PeriodicBox((L, L, L)),
        types=[t],
        positions=pos,
        type_ids=np.zeros(n, np.int32),
        params=RunParams(cutoff=30.0, pairlist_cutoff=35.0, rebuild_period=20, timestep=1e-4, kT=KT,
                         integrator=api.INTEGRATOR_BROWNIAN),
        nb_tables=[pair_table(qq=1.0)],
        seed=seed,
        name=f"bd-{n}-beads-L{L:g}",
    )
    spec.steps = steps
    return spec


def config3(replicas: int = 512, n_min: int = 500, n_max: int = 20_000, density: float = 1e-4,
            seed: int = 42) -> List[SystemSpec]:
    """Independent replicas of the config-0 system: N_r log-uniform in [n_min, n_max]
    (seeded), L_r = (N_r / density)^(1/3), the config-0 table and BD; replica r has patch
    id r (its seed-mix word) and positions of its own (uniform, rejection at 3 A); all
    replicas share the base seed."""
    u = rng.uniform(seed, rng.STREAM_POSITIONS ^ 0x3, replicas)[:, 0].astype(np.float64)
    ns = np.round(np.exp(np.log(n_min) + u * (np.log(n_max) - np.log(n_min)))).astype(np.int64)
    specs = []
    for r, n in enumerate(ns):
        L = float((n / density) ** (1.0 / 3.0))
        spec = config0(n=int(n), L=L, seed=seed)
        spec.positions = uniform_positions(int(n), L, seed + 7919 * (r + 1))
        spec.patch_id = r
        spec.name = f"replica{r}-bd-{n}-beads-L{L:.1f}"
        specs.append(spec)
    return specs


def config1(chains: int = 200, beads: int = 500, L: float = 800.0, seed: int = 42, steps: int = 1000,
            dihedrals: bool = False, restraints: bool = False) -> SystemSpec:
    """Bead-spring polymers: random-walk chains (step r0 = 5 A, a bead is redrawn when it
    lands within 3 A of the bead two back), two types in blocks of 50 beads, harmonic
    bond table U = k/2 (r - r0)^2 (512 knots on [0, 20]), cosine angle table k(1 + cos)
    (256 knots, degrees), WCA+DH for 3 type pairs with charges (+0.5, -0.5), 1-2 and 1-3
    exclusions, BD, reorder period 20. Optionally dihedrals along every chain (k = 1,
    n = 3 cosine table, 361 knots) and harmonic restraints (k = 0.5) pinning every 50th
    bead to its initial position."""
    n = chains * beads
    r0 = 5.0
    starts = rng.uniform(seed, rng.STREAM_CHAINS, chains, draw=0)[:, :3].astype(np.float64) * L
    pos = np.zeros((chains, beads, 3))
    pos[:, 0] = starts
    for b in range(1, beads):
        d = rng.normal(seed, rng.STREAM_CHAINS, chains, draw=b, offset=0)
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        cand = pos[:, b - 1] + r0 * d
        if b >= 2:
            bad = np.linalg.norm(cand - pos[:, b - 2], axis=1) < 3.0
            tries = 0
            while bad.any():
                tries += 1
                d2 = rng.normal(seed, rng.STREAM_CHAINS, chains, draw=b, offset=(tries * chains))
                d2 /= np.linalg.norm(d2, axis=1, keepdims=True)
                cand[bad] = pos[bad, b - 1] + r0 * d2[bad]
                bad = np.linalg.norm(cand - pos[:, b - 2], axis=1) < 3.0
        pos[:, b] = cand
    pos = np.mod(pos.reshape(n, 3), L)
    pos = np.minimum(pos, np.nextafter(L, 0)).astype(np.float32)
    type_ids = np.tile((np.arange(beads) // 50) % 2, chains).astype(np.int32)
    idx = np.arange(n).reshape(chains, beads)
    bonds = np.stack([idx[:, :-1].ravel(), idx[:, 1:].ravel()], axis=1).astype(np.int32)
    angles = np.stack([idx[:, :-2].ravel(), idx[:, 1:-1].ravel(), idx[:, 2:].ravel()], axis=1).astype(np.int32)
    excl = np.vstack([bonds, angles[:, [0, 2]]]).astype(np.int32)
    q = [0.5, -0.5]
    nb = [pair_table(qq=q[0] * q[0]), pair_table(qq=q[0] * q[1]), pair_table(qq=q[1] * q[1])]
    types = [ParticleType(mass=1.0, charge=qi, diffusion=(150.0, 150.0, 150.0)) for qi in q]
    spec = SystemSpec(
        box=PeriodicBox((L, L, L)),
        types=types,
        positions=pos,
        type_ids=type_ids,
        params=RunParams(cutoff=30.0, pairlist_cutoff=35.0, rebuild_period=20, reorder_period=20, timestep=1e-4,
                         kT=KT, integrator=api.INTEGRATOR_BROWNIAN),
        nb_tables=nb,
        bonds=bonds,
        bond_table=np.zeros(len(bonds), np.int32),
        bond_tables=[harmonic_bond_table()],
        angles=angles,
        angle_table=np.zeros(len(angles), np.int32),
        angle_tables=[cosine_angle_table()],
        exclusions=excl,
        seed=seed,
        name=f"bd-polymers-{chains}x{beads}-L{L:g}",
    )
    if dihedrals:
        spec.dihedrals = np.stack([idx[:, :-3].ravel(), idx[:, 1:-2].ravel(), idx[:, 2:-1].ravel(),
                                   idx[:, 3:].ravel()], axis=1).astype(np.int32)
        spec.dihedral_table = np.zeros(len(spec.dihedrals), np.int32)
        spec.dihedral_tables = [cosine_dihedral_table()]
        spec.name += "-dihedrals"
    if restraints:
        spec.restraint_ids = np.arange(0, n, 50, dtype=np.int32)
        spec.restraint_anchors = pos[spec.restraint_ids].astype(np.float32)
        spec.restraint_k = np.full(len(spec.restraint_ids), 0.5, np.float32)
        spec.name += "-restraints"
    spec.steps = steps
    return spec


def membrane_grid(L: float = 2000.0, n: int = 256, height: float = 2.0, thickness: float = 40.0,
                  pore_radius: float = 20.0, width: float = 5.0) -> Grid:
    """Slab membrane centred at z = L/2 (thickness 40 A, tanh edges of width 5 A) with a
    cylindrical pore of radius 20 A (40 A diameter) on the box axis; V = height inside
    the slab away from the pore."""
    h = L / n
    c = np.arange(n) * h
    x = c[:, None, None]
    y = c[None, :, None]
    z = c[None, None, :]
    z0 = L / 2
    slab = 0.5 * (np.tanh((z - z0 + thickness / 2) / width) - np.tanh((z - z0 - thickness / 2) / width))
    r = np.sqrt((x - L / 2) ** 2 + (y - L / 2) ** 2)
    outside = 0.5 * (1.0 + np.tanh((r - pore_radius) / width))
    v = (height * slab * outside).astype(np.float32)
    return Grid(v, origin=(0.0, 0.0, 0.0), spacing=(h, h, h), boundary_condition=api.BC_PERIODIC)


def config2(n: int = 1_000_000, L: float = 2000.0, seed: int = 42, steps: int = 500,
            grid_n: int = 256) -> SystemSpec:
    """Large system with an external field: 3 types (charges +1, -1, +0.5), 6 WCA+DH
    pair tables (1024 knots), 256^3 membrane grid sampled tricubically (type scales
    +1, -1, +0.5 give the +-2 kcal/mol membrane), Langevin gamma = 1/ps."""
    pos = uniform_positions(n, L, seed)
    ids = rng.uniform(seed, rng.STREAM_POSITIONS ^ 0x7, n)[:, 0]
    type_ids = np.minimum((ids * 3).astype(np.int32), 2)
    q = [1.0, -1.0, 0.5]
    scales = [1.0, -1.0, 0.5]
    mass = 100.0
    gamma = 1000.0
    types = [ParticleType(mass=mass, charge=q[i], diffusion=(KT / (mass * gamma),) * 3,
                          trans_damping=(gamma, gamma, gamma), pmf_terms=[(0, scales[i], -1)]) for i in range(3)]
    nb = []
    tags = np.zeros((3, 3), np.uint32)
    k = 0
    for i in range(3):
        for j in range(i, 3):
            nb.append(pair_table(qq=q[i] * q[j]))
            tags[i, j] = tags[j, i] = make_pair_tag(api.PAIR_TERM_TABULATED, k)
            k += 1
    masses = np.full(n, mass)
    mom = maxwell_momenta(n, masses, KT, seed)
    spec = SystemSpec(
        box=PeriodicBox((L, L, L)),
        types=types,
        positions=pos,
        type_ids=type_ids,
        momenta=mom,
        params=RunParams(cutoff=30.0, pairlist_cutoff=35.0, rebuild_period=20, timestep=1e-4, kT=KT,
                         integrator=api.INTEGRATOR_LANGEVIN, interpolation=api.INTERP_CUBIC),
        nb_tables=nb,
        term_matrix=tags,
        grids=[membrane_grid(L, grid_n)],
        seed=seed,
        name=f"langevin-{n}-beads-3types-L{L:g}-grid{grid_n}",
    )
    spec.steps = steps
    return spec


@dataclass
class FMWorkload:
    """A force-matching problem: snapshots, the starting tables and the tables whose
    forces are the targets."""
    name: str
    box: PeriodicBox
    positions: np.ndarray  # (S, n, 3) float32
    type_ids: np.ndarray  # (n,) int32
    tables: List[TabulatedPotential]
    target_tables: List[TabulatedPotential]
    term_matrix: np.ndarray  # (T, T) uint32 pair tags
    cutoff: float


def fm_tables(knots: int = 512, rmax: float = 30.0):
    """4 types (charges +1, -1, +0.5, -0.5) and one WCA+DH table of `knots` knots on
    [0, rmax] per type pair (10 tables); returns (tables, tag matrix)."""
    q = [1.0, -1.0, 0.5, -0.5]
    tables = []
    tm = np.zeros((4, 4), np.uint32)
    k = 0
    for i in range(4):
        for j in range(i, 4):
            tables.append(pair_table(qq=q[i] * q[j], rmax=rmax, knots=knots))
            tm[i, j] = tm[j, i] = make_pair_tag(api.PAIR_TERM_TABULATED, k)
            k += 1
    return tables, tm


def perturbed_tables(tables, sigma: float, seed: int):
    """Every knot but the last (the force must still vanish past the cutoff) plus a
    seeded Gaussian of standard deviation sigma (kcal/mol)."""
    g = np.random.default_rng(seed)
    out = []
    for t in tables:
        noise = g.normal(0.0, sigma, len(t.pot))
        noise[-1] = 0.0
        out.append(TabulatedPotential((t.pot.astype(np.float64) + noise).astype(np.float32), t.start, t.step))
    return out


def config4(n: int = 100_000, L: float = 1000.0, snapshots: int = 64, knots: int = 512, seed: int = 42,
            perturbation: float = 0.2, frames: Optional[np.ndarray] = None) -> FMWorkload:
    """Force matching of 10 tables x 512 knots over 64 snapshots of 100k beads (4 types,
    L = 1000 A, cutoff 30 A). `frames` (S, n, 3) supplies the snapshots (BD frames);
    without it they are independent uniform configurations (rejection at 3 A). Targets
    are the forces of the tables plus a seeded 0.2 kcal/mol Gaussian on every knot."""
    tables, tm = fm_tables(knots)
    type_ids = np.minimum((rng.uniform(seed, rng.STREAM_POSITIONS ^ 0x4, n)[:, 0] * 4).astype(np.int32), 3)
    if frames is None:
        frames = np.stack([uniform_positions(n, L, seed + 1000 * (s + 1)) for s in range(snapshots)])
    frames = np.asarray(frames, np.float32)[..., :3]
    return FMWorkload(name=f"force-matching-{frames.shape[0]}x{n}-beads-4types-L{L:g}-10x{knots}-knots",
                      box=PeriodicBox((L, L, L)), positions=frames, type_ids=type_ids, tables=tables,
                      target_tables=perturbed_tables(tables, perturbation, seed), term_matrix=tm, cutoff=30.0)


def bd_snapshots(n: int = 100_000, L: float = 1000.0, snapshots: int = 64, seed: int = 42, warmup: int = 100,
                 stride: int = 10):
    """Brownian-dynamics snapshots of the 4-type WCA+DH system of the force-matching
    workload (the fm_tables() pair tables, D = 150, dt = 100 fs, cutoff 30 + 5 A skin),
    stepped by a GPU Patch from uniform positions: `warmup` steps, then one frame every
    `stride` steps. Returns (frames (S, n, 3) float32 in global-id order, type ids)."""
    from .patch import Patch

    tables, tm = fm_tables()
    type_ids = np.minimum((rng.uniform(seed, rng.STREAM_POSITIONS ^ 0x4, n)[:, 0] * 4).astype(np.int32), 3)
    q = [1.0, -1.0, 0.5, -0.5]
    spec = SystemSpec(box=PeriodicBox((L, L, L)),
                      types=[ParticleType(mass=1.0, charge=qi, diffusion=(150.0,) * 3) for qi in q],
                      positions=uniform_positions(n, L, seed), type_ids=type_ids,
                      params=RunParams(cutoff=30.0, pairlist_cutoff=35.0, rebuild_period=20, timestep=1e-4, kT=KT,
                                       integrator=api.INTEGRATOR_BROWNIAN),
                      nb_tables=tables, term_matrix=tm, seed=seed, name="fm-bd")
    patch = Patch(spec)
    patch.run(warmup)
    out = np.zeros((snapshots, n, 3), np.float32)
    step = warmup
    for s in range(snapshots):
        patch.run(stride, first_step=step)
        step += stride
        h = patch.copy_particles_to_host()
        out[s, h["id"]] = h["pos"][:, :3]
    return out, type_ids

