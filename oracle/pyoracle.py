"""Python front end of the parity oracle (liboracle.so, built from oracle_capi.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker. The product path never imports it.
"""
from __future__ import annotations

import ctypes as C
import sys
from pathlib import Path
from typing import Optional

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
_lib = None

sys.path.insert(0, str(HERE.parent / "mars-md_amd"))
from marsmd import _capi as api  # noqa: E402  (struct layouts only)
from marsmd.system import RunParams, SystemSpec  # noqa: E402


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} missing: run `make -C oracle`")
        _lib = C.CDLL(str(LIB))
        L = _lib
        P, U64, I32, U32, F = C.c_void_p, C.c_uint64, C.c_int32, C.c_uint32, C.c_float
        sig = {
            "orc_last_error": (C.c_char_p, []),
            "orc_philox_draw": (None, [U64, U32, U32, U32, U32, P]),
            "orc_u01": (F, [U32]),
            "orc_box_muller": (None, [P, P]),
            "orc_mix_seed": (U64, [U64, U64, U64]),
            "orc_morton_encode": (None, [P, U64, P, P, P]),
            "orc_morton_from_coords": (U32, [U32, U32, U32]),
            "orc_morton_decode": (None, [U32, P]),
            "orc_stable_sort_pairs": (None, [P, U64, P, P]),
            "orc_pairlist_domain": (I32, [P, U64, C.POINTER(api.Box), P, P]),
            "orc_select_cell_grid": (None, [F, U64, P, U64, P, P, P]),
            "orc_cell_tables": (None, [P, U64, I32, P, I32, P, P, P, P]),
            "orc_build_pairlist": (C.c_int64, [P, U64, C.POINTER(api.Box), F, C.POINTER(api.ExclusionView), U64, I32, P, U64,
                                               P, P, P, P]),
            "orc_exclusion_csr": (C.c_int64, [P, U64, P, P]),
            "orc_pair_forces": (I32, [P, U64, P, P, P, I32, P, I32, C.POINTER(api.Box), F, U32, I32, U64, P,
                                      P, P]),
            "orc_table_compute": (None, [C.POINTER(api.Table), F, P]),
            "orc_table_compute64": (None, [C.POINTER(api.Table), C.c_double, P]),
            "orc_bonded_forces": (I32, [P, P, U64, P, P, U64, P, P, I32, P, I32, C.POINTER(api.Box), I32, U64, P]),
            "orc_grid_sample": (None, [C.POINTER(api.GridView), P, U64, I32, I32, I32, P]),
            "orc_sim_create": (P, [C.POINTER(api.PatchDesc)]),
            "orc_sim_destroy": (None, [P]),
            "orc_sim_set_base_seed": (None, [P, U64]),
            "orc_sim_set_particles": (I32, [P, P, P, P, U64]),
            "orc_sim_get_particles": (I32, [P, P, P, P, P, P, P, U64]),
            "orc_sim_set_state": (I32, [P, P, P, P, P, P, U64, I32]),
            "orc_sim_set_field": (I32, [P, P, I32]),
            "orc_acos": (F, [F]),
            "orc_atan2": (F, [F, F]),
            "orc_fm_loss_grad": (I32, [U64, U64, P, P, I32, P, C.c_uint32, I32, P, P, C.POINTER(api.Box), F, P, P,
                                       C.POINTER(C.c_double), P]),
            "orc_dihedral_restraint_forces": (I32, [P, P, U64, P, P, P, P, U64, P, C.POINTER(api.Box), I32, U64, P,
                                                    P]),
            "orc_sim_calculate_nonbonded_forces": (I32, [P, F, F, U64, U64, P, I32, I32]),
            "orc_sim_calculate_bonded_forces": (I32, [P, I32]),
            "orc_sim_integrate_motion": (I32, [P, F, F, I32, U64]),
            "orc_sim_finish_deferred_kick": (I32, [P, F]),
            "orc_sim_reorder_particles": (I32, [P]),
            "orc_sim_run": (I32, [P, C.POINTER(api.RunParams), U64, U64]),
            "orc_sim_set_threads": (I32, [P, I32]),
            "orc_sim_get_pairs": (C.c_int64, [P, P, U64]),
            "orc_sim_get_topology": (I32, [P, P, P, P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().orc_last_error().decode())


# --------------------------------------------------------------------------- RNG
def philox_draw(seed, ctr0, global_seed, ctr1, draws=1) -> np.ndarray:
    out = np.zeros((draws, 4), np.uint32)
    lib().orc_philox_draw(seed, ctr0, global_seed, ctr1, draws, out.ctypes.data)
    return out


def u01(x: int) -> float:
    return lib().orc_u01(x)


def box_muller(u) -> np.ndarray:
    u = np.ascontiguousarray(u, np.float32)
    out = np.zeros(3, np.float32)
    lib().orc_box_muller(u.ctypes.data, out.ctypes.data)
    return out


def mix_seed(seed, patch_id, step) -> int:
    return lib().orc_mix_seed(seed, patch_id, step)


def acos32(x: float) -> float:
    """The FP32 arccosine shared by the angle kernels and the oracle."""
    return lib().orc_acos(x)


# --------------------------------------------------------------------------- Morton / sort
def morton_encode(pos4, bmin, bmax) -> np.ndarray:
    pos4 = np.ascontiguousarray(pos4, np.float32)
    lo = np.ascontiguousarray(bmin, np.float32)
    hi = np.ascontiguousarray(bmax, np.float32)
    out = np.zeros(len(pos4), np.uint32)
    lib().orc_morton_encode(pos4.ctypes.data, len(pos4), lo.ctypes.data, hi.ctypes.data, out.ctypes.data)
    return out


def morton_from_coords(x, y, z) -> int:
    return lib().orc_morton_from_coords(x, y, z)


def morton_decode(code) -> tuple:
    out = np.zeros(3, np.uint32)
    lib().orc_morton_decode(code, out.ctypes.data)
    return tuple(int(v) for v in out)


def stable_sort_pairs(keys):
    keys = np.ascontiguousarray(keys, np.uint32)
    sk = np.zeros_like(keys)
    sv = np.zeros_like(keys)
    lib().orc_stable_sort_pairs(keys.ctypes.data, len(keys), sk.ctypes.data, sv.ctypes.data)
    return sk, sv


def box_c(box) -> api.Box:
    return box.to_c() if hasattr(box, "to_c") else box


def pairlist_domain(pos4, box):
    pos4 = np.ascontiguousarray(pos4, np.float32)
    lo = np.zeros(3, np.float32)
    hi = np.zeros(3, np.float32)
    b = box_c(box)
    lib().orc_pairlist_domain(pos4.ctypes.data, len(pos4), C.byref(b), lo.ctypes.data, hi.ctypes.data)
    return lo, hi


def select_cell_grid(cutoff, n, extent, max_pairs):
    ext = np.ascontiguousarray(extent, np.float32)
    m = C.c_int32()
    r = np.zeros(3, np.int32)
    slots = C.c_int32()
    lib().orc_select_cell_grid(cutoff, n, ext.ctypes.data, max_pairs, C.byref(m), r.ctypes.data, C.byref(slots))
    return m.value, tuple(int(v) for v in r), slots.value


def cell_tables(sorted_codes, m, r, slots, periodic):
    codes = np.ascontiguousarray(sorted_codes, np.uint32)
    nc = 1 << (3 * m)
    b = np.zeros(nc, np.uint32)
    e = np.zeros(nc, np.uint32)
    nb = np.zeros(nc * slots, np.uint32)
    rr = np.ascontiguousarray(r, np.int32)
    per = np.ascontiguousarray(periodic, np.int32)
    lib().orc_cell_tables(codes.ctypes.data, len(codes), m, rr.ctypes.data, slots, per.ctypes.data,
                          b.ctypes.data, e.ctypes.data, nb.ctypes.data)
    return b, e, nb.reshape(nc, slots)


def exclusion_view(offsets=None, excluded=None, num_particles=0, rigid_body_id=None):
    v = api.ExclusionView()
    v.offsets = _p(offsets)
    v.excluded = _p(excluded)
    v.num_particles = num_particles
    v.rigid_body_id = _p(rigid_body_id)
    return v


def build_pairlist(pos4, box, cutoff, offsets=None, excluded=None, num_excl=0, rigid_body_id=None,
                   max_pairs=None, brute_force=False):
    """Returns dict(pairs (P,2) sorted, codes, sorted_idx, inverse, grid=(m, r, slots))."""
    pos4 = np.ascontiguousarray(pos4, np.float32)
    n = len(pos4)
    offsets = None if offsets is None else np.ascontiguousarray(offsets, np.int32)
    excluded = None if excluded is None else np.ascontiguousarray(excluded, np.int32)
    body = None if rigid_body_id is None else np.ascontiguousarray(rigid_body_id, np.int32)
    ex = exclusion_view(offsets, excluded, num_excl, body)
    b = box_c(box)
    codes = np.zeros(n, np.uint32)
    sidx = np.zeros(n, np.uint32)
    inv = np.zeros(n, np.uint32)
    gi = np.zeros(5, np.int32)
    cap = 0
    mp = max_pairs if max_pairs is not None else max(n, 1)
    cnt = lib().orc_build_pairlist(pos4.ctypes.data, n, C.byref(b), cutoff, C.byref(ex), mp, 1 if brute_force else 0, None, 0,
                                   codes.ctypes.data, sidx.ctypes.data, inv.ctypes.data, gi.ctypes.data)
    if cnt < 0:
        raise RuntimeError(lib().orc_last_error().decode())
    pairs = np.zeros((max(cnt, 1), 2), np.int32)
    cnt2 = lib().orc_build_pairlist(pos4.ctypes.data, n, C.byref(b), cutoff, C.byref(ex), mp, 1 if brute_force else 0,
                                    pairs.ctypes.data, cnt, None, None, None, None)
    assert cnt2 == cnt
    return {"pairs": pairs[:cnt], "codes": codes, "sorted_idx": sidx, "inverse": inv,
            "grid": (int(gi[0]), (int(gi[1]), int(gi[2]), int(gi[3])), int(gi[4]))}


def exclusion_csr(pairs):
    pairs = np.ascontiguousarray(pairs, np.int32).reshape(-1, 2)
    n = int(pairs.max()) + 1 if len(pairs) else 0
    offsets = np.zeros(n + 1, np.int32)
    excluded = np.zeros(2 * len(pairs), np.int32)
    np_ = lib().orc_exclusion_csr(pairs.ctypes.data, len(pairs), offsets.ctypes.data, excluded.ctypes.data)
    return offsets, excluded, int(np_)


# --------------------------------------------------------------------------- forces
def table_records(tables):
    recs = (api.Table * max(1, len(tables)))()
    keep = []
    for i, t in enumerate(tables):
        k = np.ascontiguousarray(t.pot, np.float32)
        keep.append(k)
        recs[i] = t.to_c(k.ctypes.data)
    return recs, keep


def pair_forces(pairs, pos4, types, tag_matrix, tables, box, cutoff_squared, mode=1,
                enabled_terms=api.PAIR_TERM_TABULATED, particle_types=None, pair_params=None):
    """mode 0: FP32 sums, 1: FP32 terms summed in FP64, 2: FP64 twin. -> (N, 4) float64.
    particle_types (list of ParticleType) and pair_params (PairParams, None: defaults)
    feed the analytical terms."""
    from marsmd.core import make_types

    pt = make_types(particle_types) if particle_types is not None else None
    pp = pair_params.to_c() if pair_params is not None else None
    pairs = np.ascontiguousarray(pairs, np.int32)
    pos4 = np.ascontiguousarray(pos4, np.float32)
    types = np.ascontiguousarray(types, np.int32)
    tm = None if tag_matrix is None else np.ascontiguousarray(tag_matrix, np.uint32)
    recs, keep = table_records(tables)
    out = np.zeros((len(pos4), 4), np.float64)
    b = box_c(box)
    nt = 0 if tm is None else int(round(np.sqrt(tm.size)))
    _check(lib().orc_pair_forces(pairs.ctypes.data, len(pairs), pos4.ctypes.data, types.ctypes.data, _p(tm), nt,
                                 C.addressof(recs), len(tables), C.byref(b), cutoff_squared, enabled_terms, mode,
                                 len(pos4), out.ctypes.data, C.byref(pt) if pt is not None else None,
                                 C.byref(pp) if pp is not None else None))
    return out


def table_compute(table, x, fp64=False):
    recs, keep = table_records([table])
    if fp64:
        out = np.zeros(2, np.float64)
        lib().orc_table_compute64(C.byref(recs[0]), float(x), out.ctypes.data)
    else:
        out = np.zeros(2, np.float32)
        lib().orc_table_compute(C.byref(recs[0]), float(x), out.ctypes.data)
    return out


def bonded_forces(pos4, box, bonds=None, bond_table=None, bond_tables=(), angles=None, angle_table=None,
                  angle_tables=(), mode=1):
    pos4 = np.ascontiguousarray(pos4, np.float32)
    bonds = np.zeros((0, 2), np.int32) if bonds is None else np.ascontiguousarray(bonds, np.int32)
    angles = np.zeros((0, 3), np.int32) if angles is None else np.ascontiguousarray(angles, np.int32)
    bt = np.zeros(len(bonds), np.int32) if bond_table is None else np.ascontiguousarray(bond_table, np.int32)
    at = np.zeros(len(angles), np.int32) if angle_table is None else np.ascontiguousarray(angle_table, np.int32)
    brec, k1 = table_records(list(bond_tables))
    arec, k2 = table_records(list(angle_tables))
    out = np.zeros((len(pos4), 4), np.float64)
    b = box_c(box)
    _check(lib().orc_bonded_forces(bonds.ctypes.data, bt.ctypes.data, len(bonds), angles.ctypes.data, at.ctypes.data,
                                   len(angles), pos4.ctypes.data, C.addressof(brec), len(bond_tables),
                                   C.addressof(arec), len(angle_tables), C.byref(b), mode, len(pos4),
                                   out.ctypes.data))
    return out


def atan2_32(y: float, x: float) -> float:
    """The FP32 arctangent shared by the dihedral kernel and the oracle."""
    return lib().orc_atan2(y, x)


def dihedral_restraint_forces(pos4, box, dihedrals=None, dihedral_table=None, dihedral_tables=(),
                              restraint_ids=None, restraint_anchors=None, restraint_k=None, mode=1):
    """launch_tabulated_dihedrals + launch_harmonic_restraints on explicit arrays.
    Returns (n x 4 float64 force/energy, per-dihedral angle)."""
    pos4 = np.ascontiguousarray(pos4, np.float32)
    dih = np.zeros((0, 4), np.int32) if dihedrals is None else np.ascontiguousarray(dihedrals, np.int32)
    dt = np.zeros(len(dih), np.int32) if dihedral_table is None else np.ascontiguousarray(dihedral_table, np.int32)
    rid = np.zeros(0, np.int32) if restraint_ids is None else np.ascontiguousarray(restraint_ids, np.int32)
    anc = np.zeros((len(rid), 4), np.float32)
    if len(rid):
        anc[:, :3] = np.asarray(restraint_anchors, np.float32)[:, :3]
    kk = np.zeros(len(rid), np.float32) if restraint_k is None else np.ascontiguousarray(restraint_k, np.float32)
    drec, keep = table_records(list(dihedral_tables))
    out = np.zeros((len(pos4), 4), np.float64)
    phi = np.zeros(max(len(dih), 1), np.float64)
    b = box_c(box)
    _check(lib().orc_dihedral_restraint_forces(dih.ctypes.data, dt.ctypes.data, len(dih), C.addressof(drec),
                                               rid.ctypes.data, anc.ctypes.data, kk.ctypes.data, len(rid),
                                               pos4.ctypes.data, C.byref(b), mode, len(pos4), out.ctypes.data,
                                               phi.ctypes.data))
    return out, phi[:len(dih)]


def fm_loss_grad(positions, type_ids, term_matrix, tables, knots64, box, cutoff, target=None):
    """The oracle's force-matching forces, loss and dL/dknot (FP32 pair geometry, FP64
    knots and sums). positions (S, n, 3|4); knots64 all tables' knots concatenated;
    target (S, n, 3) or None. Returns (forces (S, n, 3), loss, grad)."""
    pos = np.asarray(positions, np.float32)
    S, n = pos.shape[0], pos.shape[1]
    p4 = np.zeros((S, n, 4), np.float32)
    p4[..., :3] = pos[..., :3]
    types = np.ascontiguousarray(type_ids, np.int32)
    tm = np.ascontiguousarray(term_matrix, np.uint32)
    recs, keep = table_records(list(tables))
    k64 = np.ascontiguousarray(knots64, np.float64)
    forces = np.zeros((S, n, 3), np.float64)
    grad = np.zeros(len(k64), np.float64)
    loss = C.c_double(0.0)
    tgt = None if target is None else np.ascontiguousarray(target, np.float64)
    b = box_c(box)
    rc = lib().orc_fm_loss_grad(S, n, p4.ctypes.data, types.ctypes.data, tm.shape[0], tm.ctypes.data,
                                api.PAIR_TERM_TABULATED, len(tables), C.addressof(recs), k64.ctypes.data, C.byref(b),
                                float(cutoff), None if tgt is None else tgt.ctypes.data, forces.ctypes.data,
                                C.byref(loss), grad.ctypes.data)
    if rc != 0:
        raise RuntimeError("orc_fm_loss_grad failed")
    return forces, loss.value, grad


def grid_sample(grid, pos4, scheme, bc=None, fp64=False):
    pos4 = np.ascontiguousarray(pos4, np.float32)
    g = grid.to_c(grid.values.ctypes.data, 0)
    out = np.zeros((len(pos4), 4), np.float64)
    lib().orc_grid_sample(C.byref(g), pos4.ctypes.data, len(pos4), scheme,
                          grid.boundary_condition if bc is None else bc, 1 if fp64 else 0, out.ctypes.data)
    return out


# --------------------------------------------------------------------------- simulation
class Sim:
    """The oracle's Patch + SimManager loop, with the interface of marsmd.Patch."""

    def __init__(self, spec: SystemSpec, capacity: Optional[int] = None):
        desc, self._keep = spec.to_desc(capacity)
        self._h = lib().orc_sim_create(C.byref(desc))
        if not self._h:
            raise RuntimeError(lib().orc_last_error().decode())
        self.spec = spec
        lib().orc_sim_set_base_seed(self._h, spec.seed)
        p = spec.pos4()
        m = spec.mom4()
        t = np.ascontiguousarray(spec.type_ids, np.int32)
        self._keep += [p, m, t]
        _check(lib().orc_sim_set_particles(self._h, p.ctypes.data, _p(m), t.ctypes.data, len(t)))
        self.n = len(t)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_sim_destroy(self._h)
            self._h = None

    def copy_particles_to_host(self):
        n = self.n
        pos = np.zeros((n, 4), np.float32)
        mom = np.zeros((n, 4), np.float32)
        fe = np.zeros((n, 4), np.float32)
        f64 = np.zeros((n, 4), np.float64)
        ty = np.zeros(n, np.int32)
        ids = np.zeros(n, np.int32)
        _check(lib().orc_sim_get_particles(self._h, pos.ctypes.data, mom.ctypes.data, fe.ctypes.data,
                                           f64.ctypes.data, ty.ctypes.data, ids.ctypes.data, n))
        return {"pos": pos, "mom": mom, "force": fe, "force64": f64, "type": ty, "id": ids}

    def set_state(self, pos=None, mom=None, force=None, type_ids=None, ids=None, deferred_kick=None):
        arrs = [None if a is None else np.ascontiguousarray(a, dt)
                for a, dt in ((pos, np.float32), (mom, np.float32), (force, np.float32), (type_ids, np.int32),
                              (ids, np.int32))]
        dk = -1 if deferred_kick is None else int(bool(deferred_kick))
        _check(lib().orc_sim_set_state(self._h, *[_p(a) for a in arrs], self.n, dk))

    def set_field(self, electric_field=(0.0, 0.0, 0.0), interpolation=api.INTERP_LINEAR):
        ef = np.ascontiguousarray(electric_field, np.float32)
        _check(lib().orc_sim_set_field(self._h, ef.ctypes.data, interpolation))

    def calculate_nonbonded_forces(self, pairlist_cutoff, interaction_cutoff, step, rebuild_period,
                                   electric_field=(0.0, 0.0, 0.0), interpolation=api.INTERP_LINEAR,
                                   compute_energy=False):
        ef = np.ascontiguousarray(electric_field, np.float32)
        _check(lib().orc_sim_calculate_nonbonded_forces(self._h, pairlist_cutoff, interaction_cutoff, step,
                                                        rebuild_period, ef.ctypes.data, interpolation,
                                                        1 if compute_energy else 0))
        self._prepared = True

    def maybe_reorder(self, step, params: Optional[RunParams] = None):
        """ParticleReorderManager::maybe_reorder (ReorderManager.cpp:8-25)."""
        P = params or self.spec.params
        rp = P.reorder_period if P.reorder_period > 0 else P.rebuild_period
        if rp > 0 and step != 0 and (step - 1) % rp == 0 and getattr(self, "_prepared", False):
            self.reorder_particles()
            return True
        return False

    def calculate_bonded_forces(self, compute_energy=False):
        _check(lib().orc_sim_calculate_bonded_forces(self._h, 1 if compute_energy else 0))

    def integrate_motion(self, dt, kT, integrator, step):
        _check(lib().orc_sim_integrate_motion(self._h, dt, kT, integrator, step))

    def finish_deferred_kick(self, dt):
        _check(lib().orc_sim_finish_deferred_kick(self._h, dt))

    def reorder_particles(self):
        _check(lib().orc_sim_reorder_particles(self._h))

    def set_threads(self, threads: int):
        """OpenMP threads for timing (1: the serial checker). With more than one the
        pair forces are summed per thread and the FP64 force twin is not kept."""
        _check(lib().orc_sim_set_threads(self._h, int(threads)))

    def run(self, num_steps, first_step=0, params: Optional[RunParams] = None):
        p = (params or self.spec.params).to_c()
        _check(lib().orc_sim_run(self._h, C.byref(p), first_step, num_steps))

    def step_unfused(self, step, params: Optional[RunParams] = None):
        P = params or self.spec.params
        self.maybe_reorder(step, P)
        energy = P.energy_period > 0 and step % P.energy_period == 0
        self.calculate_nonbonded_forces(P.pairlist_cutoff, P.cutoff, step, P.rebuild_period, P.electric_field,
                                        P.interpolation, energy)
        self.calculate_bonded_forces(energy)
        self.integrate_motion(P.timestep, P.kT, P.integrator, step)

    def get_pairs(self) -> np.ndarray:
        cnt = lib().orc_sim_get_pairs(self._h, None, 0)
        out = np.zeros((max(cnt, 1), 2), np.int32)
        lib().orc_sim_get_pairs(self._h, out.ctypes.data, cnt)
        return out[:cnt]
