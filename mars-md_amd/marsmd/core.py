"""Reference-shaped Python classes over the C ABI (libmarsmd_hip.so).

Names and argument meaning follow the reference's C++ interface for the hot path:
PeriodicBox (src/System/PeriodicBox.h), ZOrderSort
(src/PatchOperation/ZOrderKernels/ZOrderSort.h), ZOrderPairlist / Pairlist
(src/PatchOperation/PairListKernels/ZOrderPairlist.h, Pairlist.h), Patch
(src/PatchOperation/Patch.h) and the launch_* free functions
(NonBondedInteraction.h, DeviceBondedInteraction.h, Integrator.h).

PyTorch provides device memory and streams only; every computation runs in the HIP
library, and a missing library raises (there is no CPU fallback).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import _capi as api


def _torch():
    import torch

    return torch


def _stream(stream=None) -> int:
    torch = _torch()
    if stream is None:
        return torch.cuda.current_stream().cuda_stream
    return getattr(stream, "cuda_stream", stream)


def _ptr(t) -> Optional[int]:
    if t is None:
        return None
    return t.data_ptr()


def _hip():
    if not hasattr(_hip, "lib"):
        h = C.CDLL("libamdhip64.so")
        h.hipMemcpy.restype = C.c_int
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.lib = h
    return _hip.lib


def copy_from_device_ptr(ptr: int, count: int, dtype) -> "torch.Tensor":
    """Copy `count` elements at a library-owned device pointer into a new CUDA tensor."""
    torch = _torch()
    out = torch.empty(count, dtype=dtype, device="cuda")
    if count:
        torch.cuda.synchronize()
        st = _hip().hipMemcpy(C.c_void_p(out.data_ptr()), C.c_void_p(ptr), out.numel() * out.element_size(), 3)
        if st != 0:
            raise RuntimeError(f"hipMemcpy failed ({st})")
    return out


# ---------------------------------------------------------------------------- box
@dataclass
class PeriodicBox:
    """Orthorhombic PeriodicBox (src/System/PeriodicBox.h:16-209)."""

    size: Sequence[float]
    periodic: Sequence[bool] = (True, True, True)
    origin: Sequence[float] = (0.0, 0.0, 0.0)

    def to_c(self) -> api.Box:
        b = api.Box()
        for a in range(3):
            b.size[a] = float(self.size[a])
            b.origin[a] = float(self.origin[a])
            b.periodic[a] = 1 if self.periodic[a] else 0
        return b


# ---------------------------------------------------------------------------- tables
@dataclass
class TabulatedPotential:
    """TabulatedPotential (src/Interactions/TabulatedPotential.h:8-13): knots U[k] at
    start + k*step. step_inv follows TablesRegistry: 1.0f / float(step)."""

    pot: np.ndarray
    start: float
    step: float
    is_periodic: bool = False

    def __post_init__(self):
        self.pot = np.ascontiguousarray(self.pot, dtype=np.float32)

    @property
    def step_inv(self) -> float:
        return float(np.float32(1.0) / np.float32(self.step))

    def to_c(self, pot_ptr: int) -> api.Table:
        t = api.Table()
        t.pot = pot_ptr
        t.step_inv = self.step_inv
        t.size = len(self.pot)
        t.start = float(np.float32(self.start))
        t.is_periodic = 1 if self.is_periodic else 0
        return t


def make_pair_tag(terms: int, table_index: int) -> int:
    """make_pair_tag (src/Interactions/Nonbonded/Pairwise.h:37-40)."""
    return (terms & 0xFF) | ((table_index << api.PAIR_TABLE_SHIFT) if table_index >= 0 else 0)


class DeviceTables:
    """Device-resident array of TabulatedPotential records (the reference's
    DeviceBuffer<TabulatedPotential>), for the launch_* functions."""

    def __init__(self, tables: Sequence[TabulatedPotential]):
        torch = _torch()
        self.knots = [torch.from_numpy(t.pot).cuda() for t in tables]
        recs = (api.Table * max(1, len(tables)))()
        for i, t in enumerate(tables):
            recs[i] = t.to_c(self.knots[i].data_ptr())
        raw = bytes(recs)
        self.records = torch.frombuffer(bytearray(raw), dtype=torch.uint8).cuda()

    @property
    def ptr(self) -> int:
        return self.records.data_ptr()


# ---------------------------------------------------------------------------- types
@dataclass
class ParticleType:
    """One row of ParticleTypeView (src/Objects/DeviceParticle.h:49-65)."""

    mass: float = 1.0
    charge: float = 0.0
    diffusion: Sequence[float] = (0.0, 0.0, 0.0)
    radius: float = 0.0  # softcore pair term
    eps: float = 0.0  # softcore pair term
    trans_damping: Sequence[float] = (0.0, 0.0, 0.0)
    pmf_terms: Sequence[tuple] = ()  # (grid_id, scale, boundary_condition)
    force_grid_id: Sequence[int] = (-1, -1, -1)
    force_grid_scale: Sequence[float] = (0.0, 0.0, 0.0)


def make_types(types: Sequence[ParticleType]) -> api.ParticleTypes:
    if not 1 <= len(types) <= api.MAX_TYPES:
        raise ValueError(f"1..{api.MAX_TYPES} particle types supported")
    out = api.ParticleTypes()
    out.num_types = len(types)
    k = 0
    for i, t in enumerate(types):
        out.mass[i] = t.mass
        out.charge[i] = t.charge
        out.radius[i] = t.radius
        out.eps[i] = t.eps
        for a in range(3):
            out.diffusion[i][a] = t.diffusion[a]
            out.trans_damping[i][a] = t.trans_damping[a]
            out.force_grid_id[i][a] = t.force_grid_id[a]
            out.force_grid_scale[i][a] = t.force_grid_scale[a]
        out.pmf_grid_offset[i] = k
        out.pmf_grid_count[i] = len(t.pmf_terms)
        for gid, scale, bc in t.pmf_terms:
            if k >= api.MAX_PMF_TERMS:
                raise ValueError("too many PMF grid terms")
            out.pmf_grid_terms[k].grid_id = gid
            out.pmf_grid_terms[k].scale = scale
            out.pmf_grid_terms[k].scale_slope = 0.0
            out.pmf_grid_terms[k].boundary_condition = bc
            k += 1
    return out


@dataclass
class PairParams:
    """Constants of the analytical pair terms: SolventParams
    (NonBondedInteraction.h:121-127; Debye-Hueckel and ONC) and GaussianPotential's
    members (Pairwise.h:53-57). The defaults are the reference's."""

    debye_length: float = 10.0
    dielectric: float = 80.0
    onck_kappa: float = 0.1
    onck_sz: float = 80.0
    onck_z: float = 6.86
    gaussian_amp: float = 1.0
    gaussian_sigma: float = 1.0
    gaussian_cutoff_squared: float = 0.0

    def to_c(self) -> api.PairParams:
        return api.PairParams(self.debye_length, self.dielectric, self.onck_kappa, self.onck_sz, self.onck_z,
                              self.gaussian_amp, self.gaussian_sigma, self.gaussian_cutoff_squared)


# ---------------------------------------------------------------------------- grids
@dataclass
class Grid:
    """A dense grid (BaseGridView<float>, src/Types/BaseGridDevice.h:24-30):
    values[ix, iy, iz] at origin + basis @ (ix, iy, iz); z is the fastest index."""

    values: np.ndarray
    origin: Sequence[float]
    spacing: Sequence[float]
    boundary_condition: int = api.BC_PERIODIC

    def __post_init__(self):
        self.values = np.ascontiguousarray(self.values, dtype=np.float32)

    def to_c(self, data_ptr: int, grid_id: int = 0) -> api.GridView:
        g = api.GridView()
        g.data = data_ptr
        for a in range(3):
            g.origin[a] = float(self.origin[a])
            g.dims[a] = self.values.shape[a]
        for k in range(9):
            g.basis[k] = 0.0
            g.basis_inv[k] = 0.0
        for a in range(3):
            g.basis[4 * a] = float(np.float32(self.spacing[a]))
            g.basis_inv[4 * a] = float(np.float32(1.0) / np.float32(self.spacing[a]))
        g.grid_id = grid_id
        g.boundary_condition = self.boundary_condition
        return g


# ---------------------------------------------------------------------------- ZOrderSort
class ZOrderSort:
    """ZOrderSort (src/PatchOperation/ZOrderKernels/ZOrderSort.h:43-313)."""

    def __init__(self, max_particles: int, _handle: Optional[int] = None):
        self._owned = _handle is None
        h = C.c_void_p(_handle)
        if self._owned:
            api.call("marsmd_zorder_create", max_particles, C.byref(h))
        self._h = h
        self.num_particles = 0

    def __del__(self):
        if getattr(self, "_owned", False) and getattr(self, "_h", None):
            api.lib().marsmd_zorder_destroy(self._h)
            self._h = None

    def sort_particles(self, positions, num_particles: int, box_min, box_max, stream=None):
        lo = (C.c_float * 3)(*map(float, box_min))
        hi = (C.c_float * 3)(*map(float, box_max))
        api.call("marsmd_zorder_sort_particles", self._h, _ptr(positions), num_particles, lo, hi, _stream(stream))
        self.num_particles = num_particles

    def _buffers(self):
        a, b, c = C.c_void_p(), C.c_void_p(), C.c_void_p()
        api.call("marsmd_zorder_get_buffers", self._h, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value

    def get_morton_codes(self):
        torch = _torch()
        return copy_from_device_ptr(self._buffers()[0], self.num_particles, torch.int32)

    def get_sorted_indices(self):
        torch = _torch()
        return copy_from_device_ptr(self._buffers()[1], self.num_particles, torch.int32)

    def get_inverse_indices(self):
        torch = _torch()
        return copy_from_device_ptr(self._buffers()[2], self.num_particles, torch.int32)

    def reorder_data(self, input_data, output_data, num_elements: int, stream=None):
        elem = input_data.element_size() * (input_data[0].numel() if input_data.dim() > 1 else 1)
        api.call("marsmd_zorder_reorder_data", self._h, _ptr(input_data), _ptr(output_data), num_elements, elem,
                 _stream(stream))

    def remap_indices(self, indices, stream=None):
        api.call("marsmd_zorder_remap_indices", self._h, _ptr(indices), indices.numel(), _stream(stream))

    def update_positions_incremental(self, positions, num_particles: int, stream=None):
        """Keep these device positions (float4 records) as the displacement reference."""
        api.call("marsmd_zorder_update_positions_incremental", self._h, _ptr(positions), num_particles,
                 _stream(stream))

    def compute_max_displacement(self, positions, num_particles: int, box_len=(0.0, 0.0, 0.0), stream=None) -> float:
        """Largest displacement since update_positions_incremental; minimum image on
        axes with box_len > 0."""
        bl = (C.c_float * 3)(*map(float, box_len))
        out = C.c_float(0.0)
        api.call("marsmd_zorder_compute_max_displacement", self._h, _ptr(positions), num_particles, bl, C.byref(out),
                 _stream(stream))
        return out.value


# ---------------------------------------------------------------------------- pairlist
@dataclass
class ExclusionCSR:
    """Symmetric per-particle exclusion CSR (ExclusionView, DeviceExclusions.h:22-75)."""

    offsets: object = None  # device int32 tensor [num_particles + 1]
    excluded: object = None  # device int32 tensor
    num_particles: int = 0
    rigid_body_id: object = None  # device int32 tensor, -1 unattached

    def to_c(self) -> api.ExclusionView:
        v = api.ExclusionView()
        v.offsets = _ptr(self.offsets)
        v.excluded = _ptr(self.excluded)
        v.num_particles = self.num_particles
        v.rigid_body_id = _ptr(self.rigid_body_id)
        return v


def exclusion_csr_from_pairs(pairs: np.ndarray):
    """The per-particle CSR DeviceBondedInteractions::copy_from_host builds
    (src/Interactions/DeviceBondedInteraction.h:111-167), on the host."""
    pairs = np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
    if len(pairs) == 0:
        return np.zeros(1, np.int32), np.zeros(0, np.int32), 0
    lo = pairs.min(axis=1)
    hi = pairs.max(axis=1)
    order = np.lexsort((hi, lo))
    lo, hi = lo[order], hi[order]
    n = int(hi.max()) + 1
    deg = np.bincount(lo, minlength=n) + np.bincount(hi, minlength=n)
    offsets = np.zeros(n + 1, np.int64)
    offsets[1:] = np.cumsum(deg)
    cursor = offsets[:-1].copy()
    excluded = np.zeros(offsets[-1], np.int32)
    for a, b in zip(lo, hi):
        excluded[cursor[a]] = b
        cursor[a] += 1
        excluded[cursor[b]] = a
        cursor[b] += 1
    return offsets.astype(np.int32), excluded, n


class ZOrderPairlist:
    """ZOrderPairlist (ZOrderPairlist.h:27-190) behind the Pairlist interface
    (Pairlist.h:56-252); create_pairlist(ZOrder, res, max_particles, max_pairs)."""

    def __init__(self, max_particles: int, max_pairs: int = 1 << 62, _handle: Optional[int] = None):
        self._owned = _handle is None
        h = C.c_void_p(_handle)
        if self._owned:
            api.call("marsmd_pairlist_create", max_particles, max_pairs, C.byref(h))
        self._h = h
        self._excl = None

    def __del__(self):
        if getattr(self, "_owned", False) and getattr(self, "_h", None):
            api.lib().marsmd_pairlist_destroy(self._h)
            self._h = None

    def set_periodic_box(self, box: PeriodicBox):
        b = box.to_c()
        api.call("marsmd_pairlist_set_periodic_box", self._h, C.byref(b))

    def set_exclusions(self, view: Optional[ExclusionCSR]):
        self._excl = view  # keep the device tensors alive
        api.call("marsmd_pairlist_set_exclusions", self._h, C.byref((view or ExclusionCSR()).to_c()))

    def set_max_pairs_ceiling(self, ceiling: int):
        api.call("marsmd_pairlist_set_max_pairs_ceiling", self._h, ceiling)

    def build_pairlist(self, positions, num_particles: int, pairlist_cutoff: float, type_ids=None, stream=None):
        api.call("marsmd_pairlist_build", self._h, _ptr(positions), _ptr(type_ids), num_particles,
                 float(pairlist_cutoff), _stream(stream))
        self.cutoff = float(pairlist_cutoff)

    def update_pairlist(self, positions, num_particles: int, stream=None):
        self.build_pairlist(positions, num_particles, self.cutoff, stream=stream)

    def needs_update(self, positions, num_particles: int, skin_distance: float, stream=None) -> bool:
        needs = C.c_int32()
        disp = C.c_float()
        api.call("marsmd_pairlist_needs_update", self._h, _ptr(positions), num_particles, float(skin_distance),
                 C.byref(needs), C.byref(disp), _stream(stream))
        self.last_max_displacement = disp.value
        return bool(needs.value)

    def get_num_pairs(self) -> int:
        n = C.c_uint64()
        api.call("marsmd_pairlist_get_num_pairs", self._h, C.byref(n))
        return n.value

    def get_neighbor_pairs(self, stream=None):
        torch = _torch()
        p = C.c_void_p()
        api.call("marsmd_pairlist_get_neighbor_pairs", self._h, C.byref(p), _stream(stream))
        n = self.get_num_pairs()
        return copy_from_device_ptr(p.value or 0, 2 * n, torch.int32).view(n, 2)

    def get_statistics(self) -> api.PairlistStats:
        s = api.PairlistStats()
        api.call("marsmd_pairlist_get_statistics", self._h, C.byref(s))
        return s

    def get_sorter(self) -> ZOrderSort:
        p = C.c_void_p()
        api.call("marsmd_pairlist_get_sorter", self._h, C.byref(p))
        z = ZOrderSort(0, _handle=p.value)
        z.num_particles = self.get_statistics().num_particles
        return z

    def get_reference_cells(self, stream=None):
        torch = _torch()
        b, e, nb = C.c_void_p(), C.c_void_p(), C.c_void_p()
        nc = C.c_uint64()
        slots = C.c_int32()
        api.call("marsmd_pairlist_get_reference_cells", self._h, C.byref(b), C.byref(e), C.byref(nb), C.byref(nc),
                 C.byref(slots), _stream(stream))
        return (copy_from_device_ptr(b.value, nc.value, torch.int32),
                copy_from_device_ptr(e.value, nc.value, torch.int32),
                copy_from_device_ptr(nb.value, nc.value * slots.value, torch.int32).view(nc.value, slots.value))


def create_pairlist(max_particles: int, max_pairs: int = 1 << 62) -> ZOrderPairlist:
    """create_pairlist(PairlistBuilderType::ZOrder, ...) (src/PatchOperation/Pairlist.cpp:7-31)."""
    return ZOrderPairlist(max_particles, max_pairs)


# ---------------------------------------------------------------------------- launchers
class ParticleViewRef(api.ParticleView):
    """ParticleView (DeviceParticle.h:18-28) that also holds the tensors it points
    into, so a view built from temporaries keeps them alive until it is dropped."""

    _keep: tuple = ()


def particle_view(pos=None, force_energy=None, type_id=None, mom=None, flags=None, ids=None) -> api.ParticleView:
    v = ParticleViewRef()
    v.id = _ptr(ids)
    v.type_id = _ptr(type_id)
    v.pos = _ptr(pos)
    v.mom = _ptr(mom)
    v.force_energy = _ptr(force_energy)
    v.flags = _ptr(flags)
    v.attached_rigid_body_id = None
    v._keep = (pos, force_energy, type_id, mom, flags, ids)
    return v


def launch_resolve_pair_tables(pairs, type_ids, term_matrix, num_types: int, pair_tag, num_pairs: int, stream=None):
    api.call("marsmd_launch_resolve_pair_tables", _ptr(pairs), _ptr(type_ids), _ptr(term_matrix), num_types,
             _ptr(pair_tag), num_pairs, _stream(stream))


def launch_pairwise_nonbonded(pairs, view: api.ParticleView, pair_tag, tables, box: PeriodicBox,
                              get_energy: bool, num_pairs: int, cutoff_squared: float,
                              enabled_terms: int = api.PAIR_TERM_TABULATED, stream=None, types=None,
                              params: Optional[PairParams] = None):
    """launch_pairwise_nonbonded (NonBondedInteraction.h:262-296): every enabled term.
    tables: DeviceTables or None; types: a list of ParticleType (charges, softcore
    radius and eps) when analytical terms are enabled."""
    t = make_types(types) if types is not None else None
    api.call("marsmd_launch_pairwise_nonbonded", _ptr(pairs), C.byref(view), _ptr(pair_tag),
             tables.ptr if tables is not None else None, C.byref(t) if t is not None else None,
             C.byref(box.to_c()), 1 if get_energy else 0, num_pairs, float(cutoff_squared), enabled_terms,
             C.byref(params.to_c()) if params is not None else None, _stream(stream))


def launch_analytical_pair_nonbonded(pairs, pos, force_energy, type_ids, pair_tag, types, box: PeriodicBox,
                                     enabled_terms: int, get_energy: bool, num_pairs: int, cutoff_squared: float,
                                     params: Optional[PairParams] = None, stream=None):
    """launch_analytical_pair_nonbonded (AnalyticalPairKernels.h:21-118)."""
    t = make_types(types)
    api.call("marsmd_launch_analytical_pair_nonbonded", _ptr(pairs), _ptr(pos), _ptr(force_energy), _ptr(type_ids),
             _ptr(pair_tag), C.byref(t), C.byref(box.to_c()), enabled_terms,
             C.byref(params.to_c()) if params is not None else None, 1 if get_energy else 0, num_pairs,
             float(cutoff_squared), _stream(stream))


def launch_tabulated_bonds(bonds, pos, force_energy, tables: DeviceTables, table_indices, forms, box: PeriodicBox,
                           get_energy: bool, num_bonds: int, stream=None):
    api.call("marsmd_launch_tabulated_bonds", _ptr(bonds), _ptr(pos), _ptr(force_energy), tables.ptr,
             _ptr(table_indices), _ptr(forms), C.byref(box.to_c()), 1 if get_energy else 0, num_bonds, _stream(stream))


def launch_tabulated_angles(angles4, pos, force_energy, tables: DeviceTables, table_indices, forms,
                            box: PeriodicBox, get_energy: bool, num_angles: int, stream=None):
    api.call("marsmd_launch_tabulated_angles", _ptr(angles4), _ptr(pos), _ptr(force_energy), tables.ptr,
             _ptr(table_indices), _ptr(forms), C.byref(box.to_c()), 1 if get_energy else 0, num_angles, _stream(stream))


def launch_tabulated_dihedrals(dihedrals4, pos, force_energy, tables: DeviceTables, table_indices, forms,
                               box: PeriodicBox, get_energy: bool, num_dihedrals: int, stream=None):
    """launch_tabulated_dihedrals (DeviceBondedInteraction.h:548-573); dihedrals4 is an
    int32 tensor of {i, j, k, l} records, the tables periodic in radians."""
    api.call("marsmd_launch_tabulated_dihedrals", _ptr(dihedrals4), _ptr(pos), _ptr(force_energy), tables.ptr,
             _ptr(table_indices), _ptr(forms), C.byref(box.to_c()), 1 if get_energy else 0, num_dihedrals,
             _stream(stream))


def launch_harmonic_restraints(particle_ids, pos, force_energy, anchors4, spring_constants, box: PeriodicBox,
                               get_energy: bool, num_restraints: int, stream=None):
    """launch_harmonic_restraints (DeviceBondedInteraction.h:595-620): anchors4 float4 records."""
    api.call("marsmd_launch_harmonic_restraints", _ptr(particle_ids), _ptr(pos), _ptr(force_energy), _ptr(anchors4),
             _ptr(spring_constants), C.byref(box.to_c()), 1 if get_energy else 0, num_restraints, _stream(stream))


def _ef(e):
    return (C.c_float * 3)(*(map(float, e) if e is not None else (0.0, 0.0, 0.0)))


def launch_BD(view: api.ParticleView, types: api.ParticleTypes, timestep: float, current_step: int, kT: float,
              num_particles: int, box: PeriodicBox, base_seed: int, base_ctr: int = 0, grids=None,
              electric_field=None, interpolation: int = api.INTERP_LINEAR, stream=None):
    api.call("marsmd_launch_bd", C.byref(view), C.byref(types), float(timestep), current_step, float(kT), num_particles,
             C.byref(box.to_c()), base_seed, base_ctr, _ptr(grids), _ef(electric_field), interpolation, _stream(stream))


def launch_BAOAB(view: api.ParticleView, types: api.ParticleTypes, box: PeriodicBox, timestep: float,
                 current_step: int, kT: float, num_particles: int, base_seed: int, base_ctr: int = 0, stream=None):
    api.call("marsmd_launch_baoab", C.byref(view), C.byref(types), C.byref(box.to_c()), float(timestep), current_step,
             float(kT),
             num_particles, base_seed, base_ctr, _stream(stream))


def launch_BAOAB_LastUpdate(view: api.ParticleView, types: api.ParticleTypes, timestep: float, num_particles: int,
                            stream=None):
    api.call("marsmd_launch_baoab_last_update", C.byref(view), C.byref(types), float(timestep), num_particles,
             _stream(stream))


def launch_PMF(view: api.ParticleView, types: api.ParticleTypes, num_particles: int, grids, electric_field=None,
               interpolation: int = api.INTERP_LINEAR, stream=None):
    api.call("marsmd_launch_pmf", C.byref(view), C.byref(types), num_particles, _ptr(grids), _ef(electric_field),
             interpolation, _stream(stream))


def philox_draw(seed: int, base_ctr: int, global_seed: int, ctr1: int, draws: int, n: int, stream=None):
    torch = _torch()
    out = torch.empty((n, draws, 4), dtype=torch.int32, device="cuda")
    api.call("marsmd_philox_draw", seed, base_ctr, global_seed, ctr1, draws, n, _ptr(out), _stream(stream))
    return out


def device_grid_views(grids: Sequence[Grid]):
    """Upload grids; returns (device record tensor, keep-alive list)."""
    torch = _torch()
    datas = [torch.from_numpy(g.values.reshape(-1)).cuda() for g in grids]
    recs = (api.GridView * max(1, len(grids)))()
    for i, g in enumerate(grids):
        recs[i] = g.to_c(datas[i].data_ptr(), i)
    dev = torch.frombuffer(bytearray(bytes(recs)), dtype=torch.uint8).cuda()
    return dev, datas
