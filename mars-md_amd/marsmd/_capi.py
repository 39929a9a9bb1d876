"""ctypes mirror of include/marsmd.h (the C ABI of libmarsmd_hip.so).

Only structure layouts and prototypes live here; marsmd.core wraps them in the
reference-shaped classes. Loading never touches the GPU.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

MAX_TYPES = 16
MAX_PMF_TERMS = 32

PAIR_TERM_COULOMB = 1 << 0
PAIR_TERM_DEBYE_HUCKEL = 1 << 1
PAIR_TERM_ONCK = 1 << 2
PAIR_TERM_GAUSSIAN = 1 << 3
PAIR_TERM_SOFTCORE = 1 << 4
PAIR_TERM_TABULATED = 1 << 5
PAIR_TABLE_SHIFT = 8
INTEGRATOR_BROWNIAN = 0
INTEGRATOR_LANGEVIN = 1
INTERP_LINEAR = 1
INTERP_CUBIC = 3
BC_DIRICHLET = 0
BC_NEUMANN = 1
BC_PERIODIC = 2
FORM_TABULATED = 1


class Box(C.Structure):
    _fields_ = [("size", C.c_float * 3), ("origin", C.c_float * 3), ("periodic", C.c_int32 * 3)]


class Table(C.Structure):
    _fields_ = [
        ("pot", C.c_void_p),
        ("step_inv", C.c_float),
        ("size", C.c_uint32),
        ("start", C.c_float),
        ("is_periodic", C.c_int32),
    ]


class ExclusionView(C.Structure):
    _fields_ = [
        ("offsets", C.c_void_p),
        ("excluded", C.c_void_p),
        ("num_particles", C.c_uint64),
        ("rigid_body_id", C.c_void_p),
    ]


class GridTerm(C.Structure):
    _fields_ = [
        ("grid_id", C.c_int32),
        ("scale", C.c_float),
        ("scale_slope", C.c_float),
        ("boundary_condition", C.c_int32),
    ]


class GridView(C.Structure):
    _fields_ = [
        ("data", C.c_void_p),
        ("origin", C.c_float * 3),
        ("basis", C.c_float * 9),
        ("basis_inv", C.c_float * 9),
        ("dims", C.c_uint64 * 3),
        ("grid_id", C.c_int32),
        ("boundary_condition", C.c_int32),
    ]


class ParticleTypes(C.Structure):
    _fields_ = [
        ("num_types", C.c_int32),
        ("mass", C.c_float * MAX_TYPES),
        ("charge", C.c_float * MAX_TYPES),
        ("radius", C.c_float * MAX_TYPES),
        ("eps", C.c_float * MAX_TYPES),
        ("diffusion", (C.c_float * 3) * MAX_TYPES),
        ("trans_damping", (C.c_float * 3) * MAX_TYPES),
        ("pmf_grid_offset", C.c_int32 * MAX_TYPES),
        ("pmf_grid_count", C.c_int32 * MAX_TYPES),
        ("pmf_grid_terms", GridTerm * MAX_PMF_TERMS),
        ("force_grid_id", (C.c_int32 * 3) * MAX_TYPES),
        ("force_grid_scale", (C.c_float * 3) * MAX_TYPES),
    ]


class PairParams(C.Structure):
    _fields_ = [
        ("debye_length", C.c_float),
        ("dielectric", C.c_float),
        ("onck_kappa", C.c_float),
        ("onck_sz", C.c_float),
        ("onck_z", C.c_float),
        ("gaussian_amp", C.c_float),
        ("gaussian_sigma", C.c_float),
        ("gaussian_cutoff_squared", C.c_float),
    ]


class ParticleView(C.Structure):
    _fields_ = [
        ("id", C.c_void_p),
        ("type_id", C.c_void_p),
        ("pos", C.c_void_p),
        ("mom", C.c_void_p),
        ("force_energy", C.c_void_p),
        ("flags", C.c_void_p),
        ("attached_rigid_body_id", C.c_void_p),
    ]


class PairlistStats(C.Structure):
    _fields_ = [
        ("num_particles", C.c_uint64),
        ("num_pairs", C.c_uint64),
        ("average_neighbors_per_particle", C.c_double),
        ("max_neighbors_per_particle", C.c_uint32),
        ("build_time_ms", C.c_double),
        ("cell_grid_bits", C.c_int32),
        ("cell_radii", C.c_int32 * 3),
        ("neighbors_per_cell", C.c_int32),
        ("build_grid_bits", C.c_int32),
        ("build_radii", C.c_int32 * 3),
        ("reference_grid_complete", C.c_int32),
        ("pair_capacity", C.c_uint64),
    ]


class PatchDesc(C.Structure):
    _fields_ = [
        ("capacity", C.c_uint64),
        ("patch_id", C.c_uint64),
        ("box", Box),
        ("types", ParticleTypes),
        ("term_matrix", C.c_void_p),
        ("enabled_terms", C.c_uint32),
        ("num_nb_tables", C.c_int32),
        ("nb_tables", C.c_void_p),
        ("num_bonds", C.c_uint64),
        ("bonds", C.c_void_p),
        ("bond_table", C.c_void_p),
        ("num_angles", C.c_uint64),
        ("angles", C.c_void_p),
        ("angle_table", C.c_void_p),
        ("num_bond_tables", C.c_int32),
        ("bond_tables", C.c_void_p),
        ("num_angle_tables", C.c_int32),
        ("angle_tables", C.c_void_p),
        ("num_exclusions", C.c_uint64),
        ("exclusions", C.c_void_p),
        ("num_grids", C.c_int32),
        ("grids", C.c_void_p),
        ("num_dihedrals", C.c_uint64),
        ("dihedrals", C.c_void_p),
        ("dihedral_table", C.c_void_p),
        ("num_dihedral_tables", C.c_int32),
        ("dihedral_tables", C.c_void_p),
        ("num_restraints", C.c_uint64),
        ("restraint_ids", C.c_void_p),
        ("restraint_anchors", C.c_void_p),
        ("restraint_k", C.c_void_p),
        ("pair_params", C.c_void_p),
    ]


class BatchDesc(C.Structure):
    _fields_ = [
        ("num_replicas", C.c_uint64),
        ("num_particles", C.c_void_p),
        ("boxes", C.c_void_p),
        ("patch_ids", C.c_void_p),
        ("types", ParticleTypes),
        ("term_matrix", C.c_void_p),
        ("enabled_terms", C.c_uint32),
        ("num_nb_tables", C.c_int32),
        ("nb_tables", C.c_void_p),
        ("pair_params", C.c_void_p),
    ]


class BatchLaunch(C.Structure):
    _fields_ = [("blocks_per_cu", C.c_int32), ("use_graph", C.c_int32)]


BATCH_BUCKETS = 32


class BatchCells(C.Structure):
    _fields_ = [("home_target", C.c_int32 * BATCH_BUCKETS), ("fine_offset", C.c_int32 * BATCH_BUCKETS)]


class PatchLaunch(C.Structure):
    _fields_ = [("blocks_per_cu", C.c_int32), ("split_pmf", C.c_int32)]


class FMDesc(C.Structure):
    _fields_ = [
        ("box", Box),
        ("num_particles", C.c_uint64),
        ("num_snapshots", C.c_uint64),
        ("positions", C.c_void_p),
        ("type_ids", C.c_void_p),
        ("num_types", C.c_int32),
        ("term_matrix", C.c_void_p),
        ("num_tables", C.c_int32),
        ("tables", C.c_void_p),
        ("cutoff", C.c_float),
        ("pairlist_cutoff", C.c_float),
    ]


class RunParams(C.Structure):
    _fields_ = [
        ("pairlist_cutoff", C.c_float),
        ("cutoff", C.c_float),
        ("rebuild_period", C.c_uint64),
        ("reorder_period", C.c_int64),
        ("timestep", C.c_float),
        ("kT", C.c_float),
        ("integrator", C.c_int32),
        ("interpolation", C.c_int32),
        ("electric_field", C.c_float * 3),
        ("energy_period", C.c_uint64),
    ]


_P = C.c_void_p
_U64 = C.c_uint64
_I32 = C.c_int32
_U32 = C.c_uint32
_F = C.c_float
_PP = C.POINTER(C.c_void_p)
_FP3 = C.POINTER(C.c_float)

# name -> (restype, argtypes)
PROTOTYPES = {
    "marsmd_last_error": (C.c_char_p, []),
    "marsmd_version": (C.c_char_p, []),
    "marsmd_zorder_create": (_I32, [_U64, _PP]),
    "marsmd_zorder_destroy": (_I32, [_P]),
    "marsmd_zorder_sort_particles": (_I32, [_P, _P, _U64, _FP3, _FP3, _P]),
    "marsmd_zorder_get_buffers": (_I32, [_P, _PP, _PP, _PP]),
    "marsmd_zorder_reorder_data": (_I32, [_P, _P, _P, _U64, _U32, _P]),
    "marsmd_zorder_remap_indices": (_I32, [_P, _P, _U64, _P]),
    "marsmd_zorder_update_positions_incremental": (_I32, [_P, _P, _U64, _P]),
    "marsmd_zorder_compute_max_displacement": (_I32, [_P, _P, _U64, _P, _P, _P]),
    "marsmd_radix_sort_pairs": (_I32, [_P, _P, _P, _P, _U64, _I32, _P, _U64, _P]),
    "marsmd_radix_sort_scratch_bytes": (_U64, [_U64]),
    "marsmd_radix_sort_pairs_static": (_I32, [_P, _P, _P, _P, _U64, _I32, _P]),
    "marsmd_radix_sort_static_scratch_bytes": (_U64, []),
    "marsmd_pairlist_create": (_I32, [_U64, _U64, _PP]),
    "marsmd_pairlist_destroy": (_I32, [_P]),
    "marsmd_pairlist_set_periodic_box": (_I32, [_P, C.POINTER(Box)]),
    "marsmd_pairlist_set_exclusions": (_I32, [_P, C.POINTER(ExclusionView)]),
    "marsmd_pairlist_set_max_pairs_ceiling": (_I32, [_P, _U64]),
    "marsmd_pairlist_build": (_I32, [_P, _P, _P, _U64, _F, _P]),
    "marsmd_pairlist_needs_update": (_I32, [_P, _P, _U64, _F, C.POINTER(_I32), C.POINTER(_F), _P]),
    "marsmd_pairlist_get_num_pairs": (_I32, [_P, C.POINTER(_U64)]),
    "marsmd_pairlist_get_neighbor_pairs": (_I32, [_P, _PP, _P]),
    "marsmd_pairlist_get_statistics": (_I32, [_P, C.POINTER(PairlistStats)]),
    "marsmd_pairlist_get_sorter": (_I32, [_P, _PP]),
    "marsmd_pairlist_get_reference_cells": (_I32, [_P, _PP, _PP, _PP, C.POINTER(_U64), C.POINTER(_I32), _P]),
    "marsmd_launch_resolve_pair_tables": (_I32, [_P, _P, _P, _U64, _P, _U64, _P]),
    "marsmd_launch_pairwise_nonbonded": (_I32, [_P, C.POINTER(ParticleView), _P, _P, C.POINTER(ParticleTypes),
                                                C.POINTER(Box), _I32, _U64, _F, _U32, C.POINTER(PairParams), _P]),
    "marsmd_launch_analytical_pair_nonbonded": (_I32, [_P, _P, _P, _P, _P, C.POINTER(ParticleTypes), C.POINTER(Box),
                                                       _U32, C.POINTER(PairParams), _I32, _U64, _F, _P]),
    "marsmd_launch_tabulated_bonds": (_I32, [_P, _P, _P, _P, _P, _P, C.POINTER(Box), _I32, _U64, _P]),
    "marsmd_launch_tabulated_angles": (_I32, [_P, _P, _P, _P, _P, _P, C.POINTER(Box), _I32, _U64, _P]),
    "marsmd_launch_tabulated_dihedrals": (_I32, [_P, _P, _P, _P, _P, _P, C.POINTER(Box), _I32, _U64, _P]),
    "marsmd_launch_harmonic_restraints": (_I32, [_P, _P, _P, _P, _P, C.POINTER(Box), _I32, _U64, _P]),
    "marsmd_launch_bd": (_I32, [C.POINTER(ParticleView), C.POINTER(ParticleTypes), _F, _U64, _F, _U64, C.POINTER(Box),
                                _U64, _U32, _P,
                                _FP3, _I32, _P]),
    "marsmd_launch_baoab": (_I32, [C.POINTER(ParticleView), C.POINTER(ParticleTypes), C.POINTER(Box), _F, _U64, _F, _U64,
                                   _U64, _U32, _P]),
    "marsmd_launch_baoab_last_update": (_I32, [C.POINTER(ParticleView), C.POINTER(ParticleTypes), _F, _U64, _P]),
    "marsmd_launch_pmf": (_I32, [C.POINTER(ParticleView), C.POINTER(ParticleTypes), _U64, _P, _FP3, _I32, _P]),
    "marsmd_philox_draw": (_I32, [_U64, _U32, _U32, _U32, _U32, _U64, _P, _P]),
    "marsmd_selftest_pair_sqrt": (_I32, [_U32, _U32, _P, _P]),
    "marsmd_patch_create": (_I32, [C.POINTER(PatchDesc), _PP]),
    "marsmd_patch_destroy": (_I32, [_P]),
    "marsmd_patch_set_base_seed": (_I32, [_P, _U64]),
    "marsmd_patch_copy_particles_from_host": (_I32, [_P, _P, _P, _P, _U64]),
    "marsmd_patch_copy_particles_to_host": (_I32, [_P, _P, _P, _P, _P, _P, _U64]),
    "marsmd_patch_calculate_nonbonded_forces": (_I32, [_P, _F, _F, _U64, _U64, _FP3, _I32, _I32, _P]),
    "marsmd_patch_calculate_bonded_forces": (_I32, [_P, _I32, _P]),
    "marsmd_patch_integrate_motion": (_I32, [_P, _F, _F, _I32, _U64, _P]),
    "marsmd_patch_finish_deferred_kick": (_I32, [_P, _F, _P]),
    "marsmd_patch_has_deferred_kick": (_I32, [_P, _P]),
    "marsmd_patch_reorder_particles": (_I32, [_P, _P]),
    "marsmd_patch_get_pairlist": (_I32, [_P, _PP]),
    "marsmd_patch_run": (_I32, [_P, C.POINTER(RunParams), _U64, _U64, _P]),
    "marsmd_patch_set_kernel_timing": (_I32, [_P, _I32]),
    "marsmd_patch_get_kernel_timing": (_I32, [_P, C.POINTER(C.c_double), C.POINTER(_U64)]),
    "marsmd_patch_set_launch": (_I32, [_P, C.POINTER(PatchLaunch)]),
    "marsmd_patch_get_launch": (_I32, [_P, C.POINTER(PatchLaunch)]),
    "marsmd_patch_set_cells": (_I32, [_P, _I32, _I32]),
    "marsmd_patch_get_cells": (_I32, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "marsmd_patch_set_pruning": (_I32, [_P, _I32, C.c_float]),
    "marsmd_patch_get_pruning": (_I32, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_float), C.POINTER(C.c_int32),
                                        C.POINTER(C.c_float)]),
    "marsmd_patch_autotune_cells": (_I32, [_P, C.POINTER(RunParams), _U64, _U64, C.POINTER(C.c_double), _P]),
    "marsmd_patch_autotune": (_I32, [_P, C.POINTER(RunParams), _U64, _U64, C.POINTER(PatchLaunch),
                                     C.POINTER(C.c_double), _P]),
    "marsmd_batch_create": (_I32, [C.POINTER(BatchDesc), C.POINTER(_P)]),
    "marsmd_batch_destroy": (_I32, [_P]),
    "marsmd_batch_set_base_seed": (_I32, [_P, _U64]),
    "marsmd_batch_copy_particles_from_host": (_I32, [_P, _P, _P, _P, _U64]),
    "marsmd_batch_copy_particles_to_host": (_I32, [_P, _P, _P, _P, _P, _P, _U64]),
    "marsmd_batch_run": (_I32, [_P, C.POINTER(RunParams), _U64, _U64, _P]),
    "marsmd_batch_finish_deferred_kick": (_I32, [_P, C.c_float, _P]),
    "marsmd_batch_get_neighbor_pairs": (_I32, [_P, _U64, C.POINTER(_P), C.POINTER(_U64), _P]),
    "marsmd_batch_set_kernel_timing": (_I32, [_P, _I32]),
    "marsmd_batch_get_kernel_timing": (_I32, [_P, C.POINTER(C.c_double), C.POINTER(_U64)]),
    "marsmd_batch_set_launch": (_I32, [_P, C.POINTER(BatchLaunch)]),
    "marsmd_batch_get_launch": (_I32, [_P, C.POINTER(BatchLaunch)]),
    "marsmd_batch_autotune": (_I32, [_P, C.POINTER(RunParams), _U64, _U64, C.POINTER(BatchLaunch),
                                     C.POINTER(C.c_double), _P]),
    "marsmd_batch_set_cells": (_I32, [_P, C.POINTER(BatchCells)]),
    "marsmd_batch_get_cells": (_I32, [_P, C.POINTER(BatchCells)]),
    "marsmd_batch_autotune_cells": (_I32, [_P, C.POINTER(RunParams), _U64, _U64, C.POINTER(C.c_double), _P]),
    "marsmd_fm_create": (_I32, [C.POINTER(FMDesc), C.POINTER(_P)]),
    "marsmd_fm_destroy": (_I32, [_P]),
    "marsmd_fm_num_pairs": (_I32, [_P, C.POINTER(_U64)]),
    "marsmd_fm_forces": (_I32, [_P, _P, _P, _P]),
    "marsmd_fm_mse": (_I32, [_P, _P, _P, _P, _P, _P]),
    "marsmd_fm_adjoint": (_I32, [_P, _P, _P, _P, _P]),
    "marsmd_fm_forces_sorted": (_I32, [_P, _P, _P, _P]),
    "marsmd_fm_loss_sorted": (_I32, [_P, _P, _P, _P, _P, _P]),
    "marsmd_fm_adjoint_sorted": (_I32, [_P, _P, _P, _P, _P]),
    "marsmd_fm_to_sorted": (_I32, [_P, _P, _I32, _P, _P]),
}

LIB_PATH = Path(__file__).resolve().parent / "libmarsmd_hip.so"
_lib = None


class MarsError(RuntimeError):
    """MARS::Exception counterpart: raised when an entry point returns non-zero."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[marsmd status {code}] {msg}")
        self.code = code


def lib():
    """Load libmarsmd_hip.so (raises if it was not built; there is no fallback)."""
    global _lib
    if _lib is None:
        path = Path(os.environ.get("MARSMD_LIB", str(LIB_PATH)))
        if not path.exists():
            raise RuntimeError(f"{path} is missing: build it with __graft_entry__.build() "
                               "or python mars-md_amd/build.py")
        handle = C.CDLL(str(path))
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(status: int) -> None:
    if status != 0:
        raise MarsError(status, lib().marsmd_last_error().decode())


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args))
