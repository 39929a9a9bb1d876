"""marsmd — MI355X (gfx950) implementation of MARS-MD's per-step particle hot path.

The compute lives in libmarsmd_hip.so (hand-written HIP kernels behind the C ABI of
include/marsmd.h); this package holds the reference-shaped Python front end.
"""
from . import _capi
from ._capi import MarsError, lib
from .core import (
    DeviceTables,
    ExclusionCSR,
    Grid,
    PairParams,
    ParticleType,
    PeriodicBox,
    TabulatedPotential,
    ZOrderPairlist,
    ZOrderSort,
    create_pairlist,
    device_grid_views,
    exclusion_csr_from_pairs,
    launch_BAOAB,
    launch_BAOAB_LastUpdate,
    launch_BD,
    launch_PMF,
    launch_analytical_pair_nonbonded,
    launch_pairwise_nonbonded,
    launch_resolve_pair_tables,
    launch_tabulated_angles,
    launch_tabulated_dihedrals,
    launch_harmonic_restraints,
    launch_tabulated_bonds,
    make_pair_tag,
    make_types,
    particle_view,
    philox_draw,
)
from .patch import Patch
from .system import RunParams, SystemSpec

__all__ = [
    "DeviceTables", "ExclusionCSR", "Grid", "MarsError", "PairParams", "ParticleType", "Patch", "PeriodicBox", "RunParams",
    "SystemSpec", "TabulatedPotential", "ZOrderPairlist", "ZOrderSort", "create_pairlist", "device_grid_views",
    "exclusion_csr_from_pairs", "launch_BAOAB", "launch_BAOAB_LastUpdate", "launch_BD", "launch_PMF",
    "launch_analytical_pair_nonbonded", "launch_pairwise_nonbonded", "launch_resolve_pair_tables", "launch_tabulated_angles", "launch_tabulated_bonds",
    "launch_tabulated_dihedrals", "launch_harmonic_restraints",
    "lib", "make_pair_tag", "make_types", "particle_view", "philox_draw",
]
