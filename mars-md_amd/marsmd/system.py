"""A system description shared by the Patch (HIP) and by whoever checks it.

SystemSpec holds what the reference's SimSystem / SystemState / TablesRegistry /
BondedInteractions / GridManager hand to a Patch for the hot path: box, particle
types, pair tables and their type-pair tag matrix, bonded topology with tables,
exclusions, grids, initial particles and the run parameters of SimManager::run.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import _capi as api
from .core import Grid, PairParams, ParticleType, PeriodicBox, TabulatedPotential, make_pair_tag, make_types


@dataclass
class RunParams:
    """SimManager per-step parameters (SimManager.cpp:271-504; ConfigParser keys
    cutoff, pairlistDistance, decompPeriod, reorderPeriod, timestep, temperature)."""

    cutoff: float
    pairlist_cutoff: float
    rebuild_period: int = 20
    reorder_period: int = 0
    timestep: float = 1e-4
    kT: float = 0.593
    integrator: int = api.INTEGRATOR_BROWNIAN
    interpolation: int = api.INTERP_LINEAR
    electric_field: Sequence[float] = (0.0, 0.0, 0.0)
    energy_period: int = 0

    def to_c(self) -> api.RunParams:
        p = api.RunParams()
        p.pairlist_cutoff = self.pairlist_cutoff
        p.cutoff = self.cutoff
        p.rebuild_period = self.rebuild_period
        p.reorder_period = self.reorder_period
        p.timestep = self.timestep
        p.kT = self.kT
        p.integrator = self.integrator
        p.interpolation = self.interpolation
        for a in range(3):
            p.electric_field[a] = float(self.electric_field[a])
        p.energy_period = self.energy_period
        return p


@dataclass
class SystemSpec:
    box: PeriodicBox
    types: List[ParticleType]
    positions: np.ndarray  # (N, 3) or (N, 4) float32
    type_ids: np.ndarray  # (N,) int32
    params: RunParams
    nb_tables: List[TabulatedPotential] = field(default_factory=list)
    term_matrix: Optional[np.ndarray] = None  # (T, T) uint32 pair tags
    enabled_terms: int = api.PAIR_TERM_TABULATED
    momenta: Optional[np.ndarray] = None
    bonds: Optional[np.ndarray] = None  # (B, 2)
    bond_table: Optional[np.ndarray] = None
    bond_tables: List[TabulatedPotential] = field(default_factory=list)
    angles: Optional[np.ndarray] = None  # (A, 3)
    angle_table: Optional[np.ndarray] = None
    angle_tables: List[TabulatedPotential] = field(default_factory=list)
    exclusions: Optional[np.ndarray] = None  # (E, 2)
    grids: List[Grid] = field(default_factory=list)
    dihedrals: Optional[np.ndarray] = None  # (D, 4)
    dihedral_table: Optional[np.ndarray] = None
    dihedral_tables: List[TabulatedPotential] = field(default_factory=list)  # periodic, radians
    restraint_ids: Optional[np.ndarray] = None  # (R,) particle per harmonic restraint
    restraint_anchors: Optional[np.ndarray] = None  # (R, 3)
    restraint_k: Optional[np.ndarray] = None  # (R,)
    pair_params: Optional[PairParams] = None  # analytical pair terms (None: the defaults)
    seed: int = 42
    patch_id: int = 0
    name: str = ""

    @property
    def n(self) -> int:
        return len(self.type_ids)

    def pos4(self) -> np.ndarray:
        p = np.zeros((self.n, 4), np.float32)
        p[:, :3] = np.asarray(self.positions, np.float32)[:, :3]
        return p

    def mom4(self) -> Optional[np.ndarray]:
        if self.momenta is None:
            return None
        m = np.zeros((self.n, 4), np.float32)
        m[:, :3] = np.asarray(self.momenta, np.float32)[:, :3]
        return m

    def tag_matrix(self) -> np.ndarray:
        if self.term_matrix is not None:
            return np.ascontiguousarray(self.term_matrix, dtype=np.uint32)
        T = len(self.types)
        # default: pair (i, j) uses table index of the upper-triangle enumeration
        m = np.zeros((T, T), np.uint32)
        k = 0
        for i in range(T):
            for j in range(i, T):
                m[i, j] = m[j, i] = make_pair_tag(api.PAIR_TERM_TABULATED, k)
                k += 1
        return m

    def to_desc(self, capacity: Optional[int] = None):
        """marsmd_patch_desc with host pointers; returns (desc, keepalive)."""
        keep = []

        def arr(a, dtype):
            a = np.ascontiguousarray(a, dtype=dtype)
            keep.append(a)
            return a.ctypes.data

        def tables(lst):
            recs = (api.Table * max(1, len(lst)))()
            for i, t in enumerate(lst):
                knots = np.ascontiguousarray(t.pot, np.float32)
                keep.append(knots)
                recs[i] = t.to_c(knots.ctypes.data)
            keep.append(recs)
            return C.addressof(recs)

        d = api.PatchDesc()
        d.capacity = capacity or max(self.n, 1)
        d.patch_id = self.patch_id
        d.box = self.box.to_c()
        d.types = make_types(self.types)
        d.term_matrix = arr(self.tag_matrix(), np.uint32)
        d.enabled_terms = self.enabled_terms
        d.num_nb_tables = len(self.nb_tables)
        d.nb_tables = tables(self.nb_tables)
        if self.bonds is not None and len(self.bonds):
            d.num_bonds = len(self.bonds)
            d.bonds = arr(self.bonds, np.int32)
            bt = self.bond_table if self.bond_table is not None else np.zeros(len(self.bonds))
            d.bond_table = arr(bt, np.int32)
        if self.angles is not None and len(self.angles):
            d.num_angles = len(self.angles)
            d.angles = arr(self.angles, np.int32)
            at = self.angle_table if self.angle_table is not None else np.zeros(len(self.angles))
            d.angle_table = arr(at, np.int32)
        d.num_bond_tables = len(self.bond_tables)
        d.bond_tables = tables(self.bond_tables)
        d.num_angle_tables = len(self.angle_tables)
        d.angle_tables = tables(self.angle_tables)
        if self.exclusions is not None and len(self.exclusions):
            d.num_exclusions = len(self.exclusions)
            d.exclusions = arr(self.exclusions, np.int32)
        if self.dihedrals is not None and len(self.dihedrals):
            d.num_dihedrals = len(self.dihedrals)
            d.dihedrals = arr(self.dihedrals, np.int32)
            dt = self.dihedral_table if self.dihedral_table is not None else np.zeros(len(self.dihedrals))
            d.dihedral_table = arr(dt, np.int32)
        d.num_dihedral_tables = len(self.dihedral_tables)
        d.dihedral_tables = tables(self.dihedral_tables)
        if self.restraint_ids is not None and len(self.restraint_ids):
            r = len(self.restraint_ids)
            d.num_restraints = r
            d.restraint_ids = arr(self.restraint_ids, np.int32)
            anchors = np.zeros((r, 4), np.float32)
            anchors[:, :3] = np.asarray(self.restraint_anchors, np.float32)[:, :3]
            d.restraint_anchors = arr(anchors, np.float32)
            d.restraint_k = arr(self.restraint_k, np.float32)
        if self.pair_params is not None:
            pp = self.pair_params.to_c()
            keep.append(pp)
            d.pair_params = C.addressof(pp)
        if self.grids:
            recs = (api.GridView * len(self.grids))()
            for i, g in enumerate(self.grids):
                keep.append(g.values)
                recs[i] = g.to_c(g.values.ctypes.data, i)
            keep.append(recs)
            d.num_grids = len(self.grids)
            d.grids = C.addressof(recs)
        keep.append(d)
        return d, keep
