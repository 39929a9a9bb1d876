"""Patch (src/PatchOperation/Patch.h:48-645) over the C ABI."""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _capi as api
from .core import ZOrderPairlist, _stream
from .system import RunParams, SystemSpec


class Patch:
    """One patch on one GPU. State lives in HBM inside the library; particle arrays
    cross the boundary in the reference's canonical order (float4 records)."""

    def __init__(self, spec: SystemSpec, capacity: Optional[int] = None, load_particles: bool = True):
        desc, keep = spec.to_desc(capacity)
        h = C.c_void_p()
        api.call("marsmd_patch_create", C.byref(desc), C.byref(h))
        self._h = h
        self.spec = spec
        self.n = 0
        api.call("marsmd_patch_set_base_seed", self._h, spec.seed)
        if load_particles:
            self.copy_particles_from_host(spec.pos4(), spec.mom4(), spec.type_ids)

    def __del__(self):
        if getattr(self, "_h", None):
            api.lib().marsmd_patch_destroy(self._h)
            self._h = None

    def set_base_seed(self, seed: int):
        api.call("marsmd_patch_set_base_seed", self._h, seed)

    def copy_particles_from_host(self, pos4: np.ndarray, mom4: Optional[np.ndarray], type_ids: np.ndarray):
        pos4 = np.ascontiguousarray(pos4, np.float32)
        mom = None if mom4 is None else np.ascontiguousarray(mom4, np.float32)
        types = np.ascontiguousarray(type_ids, np.int32)
        api.call("marsmd_patch_copy_particles_from_host", self._h, pos4.ctypes.data,
                 None if mom is None else mom.ctypes.data, types.ctypes.data, len(types))
        self.n = len(types)

    def copy_particles_to_host(self):
        n = self.n
        pos = np.zeros((n, 4), np.float32)
        mom = np.zeros((n, 4), np.float32)
        fe = np.zeros((n, 4), np.float32)
        ty = np.zeros(n, np.int32)
        ids = np.zeros(n, np.int32)
        api.call("marsmd_patch_copy_particles_to_host", self._h, pos.ctypes.data, mom.ctypes.data, fe.ctypes.data,
                 ty.ctypes.data, ids.ctypes.data, n)
        return {"pos": pos, "mom": mom, "force": fe, "type": ty, "id": ids}

    def calculate_nonbonded_forces(self, pairlist_cutoff: float, interaction_cutoff: float, step: int,
                                   rebuild_period: int, electric_field=(0.0, 0.0, 0.0),
                                   interpolation: int = api.INTERP_LINEAR, compute_energy: bool = False,
                                   stream=None):
        ef = (C.c_float * 3)(*map(float, electric_field))
        api.call("marsmd_patch_calculate_nonbonded_forces", self._h, float(pairlist_cutoff),
                 float(interaction_cutoff), step, rebuild_period, ef, interpolation, 1 if compute_energy else 0,
                 _stream(stream))

    def calculate_bonded_forces(self, compute_energy: bool = False, stream=None):
        api.call("marsmd_patch_calculate_bonded_forces", self._h, 1 if compute_energy else 0, _stream(stream))

    def integrate_motion(self, dt: float, kT: float, integrator: int, step: int, stream=None):
        api.call("marsmd_patch_integrate_motion", self._h, float(dt), float(kT), integrator, step, _stream(stream))

    def finish_deferred_kick(self, dt: float, stream=None):
        api.call("marsmd_patch_finish_deferred_kick", self._h, float(dt), _stream(stream))

    def has_deferred_kick(self) -> bool:
        """Patch::has_deferred_kick (Patch.h:300-304)."""
        v = C.c_int32(0)
        api.call("marsmd_patch_has_deferred_kick", self._h, C.byref(v))
        return bool(v.value)

    def reorder_particles(self, stream=None):
        api.call("marsmd_patch_reorder_particles", self._h, _stream(stream))

    def get_pairlist(self) -> ZOrderPairlist:
        p = C.c_void_p()
        api.call("marsmd_patch_get_pairlist", self._h, C.byref(p))
        return ZOrderPairlist(0, _handle=p.value)

    def run(self, num_steps: int, first_step: int = 0, params: Optional[RunParams] = None, stream=None):
        p = (params or self.spec.params).to_c()
        api.call("marsmd_patch_run", self._h, C.byref(p), first_step, num_steps, _stream(stream))
        if num_steps > 0:
            self._prepared = True  # the first step prepared the bonded topology

    def set_kernel_timing(self, enable: bool = True):
        """Record HIP events around each fused step kernel launched by run()."""
        api.call("marsmd_patch_set_kernel_timing", self._h, 1 if enable else 0)

    def kernel_timing(self):
        """(summed duration in ms, launches) of the fused step kernels timed so far."""
        ms, k = C.c_double(0.0), C.c_uint64(0)
        api.call("marsmd_patch_get_kernel_timing", self._h, C.byref(ms), C.byref(k))
        return ms.value, k.value

    def set_launch(self, blocks_per_cu: int = 0, split_pmf: bool = False):
        """Launch configuration of run()'s step: persistent workgroups per CU (0: as many
        as fit) and, with grids, the position-dependent force from a kernel of its own.
        Results are the same bits whatever the configuration."""
        lc = api.PatchLaunch(int(blocks_per_cu), 1 if split_pmf else 0)
        api.call("marsmd_patch_set_launch", self._h, C.byref(lc))

    def get_launch(self) -> dict:
        lc = api.PatchLaunch()
        api.call("marsmd_patch_get_launch", self._h, C.byref(lc))
        return {"blocks_per_cu": lc.blocks_per_cu, "split_pmf": bool(lc.split_pmf)}

    def autotune(self, num_steps: int, first_step: int = 0, params: Optional[RunParams] = None, stream=None) -> dict:
        """Steps first_step+1 .. first_step+num_steps (real steps, as run) while timing
        every launch candidate on the step kernels, and keeps the fastest. Returns it
        with its ms per step."""
        p = (params or self.spec.params).to_c()
        lc = api.PatchLaunch()
        ms = C.c_double(0.0)
        api.call("marsmd_patch_autotune", self._h, C.byref(p), first_step, num_steps, C.byref(lc), C.byref(ms),
                 _stream(stream))
        return {"blocks_per_cu": lc.blocks_per_cu, "split_pmf": bool(lc.split_pmf), "ms_per_step": ms.value}

    def set_cells(self, home_target: int = 24, fine_offset: int = 2):
        """Cell configuration of the pairlist walk (home cells holding home_target
        particles on average, candidate cells fine_offset levels finer); the pair set
        does not depend on it. Applies from the next build."""
        api.call("marsmd_patch_set_cells", self._h, int(home_target), int(fine_offset))

    def get_cells(self) -> tuple:
        a, b = C.c_int32(0), C.c_int32(0)
        api.call("marsmd_patch_get_cells", self._h, C.byref(a), C.byref(b))
        return a.value, b.value

    def set_pruning(self, mode: int = 3, margin: float = -1.0):
        """Pruned rows of the fused step: 0 off (the default), 1 rows placed by length
        within each 256-slot block, 2 placed so and pruned to cutoff + 2 margin, 3
        pruned in place (margin < 0: adapted to the displacement of the previous list
        life). The forces are the full rows' bits in every mode."""
        api.call("marsmd_patch_set_pruning", self._h, int(mode), float(margin))

    def get_pruning(self) -> dict:
        """mode, margin of the live pruned rows, whether they keep every entry, and the
        largest displacement since the prune recorded by the fused steps."""
        mode, margin, keep_all, disp = C.c_int32(0), C.c_float(0.0), C.c_int32(0), C.c_float(0.0)
        api.call("marsmd_patch_get_pruning", self._h, C.byref(mode), C.byref(margin), C.byref(keep_all),
                 C.byref(disp))
        return {"mode": mode.value, "margin": margin.value, "keep_all": bool(keep_all.value),
                "displacement": disp.value}

    def autotune_cells(self, num_steps: int, first_step: int = 0, params: Optional[RunParams] = None,
                       stream=None) -> dict:
        """Steps first_step+1 .. first_step+num_steps after choosing the walk's cell
        configuration, timed re-walking the first build's positions."""
        p = (params or self.spec.params).to_c()
        ms = C.c_double(0.0)
        api.call("marsmd_patch_autotune_cells", self._h, C.byref(p), first_step, num_steps, C.byref(ms),
                 _stream(stream))
        if num_steps > 0:
            self._prepared = True
        t, f = self.get_cells()
        return {"home_target": t, "fine_offset": f, "walk_ms": ms.value}

    def maybe_reorder(self, step: int, params: Optional[RunParams] = None, stream=None) -> bool:
        """ParticleReorderManager::maybe_reorder (ReorderManager.cpp:8-25) with the
        reorderPeriod fallback of SimManager::execute_force_calculation (:343-349)."""
        P = params or self.spec.params
        rp = P.reorder_period if P.reorder_period > 0 else P.rebuild_period
        if rp > 0 and step != 0 and (step - 1) % rp == 0 and getattr(self, "_prepared", False):
            self.reorder_particles(stream)
            return True
        return False

    def step_unfused(self, step: int, params: Optional[RunParams] = None, stream=None):
        """One SimManager step through the separate reference-shaped calls."""
        P = params or self.spec.params
        self.maybe_reorder(step, P, stream)
        energy = P.energy_period > 0 and step % P.energy_period == 0
        self.calculate_nonbonded_forces(P.pairlist_cutoff, P.cutoff, step, P.rebuild_period, P.electric_field,
                                        P.interpolation, energy, stream)
        self._prepared = True
        self.calculate_bonded_forces(energy, stream)
        self.integrate_motion(P.timestep, P.kT, P.integrator, step, stream)
