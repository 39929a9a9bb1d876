"""Many replicas stepped together on one GPU (marsmd_batch).

Each replica is a periodic nonbonded system that the reference would step as its own
Patch (Patch.h:48-645) under SimManager::run's sequence; the replica's index (or the
given patch id) enters its seed mix as the patch id (Patch.cpp:252-254). The batch
runs one launch per step over every replica and one batched rebuild for all of them.

Launch configuration: the persistent step kernel's workgroups per CU and whether the
steps between host decisions replay as hipGraph segments. `Batch.autotune` times the
candidates on real steps and keeps the fastest, and remembers the choice per workload
and device in a JSON file (MARSMD_AUTOTUNE_CACHE, default
~/.cache/marsmd/autotune.json), so a later batch of the same shape starts tuned.
Results do not depend on the configuration.
"""
from __future__ import annotations

import ctypes as C
import json
import os
from pathlib import Path
from typing import List, Optional, Sequence

import numpy as np

from . import _capi as api
from .core import _stream, make_types
from .system import RunParams, SystemSpec


class Batch:
    def __init__(self, specs: Sequence[SystemSpec], patch_ids: Optional[Sequence[int]] = None):
        if not specs:
            raise ValueError("a batch needs at least one replica")
        first = specs[0]
        if any(s.seed != first.seed for s in specs):
            raise ValueError("the replicas of a batch share the base seed (their patch ids tell their random "
                             "streams apart)")
        for s in specs[1:]:
            if (len(s.types) != len(first.types) or len(s.nb_tables) != len(first.nb_tables)
                    or not np.array_equal(s.tag_matrix(), first.tag_matrix()) or s.pair_params != first.pair_params):
                raise ValueError("the replicas of a batch share types, pair tables, pair tags and pair constants")
        self.specs = list(specs)
        self.counts = np.array([s.n for s in specs], np.uint64)
        self.offsets = np.concatenate([[0], np.cumsum(self.counts)]).astype(np.int64)
        self.total = int(self.offsets[-1])
        keep = []
        counts = np.ascontiguousarray(self.counts, np.uint64)
        boxes = (api.Box * len(specs))(*[s.box.to_c() for s in specs])
        pids = np.ascontiguousarray(patch_ids if patch_ids is not None else [s.patch_id for s in specs], np.uint64)
        tm = np.ascontiguousarray(first.tag_matrix(), np.uint32)
        recs = (api.Table * max(1, len(first.nb_tables)))()
        for i, t in enumerate(first.nb_tables):
            kn = np.ascontiguousarray(t.pot, np.float32)
            keep.append(kn)
            recs[i] = t.to_c(kn.ctypes.data)
        d = api.BatchDesc()
        d.num_replicas = len(specs)
        d.num_particles = counts.ctypes.data
        d.boxes = C.addressof(boxes)
        d.patch_ids = pids.ctypes.data
        d.types = make_types(first.types)
        d.term_matrix = tm.ctypes.data
        d.enabled_terms = first.enabled_terms
        d.num_nb_tables = len(first.nb_tables)
        d.nb_tables = C.addressof(recs)
        if first.pair_params is not None:
            pp = first.pair_params.to_c()
            keep.append(pp)
            d.pair_params = C.addressof(pp)
        h = C.c_void_p()
        api.call("marsmd_batch_create", C.byref(d), C.byref(h))
        self._h = h
        api.call("marsmd_batch_set_base_seed", self._h, first.seed)
        pos = np.concatenate([s.pos4() for s in specs])
        moms = [s.mom4() for s in specs]
        mom = None if any(m is None for m in moms) else np.concatenate(moms)
        types = np.concatenate([np.asarray(s.type_ids, np.int32) for s in specs])
        api.call("marsmd_batch_copy_particles_from_host", self._h, pos.ctypes.data,
                 None if mom is None else mom.ctypes.data, types.ctypes.data, self.total)

    def __del__(self):
        if getattr(self, "_h", None):
            api.lib().marsmd_batch_destroy(self._h)
            self._h = None

    def run(self, num_steps: int, first_step: int = 0, params: Optional[RunParams] = None, stream=None):
        p = (params or self.specs[0].params).to_c()
        api.call("marsmd_batch_run", self._h, C.byref(p), first_step, num_steps, _stream(stream))

    def finish_deferred_kick(self, dt: Optional[float] = None, stream=None):
        api.call("marsmd_batch_finish_deferred_kick", self._h,
                 float(dt if dt is not None else self.specs[0].params.timestep), _stream(stream))

    def copy_particles_to_host(self) -> List[dict]:
        """Per replica: {pos, mom, force, type, id} in its canonical order."""
        n = self.total
        pos = np.zeros((n, 4), np.float32)
        mom = np.zeros((n, 4), np.float32)
        fe = np.zeros((n, 4), np.float32)
        ty = np.zeros(n, np.int32)
        ids = np.zeros(n, np.int32)
        api.call("marsmd_batch_copy_particles_to_host", self._h, pos.ctypes.data, mom.ctypes.data, fe.ctypes.data,
                 ty.ctypes.data, ids.ctypes.data, n)
        out = []
        for r in range(len(self.specs)):
            a, b = self.offsets[r], self.offsets[r + 1]
            out.append({"pos": pos[a:b], "mom": mom[a:b], "force": fe[a:b], "type": ty[a:b], "id": ids[a:b]})
        return out

    def get_neighbor_pairs(self, replica: int, stream=None) -> np.ndarray:
        """Replica `replica`'s current pair set as (num_pairs, 2) int32, x < y, in its
        canonical indices."""
        ptr, np_ = C.c_void_p(), C.c_uint64(0)
        api.call("marsmd_batch_get_neighbor_pairs", self._h, replica, C.byref(ptr), C.byref(np_), _stream(stream))
        if np_.value == 0:
            return np.zeros((0, 2), np.int32)
        from .core import _torch, copy_from_device_ptr

        t = copy_from_device_ptr(ptr.value, 2 * np_.value, _torch().int32)
        return t.cpu().numpy().reshape(-1, 2)

    def set_kernel_timing(self, enable: bool = True):
        api.call("marsmd_batch_set_kernel_timing", self._h, 1 if enable else 0)

    def kernel_timing(self):
        ms, k = C.c_double(0.0), C.c_uint64(0)
        api.call("marsmd_batch_get_kernel_timing", self._h, C.byref(ms), C.byref(k))
        return ms.value, k.value

    # ------------------------------------------------------------------ launch configuration
    def set_launch(self, blocks_per_cu: int = 0, use_graph: bool = False):
        lc = api.BatchLaunch(int(blocks_per_cu), 1 if use_graph else 0)
        api.call("marsmd_batch_set_launch", self._h, C.byref(lc))

    def get_launch(self) -> dict:
        lc = api.BatchLaunch()
        api.call("marsmd_batch_get_launch", self._h, C.byref(lc))
        return {"blocks_per_cu": lc.blocks_per_cu, "use_graph": bool(lc.use_graph)}

    def workload_key(self, params: Optional[RunParams] = None) -> str:
        """What the fastest launch depends on: device, replica sizes, tables, integrator
        and the host-decision cadence."""
        from .core import _torch

        p = params or self.specs[0].params
        dev = _torch().cuda.get_device_name(0)
        knots = sum(len(t.pot) for t in self.specs[0].nb_tables)
        return (f"{dev}|R{len(self.specs)}|N{self.total}|max{int(self.counts.max())}|knots{knots}"
                f"|int{p.integrator}|rb{p.rebuild_period}|ro{p.reorder_period}|e{p.energy_period}")

    @staticmethod
    def _cache_path(cache) -> Path:
        if cache is not None:
            return Path(cache)
        env = os.environ.get("MARSMD_AUTOTUNE_CACHE")
        return Path(env) if env else Path.home() / ".cache" / "marsmd" / "autotune.json"

    # ------------------------------------------------------------------ cell configuration
    @staticmethod
    def size_bucket(n: int) -> int:
        """The size bucket of a replica of n particles: 2^b <= n < 2^(b+1)."""
        return max(0, min(api.BATCH_BUCKETS - 1, int(n).bit_length() - 1))

    def set_cells(self, cells: dict):
        """Cell configuration per size bucket, {bucket: (home_target, fine_offset)}; the
        buckets not named keep theirs. Applies from the next build."""
        cur = self.get_cells(all_buckets=True)
        cur.update({int(b): tuple(v) for b, v in cells.items()})
        c = api.BatchCells()
        for b in range(api.BATCH_BUCKETS):
            c.home_target[b], c.fine_offset[b] = cur[b]
        api.call("marsmd_batch_set_cells", self._h, C.byref(c))

    def get_cells(self, all_buckets: bool = False) -> dict:
        """{bucket: (home_target, fine_offset)} for the buckets this batch's replicas fall
        in (every bucket with all_buckets)."""
        c = api.BatchCells()
        api.call("marsmd_batch_get_cells", self._h, C.byref(c))
        buckets = range(api.BATCH_BUCKETS) if all_buckets else sorted({self.size_bucket(n) for n in self.counts})
        return {b: (c.home_target[b], c.fine_offset[b]) for b in buckets}

    def autotune_cells(self, num_steps: int, first_step: int = 0, params: Optional[RunParams] = None, stream=None,
                       cache=None, use_cache: bool = True) -> dict:
        """Steps first_step+1 .. first_step+num_steps after choosing the walk's cell
        configuration per size bucket (marsmd_batch_autotune_cells), or with the cached
        choice for this workload. Returns {"cells", "walk_ms", "cached"}."""
        key = self.workload_key(params) + "|cells"
        path = None if cache is False else self._cache_path(cache)
        table = {}
        if path is not None and path.exists():
            try:
                table = json.loads(path.read_text())
            except (OSError, ValueError):
                table = {}
        if use_cache and key in table:
            hit = table[key]
            self.set_cells({int(b): tuple(v) for b, v in hit["cells"].items()})
            self.run(num_steps, first_step, params, stream)
            return {"cells": self.get_cells(), "walk_ms": hit["walk_ms"], "cached": True}
        p = (params or self.specs[0].params).to_c()
        ms = C.c_double(0.0)
        api.call("marsmd_batch_autotune_cells", self._h, C.byref(p), first_step, num_steps, C.byref(ms),
                 _stream(stream))
        cells = self.get_cells()
        if ms.value > 0.0 and path is not None:
            table[key] = {"cells": {str(b): list(v) for b, v in cells.items()}, "walk_ms": ms.value}
            try:
                path.parent.mkdir(parents=True, exist_ok=True)
                tmp = path.with_suffix(".tmp")
                tmp.write_text(json.dumps(table, indent=1, sort_keys=True))
                tmp.replace(path)
            except OSError:
                pass  # an unwritable cache only costs the next batch a tuning run
        return {"cells": cells, "walk_ms": ms.value, "cached": False}

    def autotune(self, num_steps: int, first_step: int = 0, params: Optional[RunParams] = None, stream=None,
                 cache=None, use_cache: bool = True) -> dict:
        """Steps first_step+1 .. first_step+num_steps (real steps, as run) while timing
        every launch candidate, and keeps the fastest. With a cached choice for this
        workload the steps run with it and nothing is timed. `cache`: a JSON file, None
        for the default one, False for none. Returns the choice with its ms per step
        ("cached": True when it came from the cache)."""
        key = self.workload_key(params)
        path = None if cache is False else self._cache_path(cache)
        table = {}
        if path is not None and path.exists():
            try:
                table = json.loads(path.read_text())
            except (OSError, ValueError):
                table = {}
        if use_cache and key in table:
            hit = table[key]
            self.set_launch(hit["blocks_per_cu"], hit["use_graph"])
            self.run(num_steps, first_step, params, stream)
            return dict(hit, cached=True)
        p = (params or self.specs[0].params).to_c()
        lc = api.BatchLaunch()
        ms = C.c_double(0.0)
        api.call("marsmd_batch_autotune", self._h, C.byref(p), first_step, num_steps, C.byref(lc), C.byref(ms),
                 _stream(stream))
        choice = {"blocks_per_cu": lc.blocks_per_cu, "use_graph": bool(lc.use_graph), "ms_per_step": ms.value}
        if ms.value > 0.0 and path is not None:
            table[key] = choice
            try:
                path.parent.mkdir(parents=True, exist_ok=True)
                tmp = path.with_suffix(".tmp")
                tmp.write_text(json.dumps(table, indent=1, sort_keys=True))
                tmp.replace(path)
            except OSError:
                pass  # an unwritable cache only costs the next batch a tuning run
        return dict(choice, cached=False)
