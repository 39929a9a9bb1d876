"""Force matching of tabulated pair potentials on the GPU, with the adjoint behind
torch.autograd.

The reference has no gradient path; this is the library's own. `ForceMatching` holds a
fixed set of configurations (snapshots) and their pairlists in HBM; `forces(knots)` is
calculate_nonbonded_forces of every snapshot for the given table knots, and
`adjoint(knots, dL/dF)` is dL/dknot accumulated in FP64 (a table force is linear in its
knots, so the adjoint is exact: see fm.hip). `TabulatedPairForces` wraps both as a
torch.autograd.Function.
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np

from . import _capi as api
from .core import PeriodicBox, TabulatedPotential, _stream, _torch


class ForceMatching:
    def __init__(self, positions, type_ids, box: PeriodicBox, tables: Sequence[TabulatedPotential], term_matrix,
                 cutoff: float, pairlist_cutoff: float | None = None):
        torch = _torch()
        pos = np.asarray(positions, np.float32)
        if pos.ndim != 3 or pos.shape[2] not in (3, 4):
            raise ValueError("positions must be (snapshots, particles, 3 or 4)")
        S, n = pos.shape[0], pos.shape[1]
        p4 = np.zeros((S, n, 4), np.float32)
        p4[..., :3] = pos[..., :3]
        self.S, self.n = S, n
        self.tables = list(tables)
        self.sizes = [len(t.pot) for t in self.tables]
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)
        self.total_knots = int(self.offsets[-1])
        self._pos = torch.from_numpy(p4).cuda()
        self._types = torch.from_numpy(np.ascontiguousarray(type_ids, np.int32)).cuda()
        tm = np.ascontiguousarray(term_matrix, np.uint32)
        recs = (api.Table * len(self.tables))()
        for i, t in enumerate(self.tables):
            recs[i] = t.to_c(0)
        d = api.FMDesc()
        d.box = box.to_c()
        d.num_particles = n
        d.num_snapshots = S
        d.positions = self._pos.data_ptr()
        d.type_ids = self._types.data_ptr()
        d.num_types = tm.shape[0]
        d.term_matrix = tm.ctypes.data
        d.num_tables = len(self.tables)
        d.tables = C.addressof(recs)
        d.cutoff = float(cutoff)
        d.pairlist_cutoff = float(pairlist_cutoff if pairlist_cutoff is not None else cutoff)
        self._h = C.c_void_p()
        api.call("marsmd_fm_create", C.byref(d), C.byref(self._h))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            api.lib().marsmd_fm_destroy(h)
            self._h = None

    def knots(self, tables: Sequence[TabulatedPotential] | None = None):
        """The concatenated FP32 knots of `tables` (default: the ones given at creation)."""
        torch = _torch()
        tabs = self.tables if tables is None else tables
        return torch.from_numpy(np.concatenate([np.asarray(t.pot, np.float32) for t in tabs])).cuda()

    def num_pairs(self) -> int:
        v = C.c_uint64(0)
        api.call("marsmd_fm_num_pairs", self._h, C.byref(v))
        return v.value

    def _knots(self, knots):
        torch = _torch()
        k = knots.detach().to(device="cuda", dtype=torch.float32).contiguous()
        if k.numel() != self.total_knots:
            raise ValueError(f"expected {self.total_knots} knots, got {k.numel()}")
        return k

    def forces(self, knots, stream=None, sorted_order: bool = False):
        """(S, n, 4) float32: force and energy of every snapshot for these knots, in
        each snapshot's particle order (or its sorted order: see to_sorted)."""
        torch = _torch()
        k = self._knots(knots)
        out = torch.empty((self.S, self.n, 4), dtype=torch.float32, device="cuda")
        api.call("marsmd_fm_forces_sorted" if sorted_order else "marsmd_fm_forces", self._h, k.data_ptr(),
                 out.data_ptr(), _stream(stream))
        return out

    def to_sorted(self, rows, stream=None):
        """(S, n, 3 or 4) float32 rows from each snapshot's particle order into its sorted
        order (the Morton order of its pairlist build)."""
        torch = _torch()
        r = rows.detach().to(device="cuda", dtype=torch.float32).contiguous()
        if r.shape[:2] != (self.S, self.n) or r.dim() != 3 or r.shape[2] not in (3, 4):
            raise ValueError(f"rows must be ({self.S}, {self.n}, 3 or 4)")
        out = torch.empty_like(r)
        api.call("marsmd_fm_to_sorted", self._h, r.data_ptr(), r.shape[2], out.data_ptr(), _stream(stream))
        return out

    def _sorted_target(self, target):
        """to_sorted(target), kept while the same tensor object is passed again without an
        in-place change since (its version counter): a new tensor, even at the address of
        an old one, is permuted again."""
        cached = getattr(self, "_target_cache", None)
        if cached is not None and cached[0] is target and cached[1] == target._version:
            return cached[2]
        t = self.to_sorted(target)
        self._target_cache = (target, target._version, t)
        return t

    def mse(self, forces, target, stream=None):
        """(loss, dL/dF) of the mean squared force error in one pass: loss a device FP64
        scalar, sum |F - target|^2 / (3 S n) (FP32 differences, FP64 squares in a fixed
        order); dL/dF (S, n, 4) float32 = (F - target) * 2 / (3 S n). forces as forces()
        returns them; target (S, n, 3)."""
        torch = _torch()
        f = forces.detach()
        if f.shape != (self.S, self.n, 4) or f.dtype != torch.float32:
            raise ValueError("forces must be the (S, n, 4) float32 output of forces()")
        t = target.detach().to(device="cuda", dtype=torch.float32).contiguous()
        if t.shape != (self.S, self.n, 3):
            raise ValueError(f"target must be ({self.S}, {self.n}, 3)")
        f = f.contiguous()
        grad = torch.empty_like(f)
        loss = torch.empty((), dtype=torch.float64, device="cuda")
        api.call("marsmd_fm_mse", self._h, f.data_ptr(), t.data_ptr(), grad.data_ptr(), loss.data_ptr(),
                 _stream(stream))
        return loss, grad

    def loss_sorted(self, knots, target_sorted, stream=None):
        """(loss, dL/dF) of forces(knots, sorted_order=True) against target_sorted
        (S, n, 3), each snapshot in its sorted order, in one pass that never stores the
        forces: dL/dF has mse()'s bits; the loss sums the same FP64 terms in another
        fixed grouping."""
        torch = _torch()
        k = self._knots(knots)
        t = target_sorted.detach().to(device="cuda", dtype=torch.float32).contiguous()
        if t.shape != (self.S, self.n, 3):
            raise ValueError(f"target must be ({self.S}, {self.n}, 3)")
        grad = torch.empty((self.S, self.n, 4), dtype=torch.float32, device="cuda")
        loss = torch.empty((), dtype=torch.float64, device="cuda")
        api.call("marsmd_fm_loss_sorted", self._h, k.data_ptr(), t.data_ptr(), grad.data_ptr(), loss.data_ptr(),
                 _stream(stream))
        return loss, grad

    def adjoint(self, knots, grad_forces, stream=None, sorted_order: bool = False):
        """dL/dknot (float64, total knots) for dL/dF = grad_forces (S, n, 3 or 4), in
        particle order (or sorted order)."""
        torch = _torch()
        k = self._knots(knots)
        g = grad_forces.detach().to(device="cuda", dtype=torch.float32)
        if g.shape[-1] == 3:
            g = torch.nn.functional.pad(g, (0, 1))
        g = g.contiguous()
        out = torch.empty(self.total_knots, dtype=torch.float64, device="cuda")
        api.call("marsmd_fm_adjoint_sorted" if sorted_order else "marsmd_fm_adjoint", self._h, k.data_ptr(),
                 g.data_ptr(), out.data_ptr(), _stream(stream))
        return out


def _function():
    torch = _torch()

    class TabulatedPairForces(torch.autograd.Function):
        """F(knots) of every snapshot; backward is the FP64 adjoint."""

        @staticmethod
        def forward(ctx, knots, fm: ForceMatching):
            ctx.fm = fm
            ctx.save_for_backward(knots)
            return fm.forces(knots)[..., :3]

        @staticmethod
        def backward(ctx, grad_f):
            (knots,) = ctx.saved_tensors
            gk = ctx.fm.adjoint(knots, grad_f)
            return gk.to(knots.dtype), None

    return TabulatedPairForces


_FN = None


def tabulated_pair_forces(knots, fm: ForceMatching):
    """Differentiable forces (S, n, 3) of the snapshots of `fm` for `knots`."""
    global _FN
    if _FN is None:
        _FN = _function()
    return _FN.apply(knots, fm)


_MSE = None


def _mse_function():
    torch = _torch()

    class ForceMSE(torch.autograd.Function):
        """mean((f - target)^2): the differences in FP32 (one pass), their squares summed
        in FP64 (one reduction); the gradient 2 (f - target) / count in one pass."""

        @staticmethod
        def forward(ctx, f, target):
            d = f - target
            ctx.save_for_backward(d)
            return torch.linalg.vector_norm(d, dtype=torch.float64).square() / d.numel()

        @staticmethod
        def backward(ctx, g):
            (d,) = ctx.saved_tensors
            return d * (2.0 * g / d.numel()).to(d.dtype), None

    return ForceMSE


_LOSS = None


def _loss_function():
    torch = _torch()

    class ForceMatchingLoss(torch.autograd.Function):
        """mean((F(knots) - target)^2) as one library call (the forces go straight into
        the loss and its dL/dF); backward is the FP64 adjoint of that dL/dF. Everything
        runs in each snapshot's sorted order (the target permuted once and kept while
        the same target tensor is passed): the loss is a sum over all rows, so its
        value does not depend on their order."""

        @staticmethod
        def forward(ctx, knots, fm: ForceMatching, target):
            loss, grad_f = fm.loss_sorted(knots, fm._sorted_target(target))
            ctx.fm = fm
            ctx.save_for_backward(knots, grad_f)
            return loss

        @staticmethod
        def backward(ctx, g):
            knots, grad_f = ctx.saved_tensors
            gk = ctx.fm.adjoint(knots, grad_f, sorted_order=True)
            return (gk * g.to(gk.dtype)).to(knots.dtype), None, None

    return ForceMatchingLoss


def force_matching_loss(knots, fm: ForceMatching, target):
    """Mean squared force error over snapshots, particles and components (FP64 sum),
    differentiable in the knots. A target that itself requires a gradient goes through
    the forces' autograd function and an elementwise loss instead."""
    global _MSE, _LOSS
    if getattr(target, "requires_grad", False):
        if _MSE is None:
            _MSE = _mse_function()
        f = tabulated_pair_forces(knots, fm)
        return _MSE.apply(f, target)
    if _LOSS is None:
        _LOSS = _loss_function()
    return _LOSS.apply(knots, fm, target)
