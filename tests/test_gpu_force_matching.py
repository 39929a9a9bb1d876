"""GPU force matching (the library's gradient path) through the C ABI and
torch.autograd, against the oracle.

Forces of every snapshot within 1e-5 of the oracle's (FP32 pair geometry, FP64 sums);
the FP64 adjoint dL/dknot within 1e-5 (normwise) of the oracle's analytic gradient,
which central differences pin exactly (the loss is quadratic in the knots); autograd's
gradient is the adjoint; a short Adam run reduces the loss.
"""
import numpy as np
import pytest

from marsmd import configs
from marsmd.fm import ForceMatching, force_matching_loss, tabulated_pair_forces
from oracle import pyoracle as orc

pytestmark = pytest.mark.gpu


def _setup(n=3000, L=300.0, S=4, knots=512, seed=5):
    frames = np.stack([configs.uniform_positions(n, L, seed + 31 * s) for s in range(S)])
    w = configs.config4(n=n, L=L, snapshots=S, knots=knots, seed=seed, frames=frames)
    fm = ForceMatching(w.positions, w.type_ids, w.box, w.tables, w.term_matrix, w.cutoff)
    return w, fm


def _k64(tables):
    return np.concatenate([t.pot.astype(np.float64) for t in tables])


def test_forces_and_adjoint_match_oracle(cuda):
    w, fm = _setup()
    print(f"case: force matching n={fm.n} S={fm.S} knots={fm.total_knots}", flush=True)
    tgt, _, _ = orc.fm_loss_grad(w.positions, w.type_ids, w.term_matrix, w.tables, _k64(w.target_tables), w.box,
                                 w.cutoff)
    f_ref, loss_ref, g_ref = orc.fm_loss_grad(w.positions, w.type_ids, w.term_matrix, w.tables, _k64(w.tables),
                                              w.box, w.cutoff, target=tgt)
    knots = fm.knots()
    f = fm.forces(knots).cpu().numpy()[..., :3]
    assert np.abs(f - f_ref).max() <= 1e-5 * np.abs(f_ref).max()
    # dL/dF of the MSE from the GPU's own forces, then the adjoint
    G = 2.0 * (f.astype(np.float64) - tgt) / f.size
    g = fm.adjoint(knots, cuda.from_numpy(G.astype(np.float32)).cuda()).cpu().numpy()
    assert np.abs(g - g_ref).max() <= 1e-5 * np.abs(g_ref).max()
    assert fm.num_pairs() > 0


def test_autograd_is_the_adjoint_and_adam_reduces_the_loss(cuda):
    w, fm = _setup(n=2000, L=260.0, S=3, knots=256, seed=9)
    torch = cuda
    target = fm.forces(fm.knots(w.target_tables))[..., :3].clone()
    knots = fm.knots().clone().requires_grad_(True)
    loss = force_matching_loss(knots, fm, target)
    loss.backward()
    # dL/dF as the loss defines it: the FP32 differences times 2 / count
    f = tabulated_pair_forces(knots.detach(), fm).detach()
    d = f - target
    G = d * torch.tensor(2.0 / d.numel(), dtype=torch.float64).to(d.dtype)
    g_adj = fm.adjoint(knots.detach(), G)
    assert torch.allclose(knots.grad.double(), g_adj.to(knots.grad.device), rtol=1e-6, atol=1e-12)
    # central differences of the GPU loss on a few used knots (FP32 forces: loose bound)
    used = torch.nonzero(knots.grad).flatten().cpu().numpy()
    rs = np.random.default_rng(2)
    for k in rs.choice(used, 8, replace=False):
        h = 2e-2
        kp = knots.detach().clone()
        km = knots.detach().clone()
        kp[k] += h
        km[k] -= h
        fd = (force_matching_loss(kp, fm, target) - force_matching_loss(km, fm, target)).item() / (2 * h)
        assert abs(fd - knots.grad[k].item()) <= 2e-2 * abs(knots.grad).max().item(), (k, fd, knots.grad[k].item())
    opt = torch.optim.Adam([knots], lr=0.02)
    first = None
    for it in range(150):
        opt.zero_grad()
        loss = force_matching_loss(knots, fm, target)
        loss.backward()
        opt.step()
        first = loss.item() if first is None else first
    assert loss.item() < 0.2 * first, (first, loss.item())


def test_fused_loss_matches_the_elementwise_loss(cuda):
    torch = cuda
    w, fm = _setup(n=2000, L=260.0, S=3, knots=256, seed=4)
    target = fm.forces(fm.knots(w.target_tables))[..., :3].clone()
    knots = fm.knots()
    f = fm.forces(knots)
    loss, grad = fm.mse(f, target)
    d = f[..., :3] - target
    ref = (d.double() ** 2).sum() / d.numel()
    assert abs(loss.item() - ref.item()) <= 1e-12 * ref.item()
    scale = torch.tensor(2.0 / d.numel(), dtype=torch.float64).to(torch.float32)
    assert torch.equal(grad[..., :3], d * scale) and torch.all(grad[..., 3] == 0)
    # the fused autograd path and the forces' own autograd function give one gradient
    k1 = knots.clone().requires_grad_(True)
    force_matching_loss(k1, fm, target).backward()
    k2 = knots.clone().requires_grad_(True)
    f2 = tabulated_pair_forces(k2, fm)
    ((f2 - target) ** 2).sum().div(f2.numel()).backward()
    assert torch.allclose(k1.grad.double(), k2.grad.double(), rtol=1e-5, atol=1e-12 * k1.grad.abs().max().item())


def test_sorted_order_entry_points(cuda):
    # forces and the adjoint in each snapshot's sorted order are the particle-order ones
    # permuted by to_sorted; the loss keeps its sorted target while the target tensor
    # is unchanged and permutes it again after an in-place change
    torch = cuda
    w, fm = _setup(n=2500, L=270.0, S=3, knots=256, seed=11)
    knots = fm.knots()
    f = fm.forces(knots)
    fs = fm.forces(knots, sorted_order=True)
    assert torch.equal(fm.to_sorted(f), fs)
    assert torch.equal(fm.to_sorted(f[..., :3].contiguous()), fs[..., :3])
    G = torch.randn((fm.S, fm.n, 4), device="cuda", dtype=torch.float32)
    g = fm.adjoint(knots, G)
    gs = fm.adjoint(knots, fm.to_sorted(G), sorted_order=True)
    assert torch.allclose(g, gs, rtol=1e-12, atol=1e-12 * g.abs().max().item())
    target = fm.forces(fm.knots(w.target_tables))[..., :3].clone()
    l1 = force_matching_loss(knots, fm, target).item()
    ref = ((f[..., :3].double() - target.double()) ** 2).sum().item() / target.numel()
    assert abs(l1 - ref) <= 1e-6 * ref
    target[0, :10] += 1.0
    l2 = force_matching_loss(knots, fm, target).item()
    ref2 = ((f[..., :3].double() - target.double()) ** 2).sum().item() / target.numel()
    assert abs(l2 - ref2) <= 1e-6 * ref2 and l2 != l1
    # a new tensor object with other values is permuted again
    other = target.clone()
    other[1, :10] -= 2.0
    l3 = force_matching_loss(knots, fm, other).item()
    ref3 = ((f[..., :3].double() - other.double()) ** 2).sum().item() / other.numel()
    assert abs(l3 - ref3) <= 1e-6 * ref3


def test_one_pass_loss_is_forces_then_mse(cuda):
    # marsmd_fm_loss_sorted never stores the forces: its dL/dF must be the two-pass
    # path's (forces in sorted order, then the loss kernel) bit for bit, and its loss
    # the same FP64 terms summed in another fixed grouping
    torch = cuda
    w, fm = _setup(n=2200, L=265.0, S=3, knots=256, seed=13)
    knots = fm.knots()
    target = fm.to_sorted(fm.forces(fm.knots(w.target_tables))[..., :3].contiguous())
    loss2, grad2 = fm.mse(fm.forces(knots, sorted_order=True), target)
    loss1, grad1 = fm.loss_sorted(knots, target)
    assert torch.equal(grad1, grad2)
    assert abs(loss1.item() - loss2.item()) <= 1e-13 * loss2.item()
    # repeated calls give the same bits
    loss3, grad3 = fm.loss_sorted(knots, target)
    assert torch.equal(grad3, grad1) and loss3.item() == loss1.item()
