"""Oracle pins of the force-matching loss and its gradient (the library's gradient
path; no reference counterpart).

A table force is linear in its knots, so the loss is exactly quadratic in them and a
central difference equals the derivative up to FP64 rounding: the analytic dL/dknot is
checked against central differences at h = 1e-3 on knots the snapshots actually use.
The target tables reproduce their own forces (zero loss and zero gradient), and the
forces of the unperturbed tables equal the oracle's pair-force path on each snapshot.
"""
import numpy as np

from marsmd import configs
from oracle import pyoracle as orc


def _small(seed=3, n=600, L=120.0, S=3, knots=128):
    frames = np.stack([configs.uniform_positions(n, L, seed + 17 * s) for s in range(S)])
    return configs.config4(n=n, L=L, snapshots=S, knots=knots, seed=seed, frames=frames)


def _knots(tables):
    return np.concatenate([t.pot.astype(np.float64) for t in tables])


def test_forces_match_the_pair_force_path():
    w = _small()
    f, _, _ = orc.fm_loss_grad(w.positions, w.type_ids, w.term_matrix, w.tables, _knots(w.tables), w.box, w.cutoff)
    for s in range(len(w.positions)):
        pos4 = np.zeros((w.positions.shape[1], 4), np.float32)
        pos4[:, :3] = w.positions[s]
        pl = orc.build_pairlist(pos4, w.box, w.cutoff)
        ref = orc.pair_forces(pl["pairs"], pos4, w.type_ids, w.term_matrix, w.tables, w.box, w.cutoff ** 2, mode=1)
        scale = np.abs(ref[:, :3]).max()
        assert np.abs(f[s] - ref[:, :3]).max() <= 1e-5 * scale


def test_target_tables_give_zero_loss_and_gradient():
    w = _small()
    tgt, _, _ = orc.fm_loss_grad(w.positions, w.type_ids, w.term_matrix, w.tables, _knots(w.target_tables), w.box,
                                 w.cutoff)
    _, loss, grad = orc.fm_loss_grad(w.positions, w.type_ids, w.term_matrix, w.tables, _knots(w.target_tables),
                                     w.box, w.cutoff, target=tgt)
    assert loss == 0.0 and not grad.any()


def test_gradient_equals_central_differences():
    w = _small()
    k_target = _knots(w.target_tables)
    tgt, _, _ = orc.fm_loss_grad(w.positions, w.type_ids, w.term_matrix, w.tables, k_target, w.box, w.cutoff)
    k0 = _knots(w.tables)
    _, loss0, grad = orc.fm_loss_grad(w.positions, w.type_ids, w.term_matrix, w.tables, k0, w.box, w.cutoff,
                                      target=tgt)
    assert loss0 > 0
    used = np.nonzero(grad)[0]
    assert len(used) > 50
    rs = np.random.default_rng(1)
    pick = rs.choice(used, 25, replace=False)
    h = 1e-3
    for k in pick:
        kp, km = k0.copy(), k0.copy()
        kp[k] += h
        km[k] -= h
        lp = orc.fm_loss_grad(w.positions, w.type_ids, w.term_matrix, w.tables, kp, w.box, w.cutoff, target=tgt)[1]
        lm = orc.fm_loss_grad(w.positions, w.type_ids, w.term_matrix, w.tables, km, w.box, w.cutoff, target=tgt)[1]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - grad[k]) <= 1e-7 * max(abs(grad).max(), 1e-30), (k, fd, grad[k])
    # knots no pair reaches carry no gradient
    assert np.all(grad[np.setdiff1d(np.arange(len(grad)), used)] == 0)
