"""Pins the oracle's Morton keys, stable sort, cell tables and pair set.

Morton: closed forms (MortonCode.h:91-180). Sort: std::stable_sort equivalence
(zorder_tests.cpp:139-151). Pair set: the stencil walk equals an O(N^2) brute force
of the same predicate; known-answer sets of Zorder_pairlist2.cpp:365-404 and
zorder_pairlist_tests.cpp:204-342.
"""
import numpy as np
import pytest

from marsmd import PeriodicBox, configs
from oracle import pyoracle as orc
from tests.conftest import seam_pair_positions


def test_morton_closed_forms():
    assert orc.morton_from_coords(0, 0, 0) == 0
    assert orc.morton_from_coords(1023, 1023, 1023) == 0x3FFFFFFF
    # x bit above y above z
    assert orc.morton_from_coords(1, 0, 0) == 4
    assert orc.morton_from_coords(0, 1, 0) == 2
    assert orc.morton_from_coords(0, 0, 1) == 1
    rs = np.random.default_rng(1)
    for x, y, z in rs.integers(0, 1024, size=(200, 3)):
        assert orc.morton_decode(orc.morton_from_coords(int(x), int(y), int(z))) == (x, y, z)


def test_morton_quantisation():
    lo = np.zeros(3, np.float32)
    hi = np.full(3, 100.0, np.float32)
    p = np.array([[0, 0, 0, 0], [100, 100, 100, 0], [-5, 50, 150, 0], [50, 50, 50, 0]], np.float32)
    codes = orc.morton_encode(p, lo, hi)
    assert codes[0] == 0
    assert codes[1] == 0x3FFFFFFF
    # clamped below and above
    assert orc.morton_decode(int(codes[2])) == (0, int(np.float32(0.5) * np.float32(1023)), 1023)
    q = int(np.float32(np.float32(50) / np.float32(100)) * np.float32(1023))
    assert orc.morton_decode(int(codes[3])) == (q, q, q)


def test_stable_sort_matches_numpy_stable():
    rs = np.random.default_rng(2)
    keys = rs.integers(0, 64, size=5000).astype(np.uint32)  # many duplicates
    sk, sv = orc.stable_sort_pairs(keys)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(sv, order)
    assert np.array_equal(sk, keys[order])


def _random_box(rs, n, L, periodic):
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rs.uniform(0, L, size=(n, 3))
    return pos, PeriodicBox((L, L, L), periodic=periodic)


@pytest.mark.parametrize("periodic", [(True, True, True), (False, False, False), (True, False, True)])
@pytest.mark.parametrize("n,L,rc", [(300, 40.0, 7.5), (800, 60.0, 12.0), (50, 20.0, 9.9)])
def test_stencil_walk_equals_brute_force(periodic, n, L, rc):
    rs = np.random.default_rng(n)
    pos, box = _random_box(rs, n, L, periodic)
    walk = orc.build_pairlist(pos, box, rc)
    brute = orc.build_pairlist(pos, box, rc, brute_force=True)
    assert np.array_equal(walk["pairs"], brute["pairs"])
    assert len(walk["pairs"]) > 0


def test_pair_set_is_the_distance_predicate():
    # every listed pair is within rc (minimum image, FP64), every pair within rc - eps is listed
    rs = np.random.default_rng(5)
    n, L, rc = 600, 50.0, 8.0
    pos, box = _random_box(rs, n, L, (True, True, True))
    pairs = orc.build_pairlist(pos, box, rc)["pairs"]
    d = pos[:, None, :3].astype(np.float64) - pos[None, :, :3]
    d -= L * np.round(d / L)
    r = np.sqrt((d ** 2).sum(-1))
    iu = np.triu_indices(n, 1)
    close = set(zip(*[a[r[iu] < rc - 1e-4] for a in iu]))
    listed = set(map(tuple, pairs.tolist()))
    assert close <= listed
    assert all(r[a, b] <= rc + 1e-4 for a, b in listed)


def _test_box():
    # set_test_box (Zorder_pairlist2.cpp:56-62): 64^3 open box at origin -4
    return PeriodicBox((64.0, 64.0, 64.0), periodic=(False, False, False), origin=(-4.0, -4.0, -4.0))


def test_known_answer_three_collinear():
    pos = np.array([[0, 0, 0, 0], [1, 0, 0, 0], [2, 0, 0, 0]], np.float32)
    assert len(orc.build_pairlist(pos, _test_box(), 1.1)["pairs"]) == 2
    assert len(orc.build_pairlist(pos, _test_box(), 2.5)["pairs"]) == 3


def _four():
    return np.array([[0, 0, 0, 0], [1, 0, 0, 0], [2, 0, 0, 0], [3, 0, 0, 0]], np.float32)


def _ebox():
    return PeriodicBox((60.0, 60.0, 60.0), periodic=(False, False, False), origin=(-30.0, -30.0, -30.0))


def _pairs(res):
    return [tuple(p) for p in res["pairs"].tolist()]


def test_known_answer_exclusions():
    # zorder_pairlist_tests.cpp:218-342
    pos = _four()
    allp = [(0, 1), (0, 2), (1, 2), (1, 3), (2, 3)]
    assert _pairs(orc.build_pairlist(pos, _ebox(), 2.5)) == allp
    off = np.array([0, 1, 3, 3, 4]); exc = np.array([1, 0, 3, 1])
    assert _pairs(orc.build_pairlist(pos, _ebox(), 2.5, off, exc, 4)) == [(0, 2), (1, 2), (2, 3)]
    off = np.array([0, 1, 2, 2, 3]); exc = np.array([1, 0, 1])  # (1,3) only in row 3: survives
    assert _pairs(orc.build_pairlist(pos, _ebox(), 2.5, off, exc, 4)) == [(0, 2), (1, 2), (1, 3), (2, 3)]
    off = np.array([0, 2, 5, 8, 10]); exc = np.array([1, 2, 0, 2, 3, 0, 1, 3, 1, 2])
    assert _pairs(orc.build_pairlist(pos, _ebox(), 2.5, off, exc, 4)) == []
    body = np.array([7, 7, 9, -1])
    assert _pairs(orc.build_pairlist(pos, _ebox(), 2.5, rigid_body_id=body)) == [(0, 2), (1, 2), (1, 3), (2, 3)]
    body = np.array([-1, -1, -1, -1])
    assert _pairs(orc.build_pairlist(pos, _ebox(), 2.5, rigid_body_id=body)) == allp
    body = np.array([5, 5, -1, -1]); off = np.array([0, 0, 0, 1, 2]); exc = np.array([3, 2])
    assert _pairs(orc.build_pairlist(pos, _ebox(), 2.5, off, exc, 4, body)) == [(0, 2), (1, 2), (1, 3)]


def test_exclusion_csr_layout():
    off, exc, n = orc.exclusion_csr(np.array([[1, 0], [1, 3]]))
    assert n == 4
    assert off.tolist() == [0, 1, 3, 3, 4]
    assert exc.tolist() == [1, 0, 3, 1]


def test_select_cell_grid_closed_form():
    # config 0 (extent 250, rc 35, n 10k): the single-cell grid has the lowest cost
    assert orc.select_cell_grid(35.0, 10000, [250, 250, 250], 10000)[0] == 0
    # config 2 (extent 2000, 1M): m = 5 (62.5 A cells, r = 1, 27 slots)
    m, r, slots = orc.select_cell_grid(35.0, 1_000_000, [2000, 2000, 2000], 1_000_000)
    assert (m, r, slots) == (5, (1, 1, 1), 27)
    # more cells than particles stops the search: 8^4 cells need 4096 particles
    assert orc.select_cell_grid(1.0, 4096, [100, 100, 100], 10**6)[0] == 4
    assert orc.select_cell_grid(1.0, 4095, [100, 100, 100], 10**6)[0] == 0
    # the neighbour table must fit in the pair buffer (ZOrderPairlist.cpp:61-64)
    assert orc.select_cell_grid(1.0, 4096, [100, 100, 100], 100)[0] == 0


def test_cell_tables_cover_each_particle_once():
    rs = np.random.default_rng(9)
    pos, box = _random_box(rs, 2000, 100.0, (True, True, True))
    res = orc.build_pairlist(pos, box, 10.0)
    m, r, slots = res["grid"]
    b, e, nb = orc.cell_tables(res["codes"], m, r, slots, [1, 1, 1])
    assert (e - b).sum() == 2000
    assert np.all(np.sort(nb, axis=1)[:, 0] != 0xFFFFFFFF) or slots > 0


def test_config0_pair_count_density():
    spec = configs.config0(n=3000, L=170.0, seed=3)
    res = orc.build_pairlist(spec.pos4(), spec.box, 35.0)
    rho = 3000 / 170.0 ** 3
    expect = 0.5 * 3000 * rho * 4.0 / 3.0 * np.pi * 35.0 ** 3
    assert abs(len(res["pairs"]) - expect) / expect < 0.05


def _compact_by3(v):
    v = v & 0x09249249
    v = (v ^ (v >> 2)) & 0x030C30C3
    v = (v ^ (v >> 4)) & 0x0300F00F
    v = (v ^ (v >> 8)) & 0x030000FF
    v = (v ^ (v >> 16)) & 0x000003FF
    return v


def reference_cells(pos, L, m):
    codes = orc.morton_encode(pos, [0, 0, 0], [L, L, L])
    return np.stack([_compact_by3(codes >> s) for s in (2, 1, 0)], axis=1).astype(np.int64) >> (10 - m), codes


def test_incomplete_reference_grid_drops_exactly_the_seam_pairs():
    # 16 cells of 10.0006 A per axis with r = 1 at rc = 10: the quantised top cell is
    # short (63 of 64 quanta), so the +-1 stencil misses pairs across the periodic seam
    # (BuildCellNeighborsKernel, ZOrderNeighbor.h:55-139). The pair set is then every
    # within-cutoff pair whose cells are at most r apart per axis, wrapped.
    n, L, rc = 20000, 160.01, 10.0
    pos = seam_pair_positions(n, L)
    box = PeriodicBox((L, L, L))
    res = orc.build_pairlist(pos, box, rc, max_pairs=70000)
    m, r, _ = res["grid"]
    assert (m, r) == (4, (1, 1, 1))
    brute = orc.build_pairlist(pos, box, rc, max_pairs=70000, brute_force=True)["pairs"]
    cell, codes = reference_cells(pos, L, m)
    assert np.array_equal(np.sort(codes), np.sort(res["codes"]))
    d = np.abs(cell[brute[:, 0]] - cell[brute[:, 1]])
    d = np.minimum(d, (1 << m) - d)
    keep = (d <= 1).all(axis=1)
    planted = {(2 * i, 2 * i + 1) for i in range(120)}
    dropped = {tuple(p) for p in brute[~keep]}
    assert planted <= dropped and len(dropped) < len(brute) // 100
    assert np.array_equal(res["pairs"], brute[keep])
