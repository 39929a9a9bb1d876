"""GPU parity: Morton keys, stable radix sort, inverse map and the pair set through
the C ABI, bit for bit against the oracle (ZOrderSort / ZOrderPairlist semantics)."""
import numpy as np
import pytest

import marsmd
from marsmd import PeriodicBox, configs
from oracle import pyoracle as orc
from tests.conftest import seam_pair_positions, sorted_pairs

pytestmark = pytest.mark.gpu


def _dev(cuda, a, dtype=None):
    t = cuda.from_numpy(np.ascontiguousarray(a))
    return t.cuda() if dtype is None else t.to(dtype).cuda()


def test_philox_bitexact(cuda):
    for seed, ctr, gs, c1 in [(0, 0, 0, 0), (42, 1000, 7, 0x5324120), (0xFFFFFFFFFFFFFFFF, 5, 99, 0x1356914)]:
        g = marsmd.philox_draw(seed, ctr, gs, c1, draws=2, n=257).cpu().numpy().view(np.uint32)
        for i in (0, 1, 100, 256):
            o = orc.philox_draw(seed, ctr + i, gs, c1, draws=2)
            assert np.array_equal(g[i], o)


@pytest.mark.parametrize("n", [1, 63, 4096, 4097, 100_000, 1_000_000])
def test_morton_sort_inverse_bitexact(cuda, n):
    rs = np.random.default_rng(n)
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rs.uniform(-10.0, 510.0, size=(n, 3))  # some outside the domain: clamped
    lo, hi = [0.0, 0.0, 0.0], [500.0, 500.0, 500.0]
    z = marsmd.ZOrderSort(n)
    z.sort_particles(_dev(cuda, pos), n, lo, hi)
    cuda.cuda.synchronize()
    keys = orc.morton_encode(pos, np.float32(lo), np.float32(hi))
    sk, sv = orc.stable_sort_pairs(keys)
    assert np.array_equal(z.get_morton_codes().cpu().numpy().view(np.uint32), sk)
    assert np.array_equal(z.get_sorted_indices().cpu().numpy().view(np.uint32), sv)
    inv = np.empty(n, np.uint32)
    inv[sv] = np.arange(n, dtype=np.uint32)
    assert np.array_equal(z.get_inverse_indices().cpu().numpy().view(np.uint32), inv)


def test_sort_stability_with_heavy_duplicates(cuda):
    # every particle in one of 8 positions: the permutation must be iota within a key
    n = 50_000
    rs = np.random.default_rng(3)
    choices = rs.uniform(0, 100, size=(8, 3)).astype(np.float32)
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = choices[rs.integers(0, 8, n)]
    z = marsmd.ZOrderSort(n)
    z.sort_particles(_dev(cuda, pos), n, [0, 0, 0], [100, 100, 100])
    cuda.cuda.synchronize()
    keys = orc.morton_encode(pos, np.zeros(3, np.float32), np.full(3, 100, np.float32))
    assert np.array_equal(z.get_sorted_indices().cpu().numpy(), np.argsort(keys, kind="stable"))


@pytest.mark.parametrize("key_bits,n", [(32, 300_000), (30, 100_003), (25, 70_001), (20, 4097), (9, 200_000)])
def test_radix_sort_orders_by_the_low_key_bits(cuda, key_bits, n):
    # 8-bit digits for 32 bits; 10-bit digits for 30, 25, 20 and 9 bits (fewer passes).
    # Keys carry bits above key_bits: the sort orders by bits [0, key_bits) only and
    # keeps the input order among equal low bits.
    rs = np.random.default_rng(key_bits)
    keys = rs.integers(0, 2**32, size=n, dtype=np.uint64).astype(np.uint32)
    keys[::7] = keys[3]  # duplicates
    vals = np.arange(n, dtype=np.uint32)
    kd, vd = _dev(cuda, keys.view(np.int32)), _dev(cuda, vals.view(np.int32))
    ko, vo = cuda.empty_like(kd), cuda.empty_like(vd)
    nbytes = marsmd.lib().marsmd_radix_sort_scratch_bytes(n)
    scratch = cuda.empty(nbytes, dtype=cuda.uint8, device="cuda")
    marsmd._capi.call("marsmd_radix_sort_pairs", kd.data_ptr(), vd.data_ptr(), ko.data_ptr(), vo.data_ptr(), n,
                      key_bits, scratch.data_ptr(), nbytes, cuda.cuda.current_stream().cuda_stream)
    cuda.cuda.synchronize()
    low = keys & np.uint32((1 << key_bits) - 1) if key_bits < 32 else keys
    order = np.argsort(low, kind="stable")
    assert np.array_equal(vo.cpu().numpy().view(np.uint32), order)
    assert np.array_equal(ko.cpu().numpy().view(np.uint32), keys[order])


def _gpu_pairs(cuda, pos, box, rc, excl=None, types=None, max_pairs=1 << 62):
    n = len(pos)
    pl = marsmd.ZOrderPairlist(max(n, 1), max_pairs)
    pl.set_periodic_box(box)
    pl.set_exclusions(excl)
    pl.build_pairlist(_dev(cuda, pos), n, rc)
    cuda.cuda.synchronize()
    return pl, sorted_pairs(pl.get_neighbor_pairs().cpu().numpy())


@pytest.mark.parametrize("periodic", [(True, True, True), (False, False, False), (True, False, True)])
@pytest.mark.parametrize("n,L,rc", [(2000, 80.0, 9.0), (5000, 60.0, 12.0), (300, 30.0, 14.0)])
def test_pair_set_bitexact_random(cuda, periodic, n, L, rc):
    rs = np.random.default_rng(n + int(rc))
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rs.uniform(0, L, size=(n, 3))
    box = PeriodicBox((L, L, L), periodic=periodic)
    pl, gp = _gpu_pairs(cuda, pos, box, rc)
    ref = orc.build_pairlist(pos, box, rc)
    assert np.array_equal(gp, ref["pairs"].astype(np.int64))
    z = pl.get_sorter()
    assert np.array_equal(z.get_sorted_indices().cpu().numpy().view(np.uint32), ref["sorted_idx"])
    assert np.array_equal(z.get_morton_codes().cpu().numpy().view(np.uint32), ref["codes"])


def test_pair_set_positions_on_box_faces(cuda):
    # particles exactly on the periodic faces and outside the box (wrap_diff paths)
    rs = np.random.default_rng(11)
    n, L = 3000, 50.0
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rs.uniform(0, L, size=(n, 3))
    pos[:100, 0] = 0.0
    pos[100:200, 1] = L
    pos[200:260, 2] = -3.0
    pos[260:300, 0] = L + 2.0
    box = PeriodicBox((L, L, L))
    _, gp = _gpu_pairs(cuda, pos, box, 10.0)
    ref = orc.build_pairlist(pos, box, 10.0)
    assert np.array_equal(gp, ref["pairs"].astype(np.int64))


def test_pairs_at_the_cutoff_bitexact(cuda):
    # pairs planted within a few 1e-7 .. 1e-4 relative of the cutoff, many of them across
    # the periodic seam: the walk decides these with the reference's exact predicate
    rs = np.random.default_rng(31)
    n_bg, n_pairs, L, rc = 40_000, 12_000, 800.0, 35.0
    a = rs.uniform(0, L, size=(n_pairs, 3))
    a[: n_pairs // 3, 0] = rs.uniform(0, 20.0, size=n_pairs // 3)  # next to the x = 0 face
    u = rs.normal(size=(n_pairs, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    rel = rs.choice([-1.0, 1.0], n_pairs) * 10.0 ** rs.uniform(-7, -4, n_pairs)
    b = np.mod(a + u * (rc * (1.0 + rel))[:, None], L)
    pos = np.zeros((n_bg + 2 * n_pairs, 4), np.float32)
    pos[:n_bg, :3] = rs.uniform(0, L, size=(n_bg, 3))
    pos[n_bg::2, :3] = a
    pos[n_bg + 1::2, :3] = b
    box = PeriodicBox((L, L, L))
    _, gp = _gpu_pairs(cuda, pos, box, rc)
    ref = orc.build_pairlist(pos, box, rc)
    assert np.array_equal(gp, ref["pairs"].astype(np.int64))


def test_known_answers_and_exclusions(cuda):
    pos = np.array([[0, 0, 0, 0], [1, 0, 0, 0], [2, 0, 0, 0]], np.float32)
    tb = PeriodicBox((64.0, 64.0, 64.0), periodic=(False, False, False), origin=(-4.0, -4.0, -4.0))
    pl, gp = _gpu_pairs(cuda, pos, tb, 1.1)
    assert pl.get_num_pairs() == 2
    pl, gp = _gpu_pairs(cuda, pos, tb, 2.5)
    assert pl.get_num_pairs() == 3
    four = np.array([[0, 0, 0, 0], [1, 0, 0, 0], [2, 0, 0, 0], [3, 0, 0, 0]], np.float32)
    eb = PeriodicBox((60.0, 60.0, 60.0), periodic=(False, False, False), origin=(-30.0, -30.0, -30.0))

    def csr(off, exc, body=None, n=4):
        return marsmd.ExclusionCSR(
            None if off is None else _dev(cuda, np.array(off, np.int32)),
            None if exc is None else _dev(cuda, np.array(exc, np.int32)),
            n if off is not None else 0,
            None if body is None else _dev(cuda, np.array(body, np.int32)))

    cases = [
        (None, [(0, 1), (0, 2), (1, 2), (1, 3), (2, 3)]),
        (csr([0, 1, 3, 3, 4], [1, 0, 3, 1]), [(0, 2), (1, 2), (2, 3)]),
        (csr([0, 1, 2, 2, 3], [1, 0, 1]), [(0, 2), (1, 2), (1, 3), (2, 3)]),
        (csr([0, 2, 5, 8, 10], [1, 2, 0, 2, 3, 0, 1, 3, 1, 2]), []),
        (csr(None, None, [7, 7, 9, -1]), [(0, 2), (1, 2), (1, 3), (2, 3)]),
        (csr(None, None, [-1, -1, -1, -1]), [(0, 1), (0, 2), (1, 2), (1, 3), (2, 3)]),
        (csr([0, 0, 0, 1, 2], [3, 2], [5, 5, -1, -1]), [(0, 2), (1, 2), (1, 3)]),
    ]
    for ex, expect in cases:
        _, gp = _gpu_pairs(cuda, four, eb, 2.5, ex)
        assert [tuple(p) for p in gp.tolist()] == expect


def test_config0_pairs_and_reference_cells(cuda):
    spec = configs.config0()
    pos = spec.pos4()
    pl, gp = _gpu_pairs(cuda, pos, spec.box, 35.0)
    ref = orc.build_pairlist(pos, spec.box, 35.0)
    assert np.array_equal(gp, ref["pairs"].astype(np.int64))
    st = pl.get_statistics()
    m, r, slots = ref["grid"]
    assert (st.cell_grid_bits, tuple(st.cell_radii), st.neighbors_per_cell) == (m, r, slots)
    b, e, nb = pl.get_reference_cells()
    ob, oe, onb = orc.cell_tables(ref["codes"], m, r, slots, [1, 1, 1])
    assert np.array_equal(b.cpu().numpy().view(np.uint32), ob)
    assert np.array_equal(e.cpu().numpy().view(np.uint32), oe)
    assert np.array_equal(nb.cpu().numpy().view(np.uint32), onb)


@pytest.mark.parametrize("n,L,rc,periodic,buffer", [
    (20000, 300.0, 12.0, (True, True, True), 100000),    # 16 cells per axis
    (50000, 500.0, 15.0, (True, False, True), 100000),   # 16, one open axis
    (30000, 300.0, 9.0, (False, False, False), 100000),  # 16, every axis open
    (60000, 1000.0, 30.0, (True, True, True), 10**7),    # 32 cells per axis
    (300000, 2000.0, 30.0, (True, True, True), 10**7),   # 64 cells per axis
])
def test_reference_cell_tables_on_finer_grids(cuda, n, L, rc, periodic, buffer):
    # select_cell_grid's finer grids (the pair buffer of the first build given as the
    # pairlist's particle capacity, as Pairlist sizes it): the cell ranges and the
    # BuildCellNeighborsKernel table (ZOrderNeighbor.h:55-139) equal the oracle's
    rs = np.random.default_rng(n)
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rs.uniform(0, L, size=(n, 3))
    box = PeriodicBox((L, L, L), periodic=periodic)
    print(f"case: n={n} L={L} rc={rc} periodic={periodic} buffer={buffer}", flush=True)
    pl = marsmd.ZOrderPairlist(buffer, 1 << 62)
    pl.set_periodic_box(box)
    pl.build_pairlist(_dev(cuda, pos), n, rc)
    cuda.cuda.synchronize()
    gp = sorted_pairs(pl.get_neighbor_pairs().cpu().numpy())
    ref = orc.build_pairlist(pos, box, rc, max_pairs=buffer)
    m, r, slots = ref["grid"]
    assert m >= 4
    assert np.array_equal(gp, ref["pairs"].astype(np.int64))
    st = pl.get_statistics()
    assert (st.cell_grid_bits, tuple(st.cell_radii), st.neighbors_per_cell) == (m, r, slots)
    b, e, nb = pl.get_reference_cells()
    ob, oe, onb = orc.cell_tables(ref["codes"], m, r, slots, [int(x) for x in periodic])
    assert np.array_equal(b.cpu().numpy().view(np.uint32), ob)
    assert np.array_equal(e.cpu().numpy().view(np.uint32), oe)
    assert np.array_equal(nb.cpu().numpy().view(np.uint32).ravel(), onb.ravel())


def test_config1_pairs_with_exclusions(cuda):
    spec = configs.config1(chains=40, beads=250, L=400.0)
    pos = spec.pos4()
    off, exc, ne = marsmd.exclusion_csr_from_pairs(spec.exclusions)
    ex = marsmd.ExclusionCSR(_dev(cuda, off), _dev(cuda, exc), ne)
    _, gp = _gpu_pairs(cuda, pos, spec.box, 35.0, ex)
    ref = orc.build_pairlist(pos, spec.box, 35.0, off, exc, ne)
    assert np.array_equal(gp, ref["pairs"].astype(np.int64))


def test_million_bead_pair_set(cuda):
    n, L = 1_000_000, 2000.0
    rs = np.random.default_rng(2024)
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rs.uniform(0, L, size=(n, 3))
    box = PeriodicBox((L, L, L))
    pl, gp = _gpu_pairs(cuda, pos, box, 35.0)
    ref = orc.build_pairlist(pos, box, 35.0)
    assert len(gp) == len(ref["pairs"])
    assert np.array_equal(gp, ref["pairs"].astype(np.int64))
    st = pl.get_statistics()
    assert st.reference_grid_complete == 1


def test_needs_update_and_overflow_bookkeeping(cuda):
    rs = np.random.default_rng(8)
    n, L = 4000, 80.0
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rs.uniform(0, L, size=(n, 3))
    box = PeriodicBox((L, L, L))
    d = _dev(cuda, pos)
    pl = marsmd.ZOrderPairlist(n, 1 << 40)
    pl.set_periodic_box(box)
    pl.build_pairlist(d, n, 12.0)
    assert not pl.needs_update(d, n, 5.0)
    moved = pos.copy()
    moved[17, 0] = (moved[17, 0] + 2.6) % L
    assert pl.needs_update(_dev(cuda, moved), n, 5.0)
    moved[17, 0] = (pos[17, 0] + 2.4) % L
    assert not pl.needs_update(_dev(cuda, moved), n, 5.0)
    # capacity ceiling: fatal past the cap (ZOrderPairlist.cpp:292-306)
    small = marsmd.ZOrderPairlist(n, 10)
    small.set_periodic_box(box)
    with pytest.raises(marsmd.MarsError):
        small.build_pairlist(d, n, 12.0)
    # a zero-size box is rejected (ZOrderPairlist.cpp:118-126)
    nobox = marsmd.ZOrderPairlist(n)
    with pytest.raises(marsmd.MarsError):
        nobox.build_pairlist(d, n, 12.0)


def test_dense_cluster_outgrows_first_row_capacity(cuda):
    # 6000 beads packed in a 45 A cube inside a sparse 500 A box: the first build's row
    # capacity (sized for the mean density) is far too small, so the walk runs again
    # with rows sized to the measured maximum; the pair set must not notice
    rs = np.random.default_rng(21)
    L = 500.0
    dense = rs.uniform(200.0, 245.0, size=(6000, 3))
    sparse = rs.uniform(0.0, L, size=(3000, 3))
    pos = np.zeros((9000, 4), np.float32)
    pos[:, :3] = np.vstack([dense, sparse])
    box = PeriodicBox((L, L, L))
    pl, gp = _gpu_pairs(cuda, pos, box, 14.0)
    ref = orc.build_pairlist(pos, box, 14.0)
    assert np.array_equal(gp, ref["pairs"].astype(np.int64))
    deg = np.bincount(ref["pairs"].ravel(), minlength=len(pos))
    assert pl.get_statistics().max_neighbors_per_particle == deg.max()


def test_rebuild_keeps_pair_set_and_degrees(cuda):
    rs = np.random.default_rng(5)
    n, L = 20000, 150.0
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rs.uniform(0, L, size=(n, 3))
    box = PeriodicBox((L, L, L))
    pl = marsmd.ZOrderPairlist(n, 1 << 40)
    pl.set_periodic_box(box)
    d = _dev(cuda, pos)
    for rebuild in range(2):  # the second build reuses the first build's row capacity
        pl.build_pairlist(d, n, 16.0)
        cuda.cuda.synchronize()
        gp = sorted_pairs(pl.get_neighbor_pairs().cpu().numpy())
        ref = orc.build_pairlist(pos, box, 16.0)
        assert np.array_equal(gp, ref["pairs"].astype(np.int64)), f"build {rebuild}"
        deg = np.bincount(ref["pairs"].ravel(), minlength=n)
        st = pl.get_statistics()
        assert st.max_neighbors_per_particle == deg.max()
        assert st.num_pairs == len(ref["pairs"])


def test_incomplete_reference_grid_keeps_the_reference_pair_set(cuda):
    # a pair buffer of 70000 selects 16 cells of 10.0006 A with r = 1 at rc = 10, whose
    # stencil misses the planted pairs across the periodic seam: the list must miss them too
    n, L, rc = 20000, 160.01, 10.0
    pos = seam_pair_positions(n, L)
    box = PeriodicBox((L, L, L))
    pl = marsmd.ZOrderPairlist(70000, 1 << 62)
    pl.set_periodic_box(box)
    for rebuild in range(2):
        pl.build_pairlist(_dev(cuda, pos), n, rc)
        cuda.cuda.synchronize()
        gp = sorted_pairs(pl.get_neighbor_pairs().cpu().numpy())
        ref = orc.build_pairlist(pos, box, rc, max_pairs=70000)
        assert ref["grid"][:2] == (4, (1, 1, 1))
        assert not {(2 * i, 2 * i + 1) for i in range(120)} & {tuple(p) for p in gp.tolist()}
        assert np.array_equal(gp, ref["pairs"].astype(np.int64)), f"build {rebuild}"
