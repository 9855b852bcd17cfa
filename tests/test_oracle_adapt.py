"""Pins of the oracle's adaptive-refinement restatement and of the library's host tree
update. Tagging is kamayan's AMRLoehner (grid_refinement.cpp:30-152); the flag
bookkeeping, tree update and remesh are Parthenon's (not vendored), restated from its
published algorithm. Pinned here by closed forms of the estimator on polynomial fields,
hand-worked tree updates (one split, a split that cascades to coarser neighbours, the
derefine count, a merge vetoed by a finer neighbour), exactness of the remesh on linear
fields, conservation, and agreement of the library's host tree update with the oracle's
on random trees and tags."""
import numpy as np
import pytest

import kamayan_amd as km
from kamayan_amd import Layout, mesh_desc
from kamayan_amd._abi import amr_adapt, amr_criterion, slot
from tests import _cases

R, S, D = 1, 0, -1


def _analytic_pack(oracle, desc, lay, leaves, fn, nv=1):
    """host pack over the leaves with every zone (ghosts included) set to fn(x, y, z)."""
    blocks = oracle.amr_blocks(desc, leaves)
    hp = oracle.HostPack(lay, nv, blocks=blocks)
    e1, e2, e3 = lay.E
    g = lay.ng
    for b in range(lay.nblocks):
        bk = blocks[b]
        x = bk.xmin[0] + (np.arange(e1) - g + 0.5) * bk.dx[0]
        y = bk.xmin[1] + (np.arange(e2) - (g if lay.ndim > 1 else 0) + 0.5) * bk.dx[1]
        z = bk.xmin[2] + (np.arange(e3) - (g if lay.ndim > 2 else 0) + 0.5) * bk.dx[2]
        Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
        for v in range(nv):
            hp.cell[b, v] = fn(X, Y, Z)
    return hp, blocks


def _tags(oracle, hp, err_threshold, derefine_tol=1e-300):
    crit = amr_criterion(0, 10, refine_tol=err_threshold, derefine_tol=derefine_tol, filter=0.01)
    tags = np.zeros(hp.pack.nblocks, np.int32)
    oracle.amr_tag_loehner(hp.pack, crit, tags, True)
    return tags


def _interior_centres(lay, blocks, b):
    g = lay.ng
    bk = blocks[b]
    x = bk.xmin[0] + (np.arange(lay.nx[0]) + 0.5) * bk.dx[0]
    y = bk.xmin[1] + (np.arange(lay.nx[1] if lay.ndim > 1 else 1) + 0.5) * bk.dx[1]
    z = bk.xmin[2] + (np.arange(lay.nx[2] if lay.ndim > 2 else 1) + 0.5) * bk.dx[2]
    return np.meshgrid(z, y, x, indexing="ij")


def _check_max_error(oracle, hp, expected):
    """tags flip from refine to same exactly at each block's expected maximum error."""
    lo = _tags(oracle, hp, 0.0)
    assert np.all(lo == R)
    for b, e in enumerate(expected):
        below = _tags(oracle, hp, e * (1 - 1e-9))
        above = _tags(oracle, hp, e * (1 + 1e-9))
        assert below[b] == R and above[b] == S, (b, e)


def test_loehner_quadratic_1d(oracle):
    # u = x^2, x > 0: d = 2x exactly, num = 4, den = (2x/h + 4 f (x^2 + h^2)/h^2)^2, so the
    # error of a zone is h^2 / (x h + 2 f (x^2 + h^2)), largest at the block's first zone
    lay = Layout(1, 2, (16, 1, 1), (4, 1, 1))
    desc = mesh_desc(lay, xmin=(1.0, 0, 0), xmax=(2.0, 1, 1), bc=("outflow",) * 6)
    leaves = np.array([(0, i, 0, 0) for i in range(4)], np.int32)
    hp, blocks = _analytic_pack(oracle, desc, lay, leaves, lambda X, Y, Z: X * X)
    f = 0.01
    exp = []
    for b in range(4):
        h = blocks[b].dx[0]
        x0 = blocks[b].xmin[0] + 0.5 * h
        exp.append(h * h / (x0 * h + 2 * f * (x0 * x0 + h * h)))
    _check_max_error(oracle, hp, exp)


def test_loehner_cross_term_anisotropic_2d(oracle):
    # u = x y on cells h x k (h != k): d_x = y, d_y = x; only the mixed second differences
    # are nonzero (1 each), so num = 2 and, for x, y > 0,
    #   e_xx = y/h + 4 f x y / h^2,  e_xy = x/k + 4 f x y / (h k),
    #   e_yx = y/h + 4 f x y / (h k), e_yy = x/k + 4 f x y / k^2
    lay = Layout(2, 2, (8, 8, 1), (2, 2, 1))
    desc = mesh_desc(lay, xmin=(1.0, 1.0, 0), xmax=(2.0, 3.0, 1), bc=("outflow",) * 6)
    leaves = np.array([(0, i, j, 0) for j in range(2) for i in range(2)], np.int32)
    hp, blocks = _analytic_pack(oracle, desc, lay, leaves, lambda X, Y, Z: X * Y)
    f = 0.01
    exp = []
    for b in range(4):
        h, k = blocks[b].dx[0], blocks[b].dx[1]
        _, Y, X = _interior_centres(lay, blocks, b)
        exx = Y / h + 4 * f * X * Y / h**2
        exy = X / k + 4 * f * X * Y / (h * k)
        eyx = Y / h + 4 * f * X * Y / (h * k)
        eyy = X / k + 4 * f * X * Y / k**2
        err = np.sqrt(2.0 / (1e-12 + exx**2 + exy**2 + eyx**2 + eyy**2))
        exp.append(float(err.max()))
    _check_max_error(oracle, hp, exp)


def test_loehner_linear_field_derefines(oracle):
    lay = Layout(3, 2, (8, 8, 8), (2, 2, 2))
    desc = mesh_desc(lay, bc=("outflow",) * 6)
    leaves = np.array([(0, i, j, k) for k in range(2) for j in range(2) for i in range(2)],
                      np.int32)
    hp, _ = _analytic_pack(oracle, desc, lay, leaves, lambda X, Y, Z: 1 + X - 2 * Y + 3 * Z)
    crit = amr_criterion(0, 3, refine_tol=0.8, derefine_tol=1e-6)
    tags = np.zeros(8, np.int32)
    oracle.amr_tag_loehner(hp.pack, crit, tags, True)
    assert np.all(tags == D)
    # a refine tag already present survives (ScatterMax), a reset does not
    tags[:] = R
    oracle.amr_tag_loehner(hp.pack, crit, tags, False)
    assert np.all(tags == R)


def _roots(nb, ndim):
    return [(0, i, j, k) for k in range(nb[2]) for j in range(nb[1]) for i in range(nb[0])]


def _split(leaf, ndim):
    L, i, j, k = leaf
    out = []
    for c in range(1 << ndim):
        out.append((L + 1, 2 * i + (c & 1), 2 * j + ((c >> 1) & 1) if ndim > 1 else 0,
                    2 * k + ((c >> 2) & 1) if ndim > 2 else 0))
    return out


def _as_set(leaves):
    return {tuple(int(x) for x in r) for r in leaves}


def _desc2d(bc="outflow"):
    lay = Layout(2, 2, (8, 8, 1), (4, 4, 1))
    return mesh_desc(lay, bc=(bc,) * 6)


@pytest.mark.parametrize("impl", ["oracle", "library"])
def test_update_one_split(oracle, impl):
    desc = _desc2d()
    leaves = np.array(_roots((4, 4, 1), 2), np.int32)
    tags = np.zeros(16, np.int32)
    tags[5] = R  # root (1, 1)
    up = oracle.amr_update_tree if impl == "oracle" else km.amr_update_tree
    nl, nd, org, ch = up(desc, 3, 10, leaves, tags, np.zeros(16, np.int32))
    assert ch
    want = (_as_set(leaves) - {(0, 1, 1, 0)}) | set(_split((0, 1, 1, 0), 2))
    assert _as_set(nl) == want
    for r, o in zip(nl, org):
        if r[0] == 1:
            assert o[0] == 1 and o[1] == 5
        else:
            assert o[0] == 0 and tuple(leaves[o[1]]) == tuple(r)


@pytest.mark.parametrize("impl", ["oracle", "library"])
def test_update_split_cascades_to_coarser_neighbours(oracle, impl):
    desc = _desc2d()
    base = [r for r in _roots((4, 4, 1), 2) if r != (0, 0, 0, 0)] + _split((0, 0, 0, 0), 2)
    leaves = np.array(base, np.int32)
    tags = np.zeros(len(leaves), np.int32)
    tags[[i for i, r in enumerate(base) if r == (1, 1, 1, 0)][0]] = R
    up = oracle.amr_update_tree if impl == "oracle" else km.amr_update_tree
    nl, _, org, ch = up(desc, 3, 10, leaves, tags, np.zeros(len(leaves), np.int32))
    assert ch
    want = set(base) - {(1, 1, 1, 0)} | set(_split((1, 1, 1, 0), 2))
    for root in [(0, 1, 0, 0), (0, 0, 1, 0), (0, 1, 1, 0)]:
        want = (want - {root}) | set(_split(root, 2))
    assert _as_set(nl) == want and len(nl) == 31
    km.amr_blocks(desc, nl)  # balanced and tiling


@pytest.mark.parametrize("impl", ["oracle", "library"])
def test_merge_waits_for_derefine_count(oracle, impl):
    desc = _desc2d()
    base = [r for r in _roots((4, 4, 1), 2) if r != (0, 0, 0, 0)] + _split((0, 0, 0, 0), 2)
    leaves = np.array(base, np.int32)
    tags = np.array([D if r[0] == 1 else S for r in base], np.int32)
    deref = np.zeros(len(leaves), np.int32)
    up = oracle.amr_update_tree if impl == "oracle" else km.amr_update_tree
    for cycle in range(3):
        nl, nd, org, ch = up(desc, 3, 3, leaves, tags, deref)
        assert ch == (cycle == 2)
    assert _as_set(nl) == set(_roots((4, 4, 1), 2))
    merged = [o for r, o in zip(nl, org) if tuple(r) == (0, 0, 0, 0)][0]
    assert merged[0] == 2
    assert [tuple(leaves[i]) for i in merged[1:5]] == _split((0, 0, 0, 0), 2)


@pytest.mark.parametrize("impl", ["oracle", "library"])
def test_merge_vetoed_by_finer_neighbour(oracle, impl):
    desc = _desc2d()
    fam = _split((0, 0, 0, 0), 2)
    other = [r for r in _split((0, 1, 0, 0), 2) if r != (1, 2, 0, 0)] + _split((1, 2, 0, 0), 2)
    base = [r for r in _roots((4, 4, 1), 2) if r not in [(0, 0, 0, 0), (0, 1, 0, 0)]]
    base += fam + other
    leaves = np.array(base, np.int32)
    km.amr_blocks(desc, leaves)
    tags = np.array([D if tuple(r) in fam else S for r in base], np.int32)
    deref = np.zeros(len(leaves), np.int32)
    up = oracle.amr_update_tree if impl == "oracle" else km.amr_update_tree
    nl, _, _, ch = up(desc, 3, 1, leaves, tags, deref)
    # (1, 1, 0) touches the level-2 leaves of (1, 2, 0): no flag, so the family stays
    assert not ch and _as_set(nl) == set(base)
    assert deref[base.index((1, 1, 0, 0))] == 1


def _random_case(rng, ndim, bc):
    nb = (4, 4, 4 if ndim == 3 else 1) if ndim > 1 else (8, 1, 1)
    lay = Layout(ndim, 2, tuple(8 if d < ndim else 1 for d in range(3)),
                 tuple(nb[d] if d < ndim else 1 for d in range(3)))
    desc = mesh_desc(lay, bc=(bc,) * 6)
    c = tuple(rng.uniform(0.1, 0.9) for _ in range(3))
    leaves = km.refine_ball(desc, 2, c, rng.uniform(0.05, 0.3))
    return desc, Layout(lay.ndim, lay.ng, lay.nx, lay.nb, nleaves=len(leaves)), leaves


@pytest.mark.parametrize("ndim,bc", [(1, "outflow"), (2, "periodic"), (2, "outflow"),
                                     (3, "periodic"), (3, "reflect")])
def test_library_tree_update_matches_oracle(oracle, ndim, bc):
    rng = np.random.default_rng(100 + 10 * ndim + len(bc))
    for trial in range(6):
        desc, lay, leaves = _random_case(rng, ndim, bc)
        da = np.zeros(len(leaves), np.int32)
        db = da.copy()
        for step in range(4):
            tags = rng.choice([D, S, R], size=len(leaves), p=[0.6, 0.25, 0.15]).astype(np.int32)
            a = oracle.amr_update_tree(desc, 4, 2, leaves, tags, da)
            b = km.amr_update_tree(desc, 4, 2, leaves, tags, db)
            assert np.array_equal(da, db)
            for x, y in zip(a, b):
                assert np.array_equal(np.asarray(x), np.asarray(y)), (trial, step)
            if a[3]:
                km.amr_blocks(desc, a[0])  # valid, balanced tree
            leaves, da = a[0], a[1].copy()
            db = da.copy()


def _linear(X, Y, Z):
    return 0.5 + 1.25 * X - 0.75 * Y + 0.5 * Z


@pytest.mark.parametrize("ndim", [2, 3])
def test_remesh_exact_on_linear_fields_and_conservative(oracle, ndim):
    rng = np.random.default_rng(7 + ndim)
    desc, lay, leaves = _random_case(rng, ndim, "outflow")
    tags = rng.choice([D, S, R], size=len(leaves), p=[0.5, 0.2, 0.3]).astype(np.int32)
    nl, _, org, ch = oracle.amr_update_tree(desc, 4, 0, leaves, tags,
                                            np.zeros(len(leaves), np.int32))
    assert ch and set(org[:, 0]) >= {0, 1}
    old, _ = _analytic_pack(oracle, desc, lay, leaves, _linear, nv=2)
    rnd = rng.uniform(0.5, 2.0, size=old.cell[:, 1].shape)
    old.cell[:, 1] = rnd
    lay2 = Layout(lay.ndim, lay.ng, lay.nx, lay.nb, nleaves=len(nl))
    blocks2 = oracle.amr_blocks(desc, nl)
    new = oracle.HostPack(lay2, 2, blocks=blocks2)
    oracle.amr_remesh(old.pack, new.pack, org, 0, 2)
    sl = lay.interior()
    for b in range(len(nl)):
        Z, Y, X = _interior_centres(lay2, blocks2, b)
        assert np.allclose(new.cell[b, 0][sl], _linear(X, Y, Z), rtol=0, atol=2e-14)
    blocks1 = oracle.amr_blocks(desc, leaves)
    t0 = sum(float(np.sum(old.cell[b, 1][sl])) * blocks1[b].vol for b in range(len(leaves)))
    t1 = sum(float(np.sum(new.cell[b, 1][sl])) * blocks2[b].vol for b in range(len(nl)))
    assert t1 == pytest.approx(t0, rel=1e-13)


def test_split_then_merge_restores_the_data(oracle):
    lay = Layout(2, 2, (8, 8, 1), (2, 2, 1), nleaves=4)
    desc = mesh_desc(lay, bc=("outflow",) * 6)
    leaves = np.array(_roots((2, 2, 1), 2), np.int32)
    rng = np.random.default_rng(3)
    old, _ = _analytic_pack(oracle, desc, lay, leaves, lambda X, Y, Z: 1.0 + 0 * X)
    old.cell[:, 0] = rng.uniform(0.5, 1.5, old.cell[:, 0].shape)
    nl, _, org, _ = oracle.amr_update_tree(desc, 2, 0, leaves, np.full(4, R, np.int32),
                                           np.zeros(4, np.int32))
    lay2 = Layout(2, 2, (8, 8, 1), (2, 2, 1), nleaves=len(nl))
    mid = oracle.HostPack(lay2, 1, blocks=oracle.amr_blocks(desc, nl))
    oracle.amr_remesh(old.pack, mid.pack, org, 0, 1)
    back, _, org2, ch = oracle.amr_update_tree(desc, 2, 0, nl, np.full(len(nl), D, np.int32),
                                               np.zeros(len(nl), np.int32))
    assert ch and _as_set(back) == _as_set(leaves)
    fin = oracle.HostPack(lay, 1, blocks=oracle.amr_blocks(desc, back))
    oracle.amr_remesh(mid.pack, fin.pack, org2, 0, 1)
    sl = lay.interior()
    order = [list(map(tuple, back)).index(tuple(r)) for r in leaves]
    for b, nb_ in enumerate(order):
        assert np.allclose(fin.cell[nb_, 0][sl], old.cell[b, 0][sl], rtol=1e-14, atol=0)


def test_adaptive_sedov_conserves_and_follows_the_blast(oracle):
    cfg, desc, lay, leaves, cell = _cases.amr_case(
        "sedov", ndim=2, nroot=4, nblock=8, ng=4, lmax=0, center=(0.0, 0.0, 0.0), radius=0.1,
        bc="outflow", recon="wenoz", limiter="minmod", riemann="hllc", integrator="rk2",
        params=[1.0, 0.05, 1.0, 1e-5], xmin=(-0.5,) * 3, xmax=(0.5,) * 3, gamma=1.4)
    sim = oracle.CpuSimulation(cfg.c, desc, lay, leaves=leaves)
    sim.set_primitive(cell)
    sim.set_adaptive(amr_adapt(3, [amr_criterion(slot("PRES"), 3)]))
    st0, _ = sim.state()
    tot0 = _cases.amr_totals(st0, sim.layout, oracle.amr_blocks(desc, sim.leaves()))
    counts = []
    for _ in range(12):
        sim.advance(1)
        counts.append(len(sim.leaves()))
    lv = sim.leaves()
    km.amr_blocks(desc, lv)
    assert counts[0] > 16 and max(lv[:, 0]) == 2
    st, _ = sim.state()
    tot = _cases.amr_totals(st, sim.layout, oracle.amr_blocks(desc, lv))
    assert tot[0] == pytest.approx(tot0[0], rel=1e-12)
    assert tot[4] == pytest.approx(tot0[4], rel=1e-12)
