"""Pins of the oracle's block-tree (AMR) restatement: ghost fill across levels
(RestrictAverage / ProlongateSharedMinMod, refinement_operations.hpp:20-268) and flux
correction. Parthenon's exchange code is not vendored, so these pin what the
mathematics fixes: hand-computed known answers, exactness on linear fields,
equivalence with the uniform ghost fill when every leaf shares one level, restriction
conservation, and whole-run conservation on a periodic mesh with level jumps."""
import ctypes as C

import numpy as np
import pytest

import kamayan_amd as km
from kamayan_amd import Layout, mesh_desc
from tests import _cases


def _pack(oracle, lay, nv, blocks, nflux=0):
    return oracle.HostPack(lay, nv, nflux=nflux, blocks=blocks)


def _leaves_uniform(nb, level, ndim):
    """all leaves at one level, x-fastest (the uniform block order of that level)."""
    n = [nb[d] << level if d < ndim else 1 for d in range(3)]
    out = []
    for k in range(n[2]):
        for j in range(n[1]):
            for i in range(n[0]):
                out.append((level, i, j, k))
    return np.array(out, dtype=np.int32)


def test_library_and_oracle_agree_on_tree_and_metrics(oracle):
    lay = Layout(3, 2, (8, 8, 8), (4, 4, 4))
    desc = mesh_desc(lay, bc=("outflow",) * 6)
    leaves = km.refine_ball(desc, 2, (0.2, 0.2, 0.2), 0.1)
    assert set(leaves[:, 0]) == {0, 1, 2}
    a = km.amr_blocks(desc, leaves)
    b = oracle.amr_blocks(desc, leaves)
    assert bytes(a) == bytes(b)
    # volumes of the leaves tile the unit cube
    vol = sum(a[i].vol * 512 for i in range(len(leaves)))
    assert all(a[i].level == leaves[i, 0] for i in range(len(leaves)))
    assert vol == pytest.approx(1.0, rel=1e-14)


def test_library_rejects_unbalanced_and_gapped_trees():
    lay = Layout(2, 2, (8, 8, 1), (2, 2, 1))
    desc = mesh_desc(lay)
    # root block (0,0) refined twice next to level-0 roots: a two-level jump
    leaves = [(0, 1, 0, 0), (0, 0, 1, 0), (0, 1, 1, 0)]
    for j in range(2):
        for i in range(2):
            if (i, j) == (1, 1):
                for jj in range(2):
                    for ii in range(2):
                        leaves.append((2, 2 * i + ii, 2 * j + jj, 0))
            else:
                leaves.append((1, i, j, 0))
    with pytest.raises(km.KmError, match="differ by more than one level"):
        km.amr_blocks(desc, np.array(leaves, dtype=np.int32))
    with pytest.raises(km.KmError, match="cover"):
        km.amr_blocks(desc, np.array([(0, 0, 0, 0)], dtype=np.int32))
    # nghost 4 needs blocks of at least 8 cells once there is a second level
    small = mesh_desc(Layout(2, 4, (6, 6, 1), (2, 2, 1)))
    level1 = np.array([(1, i, j, 0) for j in range(4) for i in range(4)], dtype=np.int32)
    with pytest.raises(km.KmError, match="at least 8"):
        km.amr_blocks(small, level1)


@pytest.mark.parametrize("bc", ["periodic", "outflow", "reflect"])
@pytest.mark.parametrize("ndim", [2, 3])
def test_single_level_tree_matches_uniform_fill(oracle, bc, ndim):
    """A tree whose leaves share one level fills ghosts exactly as the uniform grid."""
    rng = np.random.default_rng(5)
    for level in (0, 1):
        nb0 = (2, 2, 2 if ndim > 2 else 1)
        nxb = (8, 8, 8 if ndim > 2 else 1)
        root = Layout(ndim, 2, nxb, nb0)
        desc = mesh_desc(root, bc=(bc,) * 6)
        leaves = _leaves_uniform(nb0, level, ndim)
        lay = Layout(ndim, 2, nxb, nb0, nleaves=len(leaves))
        fine_nb = tuple(nb0[d] << level if d < ndim else 1 for d in range(3))
        blocks = oracle.amr_blocks(desc, leaves)
        a = _pack(oracle, lay, 12, blocks)
        a.cell[:] = rng.standard_normal(a.cell.shape)
        b = _pack(oracle, lay, 12, blocks)
        b.cell[:] = a.cell
        cfg = km.Config()
        oracle.call("km_cpu_amr_ghost_fill", C.byref(cfg.c), C.byref(a.pack), C.byref(desc),
                    oracle._i32(leaves), len(leaves), 0, 11)
        nb = (C.c_int32 * 3)(*fine_nb)
        bcs = (C.c_int32 * 6)(*[desc.bc[d] for d in range(6)])
        oracle.call("km_cpu_ghost_fill_uniform", C.byref(cfg.c), C.byref(b.pack), nb, bcs, 0, 11,
                    0)
        assert np.array_equal(a.cell, b.cell)


def _centroids(blk, lay, d):
    n = lay.E[d]
    return blk.xmin[d] + (np.arange(n) - lay.lo(d) + 0.5) * blk.dx[d]


@pytest.mark.parametrize("ndim", [2, 3])
def test_linear_field_is_reproduced_across_levels(oracle, ndim):
    """Restriction averages a linear field to its centroid value and minmod prolongation
    of a linear field is exact, so every ghost zone away from the physical boundary
    holds the field at its own centroid."""
    nb0 = (4, 4, 4 if ndim > 2 else 1)
    nxb = (8, 8, 8 if ndim > 2 else 1)
    root = Layout(ndim, 2, nxb, nb0)
    desc = mesh_desc(root, bc=("outflow",) * 6)
    leaves = km.refine_ball(desc, 2, (0.45, 0.55, 0.5), 0.12)
    assert set(leaves[:, 0]) == {0, 1, 2}
    lay = Layout(ndim, 2, nxb, nb0, nleaves=len(leaves))
    blocks = oracle.amr_blocks(desc, leaves)
    h = _pack(oracle, lay, 12, blocks)
    coef = np.array([0.7, -0.4, 0.25])

    def field(b):
        blk = blocks[b]
        x = [_centroids(blk, lay, d) if d < ndim else np.zeros(1) for d in range(3)]
        return (1.0 + coef[0] * x[0][None, None, :] + coef[1] * x[1][None, :, None] +
                coef[2] * x[2][:, None, None])

    sl = lay.interior()
    for b in range(lay.nblocks):
        f = field(b)
        for v in range(12):
            h.cell[b, v][sl] = (f * (v + 1))[sl]
    oracle.call("km_cpu_amr_ghost_fill", C.byref(km.Config().c), C.byref(h.pack), C.byref(desc),
                oracle._i32(leaves), len(leaves), 0, 11)
    margin = 3 * (1.0 / 32)  # three level-0 zones from the boundary
    checked = 0
    for b in range(lay.nblocks):
        blk = blocks[b]
        f = field(b)
        x = [_centroids(blk, lay, d) if d < ndim else np.full(1, 0.5) for d in range(3)]
        inside = np.ones(f.shape, dtype=bool)
        for d in range(ndim):
            ok = (x[d] > margin) & (x[d] < 1 - margin)
            shape = [1, 1, 1]
            shape[2 - d] = -1
            inside &= ok.reshape(shape)
        for v in (0, 1, 2):
            got = h.cell[b, v]
            want = f * (v + 1)
            assert np.allclose(got[inside], want[inside], rtol=2e-14, atol=2e-14), (b, v)
        checked += int(inside.sum())
    assert checked > 0


def test_prolongation_and_restriction_known_answers(oracle):
    """1-D: block 0 at level 0 on [0, 1/2] (dx 1/16), two level-1 blocks on [1/2, 1]
    (dx 1/32), nghost 2, outflow."""
    root = Layout(1, 2, (8, 1, 1), (2, 1, 1))
    desc = mesh_desc(root, bc=("outflow",) * 6)
    leaves = np.array([(0, 0, 0, 0), (1, 2, 0, 0), (1, 3, 0, 0)], dtype=np.int32)
    lay = Layout(1, 2, (8, 1, 1), (2, 1, 1), nleaves=3)
    blocks = oracle.amr_blocks(desc, leaves)
    for fine01, want in (((3.5, 4.5), (1.75, 2.25)), ((1.0, 2.0), (2.0, 2.0))):
        h = _pack(oracle, lay, 12, blocks)
        # coarse zones 6, 7 of block 0 (local 8, 9) and fine zones 0, 1 of block 1
        h.cell[0, 0, 0, 0, 8] = 1.0
        h.cell[0, 0, 0, 0, 9] = 2.0
        h.cell[1, 0, 0, 0, 2] = fine01[0]
        h.cell[1, 0, 0, 0, 3] = fine01[1]
        oracle.call("km_cpu_amr_ghost_fill", C.byref(km.Config().c), C.byref(h.pack),
                    C.byref(desc), oracle._i32(leaves), 3, 0, 1)
        # ghost zones 0, 1 of the fine block: children of coarse zone 7 (value 2) with
        # minmod slope min(16, 32) = 16 (or 0 at an extremum), offsets -+ dx/2 = -+1/64
        assert h.cell[1, 0, 0, 0, 0] == want[0]
        assert h.cell[1, 0, 0, 0, 1] == want[1]
        # ghost zone 10 of the coarse block: restriction of the fine zones 0, 1
        assert h.cell[0, 0, 0, 0, 10] == 0.5 * (fine01[0] + fine01[1])


def test_restriction_conserves(oracle):
    """A coarse ghost zone restricted from finer leaves carries the fine volume sum."""
    root = Layout(2, 2, (8, 8, 1), (2, 2, 1))
    desc = mesh_desc(root)
    leaves = km.refine_ball(desc, 1, (0.25, 0.25, 0.5), 0.05)
    lay = Layout(2, 2, (8, 8, 1), (2, 2, 1), nleaves=len(leaves))
    blocks = oracle.amr_blocks(desc, leaves)
    h = _pack(oracle, lay, 12, blocks)
    rng = np.random.default_rng(11)
    h.cell[:] = rng.uniform(0.5, 2.0, h.cell.shape)
    oracle.call("km_cpu_amr_ghost_fill", C.byref(km.Config().c), C.byref(h.pack), C.byref(desc),
                oracle._i32(leaves), len(leaves), 0, 1)
    # level-0 block (1, 0) (x in [1/2, 1], y in [0, 1/2]) sees the refined block (0, 0)
    # on its low-x side: ghost zone (j, i=1) restricts fine zones of the leaf covering it
    lv = leaves[:, 0]
    b0 = int(np.nonzero((lv == 0) & (leaves[:, 1] == 1) & (leaves[:, 2] == 0))[0][0])
    fine = {(int(l[1]), int(l[2])): n for n, l in enumerate(leaves) if l[0] == 1}
    for j in range(8):
        # level-0 zone (x 7, y j) has children x 14-15, y 2j..2j+1 at level 1, in leaf
        # (1, j // 4) at local x 8-9, y 2 (j % 4) + 2 ..
        fb = fine[(1, j // 4)]
        y0 = 2 + 2 * (j % 4)
        kids = h.cell[fb, 0, 0, y0:y0 + 2, 8:10]
        got = h.cell[b0, 0, 0, 2 + j, 1]
        assert got * blocks[b0].vol == pytest.approx(float(kids.sum()) * blocks[fb].vol,
                                                     rel=1e-14), j


@pytest.mark.parametrize("problem,ndim", [("smooth3d", 3), ("smooth3d", 2)])
def test_periodic_run_conserves_across_level_jumps(oracle, problem, ndim):
    """Flux correction makes the coarse side of a level jump use the restricted fine
    flux, so sum V u over the leaves is conserved to round-off."""
    cfg, desc, lay, leaves, cell = _cases.amr_case(problem, ndim=ndim, nroot=2, nblock=8, ng=2,
                                                   lmax=1, center=(0.3, 0.6, 0.5), radius=0.2)
    assert set(leaves[:, 0]) == {0, 1}
    blocks = oracle.amr_blocks(desc, leaves)
    sim = oracle.CpuSimulation(cfg.c, desc, lay, leaves=leaves)
    sim.set_primitive(cell)
    before, _ = sim.state()
    sim.advance(4)
    after, _ = sim.state()
    t0 = _cases.amr_totals(before, lay, blocks)
    t1 = _cases.amr_totals(after, lay, blocks)
    scale = _cases.amr_totals(np.abs(before), lay, blocks) + 1e-300
    assert np.all(np.abs(t1 - t0) <= 1e-13 * scale), (t0, t1)


def test_refined_everywhere_equals_uniform_run(oracle):
    """A tree refined once everywhere is the uniform mesh of twice the blocks."""
    lay0 = Layout(3, 2, (8, 8, 8), (1, 1, 1))
    desc = mesh_desc(lay0)
    leaves = _leaves_uniform((1, 1, 1), 1, 3)
    cfg, _, _, _, _ = _cases.amr_case("smooth3d", ndim=3, nroot=1, nblock=8, ng=2, lmax=0)
    lay = Layout(3, 2, (8, 8, 8), (1, 1, 1), nleaves=8)
    cell, _ = km.problem_fill("smooth3d", cfg, desc, lay, seed=3, leaves=leaves)
    a = oracle.CpuSimulation(cfg.c, desc, lay, leaves=leaves)
    a.set_primitive(cell)
    a.advance(3)
    ua, _ = a.state()
    lay2 = Layout(3, 2, (8, 8, 8), (2, 2, 2))
    desc2 = mesh_desc(lay2)
    b = oracle.CpuSimulation(cfg.c, desc2, lay2)
    b.set_primitive(cell)
    b.advance(3)
    ub, _ = b.state()
    assert np.array_equal(ua, ub)
    assert a.info() == b.info()


def _random_tree(ndim, nb0, lmax, bc, rng, nseeds=3):
    """A random 2:1-balanced leaf set: a few finest-level positions want lmax, then every
    position next to a leaf two levels finer is raised until touching leaves differ by at
    most one level (faces, edges, corners, periodic wraps)."""
    nbf = [nb0 << lmax if d < ndim else 1 for d in range(3)]
    des = np.zeros(nbf[::-1], dtype=int)
    for _ in range(nseeds):
        pos = tuple(int(rng.integers(0, nbf[2 - a])) for a in range(3))
        des[pos] = lmax
    periodic = bc == "periodic"

    def leaves_of(des):
        out = []

        def visit(L, l):
            s = lmax - L
            sl = tuple(slice(l[2 - a] << s, (l[2 - a] + 1) << s) if 2 - a < ndim else slice(0, 1)
                       for a in range(3))
            if des[sl].max() <= L:
                out.append((L, l[0], l[1], l[2]))
                return
            for ck in range(2 if ndim > 2 else 1):
                for cj in range(2 if ndim > 1 else 1):
                    for ci in range(2):
                        visit(L + 1, (2 * l[0] + ci, 2 * l[1] + cj if ndim > 1 else 0,
                                      2 * l[2] + ck if ndim > 2 else 0))
        for k in range(nb0 if ndim > 2 else 1):
            for j in range(nb0 if ndim > 1 else 1):
                for i in range(nb0):
                    visit(0, (i, j, k))
        return out

    for _ in range(64):
        leaves = leaves_of(des)
        lev = np.zeros_like(des)
        for L, i, j, k in leaves:
            s = lmax - L
            sl = tuple(slice(x << s, (x + 1) << s) if d < ndim else slice(0, 1)
                       for d, x in ((2, k), (1, j), (0, i)))
            lev[sl] = L
        changed = False
        for off in np.ndindex(*(3,) * 3):
            o = [off[2] - 1, off[1] - 1, off[0] - 1]  # x, y, z offsets
            if any(o[d] != 0 and d >= ndim for d in range(3)) or o == [0, 0, 0]:
                continue
            nbr = lev
            valid = np.ones_like(lev, dtype=bool)
            for d in range(ndim):
                ax = 2 - d
                nbr = np.roll(nbr, -o[d], axis=ax)
                if not periodic and o[d] != 0:
                    idx = [slice(None)] * 3
                    idx[ax] = slice(-1, None) if o[d] > 0 else slice(0, 1)
                    valid[tuple(idx)] = False
            need = valid & (nbr > lev + 1) & (des < nbr - 1)
            if need.any():
                des = np.where(need, nbr - 1, des)
                changed = True
        if not changed:
            return np.array(leaves, dtype=np.int32)
    raise AssertionError("no balance")


@pytest.mark.parametrize("ng,nx", [(4, 8), (3, 6), (2, 4), (1, 4)])
@pytest.mark.parametrize("ndim", [2, 3])
def test_smallest_blocks_fill_without_deep_reads(oracle, ng, nx, ndim):
    """At the smallest accepted block size every read of a ghost fill lands one level
    away at most (the oracle fails loudly otherwise), on random balanced trees."""
    rng = np.random.default_rng(100 * ng + ndim)
    nxb = tuple(nx if d < ndim else 1 for d in range(3))
    for trial in range(6):
        bc = ("periodic", "outflow", "reflect")[trial % 3]
        lmax = 2 + (trial % 2) if ndim == 2 else 2
        nb0 = 2
        root = Layout(ndim, ng, nxb, tuple(nb0 if d < ndim else 1 for d in range(3)))
        desc = mesh_desc(root, bc=(bc,) * 6)
        leaves = _random_tree(ndim, nb0, lmax, bc, rng)
        blocks = km.amr_blocks(desc, leaves)  # the library accepts the tree
        lay = Layout(ndim, ng, nxb, root.nb, nleaves=len(leaves))
        h = _pack(oracle, lay, 12, oracle.amr_blocks(desc, leaves))
        h.cell[:] = rng.uniform(-1, 1, h.cell.shape)
        oracle.call("km_cpu_amr_ghost_fill", C.byref(km.Config().c), C.byref(h.pack),
                    C.byref(desc), oracle._i32(leaves), len(leaves), 0, 11)
        assert len(blocks) == len(leaves)


def _interface_mismatch(faces, lay, leaves, blocks):
    """max over coarse block faces shared with finer leaves of |A_c B_c - sum A_f B_f|,
    relative to max |B| A_c."""
    ndim, nx, ng = lay.ndim, lay.nx, lay.ng
    by_loc = {(int(l[0]), int(l[1]), int(l[2]), int(l[3])): n for n, l in enumerate(leaves)}
    worst, scale = 0.0, 0.0
    for b, lf in enumerate(leaves):
        L = int(lf[0])
        loc = [int(lf[1]), int(lf[2]), int(lf[3])]
        for d in range(ndim):
            for side in (0, 1):
                nloc = [2 * x for x in loc]
                nloc[d] = 2 * loc[d] - 1 if side == 0 else 2 * loc[d] + 2
                trans = [e for e in range(ndim) if e != d]
                # the 2^(ndim-1) finer leaves behind this face
                for q in range(1 << len(trans)):
                    fl = list(nloc)
                    for n, e in enumerate(trans):
                        fl[e] += (q >> n) & 1
                    key = (L + 1, fl[0], fl[1] if ndim > 1 else 0, fl[2] if ndim > 2 else 0)
                    if key not in by_loc:
                        continue
                    fb = by_loc[key]
                    ci = ng + (nx[d] if side else 0)      # coarse face index along d
                    fi = ng + (0 if side else nx[d])      # fine face index along d
                    half = [nx[e] // 2 for e in range(3)]
                    rng_t = [range(half[e]) if e in trans else range(1) for e in range(3)]
                    for c2 in rng_t[trans[1]] if len(trans) > 1 else [0]:
                        for c1 in rng_t[trans[0]]:
                            cidx = [0, 0, 0]
                            cidx[d] = ci
                            cidx[trans[0]] = ng + ((q >> 0) & 1) * half[trans[0]] + c1
                            if len(trans) > 1:
                                cidx[trans[1]] = ng + ((q >> 1) & 1) * half[trans[1]] + c2
                            Bc = faces[d][b][cidx[2], cidx[1], cidx[0]]
                            tot = 0.0
                            for s2 in (range(2) if len(trans) > 1 else [0]):
                                for s1 in range(2):
                                    fidx = [0, 0, 0]
                                    fidx[d] = fi
                                    fidx[trans[0]] = ng + 2 * c1 + s1
                                    if len(trans) > 1:
                                        fidx[trans[1]] = ng + 2 * c2 + s2
                                    tot += faces[d][fb][fidx[2], fidx[1], fidx[0]] * \
                                        blocks[fb].area[d]
                            worst = max(worst, abs(Bc * blocks[b].area[d] - tot))
                            scale = max(scale, abs(Bc) * blocks[b].area[d])
    return worst / max(scale, 1e-300)


@pytest.mark.parametrize("ndim", [2, 3])
def test_ct_blast_on_tree_keeps_divb_and_interfaces(oracle, ndim):
    """CT MHD blast on a three-level (2-D) / two-level (3-D) tree with nghost 4 and 8-cell
    blocks: divB stays at round-off in every interior and in every in-domain ghost cell
    (shared-face prolongation plus Toth & Roe), and each coarse face at a level jump
    keeps carrying the area sum of the fine faces on it (EMF correction)."""
    if ndim == 2:
        case = _cases.amr_mhd_case(ndim=2, nroot=4, nblock=8, ng=4, lmax=2, radius=0.12)
        ncyc = 8
    else:
        case = _cases.amr_mhd_case(ndim=3, nroot=2, nblock=8, ng=4, lmax=1,
                                   center=(0.1, 0.15, 0.2), radius=0.1)
        ncyc = 3
    cfg, desc, lay, leaves, cell, faces = case
    assert len(set(leaves[:, 0])) >= 2
    blocks = oracle.amr_blocks(desc, leaves)
    sim = oracle.CpuSimulation(cfg.c, desc, lay, leaves=leaves)
    sim.set_primitive(cell, faces)
    sim.advance(ncyc)
    out, fo = sim.state()
    div = _cases.divb_cells(fo, lay, blocks)
    sl = lay.interior()
    bmax = max(float(np.abs(f).max()) for f in fo[:ndim])
    dxmin = min(blocks[b].dx[0] for b in range(lay.nblocks))
    interior = max(float(np.abs(dv[sl]).max()) for dv in div) * dxmin / bmax
    assert interior < 1e-12, interior
    # ghost cells inside the domain
    worst = 0.0
    for b in range(lay.nblocks):
        bk = blocks[b]
        inside = np.ones(div[b].shape, dtype=bool)
        for d in range(ndim):
            x = bk.xmin[d] + (np.arange(lay.E[d]) - lay.lo(d) + 0.5) * bk.dx[d]
            ok = (x > -0.5) & (x < 0.5)
            shape = [1, 1, 1]
            shape[2 - d] = -1
            inside &= ok.reshape(shape)
        worst = max(worst, float(np.abs(div[b][inside]).max()) * bk.dx[0] / bmax)
    assert worst < 1e-12, worst
    assert _interface_mismatch(fo, lay, leaves, blocks) < 1e-13
    assert np.isfinite(out).all()


def test_face_fill_reproduces_linear_divergence_free_field(oracle):
    """Shared-face minmod prolongation and Toth & Roe are exact for a linear
    divergence-free B; every in-domain ghost face equals the field at its centre."""
    ndim = 3
    root = Layout(3, 4, (8, 8, 8), (2, 2, 2))
    desc = mesh_desc(root, bc=("outflow",) * 6)
    leaves = km.refine_ball(desc, 1, (0.25, 0.3, 0.7), 0.1)
    assert set(leaves[:, 0]) == {0, 1}
    lay = Layout(3, 4, (8, 8, 8), (2, 2, 2), nleaves=len(leaves))
    blocks = oracle.amr_blocks(desc, leaves)
    h = oracle.HostPack(lay, 16, faces=True, blocks=blocks)
    # B = (1 + 0.3x - 0.2y + 0.1z, -0.5 + 0.4x + 0.2y - 0.3z, 0.25 - 0.1x + 0.6y - 0.1z)
    C = np.array([[1.0, 0.3, -0.2, 0.1], [-0.5, 0.4, 0.2, -0.3], [0.25, -0.1, 0.6, -0.1]])

    def coords(bk, d, comp):
        n = lay.S[d]
        off = 0.0 if d == comp else 0.5
        return bk.xmin[d] + (np.arange(n) - lay.lo(d) + off) * bk.dx[d]

    def field(b, comp):
        bk = blocks[b]
        x, y, z = (coords(bk, d, comp) for d in range(3))
        c = C[comp]
        return c[0] + c[1] * x[None, None, :] + c[2] * y[None, :, None] + c[3] * z[:, None, None]

    lo, S = lay.ng, lay.S
    for b in range(lay.nblocks):
        for comp in range(3):
            f = field(b, comp)
            own = [slice(lo, lo + 8 + (1 if d == comp else 0)) for d in range(3)]
            h.face[comp][b][own[2], own[1], own[0]] = f[own[2], own[1], own[0]]
    oracle.call("km_cpu_amr_ghost_faces", C_byref(km.Config(**{"physics/MHD": "ct"}).c),
                C_byref(h.pack), C_byref(desc), oracle._i32(leaves), len(leaves))
    checked = 0
    margin = 2.0 / 16  # two level-0 cells: stencils there see the outflow boundary
    for b in range(lay.nblocks):
        bk = blocks[b]
        for comp in range(3):
            f = field(b, comp)
            inside = np.ones(f.shape, dtype=bool)
            for d in range(3):
                x = coords(bk, d, comp)
                ok = (x > margin) & (x < 1.0 - margin)
                if d != comp:  # the last row across a face array is padding
                    ok[lay.E[d]:] = False
                shape = [1, 1, 1]
                shape[2 - d] = -1
                inside &= ok.reshape(shape)
            got = h.face[comp][b]
            assert np.allclose(got[inside], f[inside], rtol=0, atol=1e-13), (b, comp)
            checked += int(inside.sum())
    assert checked > 0


def C_byref(x):
    return C.byref(x)


def _curl_faces(lay, leaves, blocks, lmax, seed):
    """Face B of every leaf as the discrete curl of edge circulations of a smooth periodic
    vector potential, each edge integral summed from the finest-level sub-edges, so every
    cell at every level is divergence free and a coarse face carries exactly the flux of
    the fine faces on it (up to summation order)."""
    rng = np.random.default_rng(seed)
    modes = []
    for _ in range(6):
        k = rng.integers(-2, 3, size=3)
        if not k.any():
            k[0] = 1
        modes.append((k, rng.uniform(-0.3, 0.3, size=3), rng.uniform(0, 2 * np.pi)))

    def A(comp, x, y, z):
        out = np.zeros(np.broadcast(x, y, z).shape)
        for k, amp, ph in modes:
            out += amp[comp] * np.sin(2 * np.pi * (k[0] * x + k[1] * y + k[2] * z) + ph)
        return out

    ng = lay.ng
    faces = [np.zeros(lay.stag_shape()) for _ in range(3)]
    for b, lf in enumerate(leaves):
        bk = blocks[b]
        nsub = 1 << (lmax - int(lf[0]))
        xn = [bk.xmin[d] + (np.arange(lay.S[d]) - ng) * bk.dx[d] for d in range(3)]

        def edge_int(t, a, c):
            # integral of A_t along t over each cell, at node coordinates a (axis p), c (q)
            p, q = [e for e in range(3) if e != t]
            h = bk.dx[t] / nsub
            tot = 0.0
            for s in range(nsub):
                xt = bk.xmin[t] + (np.arange(lay.S[t]) - ng) * bk.dx[t] + (s + 0.5) * h
                X = [None, None, None]
                X[t] = xt.reshape([-1 if e == t else 1 for e in (2, 1, 0)])
                X[p] = a.reshape([-1 if e == p else 1 for e in (2, 1, 0)])
                X[q] = c.reshape([-1 if e == q else 1 for e in (2, 1, 0)])
                tot = tot + A(t, X[0], X[1], X[2]) * h
            return tot  # shape broadcast (S3, S2, S1)

        I = [edge_int(t, *[xn[e] for e in range(3) if e != t]) for t in range(3)]
        for d in range(3):
            p, q = [e for e in range(3) if e != d]
            Ip, Iq = I[p], I[q]

            def sh(a, e):  # value at index + 1 along axis e
                out = np.zeros_like(a)
                sl = [slice(None)] * 3
                sl2 = [slice(None)] * 3
                sl[2 - e] = slice(0, -1)
                sl2[2 - e] = slice(1, None)
                out[tuple(sl)] = a[tuple(sl2)]
                return out
            # flux_d = (I_q at p+1 - I_q at p) - (I_p at q+1 - I_p at q)  (cyclic d, p, q)
            if (d, p, q) in ((0, 1, 2), (1, 2, 0), (2, 0, 1)):
                circ = (sh(Iq, p) - Iq) - (sh(Ip, q) - Ip)
            else:
                circ = (sh(Ip, q) - Ip) - (sh(Iq, p) - Iq)
            faces[d][b] = circ / bk.area[d]
    return faces


def test_face_fill_keeps_curl_fields_divergence_free(oracle):
    """Ghost faces of a periodic three-level tree, filled from a discrete-curl field that is
    divergence free and flux-consistent across levels, leave every ghost cell divergence
    free: shared-face prolongation keeps each coarse face's flux and Toth & Roe closes each
    fine cell."""
    root = Layout(3, 4, (8, 8, 8), (2, 2, 2))
    desc = mesh_desc(root)
    leaves = km.refine_ball(desc, 2, (0.3, 0.6, 0.4), 0.08)
    assert set(leaves[:, 0]) == {1, 2}
    lay = Layout(3, 4, (8, 8, 8), (2, 2, 2), nleaves=len(leaves))
    blocks = oracle.amr_blocks(desc, leaves)
    faces = _curl_faces(lay, leaves, blocks, 2, seed=9)
    h = oracle.HostPack(lay, 16, faces=True, blocks=blocks)
    for d in range(3):
        h.face[d][:] = faces[d]
    div0 = _cases.divb_cells(h.face, lay, blocks)
    sl = lay.interior()
    bmax = max(float(np.abs(f).max()) for f in faces)
    assert max(float(np.abs(dv[sl]).max()) * blocks[b].dx[0]
               for b, dv in enumerate(div0)) < 1e-12 * bmax
    oracle.call("km_cpu_amr_ghost_faces", C.byref(km.Config(**{"physics/MHD": "ct"}).c),
                C.byref(h.pack), C.byref(desc), oracle._i32(leaves), len(leaves))
    div = _cases.divb_cells(h.face, lay, blocks)
    worst = max(float(np.abs(dv).max()) * blocks[b].dx[0] for b, dv in enumerate(div)) / bmax
    assert worst < 1e-12, worst
