"""Pins of the oracle's block-tree operators in cylindrical (r, z) geometry: kamayan's
geometry-templated refinement operators (refinement_operations.hpp:52-398 with
Coordinates<cylindrical>, geometry.hpp:35-191) and the axisymmetric boundary across levels
(boundary_conditions.cpp:23-71). Pinned by: a one-level tree filling its ghost zones and
faces exactly as the uniform axisymmetric fill does, restriction conserving the
cylindrical volume integral, prolongation exact on linear fields, split and merge of a
flux-function field keeping every cell divergence free in the cylindrical metric, exact
mass conservation of a cylindrical run across level jumps (flux correction with
cylindrical face areas), and the mhd_blast suite's adaptive cylindrical run keeping divB
below its 1e-10 (mhd_blast.py:66-73,137-155)."""
import ctypes as C

import numpy as np
import pytest

import kamayan_amd as km
from kamayan_amd import Config, Layout, mesh_desc
from kamayan_amd._abi import amr_adapt, amr_criterion, slot
from tests import _cases
from tests.test_oracle_cyl import _psi_faces

R, S, D = 1, 0, -1
TP = 2.0 * np.pi


def _rz_tree(lmax=1, center=(0.1, 0.1, 0.0), radius=0.12, bc_out="outflow", ng=4):
    # the mhd_blast cylindrical shape: r in [0, 0.5] (16 cells), z in [-0.5, 0.5] (32), 8^2
    root = Layout(2, ng, (8, 8, 1), (2, 4, 1))
    desc = mesh_desc(root, xmin=(0.0, -0.5, -0.5), xmax=(0.5, 0.5, 0.5),
                     bc=("axisymmetric", bc_out, bc_out, bc_out, "periodic", "periodic"))
    leaves = km.refine_ball(desc, lmax, center, radius)
    lay = Layout(2, ng, (8, 8, 1), (2, 4, 1), nleaves=len(leaves))
    return desc, lay, leaves


def _cfg(mhd=True):
    kw = {"geometry/geometry": "cylindrical", "hydro/reconstruction": "plm"}
    if mhd:
        kw["physics/MHD"] = "ct"
    return Config(**kw)


def _cyl_vol(bk, lay, i):
    g = lay.ng
    rm = bk.xmin[0] + (i - g) * bk.dx[0]
    rp = rm + bk.dx[0]
    return 0.5 * TP * (rp * rp - rm * rm) * bk.dx[1]


def test_one_level_tree_fills_like_the_uniform_axisymmetric_fill(oracle):
    desc, lay, leaves = _rz_tree(lmax=0)
    assert set(leaves[:, 0]) == {0}
    cfg = _cfg()
    rng = np.random.default_rng(2)
    nv = 16
    blocks = oracle.amr_blocks(desc, leaves)
    a = oracle.HostPack(lay, nv, faces=True, blocks=blocks)
    b = oracle.HostPack(lay, nv, faces=True, blocks=oracle.block_table(desc, lay.nblocks))
    cell = rng.uniform(-1, 1, lay.cell_shape(nv))
    # face values as a function of the global face position, so that the two blocks
    # holding a shared face agree, as they do in a CT state
    faces = [np.zeros(lay.stag_shape()) for _ in range(3)]
    g = lay.ng
    for bb in range(lay.nblocks):
        lx = leaves[bb]
        gi = lx[1] * 8 + np.arange(lay.S[0]) - g
        gj = lx[2] * 8 + np.arange(lay.S[1]) - g
        for d in range(3):
            faces[d][bb, 0] = np.sin(1.7 * gi[None, :] + 2.3 * gj[:, None] + d)
    a.cell[:] = cell
    b.cell[:] = cell
    for d in range(3):
        a.face[d][:] = faces[d]
        b.face[d][:] = faces[d]
    oracle.call("km_cpu_amr_ghost_fill", C.byref(cfg.c), C.byref(a.pack), C.byref(desc),
                oracle._i32(leaves), len(leaves), 0, 11)
    oracle.call("km_cpu_amr_ghost_faces", C.byref(cfg.c), C.byref(a.pack), C.byref(desc),
                oracle._i32(leaves), len(leaves))
    nb = (C.c_int32 * 3)(2, 4, 1)
    bc = (C.c_int32 * 6)(*[desc.bc[d] for d in range(6)])
    oracle.call("km_cpu_ghost_fill_uniform", C.byref(cfg.c), C.byref(b.pack), nb, bc, 0, 11, 1)
    assert np.array_equal(a.cell[:, :11], b.cell[:, :11])
    for d in range(2):  # the r and z faces (the 2-D fill leaves phi faces alone)
        assert np.array_equal(a.face[d], b.face[d]), d


def _remesh(oracle, desc, lay, leaves, cell, faces, tags, nv):
    nl, _, org, ch = oracle.amr_update_tree(desc, 3, 0, leaves, np.asarray(tags, np.int32),
                                            np.zeros(len(leaves), np.int32))
    assert ch
    old = oracle.HostPack(lay, nv, faces=faces is not None, blocks=oracle.amr_blocks(desc, leaves))
    old.cell[:] = cell
    if faces is not None:
        for d in range(3):
            old.face[d][:] = faces[d]
    lay2 = Layout(2, lay.ng, lay.nx, lay.nb, nleaves=len(nl))
    blocks2 = oracle.amr_blocks(desc, nl)
    new = oracle.HostPack(lay2, nv, faces=faces is not None, blocks=blocks2)
    oracle.amr_remesh(old.pack, new.pack, org, 0, nv, desc=desc, geometry=1)
    return nl, org, lay2, blocks2, new.cell.copy(), [None if f is None else f.copy() for f in new.face]


def test_merge_conserves_the_cylindrical_volume_integral(oracle):
    desc, lay, leaves = _rz_tree(lmax=1)
    rng = np.random.default_rng(4)
    cell = rng.uniform(0.5, 2.0, lay.cell_shape(1))
    tags = [D if lf[0] == 1 else S for lf in leaves]
    nl, org, lay2, blocks2, new, _ = _remesh(oracle, desc, lay, leaves, cell, None, tags, 1)
    assert set(org[:, 0]) >= {2}
    blocks = oracle.amr_blocks(desc, leaves)

    def total(c, ly, bl, n):
        g = ly.ng
        return sum(float(np.sum(c[b, 0, 0, g:g + 8, g:g + 8] *
                                np.array([_cyl_vol(bl[b], ly, i) for i in range(g, g + 8)])))
                   for b in range(n))
    t0 = total(cell, lay, blocks, len(leaves))
    t1 = total(new, lay2, blocks2, len(nl))
    assert t1 == pytest.approx(t0, rel=1e-13)


def test_split_is_exact_on_linear_fields(oracle):
    desc, lay, leaves = _rz_tree(lmax=0)
    blocks = oracle.amr_blocks(desc, leaves)
    cell = np.zeros(lay.cell_shape(1))
    g = lay.ng
    for b in range(lay.nblocks):
        bk = blocks[b]
        r = bk.xmin[0] + (np.arange(lay.E[0]) - g + 0.5) * bk.dx[0]
        z = bk.xmin[1] + (np.arange(lay.E[1]) - g + 0.5) * bk.dx[1]
        cell[b, 0, 0] = 0.5 + 1.5 * r[None, :] - 0.75 * z[:, None]
    tags = [R if (lf[1], lf[2]) == (0, 1) else S for lf in leaves]
    nl, org, lay2, blocks2, new, _ = _remesh(oracle, desc, lay, leaves, cell, None, tags, 1)
    for b in range(len(nl)):
        if org[b, 0] != 1:
            continue
        bk = blocks2[b]
        r = bk.xmin[0] + (np.arange(8) + 0.5) * bk.dx[0]
        z = bk.xmin[1] + (np.arange(8) + 0.5) * bk.dx[1]
        want = 0.5 + 1.5 * r[None, :] - 0.75 * z[:, None]
        assert np.allclose(new[b, 0, 0, g:g + 8, g:g + 8], want, rtol=0, atol=1e-14)


def _divb_cyl(faces, lay, blocks, n):
    """cylindrical discrete divergence (A+ B+ - A- B-) / V of every interior cell"""
    g = lay.ng
    out = []
    for b in range(n):
        bk = blocks[b]
        r = bk.xmin[0] + (np.arange(g, g + 9) - g) * bk.dx[0]  # faces of the interior
        br = faces[0][b, 0, g:g + 8, g:g + 9]
        bz = faces[1][b, 0, g:g + 9, g:g + 8]
        ar = r * bk.dx[1] * TP
        az = 0.5 * (r[1:] ** 2 - r[:-1] ** 2) * TP
        vol = 0.5 * TP * (r[1:] ** 2 - r[:-1] ** 2) * bk.dx[1]
        div = ((ar[None, 1:] * br[:, 1:] - ar[None, :-1] * br[:, :-1]) +
               (az[None, :] * (bz[1:, :] - bz[:-1, :]))) / vol[None, :]
        out.append(div)
    return out


def test_split_and_merge_keep_the_cylindrical_divergence_zero(oracle):
    desc, lay, leaves = _rz_tree(lmax=1)
    blocks = oracle.amr_blocks(desc, leaves)

    def psi(r, z):
        return 0.25 * r * r * (1.0 + 0.3 * np.sin(2 * np.pi * z)) * np.exp(-4 * r * r)
    faces = _psi_faces(lay, blocks, psi)
    div0 = _divb_cyl(faces, lay, blocks, len(leaves))
    bmax = max(float(np.abs(f).max()) for f in faces[:2])
    assert max(float(np.abs(d).max()) for d in div0) * blocks[0].dx[0] / bmax < 1e-12
    cfg = _cfg()
    h = oracle.HostPack(lay, 16, faces=True, blocks=blocks)
    for d in range(3):
        h.face[d][:] = faces[d]
    oracle.call("km_cpu_amr_ghost_faces", C.byref(cfg.c), C.byref(h.pack), C.byref(desc),
                oracle._i32(leaves), len(leaves))
    tags = [R if lf[0] == 0 else D for lf in leaves]
    nl, org, lay2, blocks2, _, nf = _remesh(oracle, desc, lay, leaves,
                                            np.zeros(lay.cell_shape(16)), h.face, tags, 16)
    assert set(org[:, 0]) >= {1, 2}
    div = _divb_cyl(nf, lay2, blocks2, len(nl))
    worst = max(float(np.abs(d).max()) * blocks2[b].dx[0] for b, d in enumerate(div)) / bmax
    assert worst < 1e-12, worst


def test_cylindrical_run_conserves_mass_across_level_jumps(oracle):
    # walls all round (reflect in r and z, the axis at r = 0): mass stays exactly
    desc, lay, leaves = _rz_tree(lmax=1, bc_out="reflect", ng=3)
    assert set(leaves[:, 0]) == {0, 1}
    cfg = Config(**{"geometry/geometry": "cylindrical", "hydro/reconstruction": "plm",
                    "hydro/slope_limiter": "mc", "hydro/riemann": "hllc",
                    "material/single/gamma": 1.4})
    cfg.c.ndim = 2
    blocks = oracle.amr_blocks(desc, leaves)
    cell = np.zeros(lay.cell_shape(12))
    g = lay.ng
    for b in range(lay.nblocks):
        bk = blocks[b]
        r = bk.xmin[0] + (np.arange(lay.E[0]) - g + 0.5) * bk.dx[0]
        z = bk.xmin[1] + (np.arange(lay.E[1]) - g + 0.5) * bk.dx[1]
        Z, Rr = np.meshgrid(z, r, indexing="ij")
        cell[b, slot("DENS"), 0] = 1.0 + 0.3 * np.exp(-20 * (Rr ** 2 + Z ** 2))
        cell[b, slot("PRES"), 0] = 1.0 + 2.0 * np.exp(-40 * ((Rr - 0.1) ** 2 + Z ** 2))
        cell[b, slot("VEL2"), 0] = 0.2 * Rr * np.exp(-10 * Rr ** 2)
    sim = oracle.CpuSimulation(cfg.c, desc, lay, leaves=leaves)
    sim.set_primitive(cell)

    def mass(st):
        return sum(float(np.sum(st[b, 0, 0, g:g + 8, g:g + 8] *
                                np.array([_cyl_vol(blocks[b], lay, i) for i in range(g, g + 8)])))
                   for b in range(lay.nblocks))
    s0, _ = sim.state()
    sim.advance(8)
    s1, _ = sim.state()
    assert mass(s1) == pytest.approx(mass(s0), rel=1e-13)
    assert np.max(np.abs(s1[:, 0] - s0[:, 0])) > 1e-4


def test_adaptive_cylindrical_ct_blast_keeps_divb(oracle):
    # mhd_blast.py:66-73: r in [0, 0.5] x z in [-0.5, 0.5] on 16 x 32 root zones in 8^2
    # blocks, nghost 4, numlevel 3, wenoz + hll, rk2, axisymmetric axis, outflow elsewhere
    desc, lay, leaves = _rz_tree(lmax=0)
    cfg = Config(**{"geometry/geometry": "cylindrical", "physics/MHD": "ct",
                    "hydro/reconstruction": "wenoz", "hydro/riemann": "hll",
                    "parthenon/time/integrator": "rk2",
                    "material/single/gamma": 1.66666667})
    cfg.c.ndim = 2
    cell, faces = km.problem_fill("mhd_blast", cfg, desc, lay, leaves=leaves)
    sim = oracle.CpuSimulation(cfg.c, desc, lay, leaves=leaves)
    sim.set_primitive(cell, faces)
    sim.set_adaptive(amr_adapt(3, [amr_criterion(slot("PRES", 1), 3)], derefine_count=2))
    seen = {len(leaves)}
    worst = 0.0
    for _ in range(8):
        sim.advance(1)
        lv = sim.leaves()
        seen.add(len(lv))
        st, _ = sim.state()
        sl = sim.layout.interior()
        worst = max(worst, float(np.abs(st[(slice(None), slot("DIVB", 1)) + sl]).max()))
    assert len(seen) > 2 and max(lv[:, 0]) == 2
    assert np.isfinite(st).all()
    assert worst < 1e-10, worst


def test_face_fill_is_exact_when_r_times_b_is_linear(oracle):
    # r B_r = c + e z and r B_z = a + b r are linear and divergence free in (r, z), so the
    # r-weighted operators reproduce them: restriction weighted by the cylindrical face
    # areas, prolongation of r B with the children divided by their radius, and Toth-Roe on
    # r B. Every in-domain ghost face of a three-level tree equals the field at its centre.
    root = Layout(2, 4, (8, 8, 1), (2, 4, 1))
    desc = mesh_desc(root, xmin=(0.5, -1.0, -0.5), xmax=(1.5, 1.0, 0.5), bc=("outflow",) * 6)
    leaves = km.refine_ball(desc, 2, (1.0, 0.2, 0.0), 0.15)
    assert set(leaves[:, 0]) == {0, 1, 2}
    lay = Layout(2, 4, (8, 8, 1), (2, 4, 1), nleaves=len(leaves))
    blocks = oracle.amr_blocks(desc, leaves)
    a_, b_, c_, e_ = 0.3, 1.2, -0.4, 0.9
    g = lay.ng

    def analytic(b, comp):
        bk = blocks[b]
        rf = bk.xmin[0] + (np.arange(lay.S[0]) - g) * bk.dx[0]
        zf = bk.xmin[1] + (np.arange(lay.S[1]) - g) * bk.dx[1]
        if comp == 0:
            zc = zf + 0.5 * bk.dx[1]
            return (c_ + e_ * zc[:, None]) / rf[None, :]
        rc = rf + 0.5 * bk.dx[0]
        return (a_ + b_ * rc[None, :]) / rc[None, :] + 0.0 * zf[:, None]

    h = oracle.HostPack(lay, 16, faces=True, blocks=blocks)
    for b in range(lay.nblocks):
        for comp in range(2):
            h.face[comp][b, 0] = analytic(b, comp)
    own = [f.copy() for f in h.face]
    for comp in range(2):  # ghost faces start wrong
        for b in range(lay.nblocks):
            m = np.ones(lay.stag_shape()[1:], bool)
            m[0, g:g + 8 + (comp == 1), g:g + 8 + (comp == 0)] = False
            h.face[comp][b][m] = 7.0
    cfg = _cfg()
    oracle.call("km_cpu_amr_ghost_faces", C.byref(cfg.c), C.byref(h.pack), C.byref(desc),
                oracle._i32(leaves), len(leaves))
    checked = 0
    margin = 2 * 0.125  # two level-0 zones from the outflow boundaries
    for b in range(lay.nblocks):
        bk = blocks[b]
        for comp in range(2):
            want = analytic(b, comp)
            rpos = bk.xmin[0] + (np.arange(lay.S[0]) - g + (0.5 if comp == 1 else 0.0)) * bk.dx[0]
            zpos = bk.xmin[1] + (np.arange(lay.S[1]) - g + (0.5 if comp == 0 else 0.0)) * bk.dx[1]
            inside = (((rpos > 0.5 + margin) & (rpos < 1.5 - margin))[None, :] &
                      ((zpos > -1.0 + margin) & (zpos < 1.0 - margin))[:, None])
            inside[lay.E[1]:, :] = inside[lay.E[1]:, :] & (comp == 1)
            inside[:, lay.E[0]:] = inside[:, lay.E[0]:] & (comp == 0)
            got = h.face[comp][b, 0]
            assert np.allclose(got[inside], want[inside], rtol=1e-13, atol=1e-13), (b, comp)
            assert np.array_equal(got[g:g + 8, g:g + 8], own[comp][b, 0][g:g + 8, g:g + 8])
            checked += int(inside.sum())
    assert checked > 0
