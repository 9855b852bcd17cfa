"""Pins of the oracle's remesh of the CT face field (Parthenon's RedistributeAndRefineMeshBlocks
with kamayan's face operators: ProlongateSharedMinMod and ProlongateInternalTothAndRoe
for new fine blocks, RestrictAverage for merged ones, refinement_operations.hpp:52-398)
and of adaptive CT runs. Pinned by: exactness of a split on a linear divergence-free B,
split-then-merge restoring the faces, zero divergence in every cell of split and merged
blocks for a discrete-curl field, and the mhd_blast suite's adaptive Cartesian run keeping
divB at round-off while the tree changes (mhd_blast.py:76-77,137-155)."""
import ctypes as C

import numpy as np
import pytest

import kamayan_amd as km
from kamayan_amd import Layout, mesh_desc
from kamayan_amd._abi import amr_adapt, amr_criterion, slot
from tests import _cases
from tests.test_oracle_amr import _curl_faces

R, S, D = 1, 0, -1
# a linear divergence-free field (dBx/dx + dBy/dy + dBz/dz = 0.3 + 0.2 - 0.5)
LIN = np.array([[1.0, 0.3, -0.2, 0.1], [-0.5, 0.4, 0.2, -0.3], [0.25, -0.1, 0.6, -0.5]])
LIN2 = np.array([[1.0, 0.3, -0.2, 0.0], [-0.5, 0.4, -0.3, 0.0], [0.25, -0.1, 0.6, 0.0]])


def _face_coords(lay, bk, d, comp):
    off = 0.0 if d == comp else 0.5
    lo = lay.ng if d < lay.ndim else 0
    return bk.xmin[d] + (np.arange(lay.S[d]) - lo + off) * bk.dx[d]


def _linear_faces(lay, blocks, C):
    out = [np.zeros(lay.stag_shape()) for _ in range(3)]
    for b in range(lay.nblocks):
        for comp in range(3):
            x, y, z = (_face_coords(lay, blocks[b], d, comp) for d in range(3))
            c = C[comp]
            out[comp][b] = (c[0] + c[1] * x[None, None, :] + c[2] * y[None, :, None] +
                            c[3] * z[:, None, None])
    return out


def _own(lay, comp):
    """slices (k, j, i) of a block's own faces of component comp"""
    sl = []
    for d in range(3):
        if d < lay.ndim:
            sl.append(slice(lay.ng, lay.ng + lay.nx[d] + (1 if d == comp else 0)))
        else:
            sl.append(slice(0, 1))
    return (sl[2], sl[1], sl[0])


def _mesh(ndim):
    nxb = tuple(8 if d < ndim else 1 for d in range(3))
    nbs = tuple(2 if d < ndim else 1 for d in range(3))
    root = Layout(ndim, 4, nxb, nbs)
    desc = mesh_desc(root, bc=("outflow",) * 6)
    leaves = np.array([(0, i, j, k) for k in range(nbs[2]) for j in range(nbs[1])
                       for i in range(nbs[0])], np.int32)
    return desc, root, leaves


def _remesh(oracle, desc, lay, leaves, faces, tags, numlevel=3):
    nl, _, org, ch = oracle.amr_update_tree(desc, numlevel, 0, leaves, np.asarray(tags, np.int32),
                                            np.zeros(len(leaves), np.int32))
    assert ch
    old = oracle.HostPack(lay, 16, faces=True, blocks=oracle.amr_blocks(desc, leaves))
    for d in range(3):
        old.face[d][:] = faces[d]
    lay2 = Layout(lay.ndim, lay.ng, lay.nx, lay.nb, nleaves=len(nl))
    blocks2 = oracle.amr_blocks(desc, nl)
    new = oracle.HostPack(lay2, 16, faces=True, blocks=blocks2)
    oracle.amr_remesh(old.pack, new.pack, org, 0, 16)
    return nl, org, lay2, blocks2, [f.copy() for f in new.face]


@pytest.mark.parametrize("ndim", [2, 3])
def test_split_reproduces_a_linear_divergence_free_field(oracle, ndim):
    desc, root, leaves = _mesh(ndim)
    lay = Layout(ndim, 4, root.nx, root.nb, nleaves=len(leaves))
    C = LIN if ndim == 3 else LIN2
    faces = _linear_faces(lay, oracle.amr_blocks(desc, leaves), C)
    tags = [R] + [S] * (len(leaves) - 1)
    nl, org, lay2, blocks2, nf = _remesh(oracle, desc, lay, leaves, faces, tags)
    want = _linear_faces(lay2, blocks2, C)
    nsplit = 0
    for b in range(len(nl)):
        for comp in range(3):
            own = _own(lay2, comp)
            got, ref = nf[comp][b][own], want[comp][b][own]
            if org[b, 0] == 0:  # kept leaves copy every face, ghosts included
                assert np.array_equal(nf[comp][b], faces[comp][org[b, 1]])
            else:
                nsplit += 1
                assert np.allclose(got, ref, rtol=0, atol=1e-13), (b, comp)
    assert nsplit == 3 * (1 << ndim)


@pytest.mark.parametrize("ndim", [2, 3])
def test_split_then_merge_restores_the_faces(oracle, ndim):
    desc, root, leaves = _mesh(ndim)
    lay = Layout(ndim, 4, root.nx, root.nb, nleaves=len(leaves))
    rng = np.random.default_rng(11)
    faces = [rng.uniform(-1, 1, lay.stag_shape()) for _ in range(3)]
    nl, org, lay2, blocks2, nf = _remesh(oracle, desc, lay, leaves, faces, [R] * len(leaves))
    back, org2, lay3, blocks3, bf = _remesh(oracle, desc, lay2, nl, nf, [D] * len(nl))
    assert sorted(map(tuple, back)) == sorted(map(tuple, leaves))
    for b in range(len(back)):
        ob = [tuple(r) for r in leaves].index(tuple(back[b]))
        for comp in range(3):
            own = _own(lay3, comp)
            assert np.allclose(bf[comp][b][own], faces[comp][ob][own], rtol=1e-14, atol=1e-15)


def _psi_faces(lay, blocks, seed):
    """2-D face B from a flux function at the cell corners, Bx = dpsi/dy, By = -dpsi/dx,
    differenced over each face: every cell is divergence free and a coarse face carries
    exactly the difference of psi across the fine faces on it (up to rounding)."""
    rng = np.random.default_rng(seed)
    modes = [(rng.integers(1, 4, size=2), rng.uniform(-0.3, 0.3), rng.uniform(0, 2 * np.pi))
             for _ in range(5)]

    def psi(x, y):
        return sum(a * np.sin(2 * np.pi * (k[0] * x + k[1] * y) + ph) for k, a, ph in modes)

    out = [np.zeros(lay.stag_shape()) for _ in range(3)]
    g = lay.ng
    for b in range(lay.nblocks):
        bk = blocks[b]
        xn = bk.xmin[0] + (np.arange(lay.S[0]) - g) * bk.dx[0]
        yn = bk.xmin[1] + (np.arange(lay.S[1]) - g) * bk.dx[1]
        P = psi(xn[None, :], yn[:, None])  # (S2, S1) corner values
        bx = np.zeros((lay.S[1], lay.S[0]))
        by = np.zeros((lay.S[1], lay.S[0]))
        bx[:-1, :] = (P[1:, :] - P[:-1, :]) / bk.dx[1]
        by[:, :-1] = -(P[:, 1:] - P[:, :-1]) / bk.dx[0]
        out[0][b, 0] = bx
        out[1][b, 0] = by
    return out


@pytest.mark.parametrize("ndim", [2, 3])
def test_split_and_merge_keep_every_cell_divergence_free(oracle, ndim):
    # a discrete-curl field on a two-level tree: split blocks (Toth & Roe closes each fine
    # cell) and merged ones (restriction sums the fine fluxes) stay divergence free
    nxb = tuple(8 if d < ndim else 1 for d in range(3))
    nbs = tuple(2 if d < ndim else 1 for d in range(3))
    root = Layout(ndim, 4, nxb, nbs)
    desc = mesh_desc(root)
    leaves = km.refine_ball(desc, 1, (0.3, 0.6, 0.4), 0.1)
    assert set(leaves[:, 0]) == {0, 1}
    lay = Layout(ndim, 4, nxb, nbs, nleaves=len(leaves))
    blocks = oracle.amr_blocks(desc, leaves)
    faces = _curl_faces(lay, leaves, blocks, 2, seed=5) if ndim == 3 else \
        _psi_faces(lay, blocks, seed=5)
    div0 = _cases.divb_cells(faces, lay, blocks)
    sl0 = lay.interior()
    assert max(float(np.abs(dv[sl0]).max()) for dv in div0) < 1e-9
    h = oracle.HostPack(lay, 16, faces=True, blocks=blocks)
    for d in range(3):
        h.face[d][:] = faces[d]
    oracle.call("km_cpu_amr_ghost_faces", C.byref(km.Config(**{"physics/MHD": "ct"}).c),
                C.byref(h.pack), C.byref(desc), oracle._i32(leaves), len(leaves))
    filled = [f.copy() for f in h.face]
    # split every level-0 leaf and merge the level-1 families
    tags = [R if lf[0] == 0 else D for lf in leaves]
    nl, org, lay2, blocks2, nf = _remesh(oracle, desc, lay, leaves, filled, tags)
    assert set(org[:, 0]) >= {1, 2}
    div = _cases.divb_cells(nf, lay2, blocks2)
    sl = lay2.interior()
    bmax = max(float(np.abs(f).max()) for f in faces[:ndim])
    worst = max(float(np.abs(div[b][sl]).max()) * blocks2[b].dx[0] / bmax
                for b in range(len(nl)))
    assert worst < 1e-12, worst


def test_adaptive_ct_blast_keeps_divb_at_round_off(oracle):
    # mhd_blast.py:76-77 (adaptive Cartesian): 32^2 root in 8^2 blocks, nghost 4, numlevel 3,
    # wenoz + hll, rk2, refinement on pressure; the suite requires divB below 1e-10
    cfg, desc, lay, leaves, cell, faces = _cases.amr_mhd_case(
        ndim=2, nroot=4, nblock=8, ng=4, lmax=0, recon="wenoz", riemann="hll")
    sim = oracle.CpuSimulation(cfg.c, desc, lay, leaves=leaves)
    sim.set_primitive(cell, faces)
    sim.set_adaptive(amr_adapt(3, [amr_criterion(slot("PRES", 1), 3)], derefine_count=2))
    seen = {len(leaves)}
    worst = 0.0
    for _ in range(10):
        sim.advance(1)
        lv = sim.leaves()
        seen.add(len(lv))
        out, fo = sim.state()
        blocks = oracle.amr_blocks(desc, lv)
        div = _cases.divb_cells(fo, sim.layout, blocks)
        sl = sim.layout.interior()
        worst = max(worst, max(float(np.abs(dv[sl]).max()) for dv in div))
    assert len(seen) > 2 and max(lv[:, 0]) == 2
    assert worst < 1e-10, worst
