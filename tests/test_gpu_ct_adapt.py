"""GPU parity of the CT face field's remesh and of adaptive CT runs against the oracle, bit
for bit: kept, split and merged leaves in 2-D and 3-D, and adaptive CT blasts whose trees
change between cycles (the mhd_blast suite's adaptive Cartesian shape, mhd_blast.py:76-77,
and a 3-D two-level run), with divB at round-off."""
import ctypes as C

import numpy as np
import pytest

from kamayan_amd import Layout, mesh_desc
from kamayan_amd._abi import amr_adapt, amr_criterion, slot
from tests import _cases

pytestmark = pytest.mark.gpu
R, S, D = 1, 0, -1


def _same(a, b, what):
    if not np.array_equal(a, b):
        bad = np.argwhere(a != b)
        i = tuple(bad[0])
        raise AssertionError(f"{what}: {len(bad)} differ, first {i}: {a[i]!r} vs {b[i]!r}")


@pytest.mark.parametrize("ndim", [2, 3])
def test_face_remesh_bitwise(gpu, oracle, ndim):
    km = gpu
    rng = np.random.default_rng(51 + ndim)
    nb = (4, 4, 1) if ndim == 2 else (2, 2, 2)
    lay0 = Layout(ndim, 4, tuple(8 if d < ndim else 1 for d in range(3)), nb)
    desc = mesh_desc(lay0, bc=("outflow",) * 6)
    leaves = km.refine_ball(desc, 2, (0.3, 0.4, 0.5), 0.2)
    lay = Layout(ndim, 4, lay0.nx, nb, nleaves=len(leaves))
    tags = np.where(leaves[:, 0] == 2, D, rng.choice([S, R], size=len(leaves), p=[0.6, 0.4]))
    tags = tags.astype(np.int32)
    nl, _, org, ch = km.amr_update_tree(desc, 4, 0, leaves, tags, np.zeros(len(leaves), np.int32))
    assert ch and set(org[:, 0].tolist()) == {0, 1, 2}
    print(f"face remesh ndim={ndim} {len(leaves)} -> {len(nl)} leaves", flush=True)
    lay2 = Layout(ndim, 4, lay0.nx, nb, nleaves=len(nl))
    cell = rng.uniform(0.5, 2.0, lay.cell_shape(16))
    faces = [rng.uniform(-1.0, 1.0, lay.stag_shape()) for _ in range(3)]
    old_d = km.DevicePack(lay, 16, faces=True, blocks=km.amr_blocks(desc, leaves))
    old_d.upload(cell=cell, face=faces)
    new_d = km.DevicePack(lay2, 16, faces=True, blocks=km.amr_blocks(desc, nl))
    org_d = C.c_void_p()
    km.check(km.lib().km_malloc(C.byref(org_d), 4 * org.size), "malloc")
    km.check(km.lib().km_memcpy_h2d(org_d, km._i32(np.ascontiguousarray(org)), 4 * org.size),
             "h2d")
    try:
        km.amr_remesh(old_d, new_d, org_d.value, 0, 16)
        got, _, gf, _ = new_d.download()
    finally:
        km.lib().km_free(org_d)
    old_h = oracle.HostPack(lay, 16, faces=True, blocks=oracle.amr_blocks(desc, leaves))
    old_h.cell[:] = cell
    for d in range(3):
        old_h.face[d][:] = faces[d]
    new_h = oracle.HostPack(lay2, 16, faces=True, blocks=oracle.amr_blocks(desc, nl))
    oracle.amr_remesh(old_h.pack, new_h.pack, org, 0, 16)
    _same(got, new_h.cell, "cells")
    for d in range(3):
        _same(gf[d], new_h.face[d], f"face {d}")
    old_d.close()
    new_d.close()


CT_RUNS = [
    # mhd_blast.py:76-77: 32^2 root in 8^2 blocks, nghost 4, numlevel 3, wenoz + hll, rk2
    ("blast2d-numlevel3", dict(ndim=2, nroot=4, nblock=8, ng=4, lmax=0, recon="wenoz",
                               riemann="hll"), 3, 16, 2),
    ("blast3d-numlevel2", dict(ndim=3, nroot=2, nblock=8, ng=4, lmax=0, recon="plm",
                               riemann="hll", center=(0.05, 0.1, 0.0), radius=0.2), 2, 4, 2),
]


@pytest.mark.parametrize("name,kw,numlevel,ncyc,dcount", CT_RUNS, ids=[r[0] for r in CT_RUNS])
def test_adaptive_ct_run_bitwise(gpu, oracle, name, kw, numlevel, ncyc, dcount):
    km = gpu
    cfg, desc, lay, leaves, cell, faces = _cases.amr_mhd_case(**kw)
    adapt = amr_adapt(numlevel, [amr_criterion(slot("PRES", 1), numlevel)],
                      derefine_count=dcount)
    print(f"adaptive CT run {name}", flush=True)
    sim = km.Simulation(cfg, desc, lay, path=km.PATH_TASKS, leaves=leaves)
    sim.set_primitive(cell, faces)
    sim.set_adaptive(adapt)
    ref = oracle.CpuSimulation(cfg.c, desc, lay, leaves=leaves)
    ref.set_primitive(cell, faces)
    ref.set_adaptive(adapt)
    seen = {len(leaves)}
    for c in range(ncyc):
        sim.advance(1)
        ref.advance(1)
        gl = sim.leaves()
        _same(gl, ref.leaves(), f"{name} leaves after cycle {c + 1}")
        seen.add(len(gl))
    g, gf = sim.state()
    r, rf = ref.state()
    _same(g, r, f"{name} cells")
    for d in range(3):
        _same(gf[d], rf[d], f"{name} face {d}")
    assert sim.info() == ref.info()
    assert len(seen) > 1
    blocks = oracle.amr_blocks(desc, sim.leaves())
    div = _cases.divb_cells(gf, sim.layout, blocks)
    sl = sim.layout.interior()
    assert max(float(np.abs(dv[sl]).max()) for dv in div) < 1e-10
    sim.close()
