"""GPU parity of block-tree meshes in cylindrical (r, z) geometry against the oracle, bit
for bit: ghost fill of cells and faces across levels with the axisymmetric axis, flux and
EMF correction with cylindrical measures, the remesh of cells and faces, and whole runs --
a static three-level hydro run, a CT run and the mhd_blast suite's adaptive cylindrical
blast (mhd_blast.py:66-73)."""
import ctypes as C

import numpy as np
import pytest

from kamayan_amd import Config, Layout
from kamayan_amd._abi import amr_adapt, amr_criterion, slot
from tests.test_oracle_cyl_amr import _rz_tree

pytestmark = pytest.mark.gpu
R, S, D = 1, 0, -1


def _same(a, b, what):
    if not np.array_equal(a, b):
        bad = np.argwhere(a != b)
        i = tuple(bad[0])
        raise AssertionError(f"{what}: {len(bad)} differ, first {i}: {a[i]!r} vs {b[i]!r}")


def _consistent_faces(lay, leaves, rng):
    """face values as a function of global position and level (shared faces agree)"""
    g = lay.ng
    ph = rng.uniform(0, 2 * np.pi, size=3)
    faces = [np.zeros(lay.stag_shape()) for _ in range(3)]
    for b, lf in enumerate(leaves):
        gi = lf[1] * lay.nx[0] + np.arange(lay.S[0]) - g
        gj = lf[2] * lay.nx[1] + np.arange(lay.S[1]) - g
        s = 1.0 / (1 << int(lf[0]))
        for d in range(3):
            faces[d][b, 0] = np.sin(0.9 * s * gi[None, :] + 1.3 * s * gj[:, None] + ph[d])
    return faces


@pytest.mark.parametrize("lmax", [1, 2])
def test_cyl_tree_ghost_fill_and_corrections_bitwise(gpu, oracle, lmax):
    km = gpu
    desc, lay, leaves = _rz_tree(lmax=lmax, center=(0.05, 0.1, 0.0), radius=0.1)
    cfg = Config(**{"geometry/geometry": "cylindrical", "physics/MHD": "ct"})
    cfg.c.ndim = 2
    print(f"cylindrical tree ghost fill, lmax {lmax}, {len(leaves)} leaves", flush=True)
    rng = np.random.default_rng(60 + lmax)
    cell = rng.uniform(-1, 1, lay.cell_shape(16))
    faces = _consistent_faces(lay, leaves, rng)
    flux = [rng.standard_normal(lay.cell_shape(8)) if d < 2 else None for d in range(3)]
    edge = [rng.standard_normal(lay.stag_shape()) for _ in range(3)]
    blocks = km.amr_blocks(desc, leaves)
    dev = km.DevicePack(lay, 16, nflux=8, faces=True, edges=True, blocks=blocks)
    dev.upload(cell=cell, face=faces, flux=flux, edge=edge)
    host = oracle.HostPack(lay, 16, nflux=8, faces=True, edges=True,
                           blocks=oracle.amr_blocks(desc, leaves))
    host.cell[:] = cell
    for d in range(3):
        host.face[d][:] = faces[d]
        host.edge[d][:] = edge[d]
        if flux[d] is not None:
            host.flux[d][:] = flux[d]
    L, p = km.lib(), C.byref
    plan = C.c_void_p()
    km.check(L.km_ghost_plan_amr_geom(p(dev.pack), p(desc), km._i32(leaves), len(leaves), 1,
                                      p(plan)), "plan")
    try:
        km.check(L.km_ghost_fill(plan, p(dev.pack), 0, 11, 1, None), "ghost")
        km.check(L.km_flux_correct(plan, p(dev.pack), None), "flux correct")
    finally:
        L.km_ghost_plan_destroy(plan)
    oracle.call("km_cpu_amr_ghost_fill", p(cfg.c), p(host.pack), p(desc), oracle._i32(leaves),
                len(leaves), 0, 11)
    oracle.call("km_cpu_amr_ghost_faces", p(cfg.c), p(host.pack), p(desc), oracle._i32(leaves),
                len(leaves))
    oracle.call("km_cpu_flux_correct", p(cfg.c), p(host.pack), p(desc), oracle._i32(leaves),
                len(leaves))
    gc, gfl, gf, ge = dev.download()
    _same(gc, host.cell, "cells")
    for d in range(3):
        _same(gf[d], host.face[d], f"face {d}")
        _same(ge[d], host.edge[d], f"edge {d}")
        if flux[d] is not None:
            _same(gfl[d], host.flux[d], f"flux {d}")
    dev.close()


def test_cyl_remesh_bitwise(gpu, oracle):
    km = gpu
    desc, lay, leaves = _rz_tree(lmax=2, center=(0.05, 0.1, 0.0), radius=0.1)
    rng = np.random.default_rng(71)
    tags = np.where(leaves[:, 0] == 2, D, rng.choice([S, R], size=len(leaves), p=[0.6, 0.4]))
    tags = tags.astype(np.int32)
    nl, _, org, ch = km.amr_update_tree(desc, 4, 0, leaves, tags, np.zeros(len(leaves), np.int32))
    assert ch and set(org[:, 0].tolist()) == {0, 1, 2}
    print(f"cylindrical remesh {len(leaves)} -> {len(nl)} leaves", flush=True)
    lay2 = Layout(2, lay.ng, lay.nx, lay.nb, nleaves=len(nl))
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
        km.amr_remesh(old_d, new_d, org_d.value, 0, 16, desc=desc, geometry=1)
        got, _, gf, _ = new_d.download()
    finally:
        km.lib().km_free(org_d)
    old_h = oracle.HostPack(lay, 16, faces=True, blocks=oracle.amr_blocks(desc, leaves))
    old_h.cell[:] = cell
    for d in range(3):
        old_h.face[d][:] = faces[d]
    new_h = oracle.HostPack(lay2, 16, faces=True, blocks=oracle.amr_blocks(desc, nl))
    oracle.amr_remesh(old_h.pack, new_h.pack, org, 0, 16, desc=desc, geometry=1)
    _same(got, new_h.cell, "cells")
    for d in range(3):
        _same(gf[d], new_h.face[d], f"face {d}")
    old_d.close()
    new_d.close()


def _blast(km, oracle, adaptive, ncyc):
    desc, lay, leaves = _rz_tree(lmax=0 if adaptive else 2, center=(0.05, 0.05, 0.0),
                                 radius=0.12)
    cfg = Config(**{"geometry/geometry": "cylindrical", "physics/MHD": "ct",
                    "hydro/reconstruction": "wenoz", "hydro/riemann": "hll",
                    "parthenon/time/integrator": "rk2", "material/single/gamma": 1.66666667})
    cfg.c.ndim = 2
    cell, faces = km.problem_fill("mhd_blast", cfg, desc, lay, leaves=leaves)
    sim = km.Simulation(cfg, desc, lay, path=km.PATH_TASKS, leaves=leaves)
    sim.set_primitive(cell, faces)
    ref = oracle.CpuSimulation(cfg.c, desc, lay, leaves=leaves)
    ref.set_primitive(cell, faces)
    if adaptive:
        adapt = amr_adapt(3, [amr_criterion(slot("PRES", 1), 3)], derefine_count=2)
        sim.set_adaptive(adapt)
        ref.set_adaptive(adapt)
    seen = {len(leaves)}
    for c in range(ncyc):
        sim.advance(1)
        ref.advance(1)
        if adaptive:
            gl = sim.leaves()
            _same(gl, ref.leaves(), f"leaves after cycle {c + 1}")
            seen.add(len(gl))
    g, gf = sim.state()
    r, rf = ref.state()
    _same(g, r, "cells")
    for d in range(3):
        _same(gf[d], rf[d], f"face {d}")
    assert sim.info() == ref.info()
    sl = sim.layout.interior()
    assert float(np.abs(g[(slice(None), slot("DIVB", 1)) + sl]).max()) < 1e-10
    sim.close()
    return seen


def test_cyl_ct_blast_on_static_tree_bitwise(gpu, oracle):
    print("cylindrical CT blast, static three-level tree", flush=True)
    _blast(gpu, oracle, False, 6)


def test_adaptive_cyl_ct_blast_bitwise(gpu, oracle):
    print("cylindrical CT blast, adaptive (mhd_blast suite shape)", flush=True)
    seen = _blast(gpu, oracle, True, 10)
    assert len(seen) > 2
