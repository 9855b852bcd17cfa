"""GPU parity of the block-tree (AMR) path: ghost fill across levels, flux correction
and whole task-path runs against the oracle's composite-grid restatement, bit for bit
(FP64, both sides round each operation in the reference's order)."""
import ctypes as C

import numpy as np
import pytest

from tests import _cases

pytestmark = pytest.mark.gpu

MESHES = [
    # (ndim, nblock, ng, lmax, center, radius, bc)
    (2, 8, 2, 2, (0.3, 0.6, 0.5), 0.15, "periodic"),
    (2, 8, 2, 1, (0.05, 0.5, 0.5), 0.2, "reflect"),
    (3, 8, 2, 1, (0.3, 0.6, 0.5), 0.2, "periodic"),
    (3, 8, 2, 2, (0.2, 0.2, 0.8), 0.1, "outflow"),
    (3, 12, 4, 1, (0.2, 0.25, 0.3), 0.1, "reflect"),
    # the smallest blocks the reference's own multilevel tests use: 8 cells, nghost 4
    (2, 8, 4, 3, (0.4, 0.55, 0.5), 0.1, "outflow"),
    (3, 8, 4, 2, (0.3, 0.3, 0.6), 0.12, "periodic"),
    # odd ghost widths: coarse parents with one child inside the ghost layer
    (2, 8, 3, 2, (0.1, 0.45, 0.5), 0.2, "reflect"),
    (3, 8, 3, 1, (0.15, 0.3, 0.8), 0.15, "outflow"),
]


def _mesh(km, ndim, nblock, ng, lmax, center, radius, bc):
    nxb = tuple(nblock if d < ndim else 1 for d in range(3))
    nbs = tuple((4 if ndim == 2 else 2) if d < ndim else 1 for d in range(3))
    root = km.Layout(ndim, ng, nxb, nbs)
    desc = km.mesh_desc(root, bc=(bc,) * 6)
    leaves = km.refine_ball(desc, lmax, center, radius)
    lay = km.Layout(ndim, ng, nxb, nbs, nleaves=len(leaves))
    return desc, leaves, lay


def _same(a, b, what):
    if not np.array_equal(a, b):
        bad = np.argwhere(a != b)
        i = tuple(bad[0])
        raise AssertionError(f"{what}: {len(bad)} differ, first {i}: {a[i]!r} vs {b[i]!r}")


@pytest.mark.parametrize("mesh", MESHES, ids=lambda m: f"{m[0]}d-nx{m[1]}-l{m[3]}-{m[6]}")
def test_ghost_fill_across_levels_bitwise(gpu, oracle, mesh):
    km = gpu
    desc, leaves, lay = _mesh(km, *mesh)
    print(f"ghost fill {mesh} leaves={len(leaves)}", flush=True)
    blocks = km.amr_blocks(desc, leaves)
    rng = np.random.default_rng(17)
    cell = rng.uniform(-1.0, 2.0, lay.cell_shape(12))
    dev = km.DevicePack(lay, 12, blocks=blocks)
    dev.upload(cell=cell)
    host = oracle.HostPack(lay, 12, blocks=oracle.amr_blocks(desc, leaves))
    host.cell[:] = cell
    L = km.lib()
    plan = C.c_void_p()
    km.check(L.km_ghost_plan_amr(C.byref(dev.pack), C.byref(desc), km._i32(leaves), len(leaves),
                                 C.byref(plan)), "plan")
    try:
        km.check(L.km_ghost_fill(plan, C.byref(dev.pack), 0, 11, 0, None), "ghost")
    finally:
        L.km_ghost_plan_destroy(plan)
    oracle.call("km_cpu_amr_ghost_fill", C.byref(km.Config().c), C.byref(host.pack),
                C.byref(desc), oracle._i32(leaves), len(leaves), 0, 11)
    got, _, _, _ = dev.download()
    _same(got, host.cell, "cells")
    dev.close()


@pytest.mark.parametrize("mesh", MESHES[:4], ids=lambda m: f"{m[0]}d-l{m[3]}-{m[6]}")
def test_flux_correction_bitwise(gpu, oracle, mesh):
    km = gpu
    desc, leaves, lay = _mesh(km, *mesh)
    ndim = mesh[0]
    blocks = km.amr_blocks(desc, leaves)
    rng = np.random.default_rng(23)
    flux = [rng.standard_normal(lay.cell_shape(5)) if d < ndim else None for d in range(3)]
    dev = km.DevicePack(lay, 12, nflux=5, blocks=blocks)
    dev.upload(flux=flux)
    host = oracle.HostPack(lay, 12, nflux=5, blocks=oracle.amr_blocks(desc, leaves))
    for d in range(ndim):
        host.flux[d][:] = flux[d]
    L = km.lib()
    plan = C.c_void_p()
    km.check(L.km_ghost_plan_amr(C.byref(dev.pack), C.byref(desc), km._i32(leaves), len(leaves),
                                 C.byref(plan)), "plan")
    try:
        km.check(L.km_flux_correct(plan, C.byref(dev.pack), None), "flux correct")
    finally:
        L.km_ghost_plan_destroy(plan)
    oracle.call("km_cpu_flux_correct", C.byref(km.Config().c), C.byref(host.pack), C.byref(desc),
                oracle._i32(leaves), len(leaves))
    _, gflux, _, _ = dev.download()
    changed = 0
    for d in range(ndim):
        _same(gflux[d], host.flux[d], f"flux[{d}]")
        changed += int(np.sum(host.flux[d] != flux[d]))
    assert changed > 0  # the mesh has level jumps, so some faces were corrected
    dev.close()


RUNS = [
    ("smooth3d-periodic-plm", lambda: _cases.amr_case("smooth3d", ndim=3, nroot=2, nblock=8,
                                                      ng=2, lmax=1, center=(0.3, 0.6, 0.5),
                                                      radius=0.2), 4),
    ("smooth2d-wenoz-l2", lambda: _cases.amr_case("smooth3d", ndim=2, nroot=2, nblock=16, ng=3,
                                                  lmax=2, center=(0.4, 0.4, 0.5), radius=0.1,
                                                  recon="wenoz", integrator="rk3"), 4),
    ("sedov-outflow-plm-mc", lambda: _cases.amr_case(
        "sedov", ndim=3, nroot=4, nblock=12, ng=4, lmax=1, center=(0.0, 0.0, 0.0), radius=0.15,
        bc="outflow", xmin=(-0.5,) * 3, xmax=(0.5,) * 3, gamma=1.4,
        params=[1.0, 2.0 / 96.0, 1.0, 1e-5]), 5),
    # sedov.py:38 shape: 32^2 root in 8^2 blocks, nghost 4, three levels, WENO-Z + HLLC
    ("sedov2d-wenoz-8b-l3", lambda: _cases.amr_case(
        "sedov", ndim=2, nroot=4, nblock=8, ng=4, lmax=2, center=(0.0, 0.0, 0.0), radius=0.12,
        bc="outflow", xmin=(-0.5,) * 3, xmax=(0.5,) * 3, gamma=1.4, recon="wenoz",
        limiter="minmod", params=[1.0, 2.0 / 128.0, 1.0, 1e-5]), 6),
]


@pytest.mark.parametrize("name,make,ncyc", RUNS, ids=[r[0] for r in RUNS])
def test_amr_task_run_bitwise(gpu, oracle, name, make, ncyc):
    km = gpu
    cfg, desc, lay, leaves, cell = make()
    print(f"amr run {name} leaves={len(leaves)} levels={sorted(set(leaves[:, 0]))}", flush=True)
    sim = km.Simulation(cfg, desc, lay, path=km.PATH_TASKS, leaves=leaves)
    sim.set_primitive(cell)
    ref = oracle.CpuSimulation(cfg.c, desc, lay, leaves=leaves)
    ref.set_primitive(cell)
    sim.advance(ncyc)
    ref.advance(ncyc)
    g, _ = sim.state()
    r, _ = ref.state()
    _same(g, r, name)
    assert sim.info() == ref.info()
    sim.close()


CT_MESHES = [
    # (ndim, nroot, nblock, ng, lmax, center, radius, bc)
    (3, 2, 8, 4, 2, (0.3, 0.6, 0.4), 0.08, "periodic"),
    (2, 4, 8, 4, 2, (0.4, 0.55, 0.5), 0.1, "outflow"),
    (3, 2, 8, 2, 1, (0.1, 0.2, 0.3), 0.1, "reflect"),
]


def _ct_mesh(km, ndim, nroot, nblock, ng, lmax, center, radius, bc):
    nxb = tuple(nblock if d < ndim else 1 for d in range(3))
    nbs = tuple(nroot if d < ndim else 1 for d in range(3))
    root = km.Layout(ndim, ng, nxb, nbs)
    desc = km.mesh_desc(root, bc=(bc,) * 6)
    leaves = km.refine_ball(desc, lmax, center, radius)
    lay = km.Layout(ndim, ng, nxb, nbs, nleaves=len(leaves))
    return desc, leaves, lay


@pytest.mark.parametrize("mesh", CT_MESHES, ids=lambda m: f"{m[0]}d-l{m[4]}-ng{m[3]}-{m[7]}")
def test_face_ghost_fill_across_levels_bitwise(gpu, oracle, mesh):
    km = gpu
    desc, leaves, lay = _ct_mesh(km, *mesh)
    print(f"face fill {mesh} leaves={len(leaves)}", flush=True)
    rng = np.random.default_rng(31)
    cell = rng.uniform(-1.0, 2.0, lay.cell_shape(16))
    faces = [rng.uniform(-1.0, 1.0, lay.stag_shape()) for _ in range(3)]
    dev = km.DevicePack(lay, 16, faces=True, blocks=km.amr_blocks(desc, leaves))
    dev.upload(cell=cell, face=faces)
    host = oracle.HostPack(lay, 16, faces=True, blocks=oracle.amr_blocks(desc, leaves))
    host.cell[:] = cell
    for d in range(3):
        host.face[d][:] = faces[d]
    L = km.lib()
    plan = C.c_void_p()
    km.check(L.km_ghost_plan_amr(C.byref(dev.pack), C.byref(desc), km._i32(leaves), len(leaves),
                                 C.byref(plan)), "plan")
    try:
        km.check(L.km_ghost_fill(plan, C.byref(dev.pack), 0, 14, 1, None), "ghost")
    finally:
        L.km_ghost_plan_destroy(plan)
    cfg = km.Config(**{"physics/MHD": "ct"})
    oracle.call("km_cpu_amr_ghost_fill", C.byref(cfg.c), C.byref(host.pack), C.byref(desc),
                oracle._i32(leaves), len(leaves), 0, 14)
    oracle.call("km_cpu_amr_ghost_faces", C.byref(cfg.c), C.byref(host.pack), C.byref(desc),
                oracle._i32(leaves), len(leaves))
    got, _, gface, _ = dev.download()
    _same(got, host.cell, "cells")
    for d in range(3):
        _same(gface[d], host.face[d], f"face[{d}]")
    dev.close()


@pytest.mark.parametrize("mesh", CT_MESHES, ids=lambda m: f"{m[0]}d-l{m[4]}-{m[7]}")
def test_emf_correction_bitwise(gpu, oracle, mesh):
    km = gpu
    desc, leaves, lay = _ct_mesh(km, *mesh)
    ndim = mesh[0]
    rng = np.random.default_rng(37)
    flux = [rng.standard_normal(lay.cell_shape(8)) if d < ndim else None for d in range(3)]
    edge = [rng.standard_normal(lay.stag_shape()) for _ in range(3)]
    dev = km.DevicePack(lay, 16, nflux=8, edges=True, blocks=km.amr_blocks(desc, leaves))
    dev.upload(flux=flux, edge=edge)
    host = oracle.HostPack(lay, 16, nflux=8, edges=True, blocks=oracle.amr_blocks(desc, leaves))
    for d in range(3):
        if flux[d] is not None:
            host.flux[d][:] = flux[d]
        host.edge[d][:] = edge[d]
    L = km.lib()
    plan = C.c_void_p()
    km.check(L.km_ghost_plan_amr(C.byref(dev.pack), C.byref(desc), km._i32(leaves), len(leaves),
                                 C.byref(plan)), "plan")
    try:
        km.check(L.km_flux_correct(plan, C.byref(dev.pack), None), "flux correct")
    finally:
        L.km_ghost_plan_destroy(plan)
    oracle.call("km_cpu_flux_correct", C.byref(km.Config(**{"physics/MHD": "ct"}).c),
                C.byref(host.pack), C.byref(desc), oracle._i32(leaves), len(leaves))
    _, gflux, _, gedge = dev.download()
    changed = 0
    for d in range(3):
        if d < ndim:
            _same(gflux[d], host.flux[d], f"flux[{d}]")
        if ndim == 3 or d == 2:
            _same(gedge[d], host.edge[d], f"edge[{d}]")
            changed += int(np.sum(host.edge[d] != edge[d]))
    assert changed > 0
    dev.close()


CT_RUNS = [
    # mhd_blast.py:72-78 shape: 32^2 root in 8^2 blocks, nghost 4, three levels
    ("blast2d-8b-l3", lambda: _cases.amr_mhd_case(ndim=2, nroot=4, nblock=8, ng=4, lmax=2,
                                                  radius=0.12), 6),
    ("blast3d-8b-l2", lambda: _cases.amr_mhd_case(ndim=3, nroot=2, nblock=8, ng=4, lmax=1,
                                                  center=(0.1, 0.15, 0.2), radius=0.1), 3),
]


@pytest.mark.parametrize("name,make,ncyc", CT_RUNS, ids=[r[0] for r in CT_RUNS])
def test_ct_amr_task_run_bitwise(gpu, oracle, name, make, ncyc):
    km = gpu
    cfg, desc, lay, leaves, cell, faces = make()
    print(f"ct amr run {name} leaves={len(leaves)}", flush=True)
    sim = km.Simulation(cfg, desc, lay, path=km.PATH_TASKS, leaves=leaves)
    sim.set_primitive(cell, faces)
    ref = oracle.CpuSimulation(cfg.c, desc, lay, leaves=leaves)
    ref.set_primitive(cell, faces)
    sim.advance(ncyc)
    ref.advance(ncyc)
    g, gf = sim.state()
    r, rf = ref.state()
    _same(g, r, name)
    for d in range(3):
        _same(gf[d], rf[d], f"{name} face[{d}]")
    assert sim.info() == ref.info()
    sim.close()
