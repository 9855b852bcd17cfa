"""GPU parity of adaptive refinement against the oracle: Loehner tags
(grid_refinement.cpp:30-152), the remesh data movement (copy / ProlongateSharedMinMod /
RestrictAverage over new leaves) and whole adaptive runs in which the tree changes
between cycles. Bitwise, like every FP64 path here; leaves and their order must match."""
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


def _dev_i32(km, n):
    ptr = C.c_void_p()
    km.check(km.lib().km_malloc(C.byref(ptr), max(4 * n, 8)), "malloc")
    return ptr.value


@pytest.mark.parametrize("ndim", [2, 3])
def test_loehner_tags_bitwise(gpu, oracle, ndim):
    km = gpu
    cfg, desc, lay, leaves, cell = _cases.amr_case(
        "sedov", ndim=ndim, nroot=4 if ndim == 2 else 2, nblock=8, ng=4, lmax=1,
        center=(0.0, 0.0, 0.0), radius=0.2, bc="outflow", xmin=(-0.5,) * 3, xmax=(0.5,) * 3,
        gamma=1.4, params=[1.0, 0.12, 1.0, 1e-5])
    print(f"tags ndim={ndim} leaves={len(leaves)}", flush=True)
    ref = oracle.CpuSimulation(cfg.c, desc, lay, leaves=leaves)
    ref.set_primitive(cell)
    ref.advance(2)
    state, _ = ref.state()  # filled ghost zones, primitives current
    blocks = km.amr_blocks(desc, leaves)
    dev = km.DevicePack(lay, 12, blocks=blocks)
    dev.upload(cell=state)
    host = oracle.HostPack(lay, 12, blocks=oracle.amr_blocks(desc, leaves))
    host.cell[:] = state
    tags_d = _dev_i32(km, len(leaves))
    try:
        for var, rtol, dtol in [(slot("PRES"), 0.8, 0.2), (slot("DENS"), 0.3, 0.05),
                                (slot("VEL0"), 0.5, 0.1)]:
            crit = amr_criterion(var, 2, refine_tol=rtol, derefine_tol=dtol)
            for reset in (True, False):
                want = np.full(len(leaves), S, np.int32)
                oracle.amr_tag_loehner(host.pack, crit, want, reset)
                km.check(km.lib().km_memcpy_h2d(C.c_void_p(tags_d), km._i32(np.full(
                    len(leaves), S, np.int32)), 4 * len(leaves)), "h2d")
                km.amr_tag_loehner(dev, crit, tags_d, reset)
                got = np.zeros(len(leaves), np.int32)
                km.check(km.lib().km_memcpy_d2h(km._i32(got), C.c_void_p(tags_d),
                                                4 * len(leaves)), "d2h")
                _same(got, want, f"tags var {var} reset {reset}")
            assert len(set(want.tolist())) > 1
    finally:
        km.lib().km_free(C.c_void_p(tags_d))
        dev.close()


@pytest.mark.parametrize("ndim", [2, 3])
def test_remesh_bitwise(gpu, oracle, ndim):
    km = gpu
    rng = np.random.default_rng(41 + ndim)
    nb = (4, 4, 1) if ndim == 2 else (2, 2, 2)
    lay0 = Layout(ndim, 4, tuple(8 if d < ndim else 1 for d in range(3)), nb)
    desc = mesh_desc(lay0, bc=("outflow",) * 6)
    leaves = km.refine_ball(desc, 2, (0.3, 0.4, 0.5), 0.2)
    lay = Layout(ndim, 4, lay0.nx, nb, nleaves=len(leaves))
    # the finest leaves coarsen, some others split: kept, split and merged leaves all occur
    tags = np.where(leaves[:, 0] == 2, D, rng.choice([S, R], size=len(leaves), p=[0.6, 0.4]))
    tags = tags.astype(np.int32)
    nl, _, org, ch = km.amr_update_tree(desc, 4, 0, leaves, tags, np.zeros(len(leaves), np.int32))
    assert ch and set(org[:, 0].tolist()) == {0, 1, 2}
    print(f"remesh ndim={ndim} {len(leaves)} -> {len(nl)} leaves", flush=True)
    lay2 = Layout(ndim, 4, lay0.nx, nb, nleaves=len(nl))
    cell = rng.uniform(0.5, 2.0, lay.cell_shape(12))
    old_d = km.DevicePack(lay, 12, blocks=km.amr_blocks(desc, leaves))
    old_d.upload(cell=cell)
    new_d = km.DevicePack(lay2, 12, blocks=km.amr_blocks(desc, nl))
    org_d = _dev_i32(km, org.size)
    km.check(km.lib().km_memcpy_h2d(C.c_void_p(org_d), km._i32(np.ascontiguousarray(org)),
                                    4 * org.size), "h2d")
    try:
        km.amr_remesh(old_d, new_d, org_d, 0, 12)
        got, _, _, _ = new_d.download()
    finally:
        km.lib().km_free(C.c_void_p(org_d))
    old_h = oracle.HostPack(lay, 12, blocks=oracle.amr_blocks(desc, leaves))
    old_h.cell[:] = cell
    new_h = oracle.HostPack(lay2, 12, blocks=oracle.amr_blocks(desc, nl))
    oracle.amr_remesh(old_h.pack, new_h.pack, org, 0, 12)
    _same(got, new_h.cell, "remesh")
    old_d.close()
    new_d.close()


ADAPT_RUNS = [
    # sedov.py:38 shape: 32^2 root in 8^2 blocks, nghost 4, three levels, refinement on
    # pres; the blast starts on the root grid and the tree follows it
    ("sedov2d-8b-numlevel3", dict(ndim=2, nroot=4, nblock=8, ng=4, recon="wenoz",
                                  limiter="minmod", params=[1.0, 0.05, 1.0, 1e-5]), 3, 24, 10),
    # merges: the run starts refined to level 2 over a wide ball, and the quiet outer
    # families coarsen after a derefine count of 2
    ("sedov2d-plm-derefine2", dict(ndim=2, nroot=4, nblock=8, ng=2, recon="plm", limiter="mc",
                                   params=[1.0, 0.08, 1.0, 1e-5], lmax=2, radius=0.45), 3, 12,
     2),
    ("sedov3d-8b-numlevel2", dict(ndim=3, nroot=2, nblock=8, ng=4, recon="plm", limiter="mc",
                                  params=[1.0, 0.15, 1.0, 1e-5]), 2, 6, 10),
]


@pytest.mark.parametrize("name,kw,numlevel,ncyc,dcount", ADAPT_RUNS,
                         ids=[r[0] for r in ADAPT_RUNS])
def test_adaptive_run_bitwise(gpu, oracle, name, kw, numlevel, ncyc, dcount):
    km = gpu
    kw = dict(kw)
    lmax, radius = kw.pop("lmax", 0), kw.pop("radius", 0.1)
    cfg, desc, lay, leaves, cell = _cases.amr_case(
        "sedov", lmax=lmax, center=(0.0, 0.0, 0.0), radius=radius, bc="outflow",
        riemann="hllc", integrator="rk2", xmin=(-0.5,) * 3, xmax=(0.5,) * 3, gamma=1.4, **kw)
    adapt = amr_adapt(numlevel, [amr_criterion(slot("PRES"), numlevel)], derefine_count=dcount)
    print(f"adaptive run {name}", flush=True)
    sim = km.Simulation(cfg, desc, lay, path=km.PATH_TASKS, leaves=leaves)
    sim.set_primitive(cell)
    sim.set_adaptive(adapt)
    ref = oracle.CpuSimulation(cfg.c, desc, lay, leaves=leaves)
    ref.set_primitive(cell)
    ref.set_adaptive(adapt)
    seen = {len(leaves)}
    for c in range(ncyc):
        sim.advance(1)
        ref.advance(1)
        gl, rl = sim.leaves(), ref.leaves()
        _same(gl, rl, f"{name} leaves after cycle {c + 1}")
        seen.add(len(gl))
        if c % 4 == 3 or c == ncyc - 1:
            g, _ = sim.state()
            r, _ = ref.state()
            _same(g, r, f"{name} state after cycle {c + 1}")
            assert sim.info() == ref.info()
    assert len(seen) > 1  # the tree changed during the run
    sim.close()


def test_initial_refinement_then_adaptive_run_bitwise(gpu, oracle):
    # Parthenon's initial loop on the sedov.py:38 shape (32^2 root in 8^2 blocks,
    # nghost 4, numlevel 3), then adaptive cycles, against the oracle's loop and run
    km = gpu
    cfg = km.Config(**{"hydro/reconstruction": "wenoz", "hydro/slope_limiter": "minmod",
                       "hydro/riemann": "hllc", "parthenon/time/integrator": "rk2",
                       "material/single/gamma": 1.4})
    cfg.c.ndim = 2
    lay = Layout(2, 4, (8, 8, 1), (4, 4, 1))
    desc = mesh_desc(lay, xmin=(-0.5,) * 3, xmax=(0.5,) * 3, bc=("outflow",) * 6)
    params = [1.0, 0.05, 1.0, 1e-5]
    adapt = amr_adapt(3, [amr_criterion(slot("PRES"), 3)])
    leaves, cell, _ = km.initial_refinement(cfg, desc, lay, "sedov", adapt, params=params)

    def fill(lv, lay_l):
        c, _ = km.problem_fill("sedov", cfg, desc, lay_l, params=params, leaves=lv)
        return c
    rleaves, rcell = oracle.initial_refinement(cfg.c, desc, lay, fill, adapt)
    print(f"initial refinement: {len(leaves)} leaves", flush=True)
    _same(leaves, rleaves, "initial leaves")
    _same(cell, rcell, "initial state")
    assert max(leaves[:, 0]) == 2
    lay_l = Layout(2, 4, (8, 8, 1), (4, 4, 1), nleaves=len(leaves))
    sim = km.Simulation(cfg, desc, lay_l, path=km.PATH_TASKS, leaves=leaves)
    sim.set_primitive(cell)
    sim.set_adaptive(adapt)
    ref = oracle.CpuSimulation(cfg.c, desc, lay_l, leaves=leaves)
    ref.set_primitive(cell)
    ref.set_adaptive(adapt)
    sim.advance(8)
    ref.advance(8)
    _same(sim.leaves(), ref.leaves(), "leaves after 8 cycles")
    g, _ = sim.state()
    r, _ = ref.state()
    _same(g, r, "state after 8 cycles")
    assert sim.info() == ref.info()
    sim.close()
