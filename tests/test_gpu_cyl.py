"""GPU parity of cylindrical (r, z) geometry against the oracle, bit for bit: every task
(fluxes with the angular-momentum flux, dU/dt with r-dependent areas and volumes, the
geometric source, the stage update, primitive recovery, dt with the r/|v_phi| limit,
prepare-conserved, the axisymmetric ghost fill) and whole task-path runs, among them the
reference suite's cylindrical vortex (reconstruction.py:136)."""
import ctypes as C

import numpy as np
import pytest

from kamayan_amd import Config, Layout, mesh_desc
from tests import _cases
from tests.test_oracle_cyl import _rz, _swirl

pytestmark = pytest.mark.gpu


def _same(a, b, what):
    if not np.array_equal(a, b):
        bad = np.argwhere(a != b)
        i = tuple(bad[0])
        raise AssertionError(f"{what}: {len(bad)} differ, first {i}: {a[i]!r} vs {b[i]!r}")


def _case(oracle, which):
    if which == "vortex1d":
        cfg, desc, lay, cell, _ = _cases.vortex_cyl()
        return cfg, desc, lay, cell
    cfg, desc, lay = _rz(bc_r="outflow")
    cell, _ = _swirl(lay, desc, oracle)
    return cfg, desc, lay, cell


@pytest.mark.parametrize("which", ["vortex1d", "swirl2d"])
def test_cyl_tasks_bitwise(gpu, oracle, which):
    km = gpu
    cfg, desc, lay, cell = _case(oracle, which)
    print(f"cylindrical tasks {which}", flush=True)
    L, p = km.lib(), C.byref
    blocks = oracle.block_table(desc, lay.nblocks)
    dev = km.DevicePack(lay, 12, nflux=5, blocks=blocks)
    host = oracle.HostPack(lay, 12, nflux=5, blocks=blocks)
    dev.upload(cell=cell)
    host.cell[:] = cell
    # PostMeshInitialization, then the axisymmetric ghost fill
    km.check(L.km_prepare_conserved(p(cfg.c), p(dev.pack), None), "cons")
    oracle.call("km_cpu_prepare_conserved", p(cfg.c), p(host.pack), None)
    nb = (C.c_int32 * 3)(*[lay.nb[d] if d < lay.ndim else 1 for d in range(3)])
    bc = (C.c_int32 * 6)(*[desc.bc[d] for d in range(6)])
    plan = C.c_void_p()
    km.check(L.km_ghost_plan_uniform(p(dev.pack), nb, bc, p(plan)), "plan")
    km.check(L.km_ghost_fill(plan, p(dev.pack), 0, 11, 0, None), "ghost")
    L.km_ghost_plan_destroy(plan)
    oracle.call("km_cpu_ghost_fill_uniform", p(cfg.c), p(host.pack), nb, bc, 0, 11, 0)
    _same(dev.download()[0], host.cell, "conserved + ghost fill")
    # fluxes, dU/dt, source
    km.check(L.km_calculate_fluxes(p(cfg.c), p(dev.pack), None), "fluxes")
    oracle.call("km_cpu_calculate_fluxes", p(cfg.c), p(host.pack), None)
    _, flux, _, _ = dev.download()
    for d in range(lay.ndim):
        _same(flux[d], host.flux[d], f"flux[{d}]")
    ddev = km.DevicePack(lay, 12, blocks=blocks)
    dhost = oracle.HostPack(lay, 12, blocks=blocks)
    km.check(L.km_fluxes_to_dudt(p(cfg.c), p(dev.pack), p(ddev.pack), None), "dudt")
    oracle.call("km_cpu_fluxes_to_dudt", p(cfg.c), p(host.pack), p(dhost.pack), None)
    km.check(L.km_add_source_terms(p(cfg.c), p(dev.pack), p(ddev.pack), None), "source")
    oracle.call("km_cpu_add_source_terms", p(cfg.c), p(host.pack), p(dhost.pack), None)
    _same(ddev.download()[0], dhost.cell, "dudt + source")
    # stage update, primitives, dt
    u1d = km.DevicePack(lay, 12, blocks=blocks)
    u1h = oracle.HostPack(lay, 12, blocks=blocks)
    km.check(L.km_apply_dudt(p(cfg.c), p(dev.pack), p(dev.pack), p(u1d.pack), p(ddev.pack),
                             0.5, 1e-3, None), "apply")
    oracle.call("km_cpu_apply_dudt", p(cfg.c), p(host.pack), p(host.pack), p(u1h.pack),
                p(dhost.pack), 0.5, 1e-3, None)
    km.check(L.km_prepare_primitive(p(cfg.c), p(u1d.pack), None), "prim")
    oracle.call("km_cpu_prepare_primitive", p(cfg.c), p(u1h.pack), None)
    _same(u1d.download()[0], u1h.cell, "u1")
    dtd = km.DevicePack(lay, 1)
    km.check(L.km_estimate_dt(p(cfg.c), p(u1d.pack), C.c_void_p(dtd.cell), None), "dt")
    got = dtd.download()[0].reshape(-1)[0]
    want = C.c_double()
    oracle.call("km_cpu_estimate_dt", p(cfg.c), p(u1h.pack), p(want), None)
    assert got == want.value
    for x in (dev, ddev, u1d, dtd):
        x.close()


@pytest.mark.parametrize("which,ncyc", [("vortex1d", 40), ("swirl2d", 8)])
def test_cyl_sim_tasks_bitwise(gpu, oracle, which, ncyc):
    km = gpu
    cfg, desc, lay, cell = _case(oracle, which)
    print(f"cylindrical run {which}", flush=True)
    sim = km.Simulation(cfg, desc, lay, path=km.PATH_TASKS)
    sim.set_primitive(cell)
    ref = oracle.CpuSimulation(cfg.c, desc, lay)
    ref.set_primitive(cell)
    sim.advance(ncyc)
    ref.advance(ncyc)
    g, _ = sim.state()
    r, _ = ref.state()
    _same(g, r, which)
    assert sim.info() == ref.info()
    sim.close()


def test_cyl_refusals(gpu):
    km = gpu
    cfg, desc, lay, _, _ = _cases.vortex_cyl()
    with pytest.raises(km.KmError, match="cylindrical geometry runs on the task path"):
        km.Simulation(cfg, desc, lay, path=km.PATH_FUSED)
    c2 = km.Config(**{"hydro/reconstruction": "plm"})
    c2.c.ndim = 2
    lay2 = Layout(2, 2, (8, 8, 1), (2, 2, 1))
    d2 = mesh_desc(lay2, bc=("axisymmetric", "outflow", "periodic", "periodic", "periodic",
                             "periodic"))
    with pytest.raises(km.KmError, match="needs cylindrical geometry"):
        km.Simulation(c2, d2, lay2, path=km.PATH_TASKS)
    c2.set("geometry/geometry", "cylindrical")
    d3 = mesh_desc(lay2, bc=("outflow", "axisymmetric", "periodic", "periodic", "periodic",
                             "periodic"))
    with pytest.raises(km.KmError, match="inner x1 boundary"):
        km.Simulation(c2, d3, lay2, path=km.PATH_TASKS)
