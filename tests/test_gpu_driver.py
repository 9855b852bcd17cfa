"""GPU parity of the whole-mesh driver's control paths against the oracle:

- one-stage integration (rk1) on the fused path, where the stage's output goes to a stage
  container and is copied back (the stage input is the base state, read across block
  faces while the sweeps write), bit for bit with every sweep kernel;
- the parthenon/time step rule (dt_factor, dt_init, dt_ceil, dt_user), bit for bit;
- eos/mode_init = dens_ener / dens_temp in PostMeshInitialization, bit for bit;
- the failure flag: a non-physical zone, or dt below dt_min for its cycle limit, makes the
  advance fail with KM_ESTATE, as the oracle's does.
"""
import ctypes as C

import numpy as np
import pytest

from kamayan_amd import Config, Layout, mesh_desc
from kamayan_amd._abi import KM_ESTATE
from tests import _cases

pytestmark = pytest.mark.gpu


def _same(a, b, what):
    diff = float(np.max(np.abs(a - b))) if a.size else 0.0
    assert np.array_equal(a, b), f"{what}: max |gpu - oracle| = {diff:.3e}"


def _case_1d():
    cfg = Config(**{"hydro/reconstruction": "plm", "hydro/slope_limiter": "mc",
                    "hydro/riemann": "hllc", "parthenon/time/integrator": "rk1",
                    "material/single/gamma": 1.4})
    cfg.c.ndim = 1
    lay = Layout(1, 4, (64, 1, 1), (1, 1, 1))  # one periodic block: the sweep wraps onto itself
    desc = mesh_desc(lay)
    import kamayan_amd as km
    cell, _ = km.problem_fill("smooth3d", cfg, desc, lay, seed=11)
    return cfg, desc, lay, cell


def _case_3d():
    cfg, desc, lay, cell, _ = _cases.smooth3d(n=32, nblock=16, integrator="rk1")
    return cfg, desc, lay, cell


@pytest.mark.parametrize("case", ["1d-one-block", "3d-2x2x2"])
@pytest.mark.parametrize("kern", [-1, 0, 1, 3])
def test_rk1_fused_bitwise(gpu, oracle, case, kern):
    km = gpu
    cfg, desc, lay, cell = _case_1d() if case == "1d-one-block" else _case_3d()
    print(f"rk1 fused {case} kernel {kern}", flush=True)
    sim = km.Simulation(cfg, desc, lay, path=km.PATH_FUSED)
    if kern >= 0:
        sim.set_fused_kernel(-1, kern)
    sim.set_primitive(cell)
    ref = oracle.CpuSimulation(cfg.c, desc, lay)
    ref.set_primitive(cell)
    sim.advance(5)
    ref.advance(5)
    g, _ = sim.state()
    r, _ = ref.state()
    sl = (slice(None), slice(0, 5)) + lay.interior()
    _same(g[sl], r[sl], f"rk1 {case} kernel {kern}")
    assert sim.info() == ref.info()
    sim.close()


@pytest.mark.parametrize("path", ["fused", "tasks"])
def test_dt_rule_bitwise(gpu, oracle, path):
    km = gpu
    cfg, desc, lay, cell, _ = _cases.smooth3d(n=32, nblock=16, integrator="rk2")
    probe = oracle.CpuSimulation(cfg.c, desc, lay)
    probe.set_primitive(cell)
    est = probe.info()[1]
    probe.close()
    for k, v in {"parthenon/time/dt_factor": 1.3, "parthenon/time/dt_init": 0.25 * est,
                 "parthenon/time/dt_init_force": "true", "parthenon/time/dt_ceil": 0.9 * est,
                 "parthenon/time/dt_user": 0.95 * est}.items():
        cfg.set(k, v)
    print(f"dt rule {path}", flush=True)
    sim = km.Simulation(cfg, desc, lay, path=km.PATH_FUSED if path == "fused" else km.PATH_TASKS)
    sim.set_primitive(cell)
    ref = oracle.CpuSimulation(cfg.c, desc, lay)
    ref.set_primitive(cell)
    seq = []
    for _ in range(8):
        assert sim.info() == ref.info()
        seq.append(sim.info()[1])
        sim.advance(1)
        ref.advance(1)
    assert seq[0] == 0.25 * est and max(seq) <= 0.9 * est and seq[1] == 1.3 * seq[0], seq
    g, _ = sim.state()
    r, _ = ref.state()
    sl = (slice(None), slice(0, 5)) + lay.interior()
    _same(g[sl], r[sl], f"dt rule {path}")
    sim.close()


@pytest.mark.parametrize("mode,given", [("dens_ener", "EINT"), ("dens_temp", "TEMP")])
def test_mode_init_bitwise(gpu, oracle, mode, given):
    km = gpu
    cfg = Config(**{"eos/mode_init": mode, "material/single/gamma": 1.4,
                    "material/single/Abar": 2.0})
    cfg.c.ndim = 2
    lay = Layout(2, 2, (8, 8, 1), (2, 1, 1))
    desc = mesh_desc(lay)
    blocks = oracle.block_table(desc, lay.nblocks)
    nv = km.nvar(0)
    rng = np.random.default_rng(7)
    cell = np.zeros(lay.cell_shape(nv))
    cell[:, km.slot("DENS")] = rng.uniform(0.5, 2.0, cell[:, 0].shape)
    for d in range(3):
        cell[:, km.slot(f"VEL{d}")] = rng.uniform(-1, 1, cell[:, 0].shape)
    cell[:, km.slot(given)] = rng.uniform(1.0, 3.0, cell[:, 0].shape) * (
        1.0 if given == "EINT" else 1e16)
    dev = km.DevicePack(lay, nv, blocks=blocks)
    dev.upload(cell=cell)
    host = oracle.HostPack(lay, nv, blocks=blocks)
    host.cell[...] = cell
    km.check(km.lib().km_prepare_conserved(C.byref(cfg.c), C.byref(dev.pack), None), "prep")
    oracle.call("km_cpu_prepare_conserved", C.byref(cfg.c), C.byref(host.pack), None)
    got = dev.download()[0]
    _same(got, host.cell, mode)
    dev.close()


@pytest.mark.parametrize("path", ["fused", "tasks"])
@pytest.mark.parametrize("what", ["negative pressure", "nan density"])
def test_non_physical_zone_fails_the_first_cycle(gpu, oracle, path, what):
    km = gpu
    cfg, desc, lay, cell, _ = _cases.smooth3d(n=32, nblock=16, integrator="rk2")
    g = lay.ng
    if what == "negative pressure":
        cell[3, km.slot("PRES"), g + 2, g + 5, g + 7] = -0.5
    else:
        cell[6, km.slot("DENS"), g + 9, g + 1, g + 4] = np.nan
    print(f"failure flag {path} {what}", flush=True)
    sim = km.Simulation(cfg, desc, lay, path=km.PATH_FUSED if path == "fused" else km.PATH_TASKS)
    sim.set_primitive(cell)
    with pytest.raises(km.KmError, match=f"code {KM_ESTATE}"):
        sim.advance(1)
    ref = oracle.CpuSimulation(cfg.c, desc, lay)
    ref.set_primitive(cell)
    with pytest.raises(RuntimeError, match=f"returned {KM_ESTATE}"):
        ref.advance(1)
    sim.close()


def test_healthy_run_does_not_fail_and_dt_min_does(gpu, oracle):
    km = gpu
    cfg, desc, lay, cell, _ = _cases.smooth3d(n=32, nblock=16, integrator="rk3")
    sim = km.Simulation(cfg, desc, lay, path=km.PATH_FUSED)
    sim.set_primitive(cell)
    sim.advance(6)  # graph-replayed cycles read the flag too
    cfg.set("parthenon/time/dt_min", 1.0)
    cfg.set("parthenon/time/dt_min_cycle_limit", 4)
    sim = km.Simulation(cfg, desc, lay, path=km.PATH_FUSED)
    sim.set_primitive(cell)
    sim.advance(3)
    with pytest.raises(km.KmError, match=f"code {KM_ESTATE}"):
        sim.advance(1)
    sim.close()
