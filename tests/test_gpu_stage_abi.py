"""km_stage_fused through the C ABI on caller-owned DevicePack buffers, driven the way
kamayan's driver runs a cycle (kamayan_driver.cpp:186-248): stage containers
{base, 1, .., n-1, base}, Parthenon's beta table, PostMeshInitialization
(km_prepare_conserved) and the time-step rule dt = min(2 dt, estimate) on the host.
Whole cycles are compared with the oracle bit for bit: a 3-D single-level mesh (every
kind of stage kernel), a 2-D ensemble of uneven members, and a three-level 2-D block tree
where each stage is followed by km_prepare_primitive and km_ghost_fill, and the last by
km_estimate_dt, as in the task order."""
import ctypes as C

import numpy as np
import pytest

from kamayan_amd import Config, Layout, mesh_desc
from tests import _cases

pytestmark = pytest.mark.gpu

BETAS = {1: [1.0], 2: [1.0, 0.5], 3: [1.0, 0.25, 2.0 / 3.0]}


def _same(a, b, what):
    if not np.array_equal(a, b):
        d = np.abs(a - b)
        i = np.unravel_index(np.argmax(d), d.shape)
        raise AssertionError(f"{what}: max |gpu - oracle| = {d[i]:.3e} at {i}")


def _packs(km, lay, blocks, nstages, nflux=0):
    nv = km.nvar(0)
    base = km.DevicePack(lay, nv, nflux=nflux, blocks=blocks)
    stages = [km.DevicePack(lay, nv, nflux=nflux, blocks=blocks) for _ in range(nstages - 1)]
    return base, stages


def _task_cycle(km, cfg, gplan, base, stages, dudt, dt, est, amr=False):
    """The first cycle on the per-task entry points, in the reference's task order
    (kamayan_driver.cpp:186-248): its first stage reconstructs the problem generator's
    primitives, which the fused kernels (they recompute primitives from the conserved
    state) would round differently. Returns the step's dt."""
    L, p = km.lib(), C.byref
    n = cfg.c.nstages
    cont = [base] + stages + [base]
    dts = dt.get()
    temp = km.slot("TEMP")
    for s in range(1, n + 1):
        u0, u1 = cont[s - 1], cont[s]
        km.check(L.km_calculate_fluxes(p(cfg.c), p(u0.pack), None), "fluxes")
        if amr:
            km.check(L.km_flux_correct(gplan, p(u0.pack), None), "flux correction")
        km.check(L.km_fluxes_to_dudt(p(cfg.c), p(u0.pack), p(dudt.pack), None), "dudt")
        km.check(L.km_apply_dudt(p(cfg.c), p(base.pack), p(u0.pack), p(u1.pack), p(dudt.pack),
                                 BETAS[n][s - 1], float(dts[0]), None), "apply")
        km.check(L.km_prepare_primitive(p(cfg.c), p(u1.pack), None), "prim")
        if s == n:
            km.check(L.km_estimate_dt(p(cfg.c), p(base.pack), C.c_void_p(est.ptr), None), "dt")
        km.check(L.km_ghost_fill(gplan, p(u1.pack), 0, temp, 0, None), "ghost")
    dt.set(np.minimum(2.0 * dts, est.get()))
    return dts


def _cycles(km, cfg, plan, base, stages, dt, est, ninst, ncyc, after_stage=None,
            last_estimate=None, time=None):
    """ncyc cycles of km_stage_fused; returns (time, dt) per instance on the host"""
    L = km.lib()
    n = cfg.c.nstages
    time = np.zeros(ninst) if time is None else time
    cont = [base] + stages + [base]
    for _ in range(ncyc):
        dts = dt.get()
        for s in range(1, n + 1):
            last = s == n
            plan.stage(cfg, base, cont[s - 1], cont[s], BETAS[n][s - 1], dt,
                       est if (last and last_estimate is None) else None)
            if after_stage:
                after_stage(cont[s])
            if last and last_estimate:
                last_estimate(base)
        km.check(L.km_synchronize(), "sync")
        e = est.get()
        time = time + dts
        dt.set(np.minimum(2.0 * dts, e))
    return time, dt.get()


@pytest.mark.parametrize("choice", ["autotune", "sweeps-march", "sweeps-tile3", "stage-7",
                                    "stage-11"])
def test_uniform_3d_cycles_bitwise(gpu, oracle, choice):
    km = gpu
    cfg = km.Config(**{"hydro/reconstruction": "wenoz", "hydro/riemann": "hllc",
                       "parthenon/time/integrator": "rk3", "material/single/gamma": 5.0 / 3.0})
    cfg.c.ndim = 3
    lay = Layout(3, 4, (32, 32, 32), (2, 1, 1))
    desc = mesh_desc(lay, xmax=(2.0, 1.0, 1.0))
    cell, _ = km.problem_fill("smooth3d", cfg, desc, lay, seed=9)
    blocks = oracle.block_table(desc, lay.nblocks)
    print(f"km_stage_fused 3-D {choice}", flush=True)
    plan = km.StagePlan(cfg, descs=[desc])
    if choice == "sweeps-march":
        plan.set_kernel(-1, 0)
    elif choice == "sweeps-tile3":
        plan.set_kernel(-1, 3)
    elif choice.startswith("stage-"):
        plan.set_kernel(3, int(choice[6:]))
    base, stages = _packs(km, lay, blocks, 3, nflux=5)
    dudt = km.DevicePack(lay, km.nvar(0), blocks=blocks)
    base.upload(cell=cell)
    L, p = km.lib(), C.byref
    gplan = C.c_void_p()
    nb, bc = (C.c_int32 * 3)(*lay.nb), (C.c_int32 * 6)(*([0] * 6))
    km.check(L.km_ghost_plan_uniform(p(base.pack), nb, bc, p(gplan)), "ghost plan")
    km.check(L.km_prepare_conserved(p(cfg.c), p(base.pack), None), "prep")
    km.check(L.km_ghost_fill(gplan, p(base.pack), 0, km.slot("TEMP"), 0, None), "ghost")
    dt, est = km.DeviceArray(1), km.DeviceArray(1)
    km.check(L.km_estimate_dt(p(cfg.c), p(base.pack), C.c_void_p(dt.ptr), None), "dt")
    ref = oracle.CpuSimulation(cfg.c, desc, lay)
    ref.set_primitive(cell)
    assert dt.get()[0] == ref.info()[1]
    t0 = _task_cycle(km, cfg, gplan, base, stages, dudt, dt, est)
    t, d = _cycles(km, cfg, plan, base, stages, dt, est, 1, 3, time=t0)
    L.km_ghost_plan_destroy(gplan)
    ref.advance(4)
    rt, rd, _ = ref.info()
    assert (t[0], d[0]) == (rt, rd)
    g = base.download()[0]
    r, _ = ref.state()
    sl = (slice(None), slice(0, 5)) + lay.interior()
    _same(g[sl], r[sl], choice)
    if choice == "autotune":
        print(f"autotuned: stage {plan.kernel(3)}, sweeps {[plan.kernel(x) for x in range(3)]}")
    plan.close()


def test_ensemble_2d_cycles_bitwise(gpu, oracle):
    km = gpu
    cfg, members, lay_all, cell = _cases.kh_ensemble([32, 64, 32])
    descs = [m[0] for m in members]
    blocks = (km._abi.KmBlock * lay_all.nblocks)()
    off = 0
    for d, lay_m, _ in members:
        bm = oracle.block_table(d, lay_m.nblocks)
        for b in range(lay_m.nblocks):
            blocks[off + b] = bm[b]
        off += lay_m.nblocks
    print("km_stage_fused 2-D ensemble", flush=True)
    plan = km.StagePlan(cfg, descs=descs)
    base, stages = _packs(km, lay_all, blocks, cfg.c.nstages)
    # start from primitives that the conserved state reproduces exactly: the fused kernels
    # recompute primitives from the conserved state, where the reference's first stage
    # reads the problem generator's
    cell = _consistent_primitives(km, oracle, cfg, lay_all, blocks, cell)
    base.upload(cell=cell)
    L = km.lib()
    km.check(L.km_prepare_conserved(C.byref(cfg.c), C.byref(base.pack), None), "prep")
    n = len(members)
    dt, est = km.DeviceArray(n), km.DeviceArray(n)
    # the first dt of each member: its own CFL estimate
    off = 0
    first = []
    for d, lay_m, _ in members:
        ref = oracle.CpuSimulation(cfg.c, d, lay_m)
        ref.set_primitive(cell[off:off + lay_m.nblocks].copy())
        first.append(ref.info()[1])
        off += lay_m.nblocks
    dt.set(first)
    t, dtn = _cycles(km, cfg, plan, base, stages, dt, est, n, 3)
    g = base.download()[0]
    off = 0
    for m, (d, lay_m, _) in enumerate(members):
        ref = oracle.CpuSimulation(cfg.c, d, lay_m)
        ref.set_primitive(cell[off:off + lay_m.nblocks].copy())
        ref.advance(3)
        rt, rd, _ = ref.info()
        assert (t[m], dtn[m]) == (rt, rd), m
        sl = (slice(None), slice(0, 5)) + lay_m.interior()
        _same(g[off:off + lay_m.nblocks][sl], ref.state()[0][sl], f"member {m}")
        off += lay_m.nblocks
    plan.close()


def _consistent_primitives(km, oracle, cfg, lay, blocks, cell):
    """Primitives P with cons2prim(prim2cons(P)) == P bit for bit on the interior: iterate
    PostMeshInitialization and PreparePrimitive on the oracle until the primitives stop
    changing (a few rounds: each round only moves values by an ulp)."""
    nv = km.nvar(0)
    host = oracle.HostPack(lay, nv, blocks=blocks)
    host.cell[...] = cell
    sl = (slice(None), slice(None)) + lay.interior()
    for _ in range(20):
        before = host.cell.copy()
        oracle.call("km_cpu_prepare_conserved", C.byref(cfg.c), C.byref(host.pack), None)
        oracle.call("km_cpu_prepare_primitive", C.byref(cfg.c), C.byref(host.pack), None)
        if np.array_equal(before[sl], host.cell[sl]):
            return host.cell.copy()
    pytest.fail("primitives did not settle")


def test_block_tree_cycles_bitwise(gpu, oracle):
    km = gpu
    cfg, desc, lay, leaves, cell = _cases.amr_case(ndim=2, nroot=2, nblock=16, ng=4, lmax=2,
                                                   recon="plm", limiter="mc", riemann="hllc",
                                                   integrator="rk2")
    blocks = oracle.amr_blocks(desc, leaves)
    nv = km.nvar(0)
    print(f"km_stage_fused block tree, {len(leaves)} leaves", flush=True)
    plan = km.StagePlan(cfg, desc=desc, leaves=leaves)
    base, stages = _packs(km, lay, blocks, cfg.c.nstages, nflux=5)
    # the block-tree stage reads the stored primitives (PostMeshInitialization's, then each
    # stage's PreparePrimitive and ghost exchange), as the reference's stages do
    base.upload(cell=cell)
    L = km.lib()
    p = C.byref
    gp = C.c_void_p()
    km.check(L.km_ghost_plan_amr(p(base.pack), p(desc), km._i32(np.ascontiguousarray(leaves)),
                                 len(leaves), p(gp)), "ghost plan")
    temp = km.slot("TEMP")

    def prim_and_ghosts(u):
        km.check(L.km_prepare_primitive(p(cfg.c), p(u.pack), None), "prim")
        km.check(L.km_ghost_fill(gp, p(u.pack), 0, temp, 0, None), "ghost")

    km.check(L.km_prepare_conserved(p(cfg.c), p(base.pack), None), "prep")
    km.check(L.km_ghost_fill(gp, p(base.pack), 0, temp, 0, None), "ghost")
    dt, est = km.DeviceArray(1), km.DeviceArray(1)
    km.check(L.km_estimate_dt(p(cfg.c), p(base.pack), C.c_void_p(dt.ptr), None), "dt")

    def estimate(u):
        km.check(L.km_estimate_dt(p(cfg.c), p(u.pack), C.c_void_p(est.ptr), None), "dt")

    t, d = _cycles(km, cfg, plan, base, stages, dt, est, 1, 4, after_stage=prim_and_ghosts,
                   last_estimate=estimate)
    ref = oracle.CpuSimulation(cfg.c, desc, lay, leaves=leaves)
    ref.set_primitive(cell)
    ref.advance(4)
    rt, rd, _ = ref.info()
    assert (t[0], d[0]) == (rt, rd)
    g = base.download()[0]
    r, _ = ref.state()
    sl = (slice(None), slice(0, 5)) + lay.interior()
    _same(g[sl], r[sl], "block tree")
    L.km_ghost_plan_destroy(gp)
    plan.close()


def test_refusals(gpu, oracle):
    km = gpu
    cfg, desc, lay, cell, _ = _cases.smooth3d(n=32, nblock=16)
    plan = km.StagePlan(cfg, descs=[desc])
    base, _ = _packs(km, lay, oracle.block_table(desc, lay.nblocks), 2)
    dt = km.DeviceArray(1)
    with pytest.raises(km.KmError):  # u1 == u0
        plan.stage(cfg, base, base, base, 1.0, dt)
    with pytest.raises(km.KmError):  # a 2-D plan has four stage-kernel variants
        cfg2, desc2 = _cases.vortex(n=64, nblock=32)[:2]
        km.StagePlan(cfg2, descs=[desc2]).set_kernel(3, 5)
    mhd = Config(**{"physics/MHD": "ct"})
    mhd.c.ndim = 3
    with pytest.raises(km.KmError):
        km.StagePlan(mhd, descs=[desc])
    plan.close()
