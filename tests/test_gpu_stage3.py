"""The one-launch RK stage (km_stage3.inc) against the oracle: every tile shape and x3
segment of the stage kernel, in the FP64 form, bit for bit (time, dt and the conserved
interior after whole cycles, the last stage's CFL minimum included); the FP64-FMA form
within 1e-12 and the FP32-mixed form within 1e-5 relative L1 per conserved variable, the
FP32-mixed stage kernel also within 1e-6 of the FP32-mixed sweeps (same FP32 expressions,
only contraction choices may differ). Cases: non-cubic 64 x 64 x 32 blocks (every shape
and segment fits) with WENO-Z/HLLC RK3, PPM-MC/HLL RK2, PLM/HLLC and FOG/HLL, periodic and
reflecting/outflow boundaries; 64^3 blocks (128^3) with the autotuner's choice; a 3-D
ensemble of uneven block counts."""
import numpy as np
import pytest

from tests import _cases

pytestmark = pytest.mark.gpu

SHAPES = [(64, 4), (64, 8), (32, 8), (32, 16), (16, 32), (16, 16)]  # x1 columns x x2 rows
SEGS = [0, 32, 16]  # x3 planes per workgroup (0: the whole block)
NKERN = len(SHAPES) * len(SEGS)


def _case(kind):
    import kamayan_amd as km
    recon, riem, integ, lim, bc = {
        "wenoz-hllc-rk3": ("wenoz", "hllc", "rk3", "minmod", "periodic"),
        "ppm-mc-hll-rk2": ("ppm", "hll", "rk2", "mc", "periodic"),
        "plm-vl-hllc-rk2-walls": ("plm", "hllc", "rk2", "van_leer", "walls"),
        "fog-hll-rk3": ("fog", "hll", "rk3", "minmod", "periodic"),
    }[kind]
    cfg = km.Config(**{"hydro/reconstruction": recon, "hydro/riemann": riem,
                       "hydro/slope_limiter": lim, "parthenon/time/integrator": integ,
                       "material/single/gamma": 5.0 / 3.0})
    cfg.c.ndim = 3
    lay = km.Layout(3, 4, (64, 64, 32), (2, 1, 2))
    bcs = ("periodic",) * 6 if bc == "periodic" else (
        "reflect", "outflow", "outflow", "reflect", "reflect", "outflow")
    desc = km.mesh_desc(lay, bc=bcs)
    cell, _ = km.problem_fill("smooth3d", cfg, desc, lay, seed=21)
    return cfg, desc, lay, cell


_REF = {}


def _oracle_run(oracle, kind, ncyc, cfg, desc, lay, cell):
    key = (kind, ncyc)
    if key not in _REF:
        ref = oracle.CpuSimulation(cfg.c, desc, lay)
        ref.set_primitive(cell)
        ref.advance(ncyc)
        _REF[key] = (ref.state()[0], ref.info())
        ref.close()
    return _REF[key]


def _gpu_run(km, cfg, desc, lay, cell, ncyc, stage_kern=None, descs=None):
    sim = km.Simulation(cfg, desc, lay, path=km.PATH_FUSED, descs=descs)
    if stage_kern is not None:
        sim.set_fused_kernel(3, stage_kern)
    sim.set_primitive(cell)
    sim.advance(ncyc)
    out = sim.state()[0], sim.info(), sim.fused_kernel(3)
    sim.close()
    return out


def _same(a, b, what):
    if not np.array_equal(a, b):
        d = np.abs(a - b)
        i = np.unravel_index(np.argmax(d), d.shape)
        raise AssertionError(f"{what}: max |gpu - oracle| = {d[i]:.3e} at {i}")


@pytest.mark.parametrize("kern", range(1, NKERN + 1))
def test_stage_kernel_fp64_bitwise(gpu, oracle, kern):
    kind = "wenoz-hllc-rk3"
    cfg, desc, lay, cell = _case(kind)
    print(f"stage kernel {kern} {kind}", flush=True)
    g, gi, used = _gpu_run(gpu, cfg, desc, lay, cell, 2, kern)
    assert used == kern
    r, ri = _oracle_run(oracle, kind, 2, cfg, desc, lay, cell)
    assert gi == ri
    sl = (slice(None), slice(0, 5)) + lay.interior()
    _same(g[sl], r[sl], f"stage kernel {kern}")


@pytest.mark.parametrize("kind", ["ppm-mc-hll-rk2", "plm-vl-hllc-rk2-walls", "fog-hll-rk3"])
@pytest.mark.parametrize("kern", [2, 9, 14])
def test_stage_kernel_methods_bitwise(gpu, oracle, kind, kern):
    cfg, desc, lay, cell = _case(kind)
    print(f"stage kernel {kern} {kind}", flush=True)
    g, gi, _ = _gpu_run(gpu, cfg, desc, lay, cell, 2, kern)
    r, ri = _oracle_run(oracle, kind, 2, cfg, desc, lay, cell)
    assert gi == ri
    sl = (slice(None), slice(0, 5)) + lay.interior()
    _same(g[sl], r[sl], f"{kind} stage kernel {kern}")


@pytest.mark.parametrize("precision,tol", [("fp64_fma", 1e-12), ("fp32_mixed", 1e-5)])
@pytest.mark.parametrize("kern", [4, 8, 13])
def test_stage_kernel_other_forms(gpu, oracle, precision, tol, kern):
    kind = "wenoz-hllc-rk3"
    cfg, desc, lay, cell = _case(kind)
    cfg.set("kamayan_amd/precision", precision)
    print(f"stage kernel {kern} {precision}", flush=True)
    g, _, _ = _gpu_run(gpu, cfg, desc, lay, cell, 2, kern)
    r, _ = _oracle_run(oracle, kind, 2, cfg, desc, lay, cell)
    l1 = _cases.rel_l1(g, r, lay, 5)
    assert np.all(l1 < tol), l1
    if precision == "fp32_mixed":
        s, _, _ = _gpu_run(gpu, cfg, desc, lay, cell, 2, 0)  # the three sweeps
        l1s = _cases.rel_l1(g, s, lay, 5)
        assert np.all(l1s < 1e-6), l1s


def test_64cube_blocks_autotuned_bitwise(gpu, oracle):
    """WENO-Z/HLLC RK3 on 64^3 blocks (128^3 zones), whatever the autotuner picks."""
    cfg, desc, lay, cell, _ = _cases.smooth3d(n=128, nblock=64)
    print("128^3 in 64^3 blocks, autotuned", flush=True)
    g, gi, used = _gpu_run(gpu, cfg, desc, lay, cell, 2)
    print(f"autotuned stage kernel {used}", flush=True)
    ref = oracle.CpuSimulation(cfg.c, desc, lay)
    ref.set_primitive(cell)
    ref.advance(2)
    r, ri = ref.state()[0], ref.info()
    assert gi == ri
    sl = (slice(None), slice(0, 5)) + lay.interior()
    _same(g[sl], r[sl], "128^3 autotuned")
    for kern in (0, 5):  # and the sweeps and a stage kernel explicitly
        g2, gi2, _ = _gpu_run(gpu, cfg, desc, lay, cell, 2, kern)
        assert gi2 == ri
        _same(g2[sl], r[sl], f"128^3 stage kernel {kern}")


def test_3d_ensemble_members_bitwise(gpu, oracle):
    """Three 3-D runs of uneven block counts in one stage launch: each member equals its
    own oracle run, time and dt included."""
    km = gpu
    cfg = km.Config(**{"hydro/reconstruction": "wenoz", "hydro/riemann": "hllc",
                       "parthenon/time/integrator": "rk3", "material/single/gamma": 1.4})
    cfg.c.ndim = 3
    members, cells = [], []
    for m, nb in enumerate([(1, 1, 1), (2, 1, 1), (1, 1, 2)]):
        lay_m = km.Layout(3, 4, (32, 32, 32), nb)
        d = km.mesh_desc(lay_m, xmax=(1.0 + 0.5 * m,) * 3)
        c, _ = km.problem_fill("smooth3d", cfg, d, lay_m, seed=40 + m)
        members.append((d, lay_m))
        cells.append(c)
    total = sum(x[1].nblocks for x in members)
    lay = km.Layout(3, 4, (32, 32, 32), (total, 1, 1))
    cell = np.concatenate(cells, axis=0)
    descs = [x[0] for x in members]
    print("3-D ensemble, stage kernel 9", flush=True)
    sim = km.Simulation(cfg, descs[0], lay, path=km.PATH_FUSED, descs=descs)
    sim.set_fused_kernel(3, 9)
    sim.set_primitive(cell)
    sim.advance(3)
    g, _ = sim.state()
    times, dts = sim.ensemble_info()
    sim.close()
    off = 0
    for m, (d, lay_m) in enumerate(members):
        ref = oracle.CpuSimulation(cfg.c, d, lay_m)
        ref.set_primitive(cells[m].copy())
        ref.advance(3)
        r, _ = ref.state()
        t, dt, _ = ref.info()
        assert times[m] == t and dts[m] == dt, (m, times[m], t, dts[m], dt)
        sl = (slice(None), slice(0, 5)) + lay_m.interior()
        _same(g[off:off + lay_m.nblocks][sl], r[sl], f"member {m}")
        off += lay_m.nblocks


def test_stage_kernel_refusals(gpu):
    km = gpu
    cfg, desc, lay, cell, _ = _cases.smooth3d(n=32, nblock=16)
    sim = km.Simulation(cfg, desc, lay, path=km.PATH_FUSED)
    sim.set_fused_kernel(3, 1)  # 64-column tiles do not divide 16-cell rows
    sim.set_primitive(cell)
    with pytest.raises(km.KmError):
        sim.advance(2)
    sim.close()
    cfg2, desc2, lay2, cell2 = _cases.vortex(n=64, nblock=32)[:4]
    sim = km.Simulation(cfg2, desc2, lay2, path=km.PATH_FUSED)
    with pytest.raises(km.KmError):  # 2-D meshes have four variants (km_stage2.inc)
        sim.set_fused_kernel(3, 5)
    sim.close()
