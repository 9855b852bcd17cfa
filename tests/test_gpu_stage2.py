"""The 2-D one-launch RK stage (km_stage2.inc) against the oracle: every variant on blocks
whose x1 cell count it takes (16, 32 and 64), in the FP64 form bit for bit (the conserved
fields with their ghost zones after the fill that precedes state reads, time and dt,
which the last stage's CFL minimum sets); the FP64-FMA form within 1e-12 and the FP32-mixed
form within 1e-5 relative L1 per conserved variable, the FP32-mixed stage kernel also
within 1e-6 of the FP32-mixed sweeps. An ensemble of uneven 2-D runs whose blocks share
workgroups: each member equals its own oracle run. Variants that do not fit the blocks
are refused."""
import numpy as np
import pytest

from tests import _cases

pytestmark = pytest.mark.gpu

SHAPES = [(16, 256), (32, 256), (32, 512), (64, 256)]  # x1 cells per block row, lanes


def _case(kind):
    import kamayan_amd as km
    if kind == "vortex-plm-mc-hllc-rk2":
        cfg, desc, lay, cell, _ = _cases.vortex(n=64, nblock=32, ng=4, recon="plm",
                                                limiter="mc", riemann="hllc", cfl=0.4)
        return cfg, desc, lay, cell
    recon, riem, integ, nx, nb, bc = {
        "smooth-wenoz-hllc-rk3-16b-walls": ("wenoz", "hllc", "rk3", 16, (3, 2), "walls"),
        "smooth-ppm-hll-rk2-64b": ("ppm", "hll", "rk2", 64, (2, 1), "periodic"),
        "smooth-fog-hll-rk3-32b": ("fog", "hll", "rk3", 32, (2, 3), "periodic"),
    }[kind]
    cfg = km.Config(**{"hydro/reconstruction": recon, "hydro/riemann": riem,
                       "hydro/slope_limiter": "van_leer", "parthenon/time/integrator": integ,
                       "material/single/gamma": 5.0 / 3.0})
    cfg.c.ndim = 2
    lay = km.Layout(2, 4, (nx, 16 if nx == 16 else 32, 1), nb + (1,))
    bcs = ("periodic",) * 6 if bc == "periodic" else (
        "reflect", "outflow", "outflow", "reflect", "periodic", "periodic")
    desc = km.mesh_desc(lay, bc=bcs)
    cell, _ = km.problem_fill("smooth3d", cfg, desc, lay, seed=31)
    return cfg, desc, lay, cell


CASES = {  # case -> the variants whose row length fits its blocks
    "smooth-wenoz-hllc-rk3-16b-walls": [1],
    "vortex-plm-mc-hllc-rk2": [2, 3],
    "smooth-fog-hll-rk3-32b": [2, 3],
    "smooth-ppm-hll-rk2-64b": [4],
}


def _gpu_run(km, cfg, desc, lay, cell, ncyc, kern=None, descs=None):
    sim = km.Simulation(cfg, desc, lay, path=km.PATH_FUSED, descs=descs)
    if kern is not None:
        sim.set_fused_kernel(3, kern)
    try:
        sim.set_primitive(cell)
        sim.advance(ncyc)
        return sim.state()[0], sim.info(), sim.fused_kernel(3)
    finally:
        sim.close()


def _same(a, b, what):
    if not np.array_equal(a, b):
        d = np.abs(a - b)
        i = np.unravel_index(np.argmax(d), d.shape)
        raise AssertionError(f"{what}: max |gpu - oracle| = {d[i]:.3e} at {i}")


@pytest.mark.parametrize("kind", list(CASES))
def test_stage2_fp64_bitwise(gpu, oracle, kind):
    cfg, desc, lay, cell = _case(kind)
    ref = oracle.CpuSimulation(cfg.c, desc, lay)
    ref.set_primitive(cell)
    ref.advance(3)
    r, ri = ref.state()[0], ref.info()
    ref.close()
    for kern in CASES[kind] + [None]:
        print(f"2-D stage kernel {kern} {kind}", flush=True)
        g, gi, used = _gpu_run(gpu, cfg, desc, lay, cell, 3, kern)
        if kern is not None:
            assert used == kern
        assert gi == ri
        sl = (slice(None), slice(0, 5)) + lay.interior()
        _same(g[sl], r[sl], f"{kind} kernel {kern}")
        _same(g[:, :5], r[:, :5], f"{kind} kernel {kern} with ghost zones")


@pytest.mark.parametrize("precision,tol", [("fp64_fma", 1e-12), ("fp32_mixed", 1e-5)])
def test_stage2_other_forms(gpu, oracle, precision, tol):
    kind = "vortex-plm-mc-hllc-rk2"
    cfg, desc, lay, cell = _case(kind)
    ref = oracle.CpuSimulation(cfg.c, desc, lay)
    ref.set_primitive(cell)
    ref.advance(3)
    r = ref.state()[0]
    ref.close()
    cfg.set("kamayan_amd/precision", precision)
    for kern in (2, 3):
        print(f"2-D stage kernel {kern} {precision}", flush=True)
        g, _, _ = _gpu_run(gpu, cfg, desc, lay, cell, 3, kern)
        l1 = _cases.rel_l1(g, r, lay, 5)
        assert np.all(l1 < tol), (kern, l1)
        if precision == "fp32_mixed":
            s, _, _ = _gpu_run(gpu, cfg, desc, lay, cell, 3, 0)  # the two sweeps
            assert np.all(_cases.rel_l1(g, s, lay, 5) < 1e-6)


def test_stage2_ensemble_members_bitwise(gpu, oracle):
    """Four KH runs of 32^2, 64^2, 32^2 and 96^2 zones (1 + 4 + 1 + 9 blocks): the 8-block
    workgroups of variant 2 mix members; each member equals its own oracle run."""
    km = gpu
    cfg, members, lay, cell = _cases.kh_ensemble([32, 64, 32, 96], nblock=32, ng=4)
    descs = [m[0] for m in members]
    print("2-D ensemble, stage kernel 2", flush=True)
    sim = km.Simulation(cfg, descs[0], lay, path=km.PATH_FUSED, descs=descs)
    sim.set_fused_kernel(3, 2)
    sim.set_primitive(cell)
    sim.advance(3)
    g, _ = sim.state()
    times, dts = sim.ensemble_info()
    sim.close()
    off = 0
    for m, (d, lay_m, _) in enumerate(members):
        ref = oracle.CpuSimulation(cfg.c, d, lay_m)
        ref.set_primitive(cell[off:off + lay_m.nblocks].copy())
        ref.advance(3)
        r, _ = ref.state()
        t, dt, _ = ref.info()
        ref.close()
        assert times[m] == t and dts[m] == dt, (m, times[m], t, dts[m], dt)
        sl = (slice(None), slice(0, 5)) + lay_m.interior()
        _same(g[off:off + lay_m.nblocks][sl], r[sl], f"member {m}")
        off += lay_m.nblocks


def test_stage2_refusals(gpu):
    km = gpu
    cfg, desc, lay, cell = _case("vortex-plm-mc-hllc-rk2")
    sim = km.Simulation(cfg, desc, lay, path=km.PATH_FUSED)
    with pytest.raises(km.KmError):  # 2-D meshes have variants 1..4
        sim.set_fused_kernel(3, len(SHAPES) + 1)
    sim.set_fused_kernel(3, 1)  # 16-cell rows: the blocks have 32
    sim.set_primitive(cell)
    with pytest.raises(km.KmError):
        sim.advance(1)
    sim.close()
