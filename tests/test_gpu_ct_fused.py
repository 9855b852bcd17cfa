"""The fused CT-MHD stage (3-D, Cartesian, single-level meshes) against the FP64 oracle.

The FP64 form is bitwise the oracle (every cell slot, the face field, time and dt) for
each reconstruction and Riemann solver, every pencil-tile shape of the sweeps, RK1-RK3,
and periodic, outflow and reflecting boundaries. The FP64-FMA and FP32-mixed forms of
configuration 3's shape (PPM-MC + HLL, RK2, 32^3 blocks, two per dimension) stay within
1e-12 and 1e-5 relative L1 per conserved variable and face-B component, with max |divB|
at round-off."""
import numpy as np
import pytest

from tests import _cases

pytestmark = pytest.mark.gpu


def _run(km, oracle, case, ncycles, kern=None, precision=None):
    cfg, desc, lay, cell, faces = case
    if precision:
        cfg.set("kamayan_amd/precision", precision)
    sim = km.Simulation(cfg, desc, lay, path=km.PATH_FUSED)
    if kern is not None:
        sim.set_fused_kernel(-1, kern)
    sim.set_primitive(cell, faces)
    cfg.set("kamayan_amd/precision", "fp64")
    ref = oracle.CpuSimulation(cfg.c, desc, lay)
    ref.set_primitive(cell, faces)
    sim.advance(ncycles)
    ref.advance(ncycles)
    g, gf = sim.state()
    r, rf = ref.state()
    out = (g, gf, sim.info()), (r, rf, ref.info())
    sim.close()
    ref.close()
    return lay, out


def _same(a, b, what):
    diff = float(np.max(np.abs(a - b))) if a.size else 0.0
    assert np.array_equal(a, b), f"{what}: max |gpu - oracle| = {diff:.3e}"


def _blast(km, bc, recon="plm", riemann="hllc", integrator="rk2"):
    """mhd_blast (mhd_blast.cpp) on a uniform 16^3 mesh of 8^3 blocks over [-0.5, 0.5]^3."""
    cfg = _cases.Config(**{"hydro/reconstruction": recon, "hydro/slope_limiter": "mc",
                           "hydro/riemann": riemann, "parthenon/time/integrator": integrator,
                           "material/single/gamma": 5.0 / 3.0, "physics/MHD": "ct"})
    cfg.c.ndim = 3
    lay = _cases.Layout(3, 4, (8, 8, 8), (2, 2, 2))
    desc = _cases.mesh_desc(lay, xmin=(-0.5,) * 3, xmax=(0.5,) * 3, bc=(bc,) * 6)
    cell, faces = km.problem_fill("mhd_blast", cfg, desc, lay)
    return cfg, desc, lay, cell, faces


CASES = [
    ("ppm-mc-hll-rk2", dict(recon="ppm", riemann="hll", integrator="rk2")),
    ("wenoz-hllc-rk3", dict(recon="wenoz", riemann="hllc", integrator="rk3")),
    ("plm-mc-hllc-rk1", dict(recon="plm", riemann="hllc", integrator="rk1")),
    ("fog-hll-rk2", dict(recon="fog", riemann="hll", integrator="rk2")),
]


@pytest.mark.parametrize("name,kw", CASES)
def test_ct_fused_bitwise(gpu, oracle, name, kw):
    print(f"CT fused FP64, random field 16^3 in 8^3 blocks, {name}", flush=True)
    lay, ((g, gf, gi), (r, rf, ri)) = _run(gpu, oracle, _cases.mhd3d(n=16, nblock=8, **kw), 2)
    assert gi == ri
    sl = (slice(None), slice(None)) + lay.interior()
    _same(g[sl], r[sl], f"{name} cells")
    # ghost zones too, for every field the exchange fills: the stage leaves some of them
    # stale or holding the edge zone's value, and the fill before the state is read must
    # restore all of them
    _same(g[:, :gpu.slot("TEMP", 1)], r[:, :gpu.slot("TEMP", 1)], f"{name} cells with ghosts")
    for d in range(3):
        _same(gf[d], rf[d], f"{name} face {d}")
    divb = g[(slice(None), gpu.slot("DIVB", 1)) + lay.interior()]
    assert float(np.max(np.abs(divb))) < 1e-11


@pytest.mark.parametrize("kern", [1, 2, 3, 4])
def test_ct_fused_tile_shapes_bitwise(gpu, oracle, kern):
    print(f"CT fused FP64, tile shape {kern}, 32^3 in 16^3 blocks", flush=True)
    case = _cases.mhd3d(n=32, nblock=16)
    lay, ((g, gf, gi), (r, rf, ri)) = _run(gpu, oracle, case, 1, kern=kern)
    assert gi == ri
    sl = (slice(None), slice(None)) + lay.interior()
    _same(g[sl], r[sl], f"shape {kern} cells")
    for d in range(3):
        _same(gf[d], rf[d], f"shape {kern} face {d}")


@pytest.mark.parametrize("bc", ["outflow", "reflect"])
def test_ct_fused_blast_boundaries_bitwise(gpu, oracle, bc):
    print(f"CT fused FP64, mhd_blast 16^3, {bc} boundaries", flush=True)
    lay, ((g, gf, gi), (r, rf, ri)) = _run(gpu, oracle, _blast(gpu, bc), 3)
    assert gi == ri
    sl = (slice(None), slice(None)) + lay.interior()
    _same(g[sl], r[sl], f"{bc} cells")
    for d in range(3):
        _same(gf[d], rf[d], f"{bc} face {d}")


def _rel_l1_faces(g, r):
    return [float(np.abs(g[d] - r[d]).sum()) / float(np.abs(r[d]).sum()) for d in range(3)]


@pytest.mark.parametrize("precision,tol", [("fp64_fma", 1e-12), ("fp32_mixed", 1e-5)])
def test_ct_fused_c3_shape_forms(gpu, oracle, precision, tol):
    print(f"CT fused {precision}, 64^3 in 32^3 blocks, PPM-MC + HLL, RK2", flush=True)
    case = _cases.mhd3d(n=64, nblock=32)
    lay, ((g, gf, _), (r, rf, _)) = _run(gpu, oracle, case, 4, precision=precision)
    l1 = _cases.rel_l1(g, r, lay, 8)
    l1f = _rel_l1_faces(gf, rf)
    divb = float(np.abs(g[(slice(None), gpu.slot("DIVB", 1)) + lay.interior()]).max())
    print(f"rel L1 cells {l1}, faces {l1f}, max |divB| {divb:.3e}")
    assert np.all(np.array(l1) < tol), l1
    assert np.all(np.array(l1f) < tol), l1f
    assert divb < 1e-10
