"""The fused path's sweep kernels: the pencil march and every pencil-tile shape, in each
direction, must give the same stage, and the autotuner's picks must be among them. In the FP64 form every
shape is compared with the oracle bit for bit. In the FP32-mixed form each shape is
compared with the march (relative L1 per conserved variable below 1e-6: both evaluate
the same FP32 expressions per value, so only the compiler's contraction choices may
differ) and with the FP64 oracle at the form's tolerance (1e-5). A shape whose LDS does
not fit the block's rows is refused with an error."""
import numpy as np
import pytest

from tests import _cases

pytestmark = pytest.mark.gpu

NSHAPES = 6  # pencil-tile shapes
SHAPES = [(4, 256), (8, 256), (8, 512), (16, 256), (32, 1024), (64, 1024)]  # (pencils, lanes)
INFO_BYTES = 336  # per-pencil table of a tile workgroup (sizeof(PencilInfo))
WIDTH = {"wenoz": 2, "ppm": 2, "plm": 1, "fog": 0}


def _fits(shape, lay, recon, size):
    """Whether tile shape `shape` fits every direction's pencils in LDS."""
    rows = SHAPES[shape - 1][0]
    w = WIDTH[recon]
    for n in lay.nx[:lay.ndim]:
        vals = rows * (7 * (n + 2 + 2 * w) + 7 * (n + 2))
        if vals * size + INFO_BYTES * rows > 160 * 1024:
            return False
    return True


def _case(kind, precision):
    case = _case_fp(kind)
    case[0].set("kamayan_amd/precision", precision)
    return case


def _case_fp(kind):
    import kamayan_amd as km
    if kind == "smooth3d-wenoz-hllc":
        return _cases.smooth3d(n=32, nblock=16)
    if kind == "smooth3d-ppm-hll-64":
        return _cases.smooth3d(n=64, nblock=64, recon="ppm", riemann="hll", integrator="rk2")
    if kind == "vortex2d-plm-mc-hllc":
        return _cases.vortex(n=64, nblock=32, ng=4, recon="plm", limiter="mc", riemann="hllc",
                             cfl=0.4)
    # 3-D, 12 x 8 x 8 cells per block (rows shorter than a wavefront and not a power of
    # two), reflecting inner and outflow outer x1 boundary, periodic elsewhere
    cfg = km.Config(**{"hydro/reconstruction": "wenoz", "hydro/riemann": "hllc",
                       "parthenon/time/integrator": "rk3", "material/single/gamma": 5.0 / 3.0})
    cfg.c.ndim = 3
    lay = km.Layout(3, 4, (12, 8, 8), (3, 2, 2))
    desc = km.mesh_desc(lay, bc=("reflect", "outflow") + ("periodic",) * 4)
    cell, faces = km.problem_fill("smooth3d", cfg, desc, lay, seed=7)
    return cfg, desc, lay, cell, faces


KINDS = ["smooth3d-wenoz-hllc", "smooth3d-ppm-hll-64", "vortex2d-plm-mc-hllc",
         "wenoz-12x8x8-reflect-outflow"]


def _run(km, case, kern, ncyc):
    """kern: one kernel for every direction, a tuple per direction, or None (autotune)."""
    cfg, desc, lay, cell, faces = case
    sim = km.Simulation(cfg, desc, lay, path=km.PATH_FUSED)
    if isinstance(kern, tuple):
        for d, k in enumerate(kern):
            sim.set_fused_kernel(d, k)
    elif kern is not None:
        sim.set_fused_kernel(-1, kern)
    try:
        sim.set_primitive(cell, faces)
        sim.advance(ncyc)
        g, _ = sim.state()
        info = sim.info()
        used = tuple(sim.fused_kernel(d) for d in range(3))
    finally:
        sim.close()
    return g, info, used


def _recon(cfg):
    return {0: "fog", 1: "plm", 2: "ppm", 3: "wenoz"}[cfg.c.recon]


@pytest.mark.parametrize("kind", KINDS)
def test_tile_shapes_fp64_bitwise(gpu, oracle, kind):
    import kamayan_amd as km
    case = _case(kind, "fp64")
    cfg, desc, lay, cell, faces = case
    ncyc = 3
    ref = oracle.CpuSimulation(cfg.c, desc, lay)
    ref.set_primitive(cell, faces)
    ref.advance(ncyc)
    r, _ = ref.state()
    rinfo = ref.info()
    sl = (slice(None), slice(0, 5)) + lay.interior()
    for shape in [0] + list(range(1, NSHAPES + 1)) + [(1, 3, 2), (2, 0, 1), None]:
        print(f"[tiles fp64] {kind} kernels {shape}", flush=True)
        if isinstance(shape, int) and shape > 0 and not _fits(shape, lay, _recon(cfg), 8):
            with pytest.raises(km.KmError):
                _run(gpu, case, shape, ncyc)
            continue
        g, info, used = _run(gpu, case, shape, ncyc)
        if shape is None:
            assert all(0 <= u <= NSHAPES for u in used)
        else:
            want = shape if isinstance(shape, tuple) else (shape,) * 3
            assert used[:lay.ndim] == want[:lay.ndim]
        assert info == rinfo
        assert np.array_equal(g[sl], r[sl]), (shape, float(np.max(np.abs(g[sl] - r[sl]))))
        # the conserved ghost zones as well: the final sweep may leave the edge zone's value
        # in u1's x1 ghost columns, which the fill before the state is read must replace
        assert np.array_equal(g[:, :5], r[:, :5]), (shape, "ghost zones")


@pytest.mark.parametrize("kind", KINDS)
def test_tile_shapes_fp32_mixed(gpu, oracle, kind):
    case = _case(kind, "fp32_mixed")
    cfg, desc, lay, cell, faces = case
    ncyc = 3
    ref = oracle.CpuSimulation(cfg.c, desc, lay)
    ref.set_primitive(cell, faces)
    ref.advance(ncyc)
    r, _ = ref.state()
    march, _, used = _run(gpu, case, 0, ncyc)
    assert used == (0, 0, 0)
    l1 = _cases.rel_l1(march, r, lay, 5)
    assert np.all(l1 < 1e-5), ("march vs oracle", l1)
    for shape in list(range(1, NSHAPES + 1)) + [(3, 1, 4)]:
        print(f"[tiles fp32] {kind} kernels {shape}", flush=True)
        if isinstance(shape, int) and not _fits(shape, lay, _recon(cfg), 4):
            with pytest.raises(gpu.KmError):
                _run(gpu, case, shape, ncyc)
            continue
        g, _, used = _run(gpu, case, shape, ncyc)
        want = shape if isinstance(shape, tuple) else (shape,) * 3
        assert used[:lay.ndim] == want[:lay.ndim]
        sl = (slice(None), slice(0, 5)) + lay.interior()
        assert np.all(np.isfinite(g[sl]))
        d = _cases.rel_l1(g, march, lay, 5)
        assert np.all(d < 1e-6), (shape, d)
        l1 = _cases.rel_l1(g, r, lay, 5)
        assert np.all(l1 < 1e-5), (shape, l1)
    auto, _, used = _run(gpu, case, None, ncyc)
    assert all(0 <= u <= NSHAPES for u in used)
    assert np.all(_cases.rel_l1(auto, march, lay, 5) < 1e-6)


def test_kernel_choice_refusals_and_one_dimension(gpu, oracle):
    import kamayan_amd as km
    case = _cases.smooth3d(n=16, nblock=8, precision="fp32_mixed")
    cfg, desc, lay, cell, faces = case
    sim = gpu.Simulation(cfg, desc, lay, path=gpu.PATH_FUSED)
    with pytest.raises(km.KmError):
        sim.set_fused_kernel(0, NSHAPES + 1)
    with pytest.raises(km.KmError):
        sim.set_fused_kernel(0, -2)
    with pytest.raises(km.KmError):
        sim.set_fused_kernel(4, 1)  # directions 0..2, and 3 for the whole stage
    with pytest.raises(km.KmError):
        sim.set_fused_kernel(3, 99)
    sim.close()
    # one dimension, RK3: x1 is both the first and the final sweep; every shape is
    # bitwise the oracle in FP64
    cfg = km.Config(**{"hydro/reconstruction": "wenoz", "hydro/riemann": "hllc",
                       "parthenon/time/integrator": "rk3", "material/single/gamma": 1.4})
    cfg.c.ndim = 1
    lay = km.Layout(1, 4, (32, 1, 1), (4, 1, 1))
    desc = km.mesh_desc(lay)
    cell, faces = km.problem_fill("smooth3d", cfg, desc, lay, seed=3)
    ref = oracle.CpuSimulation(cfg.c, desc, lay)
    ref.set_primitive(cell, faces)
    ref.advance(3)
    r, _ = ref.state()
    sl = (slice(None), slice(0, 5)) + lay.interior()
    for shape in range(0, NSHAPES + 1):
        if shape > 0 and not _fits(shape, lay, "wenoz", 8):
            with pytest.raises(km.KmError):
                _run(gpu, (cfg, desc, lay, cell, faces), shape, 3)
            continue
        g, info, used = _run(gpu, (cfg, desc, lay, cell, faces), shape, 3)
        assert used[0] == shape
        assert info == ref.info()
        assert np.array_equal(g[sl], r[sl]), shape
