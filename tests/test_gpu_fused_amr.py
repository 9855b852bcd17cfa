"""The fused path on block-tree meshes: pencil-tile sweeps that read the stored
primitives (ghost zones filled across levels), the fluxes of faces where a block borders
a coarser leaf computed first and restricted onto the coarse face, and the task path's
PreparePrimitive, CFL estimate and ghost exchange. In FP64 every run must equal the
oracle bit for bit (conserved variables and the exchanged primitives in every zone, the
temperature in the interior, time, dt and cycle), for the autotuner's kernels and for
every tile shape; the FP32-mixed form to 1e-5 relative L1 per conserved variable."""
import numpy as np
import pytest

from kamayan_amd._abi import amr_adapt, amr_criterion, slot
from tests import _cases

pytestmark = pytest.mark.gpu

TEMP = 11  # hydro slot of the temperature (not a ghost-exchanged field)


def _same_state(g, r, lay, what):
    sl = (slice(None), slice(0, TEMP))
    for a, b, part in ((g[sl], r[sl], "exchanged fields"),
                       (g[(slice(None), TEMP) + lay.interior()],
                        r[(slice(None), TEMP) + lay.interior()], "interior temperature")):
        if not np.array_equal(a, b):
            bad = np.argwhere(a != b)
            i = tuple(bad[0])
            raise AssertionError(f"{what} ({part}): {len(bad)} differ, first {i}: "
                                 f"{a[i]!r} vs {b[i]!r}")


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
    ("smooth3d-reflect-ppm-hll", lambda: _cases.amr_case(
        "smooth3d", ndim=3, nroot=2, nblock=8, ng=3, lmax=1, center=(0.2, 0.7, 0.4),
        radius=0.25, bc="reflect", recon="ppm", riemann="hll"), 3),
]


def _run(km, cfg, desc, lay, leaves, cell, ncyc, kern=None):
    sim = km.Simulation(cfg, desc, lay, path=km.PATH_FUSED, leaves=leaves)
    if kern is not None:
        sim.set_fused_kernel(-1, kern)
    sim.set_primitive(cell)
    sim.advance(ncyc)
    g, _ = sim.state()
    info = sim.info()
    used = tuple(sim.fused_kernel(d) for d in range(cfg.c.ndim))
    sim.close()
    return g, info, used


@pytest.mark.parametrize("name,make,ncyc", RUNS, ids=[r[0] for r in RUNS])
def test_fused_amr_run_bitwise(gpu, oracle, name, make, ncyc):
    cfg, desc, lay, leaves, cell = make()
    print(f"fused amr run {name} leaves={len(leaves)} levels={sorted(set(leaves[:, 0]))}",
          flush=True)
    ref = oracle.CpuSimulation(cfg.c, desc, lay, leaves=leaves)
    ref.set_primitive(cell)
    ref.advance(ncyc)
    r, _ = ref.state()
    g, info, used = _run(gpu, cfg, desc, lay, leaves, cell, ncyc)
    assert all(1 <= u <= 6 for u in used), used
    _same_state(g, r, lay, name)
    assert info == ref.info()


def test_fused_amr_every_tile_shape_bitwise(gpu, oracle):
    cfg, desc, lay, leaves, cell = RUNS[1][1]()
    ref = oracle.CpuSimulation(cfg.c, desc, lay, leaves=leaves)
    ref.set_primitive(cell)
    ref.advance(3)
    r, _ = ref.state()
    import kamayan_amd as km
    shapes = [(4, 256), (8, 256), (8, 512), (16, 256), (32, 1024), (64, 1024)]
    w = {0: 0, 1: 1, 2: 2, 3: 2}[cfg.c.recon]
    for shape in range(1, 7):
        print(f"[fused amr] tile shape {shape}", flush=True)
        rows = shapes[shape - 1][0]
        n = lay.nx[0]
        # LDS of a tile (FP64): primitives and face values of its pencils, plus the
        # 336-byte per-pencil table; a shape past 160 KiB is refused
        if rows * (7 * (n + 2 + 2 * w) + 7 * (n + 2)) * 8 + 336 * rows > 160 * 1024:
            with pytest.raises(km.KmError):
                _run(gpu, cfg, desc, lay, leaves, cell, 3, kern=shape)
            continue
        g, info, used = _run(gpu, cfg, desc, lay, leaves, cell, 3, kern=shape)
        assert used == (shape,) * 2
        _same_state(g, r, lay, f"shape {shape}")
        assert info == ref.info()
    # the march reads face neighbours of a single-level mesh: refused on a block tree
    with pytest.raises(km.KmError):
        _run(gpu, cfg, desc, lay, leaves, cell, 3, kern=0)


def test_fused_amr_fp32_mixed(gpu, oracle):
    cfg, desc, lay, leaves, cell = RUNS[3][1]()
    ref = oracle.CpuSimulation(cfg.c, desc, lay, leaves=leaves)
    ref.set_primitive(cell)
    ref.advance(6)
    r, _ = ref.state()
    cfg.set("kamayan_amd/precision", "fp32_mixed")
    g, _, _ = _run(gpu, cfg, desc, lay, leaves, cell, 6)
    l1 = _cases.rel_l1(g, r, lay, 5)
    l1 = [x for x, v in zip(l1, range(5)) if v != 3]  # no z-momentum in 2-D
    assert np.all(np.isfinite(l1)) and np.all(np.array(l1) < 1e-5), l1


def test_fused_adaptive_sedov_bitwise(gpu, oracle):
    """sedov.py:38 shape with refinement on pres: the tree follows the blast while the
    fused stages advance it; leaves, state, time and dt match the oracle."""
    km = gpu
    cfg, desc, lay, leaves, cell = _cases.amr_case(
        "sedov", ndim=2, nroot=4, nblock=8, ng=4, recon="wenoz", limiter="minmod",
        center=(0.0, 0.0, 0.0), radius=0.1, lmax=0, bc="outflow", riemann="hllc",
        integrator="rk2", xmin=(-0.5,) * 3, xmax=(0.5,) * 3, gamma=1.4,
        params=[1.0, 0.05, 1.0, 1e-5])
    adapt = amr_adapt(3, [amr_criterion(slot("PRES"), 3)], derefine_count=10)
    sim = km.Simulation(cfg, desc, lay, path=km.PATH_FUSED, leaves=leaves)
    sim.set_primitive(cell)
    sim.set_adaptive(adapt)
    ref = oracle.CpuSimulation(cfg.c, desc, lay, leaves=leaves)
    ref.set_primitive(cell)
    ref.set_adaptive(adapt)
    seen = {len(leaves)}
    for c in range(16):
        sim.advance(1)
        ref.advance(1)
        gl, rl = sim.leaves(), ref.leaves()
        assert np.array_equal(gl, rl), f"leaves after cycle {c + 1}"
        seen.add(len(gl))
        if c % 4 == 3:
            g, _ = sim.state()
            r, _ = ref.state()
            lay_c = km.Layout(2, 4, (8, 8, 1), (4, 4, 1), nleaves=len(gl))
            _same_state(g, r, lay_c, f"state after cycle {c + 1}")
            assert sim.info() == ref.info()
    assert len(seen) > 1  # the tree changed
    sim.close()
