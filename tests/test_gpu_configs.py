"""GPU parity of the benchmark configurations' FP32-mixed forms against the FP64 oracle,
each at the configuration's shape (the plain FP64 forms are bitwise elsewhere):

- CT MHD as configuration 3 runs it: 32^3 blocks, two per dimension, PPM-MC + HLL, RK2,
  face B from the curl of the seeded random vector potential; FP32 reconstruction and
  Riemann fluxes, FP64 stage sums, energy recovery and CT face update. Relative L1 per
  conserved variable and per face-B component below 1e-5; max |divB| below 1e-10
  (mhd_blast.py:143-158 asserts 1e-10 in L2);
- the Kelvin-Helmholtz ensemble of configuration 5 with 256^2 and 512^2 members among
  smaller ones, WENO-Z + HLLC, RK3, one fused launch per stage: each member within 1e-5
  of its own FP64 oracle run;
- the 3-D Sedov blast of configuration 4 on a three-level tree of 16^3 blocks, PLM-MC +
  HLLC, RK2, fused path: within 1e-5 of the FP64 oracle (the FP64 fused tree is bitwise,
  test_gpu_fused_amr.py)."""
import numpy as np
import pytest

from tests import _cases

pytestmark = pytest.mark.gpu

TOL = 1e-5  # FP32-mixed against the FP64 oracle, relative L1 per variable


def _rel_l1_faces(g, r, lay):
    out = []
    for d in range(lay.ndim):
        num = float(np.abs(g[d] - r[d]).sum())
        den = float(np.abs(r[d]).sum())
        out.append(num / den)
    return out


def test_c3_ct_mhd_fp32_mixed(gpu, oracle):
    km = gpu
    cfg, desc, lay, cell, faces = _cases.mhd3d(n=64, nblock=32)
    ref = oracle.CpuSimulation(cfg.c, desc, lay)
    ref.set_primitive(cell, faces)
    ref.advance(5)
    r, rf = ref.state()
    cfg.set("kamayan_amd/precision", "fp32_mixed")
    print("C3 shape, FP32-mixed, 64^3 in 32^3 blocks", flush=True)
    sim = km.Simulation(cfg, desc, lay, path=km.PATH_TASKS)
    sim.set_primitive(cell, faces)
    sim.advance(5)
    g, gf = sim.state()
    sim.close()
    l1 = _cases.rel_l1(g, r, lay, 8)
    l1f = _rel_l1_faces(gf, rf, lay)
    divb = float(np.abs(g[(slice(None), km.slot("DIVB", 1)) + lay.interior()]).max())
    print(f"rel L1 cells {l1}, faces {l1f}, max |divB| {divb:.3e}")
    assert np.all(np.array(l1) < TOL), l1
    assert np.all(np.array(l1f) < TOL), l1f
    assert divb < 1e-10


def test_c5_kh_ensemble_fp32_mixed_large_members(gpu, oracle):
    km = gpu
    cfg, members, lay_all, cell = _cases.kh_ensemble([256, 64, 512, 128],
                                                     precision="fp32_mixed")
    descs = [m[0] for m in members]
    print("C5 shape, FP32-mixed ensemble with 256^2 and 512^2 members", flush=True)
    sim = km.Simulation(cfg, descs[0], lay_all, path=km.PATH_FUSED, descs=descs)
    sim.set_primitive(cell)
    sim.advance(3)
    g, _ = sim.state()
    sim.close()
    cfg.set("kamayan_amd/precision", "fp64")
    off = 0
    for m, (d, lay_m, _) in enumerate(members):
        ref = oracle.CpuSimulation(cfg.c, d, lay_m)
        ref.set_primitive(cell[off:off + lay_m.nblocks].copy())
        ref.advance(3)
        r, _ = ref.state()
        l1 = _cases.rel_l1(g[off:off + lay_m.nblocks], r, lay_m, 5)
        l1 = [x for v, x in enumerate(l1) if v != 3]  # no z-momentum in 2-D
        print(f"member {m} ({lay_m.nx[0] * lay_m.nb[0]}^2): rel L1 {l1}")
        assert np.all(np.array(l1) < TOL), (m, l1)
        off += lay_m.nblocks


def test_c4_sedov3d_three_levels_fused_fp32_mixed(gpu, oracle):
    km = gpu
    cfg, desc, lay, leaves, cell = _cases.amr_case(
        "sedov", ndim=3, nroot=4, nblock=16, ng=4, lmax=2, center=(0.0, 0.0, 0.0), radius=0.1,
        bc="outflow", xmin=(-0.5,) * 3, xmax=(0.5,) * 3, gamma=1.4, recon="plm", limiter="mc",
        riemann="hllc", integrator="rk2", params=[1.0, 2.0 / 256.0, 1.0, 1e-5])
    levels = np.bincount(leaves[:, 0], minlength=3)
    assert (levels > 0).all(), levels  # three levels present
    ref = oracle.CpuSimulation(cfg.c, desc, lay, leaves=leaves)
    ref.set_primitive(cell)
    ref.advance(4)
    r, _ = ref.state()
    cfg.set("kamayan_amd/precision", "fp32_mixed")
    print(f"C4 shape, fused FP32-mixed, {len(leaves)} leaves {levels.tolist()}", flush=True)
    sim = km.Simulation(cfg, desc, lay, path=km.PATH_FUSED, leaves=leaves)
    sim.set_primitive(cell)
    sim.advance(4)
    g, _ = sim.state()
    sim.close()
    l1 = _cases.rel_l1(g, r, lay, 5)
    print(f"rel L1 {l1}")
    assert np.all(np.isfinite(l1)) and np.all(np.array(l1) < TOL), l1
