"""Writes sedov3d_density_tree.npz: the block tree of a 3-D Sedov blast refined on the
density gradient, and the radial profiles of its state, for the block-tree benchmark.

The run: the domain [-0.5, 0.5]^3 with 64^3 root zones in 16^3 blocks (nghost 4,
outflow), numlevel 4 (three refinement levels above the root grid), energy 1 deposited
in a sphere of radius two finest zones (2/512) around the origin, ambient density 1 and
pressure 1e-5, gamma 1.4, PLM-MC + HLLC, RK2, CFL 0.8, FP64 (the sedov problem,
sedov.cpp / sedov.in). The start refines the root grid to the finest level where blocks
meet a ball of radius 0.02 (so the deposit is resolved); from then on the tree adapts
every cycle to the Loehner estimator on density (kamayan/refinement0 with field = dens,
grid.cpp defaults refine_tol 0.8, derefine_tol 0.2, filter 0.01, derefine_count 10).
The run stops at the first check (every 20 cycles) where the tree holds 3,500 leaves or
more. Saved: the leaves (level, lx1, lx2, lx3), the simulation time and cycle, and the
volume-weighted radial profiles of density, radial velocity and pressure over the leaf
interiors in shells of one finest zone width, which the benchmark interpolates at the
zone centres of the tree to recreate the blast state on it.

Needs a GPU and the built library: python tests/golden/make_sedov_density_tree.py [out.npz]"""
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
import kamayan_amd as km  # noqa: E402
from kamayan_amd._abi import amr_adapt, amr_criterion  # noqa: E402

NUMLEVEL = 4
NX, NB, NG = 16, 4, 4
TARGET_LEAVES = 3500


def setup():
    cfg = km.Config(**{"hydro/reconstruction": "plm", "hydro/slope_limiter": "mc",
                       "hydro/riemann": "hllc", "parthenon/time/integrator": "rk2",
                       "material/single/gamma": 1.4})
    cfg.c.ndim = 3
    root = km.Layout(3, NG, (NX,) * 3, (NB,) * 3)
    desc = km.mesh_desc(root, xmin=(-0.5,) * 3, xmax=(0.5,) * 3, bc=("outflow",) * 6)
    return cfg, desc


def zone_centres(leaves):
    """x1, x2, x3 zone centres [nleaves, NX, NX, NX] (x3 slowest) and the zone volumes."""
    L = leaves[:, 0].astype(np.float64)
    dx = 1.0 / (NB * NX * 2.0 ** L)
    xs = [-0.5 + (leaves[:, 1 + d][:, None] * NX + np.arange(NX)[None, :] + 0.5) * dx[:, None]
          for d in range(3)]
    x1 = xs[0][:, None, None, :]
    x2 = xs[1][:, None, :, None]
    x3 = xs[2][:, :, None, None]
    return x1, x2, x3, dx ** 3


def main():
    cfg, desc = setup()
    lmax = NUMLEVEL - 1
    leaves = km.refine_ball(desc, lmax, (0.0, 0.0, 0.0), 0.02)
    lay = km.Layout(3, NG, (NX,) * 3, (NB,) * 3, nleaves=len(leaves))
    cell, _ = km.problem_fill("sedov", cfg, desc, lay,
                              params=[1.0, 2.0 / (NB * NX << lmax), 1.0, 1e-5], leaves=leaves)
    sim = km.Simulation(cfg, desc, lay, path=km.PATH_FUSED, leaves=leaves)
    sim.set_primitive(cell)
    sim.set_adaptive(amr_adapt(NUMLEVEL, [amr_criterion(km.slot("DENS"), lmax)],
                               derefine_count=10))
    t0 = time.perf_counter()
    cycles = 0
    while True:
        sim.advance(20)
        cycles += 20
        leaves = sim.leaves()
        if cycles % 500 == 0 or len(leaves) >= TARGET_LEAVES:
            t, dt, _ = sim.info()
            print(f"cycle {cycles} t {t:.6f} leaves {len(leaves)} "
                  f"{np.bincount(leaves[:, 0], minlength=NUMLEVEL).tolist()} "
                  f"{time.perf_counter() - t0:.0f} s", flush=True)
        if len(leaves) >= TARGET_LEAVES:
            break
    t, _, _ = sim.info()
    g, _ = sim.state()
    sim.close()
    inner = (slice(None),) + (slice(NG, NG + NX),) * 3
    dens = g[(slice(None), km.slot("DENS")) + inner[1:]]
    pres = g[(slice(None), km.slot("PRES")) + inner[1:]]
    v = [g[(slice(None), km.slot(f"VEL{d}")) + inner[1:]] for d in range(3)]
    x1, x2, x3, vol = zone_centres(leaves)
    r = np.sqrt(x1 ** 2 + x2 ** 2 + x3 ** 2)
    vr = (v[0] * x1 + v[1] * x2 + v[2] * x3) / np.maximum(r, 1e-300)
    h = 1.0 / (NB * NX << lmax)
    nbin = int(np.ceil(np.sqrt(3) * 0.5 / h)) + 1
    k = np.minimum((r / h).astype(np.int64), nbin - 1).ravel()
    w = np.broadcast_to(vol[:, None, None, None], r.shape).ravel()
    wsum = np.bincount(k, weights=w, minlength=nbin)
    prof = {}
    for name, f in (("dens", dens), ("vrad", vr), ("pres", pres)):
        prof[name] = np.bincount(k, weights=w * f.ravel(), minlength=nbin) / np.maximum(wsum, 1e-300)
    filled = wsum > 0
    rc = (np.arange(nbin) + 0.5) * h
    out = sys.argv[1] if len(sys.argv) > 1 else os.path.join(
        os.path.dirname(os.path.abspath(__file__)), "sedov3d_density_tree.npz")
    np.savez_compressed(out, leaves=leaves.astype(np.int32), time=np.float64(t),
                        cycle=np.int64(cycles), r=rc[filled], dens=prof["dens"][filled],
                        vrad=prof["vrad"][filled], pres=prof["pres"][filled],
                        numlevel=np.int64(NUMLEVEL))
    print("wrote", out, len(leaves), "leaves at t", t, flush=True)


if __name__ == "__main__":
    main()
