"""Pins of the oracle's history variables of the isentropic vortex (isentropic_vortex.cpp:
74-101, ErrorHistory isentropic_vortex.hpp:150-200): zero on the initial state, the closed
form of a uniform offset (delta times the volume over the x1-x2 extent), and agreement
with the L1 helpers the reference suites' ceilings are checked with."""
import numpy as np
import pytest

import kamayan_amd as km
from tests import _cases


def _errors(oracle, cfg, lay, desc, cell, nv, params=None):
    """the oracle's history variables of a cell array (the host pack lives through the call)"""
    h = oracle.HostPack(lay, nv, blocks=oracle.block_table(desc, lay.nblocks))
    h.cell[:] = cell
    return oracle.history_vortex_errors(cfg.c, h.pack, desc, params)


@pytest.mark.parametrize("case", ["hydro2d", "mhd2d", "cyl1d", "cylct1d"])
def test_initial_state_has_no_error(oracle, case):
    if case == "hydro2d":
        cfg, desc, lay, cell, _ = _cases.vortex()
    elif case == "mhd2d":
        cfg, desc, lay, cell, _ = _cases.vortex(ng=4)
        cfg.set("physics/MHD", "ct")
        cell, _ = km.problem_fill("isentropic_vortex", cfg, desc, lay)
    else:
        cfg, desc, lay, cell, _ = _cases.vortex_cyl(mhd="ct" if case == "cylct1d" else "off")
    err = _errors(oracle, cfg, lay, desc, cell, km.nvar(cfg.c.mhd))
    assert len(err) == (8 if cfg.c.mhd else 5)
    assert np.all(err == 0.0)


def test_uniform_offset_gives_delta_times_volume_over_area(oracle):
    # Cartesian 2-D on [-5,5]^3 (one x3 cell of depth 10): sum V / (10 * 10) = 10
    cfg, desc, lay, cell, _ = _cases.vortex()
    d = 1.0e-3
    sl = (slice(None), km.slot("PRES")) + lay.interior()
    cell[sl] += d
    err = _errors(oracle, cfg, lay, desc, cell, 12)
    assert err[4] == pytest.approx(10.0 * d, rel=1e-12)
    assert np.all(err[:4] == 0.0)
    # cylindrical r in [0, 5], z in [-5, 5]: sum V = pi 5^2 10, over 5 * 10 -> 5 pi
    cfg, desc, lay, cell, _ = _cases.vortex_cyl()
    cell[(slice(None), km.slot("VEL2")) + lay.interior()] += d
    err = _errors(oracle, cfg, lay, desc, cell, 12)
    assert err[3] == pytest.approx(5.0 * np.pi * d, rel=1e-12)
    assert np.all(err[[0, 1, 2, 4]] == 0.0)


def test_history_equals_the_suite_helpers(oracle):
    cfg, desc, lay, cell, _ = _cases.vortex(n=64, nblock=32, recon="plm", limiter="mc",
                                            riemann="hllc")
    sim = oracle.CpuSimulation(cfg.c, desc, lay)
    sim.set_primitive(cell)
    sim.advance(30)
    st, _ = sim.state()
    err = _errors(oracle, cfg, lay, desc, st, 12)
    for k, name in ((0, "DENS"), (1, "VEL0"), (2, "VEL1"), (4, "PRES")):
        want = _cases.vortex_l1(st, lay, cell, km.slot(name))
        assert err[k] == pytest.approx(want, rel=1e-12), name
    cfg, desc, lay, cell, _ = _cases.vortex_cyl()
    sim = oracle.CpuSimulation(cfg.c, desc, lay)
    sim.set_primitive(cell)
    sim.advance(30)
    st, _ = sim.state()
    err = _errors(oracle, cfg, lay, desc, st, 12)
    blocks = oracle.block_table(desc, lay.nblocks)
    want = _cases.vortex_cyl_l1(st, lay, blocks, km.slot("VEL2"))
    assert err[3] == pytest.approx(want, rel=1e-12)
