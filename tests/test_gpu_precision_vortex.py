"""Accuracy of the fused path's precision forms on the smooth isentropic vortex (the
reference's convergence suite, isentropic_vortex.py:13-60: WENO-Z, RK2, nghost 3, one
period to t = 10, density L1 error against the initial state).

Each form must converge with slope > 1.5 from 32^2 to 64^2, stay under the suite's WENO-Z
ceiling at 64^2 (0.005, reconstruction.py:135), and the FP32-mixed form must reach the
same L1 error as the FP64 form: relative difference below 1e-3 at both resolutions. The
FP64 form is also compared with the oracle's error to 1e-12."""
import math

import pytest

from tests import _cases

pytestmark = pytest.mark.gpu


def _vortex_error(km, precision, n, nblock):
    cfg, desc, lay, cell, _ = _cases.vortex(n=n, nblock=nblock)
    cfg.set("kamayan_amd/precision", precision)
    sim = km.Simulation(cfg, desc, lay, path=km.PATH_FUSED)
    sim.set_primitive(cell)
    sim.advance(100000, tlim=10.0)
    t, _, ncyc = sim.info()
    out, _ = sim.state()
    sim.close()
    assert t == pytest.approx(10.0, abs=1e-12)
    return _cases.vortex_l1(out, lay, cell), ncyc


def test_fp32_mixed_converges_at_the_fp64_error(gpu, oracle):
    errs = {}
    for prec in ("fp64", "fp32_mixed"):
        e32, _ = _vortex_error(gpu, prec, 32, 16)
        e64, ncyc = _vortex_error(gpu, prec, 64, 32)
        slope = -math.log(e64 / e32) / math.log(2.0)
        print(f"[vortex] {prec}: L1(32^2) = {e32:.6e}, L1(64^2) = {e64:.6e} "
              f"({ncyc} cycles), slope {slope:.3f}", flush=True)
        assert slope > 1.5, (prec, e32, e64, slope)
        assert e64 < 0.005, (prec, e64)
        errs[prec] = (e32, e64)
    for a, b in zip(errs["fp32_mixed"], errs["fp64"]):
        assert abs(a / b - 1.0) < 1e-3, (errs["fp32_mixed"], errs["fp64"])
    # the FP64 form against the oracle's own run
    cfg, desc, lay, cell, _ = _cases.vortex(n=64, nblock=32)
    ref = oracle.CpuSimulation(cfg.c, desc, lay)
    ref.set_primitive(cell)
    ref.advance(100000, tlim=10.0)
    out, _ = ref.state()
    e_ref = _cases.vortex_l1(out, lay, cell)
    assert abs(errs["fp64"][1] / e_ref - 1.0) < 1e-12, (errs["fp64"][1], e_ref)
