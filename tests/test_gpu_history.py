"""GPU history output (Parthenon's hst file type) of the isentropic vortex: the device
reduction of the history variables against the oracle's, and history files read the way
the reference suites read them -- np.loadtxt with the suites' column (isentropic_vortex.py:
45, reconstruction.py:123-127,199-201, mhd_vortex.py:65) -- checked against the suites'
ceilings, with the last row equal to the oracle's errors of the final state."""
import numpy as np
import pytest

from kamayan_amd import Config, Layout, mesh_desc, problem_fill
from tests import _cases

pytestmark = pytest.mark.gpu


def _errors(oracle, cfg, lay, desc, cell, nv, params=None):
    """the oracle's history variables of a cell array (the host pack lives through the call)"""
    h = oracle.HostPack(lay, nv, blocks=oracle.block_table(desc, lay.nblocks))
    h.cell[:] = cell
    return oracle.history_vortex_errors(cfg.c, h.pack, desc, params)


def _mhd_vortex(n=64, strength=1.0):
    # mhd_vortex.py:40-56 on isentropic_vortex.in: 64^2 in 32^2 blocks, nghost 4, PPM-MC,
    # RK2, CT, strength 1
    cfg = Config(**{"hydro/reconstruction": "ppm", "hydro/slope_limiter": "mc",
                    "parthenon/time/integrator": "rk2", "hydro/cfl": 0.8,
                    "material/single/gamma": 1.4, "physics/MHD": "ct"})
    cfg.c.ndim = 2
    lay = Layout(2, 4, (n // 2, n // 2, 1), (2, 2, 1))
    desc = mesh_desc(lay, xmin=(-5.0,) * 3, xmax=(5.0,) * 3)
    params = [1.0, 1.0, 1.0, 1.0, strength, 1.0]
    cell, faces = problem_fill("isentropic_vortex", cfg, desc, lay, params=params)
    return cfg, desc, lay, cell, faces, params


@pytest.mark.parametrize("case", ["hydro2d", "cyl1d", "mhd2d", "cylct1d"])
def test_history_reduction_matches_oracle(gpu, oracle, case):
    km = gpu
    params = None
    faces = None
    if case == "hydro2d":
        cfg, desc, lay, cell, _ = _cases.vortex(recon="plm", limiter="mc", riemann="hllc")
    elif case == "mhd2d":
        cfg, desc, lay, cell, faces, params = _mhd_vortex(n=32)
    else:
        cfg, desc, lay, cell, faces = _cases.vortex_cyl(mhd="ct" if case == "cylct1d" else "off")
    print(f"history reduction {case}", flush=True)
    sim = km.Simulation(cfg, desc, lay, path=km.PATH_TASKS)
    sim.set_primitive(cell, faces)
    sim.advance(15)
    st, _ = sim.state()
    sim.close()
    nv = km.nvar(cfg.c.mhd)
    dev = km.DevicePack(lay, nv, blocks=oracle.block_table(desc, lay.nblocks))
    dev.upload(cell=st)
    got = km.history_vortex_errors(cfg, dev, desc, params)
    dev.close()
    want = _errors(oracle, cfg, lay, desc, st, nv, params)
    assert np.all(want[[0, 4]] > 0)
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-17)


def _read(path, ncols):
    data = np.loadtxt(path)
    assert data.ndim == 2 and data.shape[1] == ncols
    with open(path) as f:
        assert f.readline().startswith("#")
    return data


@pytest.mark.parametrize("path", ["fused", "tasks"])
def test_reconstruction_suite_history_file(gpu, oracle, tmp_path, path):
    # reconstruction.py:125-135: plm/mc at 64^2, rk2, output dt 1, tlim 10; the suite reads
    # column 4 (density error) of the last row and requires <= 0.02
    km = gpu
    cfg, desc, lay, cell, _ = _cases.vortex(n=64, nblock=32, recon="plm", limiter="mc")
    print(f"reconstruction suite history, {path} path", flush=True)
    sim = km.Simulation(cfg, desc, lay, path=km.PATH_FUSED if path == "fused" else km.PATH_TASKS)
    sim.set_primitive(cell)
    hst = tmp_path / "isentropic_vortex_plm_mc.out0.hst"
    sim.set_history(hst, 1.0)
    sim.advance(100000, 10.0)
    st, _ = sim.state()
    t, _, ncyc = sim.info()
    sim.close()
    data = _read(hst, 9)
    assert t == 10.0 and data[-1, 0] == 10.0 and data[-1, 2] == ncyc
    assert data[0, 0] == 0.0 and data[0, 2] == 0 and np.all(data[:, 3] == lay.nblocks)
    # one row at t = 0 and one at the first cycle reaching each of 1, 2, ..., 10
    assert len(data) == 11
    assert np.all(data[1:, 0] >= np.arange(1, 11)) and np.all(np.diff(data[:, 2]) > 0)
    err = float(np.loadtxt(hst, usecols=4)[-1])
    print(f"density error {err:.4e}")
    assert err <= 0.02
    want = _errors(oracle, cfg, lay, desc, st, 12)
    np.testing.assert_allclose(data[-1, 4:], want, rtol=1e-12, atol=1e-16)


@pytest.mark.parametrize("mhd,ceiling", [("off", 0.3), ("ct", 0.35)])
def test_cylindrical_suite_history_file(gpu, tmp_path, mhd, ceiling):
    # reconstruction.py:136-137: wenoz, cylindrical, column 7 (azimuthal velocity error)
    km = gpu
    cfg, desc, lay, cell, faces = _cases.vortex_cyl(mhd=mhd)
    print(f"cylindrical suite history, MHD {mhd}", flush=True)
    sim = km.Simulation(cfg, desc, lay, path=km.PATH_TASKS)
    sim.set_primitive(cell, faces)
    hst = tmp_path / "cyl.out0.hst"
    sim.set_history(hst, 1.0)
    sim.advance(100000, 10.0)
    sim.close()
    data = _read(hst, 12 if mhd == "ct" else 9)
    err = float(np.loadtxt(hst, usecols=7)[-1])
    print(f"v_phi error {err:.4e}")
    assert data[-1, 0] == 10.0 and err <= ceiling


def test_mhd_vortex_suite_history_file(gpu, tmp_path):
    # mhd_vortex.py:26-28,51-70: ppm/mc CT at 64^2, column 9 (magx error) <= 0.05
    km = gpu
    cfg, desc, lay, cell, faces, params = _mhd_vortex()
    print("mhd vortex suite history", flush=True)
    sim = km.Simulation(cfg, desc, lay, path=km.PATH_TASKS)
    sim.set_primitive(cell, faces)
    hst = tmp_path / "mhd_vortex.out0.hst"
    sim.set_history(hst, 1.0, params=params)
    sim.advance(100000, 10.0)
    sim.close()
    data = _read(hst, 12)
    err = float(np.loadtxt(hst, usecols=9)[-1])
    print(f"magx error {err:.4e}")
    assert data[-1, 0] == 10.0 and err <= 0.05


def test_history_refusals(gpu, tmp_path):
    km = gpu
    cfg, desc, lay, cell, _ = _cases.vortex(n=64, nblock=32)
    sim = km.Simulation(cfg, desc, lay, path=km.PATH_TASKS)
    sim.set_primitive(cell)
    with pytest.raises(km.KmError, match="no history variables"):
        sim.set_history(tmp_path / "x.hst", 1.0, problem="sedov")
    with pytest.raises(km.KmError):
        sim.set_history(tmp_path / "x.hst", 0.0)
    sim.close()
    cfg2, members, lay_all, cells = _cases.kh_ensemble([64, 64], nblock=32)
    descs = [m[0] for m in members]
    ens = km.Simulation(cfg2, descs[0], lay_all, path=km.PATH_FUSED, descs=descs)
    with pytest.raises(km.KmError, match="single runs"):
        ens.set_history(tmp_path / "y.hst", 1.0, problem="")
    ens.close()
