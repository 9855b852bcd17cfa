"""The option surface against the reference's own decks and deck writer, without a GPU:
physics/fluid takes 1t / 3t (physics.cpp:20-21; 3t is refused at dispatch, as only oneT
is instantiated, physics_types.hpp:10, eos.cpp:115-118), eos/mode_init takes dens_pres /
dens_ener / dens_temp (eos.cpp:46-62), the EOS takes gamma and Abar of the first species
(MakeEos, eos.cpp:170-181), and the parthenon/time step keys are read
(kamayan_driver.cpp:28-53). The deck below has the block layout and value spelling of
python/kamayan/code_units (physics.py:26-28, eos.py:25-26, Hydro.py:44-51,
driver.py:13-30): Python's str() of each value, True/False for booleans."""
import ctypes as C

import numpy as np
import pytest

import kamayan_amd as km
from kamayan_amd import Config, Layout, mesh_desc
from kamayan_amd._abi import KmPack

CODE_UNITS_DECK = """
<parthenon/mesh>
nx1 = 128
nx2 = 128
nx3 = 1
nghost = 4

<parthenon/time>
integrator = rk3
nlim = -1
tlim = 10.0
ncycle_out_mesh = -10000
dt_force = -1e+300
dt_factor = 1.5
dt_ceil = 1e+300
dt_min = 0.0
dt_min_cycle_limit = 10
dt_init = 1e+300
dt_init_force = False

<physics>
fluid = 1t
MHD = off

<eos>
mode_init = dens_ener

<hydro>
reconstruction = wenoz
slope_limiter = mc
riemann = hllc
ReconstructionVars = primitive
EMF_averaging = arithmetic
cfl = 0.4

<material/air>
gamma = 1.3
Abar = 2.0

<material/he>
gamma = 1.1

<material>
species = air,he
"""


def test_code_units_deck():
    cfg = Config(deck=CODE_UNITS_DECK).c
    assert (cfg.recon, cfg.limiter, cfg.riemann) == (3, 2, 1)
    assert cfg.nstages == 3 and cfg.cfl == 0.4 and cfg.ndim == 2
    assert cfg.fluid == 0 and cfg.mhd == 0
    assert cfg.mode_init == km._abi.EOS_INIT["dens_ener"]
    # the EOS is the first species' even though <material/air> comes before the species list
    assert cfg.nspecies == 2 and cfg.species0 == b"air"
    assert cfg.gamma == 1.3 and cfg.abar == 2.0
    assert cfg.dt_factor == 1.5 and cfg.dt_init_force == 0
    assert cfg.dt_force == -1e300 and cfg.dt_ceil == 1e300 and cfg.dt_min == 0.0
    assert cfg.dt_min_cycle_limit == 10


def test_reference_problem_deck_defaults():
    # src/problems/isentropic_vortex.in: <eos> mode_init, <material/single> gamma
    deck = """
<eos>
mode_init = dens_pres
<hydro>
cfl = 0.8
reconstruction = wenoz
slope_limiter = mc
<material/single>
gamma = 1.4
"""
    cfg = Config(deck=deck).c
    assert cfg.gamma == 1.4 and cfg.mode_init == 0 and cfg.species0 == b"single"
    assert cfg.fluid == 0 and cfg.dt_factor == 2.0 and cfg.dt_init_force == 1


@pytest.mark.parametrize("key,val", [("physics/fluid", "oneT"), ("physics/fluid", "2t"),
                                     ("eos/mode_init", "dens_entropy"),
                                     ("parthenon/time/dt_init_force", "maybe"),
                                     ("parthenon/time/dt_factor", "fast"),
                                     ("material/single/eos_type", "tabulated")])
def test_values_the_reference_rejects(key, val):
    with pytest.raises(km.KmError):
        Config().set(key, val)


def test_threet_is_refused_at_dispatch():
    cfg = Config(**{"physics/fluid": "3t"})
    assert cfg.c.fluid == 1
    # a host-only check: the dispatcher refuses before anything reaches the device
    pack = KmPack()
    pack.nblocks, pack.ndim, pack.ng, pack.nvar = 0, 1, 4, 12
    pack.nx[0] = pack.nx[1] = pack.nx[2] = 1
    rc = km.lib().km_calculate_fluxes(C.byref(cfg.c), C.byref(pack), None)
    assert rc == km._abi.KM_EINVAL
    km.lib().km_last_error.restype = C.c_char_p
    assert b"ThreeT" in km.lib().km_last_error()


def _eos_pack(oracle, mode, gamma=1.4, abar=2.0):
    cfg = Config(**{"eos/mode_init": mode, "material/single/gamma": gamma,
                    "material/single/Abar": abar})
    cfg.c.ndim = 1
    lay = Layout(1, 2, (8, 1, 1), (1, 1, 1))
    desc = mesh_desc(lay)
    host = oracle.HostPack(lay, km.nvar(0), blocks=oracle.block_table(desc, 1))
    rng = np.random.default_rng(3)
    c = host.cell
    c[:, km.slot("DENS")] = rng.uniform(0.5, 2.0, c[:, 0].shape)
    c[:, km.slot("VEL0")] = rng.uniform(-1, 1, c[:, 0].shape)
    return cfg, lay, host


def test_mode_init_dens_ener(oracle):
    gamma = 1.4
    cfg, lay, host = _eos_pack(oracle, "dens_ener", gamma)
    c = host.cell
    e = np.random.default_rng(4).uniform(1.0, 3.0, c[:, 0].shape)
    c[:, km.slot("EINT")] = e
    oracle.call("km_cpu_prepare_conserved", C.byref(cfg.c), C.byref(host.pack), None)
    sl = (slice(None),) + lay.interior()
    rho, v = c[:, 0][sl], c[:, km.slot("VEL0")][sl]
    np.testing.assert_allclose(c[:, km.slot("PRES")][sl], (gamma - 1) * rho * e[sl], rtol=1e-15)
    np.testing.assert_allclose(c[:, km.slot("BMOD")][sl], gamma * (gamma - 1) * rho * e[sl],
                               rtol=1e-15)
    np.testing.assert_allclose(c[:, km.slot("ENER")][sl], rho * e[sl] + 0.5 * rho * v * v,
                               rtol=1e-15)


def test_mode_init_dens_temp(oracle):
    gamma, abar = 1.4, 2.0
    cfg, lay, host = _eos_pack(oracle, "dens_temp", gamma, abar)
    c = host.cell
    cv = gamma * 1.380649e-16 / abar  # equation_of_state.hpp:44-46
    T = np.random.default_rng(5).uniform(1e15, 3e16, c[:, 0].shape)
    c[:, km.slot("TEMP")] = T
    oracle.call("km_cpu_prepare_conserved", C.byref(cfg.c), C.byref(host.pack), None)
    sl = (slice(None),) + lay.interior()
    rho = c[:, 0][sl]
    np.testing.assert_allclose(c[:, km.slot("EINT")][sl], cv * T[sl], rtol=1e-15)
    np.testing.assert_allclose(c[:, km.slot("PRES")][sl], (gamma - 1) * rho * cv * T[sl],
                               rtol=1e-15)
    # and dens_pres recovers the same temperature from that pressure
    cfg2, lay2, host2 = _eos_pack(oracle, "dens_pres", gamma, abar)
    host2.cell[...] = c
    oracle.call("km_cpu_prepare_conserved", C.byref(cfg2.c), C.byref(host2.pack), None)
    np.testing.assert_allclose(host2.cell[:, km.slot("TEMP")][sl], T[sl], rtol=1e-14)
