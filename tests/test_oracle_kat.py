"""Pins of the CPU oracle against the reference's own known-answer tests and against
what the mathematics fixes (closed forms, symmetries, consistency)."""
import math

import numpy as np
import pytest

RECON = {"fog": 0, "plm": 1, "ppm": 2, "wenoz": 3}
LIM = {"minmod": 0, "van_leer": 1, "mc": 2}

# test_reconstruction.cpp:13-14
COEFFS = [10.0, 3.0, 1.0, 0.5]
XC = 20.0


def polynomial_avg(order, idx):
    # test_reconstruction.cpp:16-29 (cell average of sum c_i (x - xc)^i)
    v = COEFFS[0]
    for i in range(1, order + 1):
        v += COEFFS[i] / (i + 1) * ((idx + 0.5 - XC) ** (i + 1) - (idx - 0.5 - XC) ** (i + 1))
    return v


def polynomial(order, x):
    # test_reconstruction.cpp:58-65
    v = COEFFS[0]
    for i in range(1, order + 1):
        v += COEFFS[i] * (x - XC) ** i
    return v


def tanh_avg(idx, steepness):
    # test_reconstruction.cpp:67-81
    n = 100
    dx = 1.0 / n
    x0 = 0.25 * dx
    s = 0.0
    for i in range(n):
        x = idx - 0.5 + (i + 0.5) * dx
        s += math.tanh(steepness * (x - x0))
    return s * dx


# the 7 typed-test traits, test_reconstruction.cpp:146-155, with (order, stencil size)
TRAITS = [("plm", "minmod", 1, 1), ("plm", "mc", 1, 1), ("plm", "van_leer", 1, 1),
          ("ppm", "minmod", 2, 2), ("ppm", "mc", 2, 2), ("ppm", "van_leer", 2, 2),
          ("wenoz", "van_leer", 3, 2)]


@pytest.mark.parametrize("recon,lim,order,size", TRAITS)
def test_reconstruction_polynomial_kat(oracle, recon, lim, order, size):
    """test_reconstruction.cpp:159-170: reproduce the polynomial at +-1/2 to 5e-6."""
    s5 = [polynomial_avg(order, o) for o in range(-2, 3)]
    vm, vp = oracle.reconstruct(RECON[recon], LIM[lim], s5)
    eps = 5.0e-6
    assert abs((vm - polynomial(order, -0.5)) / polynomial(order, -0.5)) < eps
    assert abs((vp - polynomial(order, 0.5)) / polynomial(order, 0.5)) < eps


@pytest.mark.parametrize("recon,lim,order,size", TRAITS)
def test_reconstruction_mirror_symmetry_kat(oracle, recon, lim, order, size):
    """test_reconstruction.cpp:172-194: reversing the stencil swaps vM and vP."""
    fwd = [tanh_avg(o, 10.0) for o in range(-2, 3)]
    rev = fwd[::-1]
    vm1, vp1 = oracle.reconstruct(RECON[recon], LIM[lim], fwd)
    vm2, vp2 = oracle.reconstruct(RECON[recon], LIM[lim], rev)
    # EXPECT_DOUBLE_EQ allows 4 ulps
    assert abs(vm1 - vp2) <= 4 * np.spacing(abs(vm1))
    assert abs(vp1 - vm2) <= 4 * np.spacing(abs(vp1))


def test_fog_is_piecewise_constant(oracle):
    vm, vp = oracle.reconstruct(0, 0, [1.0, 2.0, 3.5, 7.0, 9.0])
    assert vm == 3.5 and vp == 3.5


@pytest.mark.parametrize("lim", ["minmod", "van_leer", "mc"])
def test_plm_limiters_closed_form(oracle, lim):
    """Limited slope of (a, b) on monotone data, and zero at an extremum."""
    s = [0.0, 1.0, 1.5, 3.5, 4.0]  # a = 0.5, b = 2.0
    a, b = 0.5, 2.0
    slope = {"minmod": min(a, b), "mc": min(2 * a, 0.5 * (a + b), 2 * b),
             "van_leer": 2 * a * b / (a + b)}[lim]
    vm, vp = oracle.reconstruct(1, LIM[lim], s)
    assert vm == pytest.approx(1.5 - 0.5 * slope, rel=1e-15)
    assert vp == pytest.approx(1.5 + 0.5 * slope, rel=1e-15)
    vm, vp = oracle.reconstruct(1, LIM[lim], [0.0, 1.0, 2.0, 1.0, 0.0])
    assert vm == 2.0 and vp == 2.0


def test_wenoz_exact_on_quadratic_cell_averages(oracle):
    """WENO-Z is exact on smooth parabolas when the smoothness indicators match."""
    s5 = [polynomial_avg(2, o) for o in range(-2, 3)]
    vm, vp = oracle.reconstruct(3, 0, s5)
    assert vm == pytest.approx(polynomial(2, -0.5), rel=1e-12)
    assert vp == pytest.approx(polynomial(2, 0.5), rel=1e-12)


def test_eos_kat(oracle):
    """test_eos.cpp:55-69: ideal gas gamma=1.4, rho=1, P=1 -> e=2.5; P from e is 1."""
    st = np.array([1.0, 0.0, 0.0, 1.0, 0.0])  # dens, temp, eint, pres, bmod
    st = oracle.eos(1, 1.4, 1.0, st)
    assert abs(st[2] - 1.0 / 0.4) < 1e-14
    st[3] = -1.0
    st = oracle.eos(0, 1.4, 1.0, st)
    assert st[3] == 1.0
    # closed forms: BMOD = gamma P, T = e / cv with cv = gamma kB / Abar
    assert st[4] == pytest.approx(1.4 * st[3], rel=1e-15)
    assert st[1] == pytest.approx(st[2] / (1.4 * 1.380649e-16), rel=1e-15)


def _prim(rho, u, v, w, p, gamma, B=None):
    e = p / ((gamma - 1.0) * rho)
    V = [rho, u, v, w, p, gamma * p, e]
    if B is not None:
        V += list(B)
    return np.array(V)


def _phys_flux(V, d, gamma, mhd=False):
    rho, vel, p = V[0], V[1:4], V[4]
    e = V[6]
    B = V[7:10] if mhd else np.zeros(3)
    E = rho * e + 0.5 * rho * vel @ vel + 0.5 * B @ B
    pt = p + 0.5 * B @ B
    F = [rho * vel[d]]
    for c in range(3):
        F.append(rho * vel[c] * vel[d] - B[d] * B[c] + (pt if c == d else 0.0))
    F.append((E + pt) * vel[d] - (vel @ B) * B[d])
    if mhd:
        F += [vel[d] * B[c] - vel[c] * B[d] for c in range(3)]
    return np.array(F)


@pytest.mark.parametrize("solver", [0, 1])
@pytest.mark.parametrize("d", [0, 1, 2])
def test_riemann_consistency(oracle, solver, d):
    """F(U, U) equals the physical flux (riemann_solver.hpp, both solvers)."""
    rng = np.random.default_rng(7 + d + 3 * solver)
    for _ in range(20):
        V = _prim(rng.uniform(0.5, 2), *rng.uniform(-1, 1, 3), rng.uniform(0.5, 2), 1.4)
        F = oracle.riemann(solver, 0, d, V, V)
        np.testing.assert_allclose(F, _phys_flux(V, d, 1.4), rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("solver", [0, 1])
def test_riemann_mirror_symmetry(oracle, solver):
    """Reflecting the problem about the face flips the odd flux components."""
    rng = np.random.default_rng(11 + solver)
    for _ in range(20):
        VL = _prim(rng.uniform(0.5, 2), *rng.uniform(-1, 1, 3), rng.uniform(0.5, 2), 1.4)
        VR = _prim(rng.uniform(0.5, 2), *rng.uniform(-1, 1, 3), rng.uniform(0.5, 2), 1.4)
        F = oracle.riemann(solver, 0, 0, VL, VR)
        mL, mR = VR.copy(), VL.copy()
        mL[1] *= -1
        mR[1] *= -1
        Fm = oracle.riemann(solver, 0, 0, mL, mR)
        sign = np.array([-1, 1, -1, -1, -1])
        np.testing.assert_allclose(Fm, sign * F, rtol=1e-12, atol=1e-13)


def test_hllc_resolves_contact(oracle):
    """HLLC keeps an isolated contact (equal u and p) exact: upwind flux."""
    for u in (0.3, -0.4):
        VL = _prim(1.0, u, 0.2, -0.1, 0.7, 1.4)
        VR = _prim(0.25, u, -0.5, 0.3, 0.7, 1.4)
        F = oracle.riemann(1, 0, 0, VL, VR)
        up = VL if u > 0 else VR
        np.testing.assert_allclose(F, _phys_flux(up, 0, 1.4), rtol=1e-13, atol=1e-14)
    # a stationary contact carries pressure only
    F = oracle.riemann(1, 0, 0, _prim(1.0, 0.0, 0.0, 0.0, 1.0, 1.4),
                       _prim(0.125, 0.0, 0.0, 0.0, 1.0, 1.4))
    np.testing.assert_allclose(F, [0, 1.0, 0, 0, 0], atol=1e-15)


@pytest.mark.parametrize("solver", [0, 1])
def test_supersonic_upwinding(oracle, solver):
    """All waves moving right: the flux is the left state's physical flux."""
    VL = _prim(1.0, 5.0, 0.1, 0.0, 1.0, 1.4)
    VR = _prim(0.5, 4.5, -0.1, 0.2, 0.8, 1.4)
    F = oracle.riemann(solver, 0, 0, VL, VR)
    # the wave speeds are clamped by +-DBL_MIN, so a zero flux may carry ~1e-308 terms
    np.testing.assert_allclose(F, _phys_flux(VL, 0, 1.4), rtol=1e-13, atol=1e-300)


def test_hll_sod_closed_form(oracle):
    """HLL flux of Sod's states from the Davis wave-speed estimates (Toro 10.5)."""
    g = 1.4
    VL = _prim(1.0, 0.0, 0.0, 0.0, 1.0, g)
    VR = _prim(0.125, 0.0, 0.0, 0.0, 0.1, g)
    cL, cR = math.sqrt(g * 1.0), math.sqrt(g * 0.1 / 0.125)
    sL, sR = min(-cL, -cR), max(cL, cR)
    UL = np.array([1.0, 0, 0, 0, 1.0 / (g - 1)])
    UR = np.array([0.125, 0, 0, 0, 0.1 / (g - 1)])
    FL, FR = _phys_flux(VL, 0, g), _phys_flux(VR, 0, g)
    ref = (sR * FL - sL * FR + sR * sL * (UR - UL)) / (sR - sL)
    np.testing.assert_allclose(oracle.riemann(0, 0, 0, VL, VR), ref, rtol=1e-14, atol=1e-15)


def test_mhd_fast_speed_closed_forms(oracle):
    rho, p, g = 1.3, 0.9, 5.0 / 3.0
    a2 = g * p / rho
    Bn = 0.7
    V = _prim(rho, 0, 0, 0, p, g, B=(Bn, 0.0, 0.0))
    assert oracle.fast_speed(1, 0, V) == pytest.approx(math.sqrt(max(a2, Bn * Bn / rho)), 1e-14)
    V = _prim(rho, 0, 0, 0, p, g, B=(0.0, 0.4, 0.3))
    assert oracle.fast_speed(1, 0, V) == pytest.approx(math.sqrt(a2 + 0.25 / rho), rel=1e-14)
    V = _prim(rho, 0, 0, 0, p, g, B=(0.5, 0.4, 0.3))
    assert oracle.fast_speed(0, 0, V) == pytest.approx(math.sqrt(a2), rel=1e-15)


@pytest.mark.parametrize("solver", [0, 1])
@pytest.mark.parametrize("d", [0, 1, 2])
def test_mhd_riemann_consistency(oracle, solver, d):
    rng = np.random.default_rng(31 + d)
    for _ in range(10):
        V = _prim(rng.uniform(0.5, 2), *rng.uniform(-1, 1, 3), rng.uniform(0.5, 2), 5 / 3,
                  B=rng.uniform(-1, 1, 3))
        F = oracle.riemann(solver, 1, d, V, V)
        np.testing.assert_allclose(F, _phys_flux(V, d, 5 / 3, mhd=True), rtol=1e-12,
                                   atol=1e-13)


def test_prim_cons_roundtrip(oracle):
    V = _prim(1.7, 0.3, -0.2, 0.9, 2.1, 1.4)
    U = oracle.prim2cons(0, V)
    assert U[0] == V[0]
    np.testing.assert_allclose(U[1:4], V[0] * V[1:4], rtol=1e-15)
    assert U[4] == pytest.approx(V[0] * V[6] + 0.5 * V[0] * (V[1:4] @ V[1:4]), rel=1e-15)
