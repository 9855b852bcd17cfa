// km_oracle.cpp -- CPU FP64 restatement of kamayan's hydro/MHD flux-and-update path.
//
// TEST INFRASTRUCTURE ONLY. This library is the checker for libkm_hydro.so: only the
// test suite, the smoke check and the CPU-baseline leg of the benchmark load it. It is
// never the thing measured on the GPU and the product never calls it.
//
// It restates the reference algorithm from its source, loop for loop, in plain C++ with
// OpenMP over blocks; every function cites the reference lines it follows (paths
// relative to the kamayan tree). Arithmetic is written in the reference's own
// operation order and the file is compiled with -ffp-contract=off, so results are the
// reference's rounding of each operation, with two documented exceptions:
//   * WENO-Z squares use x*x where the reference calls Kokkos::pow(x, 2)
//     (reconstruction.hpp:109-135);
//   * the ideal-gas closure restates singularity-eos IdealGas (version unrecorded):
//     P = (g-1) rho e, BMOD = max(0, g (g-1) rho e), T = max(0, e / cv).
//
// Pins (tests/test_oracle_*.py): reconstruction polynomial KAT and mirror-symmetry KAT
// (test_reconstruction.cpp:13-194), EOS KAT (test_eos.cpp:55-85), Riemann consistency /
// symmetry / contact / Toro star states, conservation, divB, vortex error ceilings
// (reconstruction.py:129-135) and convergence slope (isentropic_vortex.py:60).
// Ghost fill, dt control and RK coefficient tables belong to Parthenon, which is not
// vendored: parity unpinned there beyond the invariants listed in the tests.

#include <algorithm>
#include <array>
#include <functional>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "../include/km_hydro.h"

namespace {

constexpr double kTiny = std::numeric_limits<double>::min();  // riemann_solver.hpp:32
constexpr double kBig = std::numeric_limits<double>::max();

// Kokkos::min / Kokkos::max semantics (== std::min / std::max).
inline double kmin(double a, double b) { return (b < a) ? b : a; }
inline double kmax(double a, double b) { return (a < b) ? b : a; }

// ---------------------------------------------------------------------------------
// Layout
// ---------------------------------------------------------------------------------
struct Shape {
  int ndim, ng, nx[3], E[3], S[3];
  int is, ie, js, je, ks, ke;
  explicit Shape(const km_pack_t &p) {
    ndim = p.ndim;
    ng = p.ng;
    for (int d = 0; d < 3; ++d) nx[d] = p.nx[d];
    E[0] = nx[0] + 2 * ng;
    E[1] = ndim > 1 ? nx[1] + 2 * ng : 1;
    E[2] = ndim > 2 ? nx[2] + 2 * ng : 1;
    S[0] = E[0] + 1;
    S[1] = E[1] + (ndim > 1 ? 1 : 0);
    S[2] = E[2] + (ndim > 2 ? 1 : 0);
    is = ng;
    ie = ng + nx[0] - 1;
    js = ndim > 1 ? ng : 0;
    je = ndim > 1 ? ng + nx[1] - 1 : 0;
    ks = ndim > 2 ? ng : 0;
    ke = ndim > 2 ? ng + nx[2] - 1 : 0;
  }
  long c(int k, int j, int i) const { return ((long)k * E[1] + j) * E[0] + i; }
  long s(int k, int j, int i) const { return ((long)k * S[1] + j) * S[0] + i; }
  long ncell() const { return (long)E[0] * E[1] * E[2]; }
  long nstag() const { return (long)S[0] * S[1] * S[2]; }
};

// recon-variable order (hydro_types.hpp:61-74,110): DENS, VELOCITY[3], PRES, BMOD, EINT,
// MAGC[3]
enum { R_DENS = 0, R_VEL = 1, R_PRES = 4, R_BMOD = 5, R_EINT = 6, R_MAGC = 7 };
// conserved order: DENS, MOMENTUM[3], ENER, MAGC[3]
enum { C_DENS = 0, C_MOM = 1, C_ENER = 4, C_MAGC = 5 };

inline int nrecon(int mhd) { return mhd == KM_MHD_CT ? 10 : 7; }
inline int ncons(int mhd) { return KM_NCONS(mhd); }
// slot in the cell pack of recon variable r
inline int recon_slot(int r, int mhd) {
  if (r == R_DENS) return KM_DENS;
  if (r < R_PRES) return KM_VEL0(mhd) + (r - R_VEL);
  if (r == R_PRES) return KM_PRES(mhd);
  if (r == R_BMOD) return KM_BMOD(mhd);
  if (r == R_EINT) return KM_EINT(mhd);
  return KM_MAGC0 + (r - R_MAGC);
}

// ---------------------------------------------------------------------------------
// Reconstruction (reconstruction.hpp)
// ---------------------------------------------------------------------------------
// LimitedSlope, reconstruction.hpp:24-39
double LimitedSlope(int lim, double a, double b) {
  if (lim == KM_LIM_MC) {
    return (std::copysign(1., a) + std::copysign(1., b)) *
           kmin(std::fabs(a), kmin(.25 * std::fabs(a + b), std::fabs(b)));
  } else if (lim == KM_LIM_VAN_LEER) {
    if (a * b > 0.) return 2. * a * b / (a + b);
    return 0.;
  } else if (lim == KM_LIM_MINMOD) {
    return 0.5 * (std::copysign(1.0, a) + std::copysign(1.0, b)) *
           kmin(std::fabs(a), std::fabs(b));
  }
  return 0.;
}

// A 1-D stencil view: s(o) = p[o * stride]
struct Stencil {
  const double *p;
  long stride;
  double operator()(int o) const { return p[o * stride]; }
};

// Slope, reconstruction.hpp:41-46
double Slope(int lim, int idx, const Stencil &s) {
  return LimitedSlope(lim, s(idx + 1) - s(idx), s(idx) - s(idx - 1));
}

// Reconstruct, reconstruction.hpp:16-142
void Reconstruct(int recon, int lim, const Stencil &s, double &vM, double &vP) {
  if (recon == KM_RECON_FOG) {  // :16-22
    vM = s(0);
    vP = s(0);
  } else if (recon == KM_RECON_PLM) {  // :48-59
    const double dvL = s(0) - s(-1);
    const double dvR = s(1) - s(0);
    const double del = LimitedSlope(lim, dvL, dvR);
    vM = s(0) - 0.5 * del;
    vP = s(0) + 0.5 * del;
  } else if (recon == KM_RECON_PPM) {  // :61-88
    const double dv_p = Slope(lim, 1, s);
    const double dv_0 = Slope(lim, 0, s);
    const double dv_m = Slope(lim, -1, s);
    vM = 0.5 * (s(-1) + s(0)) - 1. / 6. * (dv_0 - dv_m);
    vP = 0.5 * (s(0) + s(1)) - 1. / 6. * (dv_p - dv_0);
    if ((vP - s(0)) * (s(0) - vM) <= 0.) {
      vM = s(0);
      vP = s(0);
    }
    if (-(vP - vM) * (vP - vM) > 6. * (vP - vM) * (s(0) - 0.5 * (vP + vM))) {
      vP = 3.0 * s(0) - 2. * vM;
    }
    if ((vP - vM) * (vP - vM) < 6. * (vP - vM) * (s(0) - 0.5 * (vP + vM))) {
      vM = 3.0 * s(0) - 2. * vP;
    }
  } else if (recon == KM_RECON_WENOZ) {  // :90-142
    auto eno = [&](int pm, double e[3]) {
      e[0] = 1. / 6. * (-s(pm * 2) + 5. * s(pm * 1) + 2. * s(0));
      e[1] = 1. / 6. * (2. * s(pm * 1) + 5. * s(0) - s(-pm * 1));
      e[2] = 1. / 6. * (11. * s(0) - 7. * s(-pm * 1) + 2. * s(-pm * 2));
    };
    double eno_plus[3], eno_minus[3];
    eno(1, eno_plus);
    eno(-1, eno_minus);
    auto sq = [](double x) { return x * x; };  // Kokkos::pow(x, 2) in the reference
    const double b0 = 13. / 12. * sq(s(-2) - 2. * s(-1) + s(0)) +
                      0.25 * sq(s(-2) - 4. * s(-1) + 3. * s(0));
    const double b1 =
        13. / 12. * sq(s(-1) - 2. * s(0) + s(1)) + 0.25 * sq(s(-1) - s(1));
    const double b2 = 13. / 12. * sq(s(0) - 2. * s(1) + s(2)) +
                      0.25 * sq(s(2) - 4. * s(1) + 3. * s(0));
    const double beta[3] = {b0, b1, b2};
    constexpr double eps = 1.e-36;
    auto weno = [&](int pm, const double e[3]) {
      double w[3] = {3., 6., 1.};
      w[0] *= (1.0 + sq(std::fabs(beta[2] - beta[0]) / (eps + beta[1 + pm])));
      w[1] *= (1.0 + sq(std::fabs(beta[2] - beta[0]) / (eps + beta[1])));
      w[2] *= (1.0 + sq(std::fabs(beta[2] - beta[0]) / (eps + beta[1 - pm])));
      const double norm = w[0] + w[1] + w[2];
      return (w[0] * e[0] + w[1] * e[1] + w[2] * e[2]) * 1. / norm;
    };
    vM = weno(-1, eno_minus);
    vP = weno(1, eno_plus);
  }
}

// ---------------------------------------------------------------------------------
// State conversions (primconsflux.hpp:21-138)
// ---------------------------------------------------------------------------------
double TotalPres(int mhd, const double *V) {  // :21-30
  double pres = V[R_PRES];
  if (mhd != KM_MHD_OFF) {
    pres += 0.5 * (V[R_MAGC] * V[R_MAGC] + V[R_MAGC + 1] * V[R_MAGC + 1] +
                   V[R_MAGC + 2] * V[R_MAGC + 2]);
  }
  return pres;
}

double FastSpeed(int mhd, int dir1, const double *V) {  // :31-53
  const double idens = 1. / V[R_DENS];
  const double a2 = idens * V[R_BMOD];
  double cfast2;
  if (mhd == KM_MHD_OFF) {
    cfast2 = a2;
  } else {
    const double bb2 = V[R_MAGC + dir1] * V[R_MAGC + dir1] * idens;
    double b2 = 0.0;
    for (int dir = 0; dir < 3; dir++) b2 += V[R_MAGC + dir] * V[R_MAGC + dir];
    b2 *= idens;
    cfast2 = 0.5 * ((a2 + b2) + std::sqrt((a2 - b2) * (a2 - b2) + 4. * a2 * (b2 - bb2)));
  }
  return std::sqrt(cfast2);
}

void Prim2Cons(int mhd, const double *V, double *U) {  // :55-74
  U[C_DENS] = V[R_DENS];
  double emag = 0.;
  double ekin = 0.;
  for (int dir = 0; dir < 3; dir++) {
    U[C_MOM + dir] = V[R_DENS] * V[R_VEL + dir];
    ekin += V[R_VEL + dir] * V[R_VEL + dir];
    if (mhd != KM_MHD_OFF) {
      U[C_MAGC + dir] = V[R_MAGC + dir];
      emag += V[R_MAGC + dir] * V[R_MAGC + dir];
    }
  }
  const double eint = V[R_EINT] * V[R_DENS];
  ekin *= 0.5 * V[R_DENS];
  emag *= 0.5;
  U[C_ENER] = eint + ekin + emag;
}

void Prim2Flux(int mhd, int dir1, const double *V, double *F) {  // :100-138
  const int dir2 = (dir1 + 1) % 3;
  const int dir3 = (dir1 + 2) % 3;
  F[C_DENS] = V[R_DENS] * V[R_VEL + dir1];
  F[C_MOM + dir1] = F[C_DENS] * V[R_VEL + dir1];
  F[C_MOM + dir2] = F[C_DENS] * V[R_VEL + dir2];
  F[C_MOM + dir3] = F[C_DENS] * V[R_VEL + dir3];
  double ptot = V[R_PRES];
  double B2 = 0.;
  double uB = 0.;
  double ekin = 0.;
  for (int dir = 0; dir < 3; dir++) {
    F[C_MOM + dir] = F[C_DENS] * V[R_VEL + dir];
    ekin += V[R_VEL + dir] * V[R_VEL + dir];
    if (mhd != KM_MHD_OFF) {
      B2 += V[R_MAGC + dir] * V[R_MAGC + dir];
      uB += V[R_VEL + dir] * V[R_MAGC + dir];
      F[C_MOM + dir] -= V[R_MAGC + dir1] * V[R_MAGC + dir];
    }
  }
  ekin *= 0.5 * V[R_DENS];
  const double emag = 0.5 * B2;
  ptot += 0.5 * B2;
  const double etot = V[R_EINT] * V[R_DENS] + ekin + emag;
  F[C_MOM + dir1] += ptot;
  F[C_ENER] = (etot + ptot) * V[R_VEL + dir1];
  if (mhd != KM_MHD_OFF) {
    F[C_ENER] -= uB * V[R_MAGC + dir1];
    F[C_MAGC + dir1] = 0.;
    F[C_MAGC + dir2] = V[R_VEL + dir1] * V[R_MAGC + dir2] - V[R_VEL + dir2] * V[R_MAGC + dir1];
    F[C_MAGC + dir3] = V[R_VEL + dir1] * V[R_MAGC + dir3] - V[R_VEL + dir3] * V[R_MAGC + dir1];
  }
}

// ---------------------------------------------------------------------------------
// Riemann solvers (riemann_solver.hpp)
// ---------------------------------------------------------------------------------
void RiemannHLL(int mhd, int dir1, const double *vL, const double *vR, double *flux) {
  // riemann_solver.hpp:20-57
  const double cfL = FastSpeed(mhd, dir1, vL);
  const double cfR = FastSpeed(mhd, dir1, vR);
  const double tiny = kTiny;
  const double sL = kmin(-tiny, kmin(vL[R_VEL + dir1] - cfL, vR[R_VEL + dir1] - cfR));
  const double sR = kmax(tiny, kmax(vL[R_VEL + dir1] + cfL, vR[R_VEL + dir1] + cfR));
  const double sRmsLi = 1.0 / (sR - sL);
  double UL[8], UR[8], FL[8], FR[8];
  Prim2Cons(mhd, vL, UL);
  Prim2Cons(mhd, vR, UR);
  Prim2Flux(mhd, dir1, vR, FR);
  Prim2Flux(mhd, dir1, vL, FL);
  const int nc = ncons(mhd);
  for (int v = 0; v < nc; ++v)
    flux[v] = sRmsLi * (sR * FL[v] - sL * FR[v] + sR * sL * (UR[v] - UL[v]));
}

void RiemannHLLC(int mhd, int dir1, const double *vL, const double *vR, double *flux) {
  // riemann_solver.hpp:59-149
  const int dir2 = (dir1 + 1) % 3;
  const int dir3 = (dir1 + 2) % 3;
  const double cfL = FastSpeed(mhd, dir1, vL);
  const double cfR = FastSpeed(mhd, dir1, vR);
  const double tiny = kTiny;
  const double sL = kmin(-tiny, kmin(vL[R_VEL + dir1] - cfL, vR[R_VEL + dir1] - cfR));
  const double sR = kmax(tiny, kmax(vL[R_VEL + dir1] + cfL, vR[R_VEL + dir1] + cfR));
  double UL[8], UR[8], FL[8], FR[8];
  Prim2Cons(mhd, vL, UL);
  Prim2Cons(mhd, vR, UR);
  Prim2Flux(mhd, dir1, vR, FR);
  Prim2Flux(mhd, dir1, vL, FL);
  const double total_presL = TotalPres(mhd, vL);
  const double total_presR = TotalPres(mhd, vR);
  const double uL = vL[R_VEL + dir1], uR = vR[R_VEL + dir1];
  double ustar = total_presR - total_presL + UL[C_MOM + dir1] * (sL - uL) -
                 UR[C_MOM + dir1] * (sR - uR);
  ustar = ustar * 1. / (vL[R_DENS] * (sL - uL) - vR[R_DENS] * (sR - uR));
  const double pstar = 0.5 * (total_presL + total_presR +
                              vL[R_DENS] * (sL - uL) * (ustar - uL) +
                              vR[R_DENS] * (sR - uR) * (ustar - uR));
  auto hllc_state = [&](double S, const double *U, double Pu, double *Ustar) {
    const double susi = 1. / (S - ustar + tiny);
    Ustar[C_DENS] = susi * (S * U[C_DENS] - U[C_MOM + dir1]);
    Ustar[C_MOM + dir1] = Ustar[C_DENS] * ustar;
    Ustar[C_MOM + dir2] = U[C_MOM + dir2] * Ustar[C_DENS] / U[C_DENS];
    Ustar[C_MOM + dir3] = U[C_MOM + dir3] * Ustar[C_DENS] / U[C_DENS];
    Ustar[C_ENER] = U[C_ENER] * Ustar[C_DENS] / U[C_DENS] + susi * (pstar * ustar - Pu);
    if (mhd != KM_MHD_OFF) {  // :111-133
      const double sRsLi = 1. / (sR - sL);
      auto hll = [&](int v) { return sRsLi * (sR * UR[v] - sL * UL[v] + FL[v] - FR[v]); };
      Ustar[C_MAGC + dir1] = hll(C_MAGC + dir1);
      Ustar[C_MAGC + dir2] = hll(C_MAGC + dir2);
      Ustar[C_MAGC + dir3] = hll(C_MAGC + dir3);
      Ustar[C_MOM + dir2] -=
          (Ustar[C_MAGC + dir1] * Ustar[C_MAGC + dir2] - U[C_MAGC + dir1] * U[C_MAGC + dir2]) *
          susi;
      Ustar[C_MOM + dir3] -=
          (Ustar[C_MAGC + dir1] * Ustar[C_MAGC + dir3] - U[C_MAGC + dir1] * U[C_MAGC + dir3]) *
          susi;
      Ustar[C_ENER] -=
          susi * (Ustar[C_MAGC + dir1] *
                      (Ustar[C_MAGC + dir1] * hll(C_MOM + dir1) +
                       Ustar[C_MAGC + dir2] * hll(C_MOM + dir2) +
                       Ustar[C_MAGC + dir3] * hll(C_MOM + dir3)) /
                      hll(C_DENS) -
                  U[C_MAGC + dir1] *
                      (U[C_MAGC + dir1] * U[C_MOM + dir1] + U[C_MAGC + dir2] * U[C_MOM + dir2] +
                       U[C_MAGC + dir3] * U[C_MOM + dir3]) /
                      U[C_DENS]);
    }
  };
  double UstarL[8], UstarR[8];
  hllc_state(sL, UL, total_presL * uL, UstarL);
  hllc_state(sR, UR, total_presR * uR, UstarR);
  const double biasL = -kmin(-tiny, std::copysign(1., ustar));
  const double biasR = kmax(tiny, std::copysign(1., ustar));
  const int nc = ncons(mhd);
  for (int v = 0; v < nc; ++v)
    flux[v] = biasR * (FL[v] + sL * (UstarL[v] - UL[v])) +
              biasL * (FR[v] + sR * (UstarR[v] - UR[v]));
}

void RiemannFlux(int riemann, int mhd, int dir1, const double *vL, const double *vR,
                 double *flux) {
  if (riemann == KM_RIEMANN_HLLC)
    RiemannHLLC(mhd, dir1, vL, vR, flux);
  else
    RiemannHLL(mhd, dir1, vL, vR, flux);
}

// ---------------------------------------------------------------------------------
// Ideal-gas closure: singularity-eos IdealGas through EquationOfState<gamma>
// (equation_of_state.hpp:40-64; eos_singularity.hpp:17-36)
// ---------------------------------------------------------------------------------
struct IdealGas {
  double gm1, cv;
  IdealGas(double gamma, double abar) {
    constexpr double kboltz = 1.380649e-16;  // equation_of_state.hpp:44
    cv = gamma * kboltz / abar;
    gm1 = gamma - 1.0;
  }
  double P(double rho, double sie) const { return gm1 * rho * sie; }
  double T(double rho, double sie) const { return kmax(0.0, sie / cv); }
  double B(double rho, double sie) const { return kmax(0.0, (gm1 + 1) * gm1 * rho * sie); }
  double sie_from_PT(double rho, double press) const { return press / (gm1 * rho); }
  double sie_from_T(double temp) const { return kmax(0.0, cv * temp); }
};

// ---------------------------------------------------------------------------------
// Pack access
// ---------------------------------------------------------------------------------
inline double *cellp(const km_pack_t &p, int b, int v) { return p.cell[(long)b * p.nvar + v]; }
inline double *fluxp(const km_pack_t &p, int d, int b, int v) {
  return p.flux[d][(long)b * p.nflux + v];
}
inline double *facep(const km_pack_t &p, int d, int b) { return p.face[d][b]; }
inline double *edgep(const km_pack_t &p, int d, int b) { return p.edge[d][b]; }

struct Range {
  int s, e;
};

// ---------------------------------------------------------------------------------
// CalculateFluxesNested::dispatch, hydro_add_flux_tasks.cpp:36-264 (the scratch-var
// strategy, :266-385, evaluates the same faces from the same states)
// ---------------------------------------------------------------------------------
// ---- cylindrical (r, z, phi) geometry: grid::Coordinates<cylindrical>
// (geometry.hpp:35-167) on Parthenon's uniform coordinates, whose x1 faces sit at
// (xmin - is dx) + i dx; Dx along phi is 2 pi (:37-40)
constexpr double kPhiSpan = 2.0 * 3.141592653589793;
inline bool Cyl(const km_cfg_t &cfg) { return cfg.geometry == KM_GEOM_CYLINDRICAL; }
inline double RFace(const km_block_t &blk, const Shape &sh, int i) {
  return (blk.xmin[0] - sh.is * blk.dx[0]) + i * blk.dx[0];
}
inline double RCentre(const km_block_t &blk, const Shape &sh, int i) {  // Xi, :67-70
  return (blk.xmin[0] - sh.is * blk.dx[0]) + (i + 0.5) * blk.dx[0];
}
inline double RCentroid(const km_block_t &blk, const Shape &sh, int i) {  // Xc, :55-63
  const double r0 = RFace(blk, sh, i), r1 = RFace(blk, sh, i + 1);
  const double r0sq = r0 * r0, r1sq = r1 * r1;
  return (2.0 / 3.0) * (r1sq * std::abs(r1) - r0sq * std::abs(r0)) / (r1sq - r0sq);
}
// FaceArea<axis> at the low face of cell i (:124-137) and CellVolume (:154-160)
inline double Area(const km_cfg_t &cfg, const km_block_t &blk, const Shape &sh, int dir, int i) {
  if (!Cyl(cfg) || dir == 2) return blk.area[dir];
  if (dir == 0) return RFace(blk, sh, i) * blk.dx[1] * kPhiSpan;
  const double rp = RFace(blk, sh, i + 1), rm = RFace(blk, sh, i);
  return 0.5 * (rp * rp - rm * rm) * kPhiSpan;
}
inline double Volume(const km_cfg_t &cfg, const km_block_t &blk, const Shape &sh, int i) {
  if (!Cyl(cfg)) return blk.vol;
  const double rp = RFace(blk, sh, i + 1), rm = RFace(blk, sh, i);
  return 0.5 * kPhiSpan * (rp * rp - rm * rm) * blk.dx[1];
}
// utils::Ratio, robust.hpp:29-32
inline double Ratio(double a, double b) {
  const int sgn = (0.0 <= b) - (b < 0.0);
  return a / (b + sgn * (10.0 * std::numeric_limits<double>::min()));
}

void CalculateFluxes(const km_cfg_t &cfg, const km_pack_t &u) {
  const Shape sh(u);
  const int mhd = cfg.mhd;
  const int nr = nrecon(mhd);
  const int ndim = sh.ndim;
  Range ib{sh.is, sh.ie}, jb{sh.js, sh.je}, kb{sh.ks, sh.ke};
  if (mhd == KM_MHD_CT) {  // :59-70
    const int k1d = ndim > 1 ? 1 : 0, k2d = ndim > 1 ? 1 : 0, k3d = ndim > 2 ? 1 : 0;
    ib.s -= k1d;
    ib.e += k1d;
    jb.s -= k2d;
    jb.e += k2d;
    kb.s -= k3d;
    kb.e += k3d;
  }
  const int nxb = sh.E[0];
  const int nb = u.nblocks;
  // reconstruct_vars = Reconstruct + MassScalars (:40-46): with more than one species the
  // mass fractions follow the primitives in the pencils, at offset nr; flux slot nc + s
  const int nsp = cfg.nspecies > 1 ? cfg.nspecies : 0;
  const int nrt = nr + nsp, nc = ncons(mhd);
  std::vector<int> slot(nrt);
  for (int r = 0; r < nr; ++r) slot[r] = recon_slot(r, mhd);
  for (int sp = 0; sp < nsp; ++sp) slot[nr + sp] = KM_MFRAC(mhd, sp);

#pragma omp parallel
  {
    std::vector<double> bufM(nrt * nxb), bufP(nrt * nxb), bufMP(nrt * nxb);
    double vL[10], vR[10], F[8];
    // X-direction, :79-136
#pragma omp for collapse(3) schedule(static)
    for (int b = 0; b < nb; ++b)
      for (int k = kb.s; k <= kb.e; ++k)
        for (int j = jb.s; j <= jb.e; ++j) {
          double *vM = bufM.data(), *vP = bufP.data();
          for (int var = 0; var < nrt; ++var) {
            const double *q = cellp(u, b, slot[var]);
            for (int i = ib.s - 1; i <= ib.e + 1; ++i) {
              Stencil st{q + sh.c(k, j, i), 1};
              Reconstruct(cfg.recon, cfg.limiter, st, vM[var * nxb + i], vP[var * nxb + i]);
            }
          }
          for (int i = ib.s; i <= ib.e + 1; ++i) {
            for (int var = 0; var < nr; ++var) {
              vL[var] = vP[var * nxb + i - 1];
              vR[var] = vM[var * nxb + i];
            }
            if (mhd == KM_MHD_CT) {
              const double bf = facep(u, 0, b)[sh.s(k, j, i)];
              vL[R_MAGC + 0] = bf;
              vR[R_MAGC + 0] = bf;
            }
            RiemannFlux(cfg.riemann, mhd, 0, vL, vR, F);
            if (Cyl(cfg)) {  // :108-117, angular momentum flux
              const double rf = RFace(u.blocks[b], sh, i);
              F[KM_MOM0 + 2] *= rf;
              if (mhd == KM_MHD_CT) F[KM_MAGC0 + 2] *= Ratio(1.0, rf);
            }
            for (int v = 0; v < ncons(mhd); ++v) fluxp(u, 0, b, v)[sh.c(k, j, i)] = F[v];
          }
          // mass scalars, :121-135: upwind mass fraction times the mass flux
          for (int sp = 0; sp < nsp; ++sp)
            for (int i = ib.s; i <= ib.e + 1; ++i) {
              const long c = sh.c(k, j, i);
              const double rf = fluxp(u, 0, b, KM_DENS)[c];
              fluxp(u, 0, b, nc + sp)[c] = rf > 0.0 ? rf * vP[(nr + sp) * nxb + i - 1]
                                                    : rf * vM[(nr + sp) * nxb + i];
            }
        }
    // Y-direction, :138-203
    if (ndim > 1) {
#pragma omp for collapse(2) schedule(static)
      for (int b = 0; b < nb; ++b)
        for (int k = kb.s; k <= kb.e; ++k) {
          double *vMP = bufMP.data(), *vM = bufM.data(), *vP = bufP.data();
          for (int j = jb.s - 1; j <= jb.e + 1; ++j) {
            for (int var = 0; var < nrt; ++var) {
              const double *q = cellp(u, b, slot[var]);
              for (int i = ib.s; i <= ib.e; ++i) {
                Stencil st{q + sh.c(k, j, i), sh.E[0]};
                Reconstruct(cfg.recon, cfg.limiter, st, vM[var * nxb + i], vP[var * nxb + i]);
              }
            }
            if (j > jb.s - 1) {
              for (int i = ib.s; i <= ib.e; ++i) {
                for (int var = 0; var < nr; ++var) {
                  vL[var] = vMP[var * nxb + i];
                  vR[var] = vM[var * nxb + i];
                }
                if (mhd == KM_MHD_CT) {
                  const double bf = facep(u, 1, b)[sh.s(k, j, i)];
                  vL[R_MAGC + 1] = bf;
                  vR[R_MAGC + 1] = bf;
                }
                RiemannFlux(cfg.riemann, mhd, 1, vL, vR, F);
                if (Cyl(cfg) && mhd == KM_MHD_CT) {  // :171-178, divided by the face's x2
                  const km_block_t &bk = u.blocks[b];
                  const double zf = (bk.xmin[1] - sh.js * bk.dx[1]) + j * bk.dx[1];
                  F[KM_MAGC0 + 2] *= Ratio(1.0, zf);
                }
                for (int v = 0; v < ncons(mhd); ++v) fluxp(u, 1, b, v)[sh.c(k, j, i)] = F[v];
              }
              // mass scalars, :180-197: the loop ends at s < upper bound, so the last
              // species has no x2 flux
              for (int sp = 0; sp + 1 < nsp; ++sp)
                for (int i = ib.s; i <= ib.e; ++i) {
                  const long c = sh.c(k, j, i);
                  const double rf = fluxp(u, 1, b, KM_DENS)[c];
                  fluxp(u, 1, b, nc + sp)[c] = rf > 0.0 ? rf * vMP[(nr + sp) * nxb + i]
                                                        : rf * vM[(nr + sp) * nxb + i];
                }
            }
            std::swap(vMP, vP);
          }
        }
    }
    // Z-direction, :205-261
    if (ndim > 2) {
#pragma omp for collapse(2) schedule(static)
      for (int b = 0; b < nb; ++b)
        for (int j = jb.s; j <= jb.e; ++j) {
          double *vMP = bufMP.data(), *vM = bufM.data(), *vP = bufP.data();
          for (int k = kb.s - 1; k <= kb.e + 1; ++k) {
            for (int var = 0; var < nrt; ++var) {
              const double *q = cellp(u, b, slot[var]);
              for (int i = ib.s; i <= ib.e; ++i) {
                Stencil st{q + sh.c(k, j, i), (long)sh.E[0] * sh.E[1]};
                Reconstruct(cfg.recon, cfg.limiter, st, vM[var * nxb + i], vP[var * nxb + i]);
              }
            }
            if (k > kb.s - 1) {
              for (int i = ib.s; i <= ib.e; ++i) {
                for (int var = 0; var < nr; ++var) {
                  vL[var] = vMP[var * nxb + i];
                  vR[var] = vM[var * nxb + i];
                }
                if (mhd == KM_MHD_CT) {
                  const double bf = facep(u, 2, b)[sh.s(k, j, i)];
                  vL[R_MAGC + 2] = bf;
                  vR[R_MAGC + 2] = bf;
                }
                RiemannFlux(cfg.riemann, mhd, 2, vL, vR, F);
                for (int v = 0; v < ncons(mhd); ++v) fluxp(u, 2, b, v)[sh.c(k, j, i)] = F[v];
              }
              // mass scalars, :239-256, again without the last species. The reference's
              // i loop runs to ib.e + 1, one x1 ghost column whose x3 flux no interior
              // update reads; it is left out here.
              for (int sp = 0; sp + 1 < nsp; ++sp)
                for (int i = ib.s; i <= ib.e; ++i) {
                  const long c = sh.c(k, j, i);
                  const double rf = fluxp(u, 2, b, KM_DENS)[c];
                  fluxp(u, 2, b, nc + sp)[c] = rf > 0.0 ? rf * vMP[(nr + sp) * nxb + i]
                                                        : rf * vM[(nr + sp) * nxb + i];
                }
            }
            std::swap(vMP, vP);
          }
        }
    }
  }
}

// ---------------------------------------------------------------------------------
// CalculateEMF, hydro_add_flux_tasks.cpp:387-442 (arithmetic averaging, Cartesian).
// Each edge family is evaluated over its own interior range (the reference loops E1/E2
// over E3's range, :420-437, which leaves E1/E2 on the upper x3 face unset in 3-D).
// ---------------------------------------------------------------------------------
void CalculateEMF(const km_cfg_t &cfg, const km_pack_t &u) {
  if (cfg.mhd != KM_MHD_CT || u.ndim < 2) return;
  const Shape sh(u);
  const int ndim = sh.ndim;
  const int bx = C_MAGC + 0, by = C_MAGC + 1, bz = C_MAGC + 2;
  const long sj = sh.E[0], sk = (long)sh.E[0] * sh.E[1];
  const double esign = Cyl(cfg) ? -1.0 : 1.0;  // GetEdgeEMF, :398
#pragma omp parallel for schedule(static)
  for (int b = 0; b < u.nblocks; ++b) {
    const double *fxBy = fluxp(u, 0, b, by), *fxBz = fluxp(u, 0, b, bz);
    const double *fyBx = fluxp(u, 1, b, bx), *fyBz = fluxp(u, 1, b, bz);
    const double *fzBx = ndim > 2 ? fluxp(u, 2, b, bx) : nullptr;
    const double *fzBy = ndim > 2 ? fluxp(u, 2, b, by) : nullptr;
    // E3 at (i-1/2, j-1/2, k): i in [is, ie+1], j in [js, je+1], k in [ks, ke]
    double *e3 = edgep(u, 2, b);
    for (int k = sh.ks; k <= sh.ke; ++k)
      for (int j = sh.js; j <= sh.je + 1; ++j)
        for (int i = sh.is; i <= sh.ie + 1; ++i) {
          const long c = sh.c(k, j, i);
          e3[sh.s(k, j, i)] =
              esign * 0.25 * (fyBx[c - 1] + fyBx[c] - fxBy[c - sj] - fxBy[c]);
        }
    if (ndim > 2) {
      // E1 at (i, j-1/2, k-1/2): F3(By)(j-1) + F3(By) - F2(Bz)(k-1) - F2(Bz)
      double *e1 = edgep(u, 0, b);
      for (int k = sh.ks; k <= sh.ke + 1; ++k)
        for (int j = sh.js; j <= sh.je + 1; ++j)
          for (int i = sh.is; i <= sh.ie; ++i) {
            const long c = sh.c(k, j, i);
            e1[sh.s(k, j, i)] =
                esign * 0.25 * (fzBy[c - sj] + fzBy[c] - fyBz[c - sk] - fyBz[c]);
          }
      // E2 at (i-1/2, j, k-1/2): F1(Bz)(k-1) + F1(Bz) - F3(Bx)(i-1) - F3(Bx)
      double *e2 = edgep(u, 1, b);
      for (int k = sh.ks; k <= sh.ke + 1; ++k)
        for (int j = sh.js; j <= sh.je; ++j)
          for (int i = sh.is; i <= sh.ie + 1; ++i) {
            const long c = sh.c(k, j, i);
            e2[sh.s(k, j, i)] =
                esign * 0.25 * (fxBz[c - sk] + fxBz[c] - fzBx[c - 1] - fzBx[c]);
          }
    }
  }
}

// ---------------------------------------------------------------------------------
// FluxesToDuDt, grid.cpp:179-211: FluxDivergence (grid_update.hpp:15-57) and
// FluxStokes (grid_update.hpp:80-137)
// ---------------------------------------------------------------------------------
void FluxesToDuDt(const km_cfg_t &cfg, const km_pack_t &u, const km_pack_t &dudt) {
  const Shape sh(u);
  const int ndim = sh.ndim;
  const int nc = ncons(cfg.mhd);
  const int nsp = cfg.nspecies > 1 ? cfg.nspecies : 0;
  const long off[3] = {1, sh.E[0], (long)sh.E[0] * sh.E[1]};
#pragma omp parallel for collapse(2) schedule(static)
  for (int b = 0; b < u.nblocks; ++b)
    for (int k = sh.ks; k <= sh.ke; ++k) {
      const km_block_t &blk = u.blocks[b];
      for (int j = sh.js; j <= sh.je; ++j)
        for (int i = sh.is; i <= sh.ie; ++i) {
          const long c = sh.c(k, j, i);
          // every WithFluxes field: the conserved ones, then the mass fractions (flux
          // slot nc + s, cell slot KM_MFRAC(s))
          for (int fv = 0; fv < nc + nsp; ++fv) {
            const int var = fv < nc ? fv : KM_MFRAC(cfg.mhd, fv - nc);
            double du = 0.;
            for (int dir = 0; dir < ndim; ++dir) {
              const double *F = fluxp(u, dir, b, fv);
              du -= (Area(cfg, blk, sh, dir, i + (dir == 0)) * F[c + off[dir]] -
                     Area(cfg, blk, sh, dir, i) * F[c]);
            }
            cellp(dudt, b, var)[c] = du / Volume(cfg, blk, sh, i);
          }
        }
    }
  if (cfg.mhd != KM_MHD_CT || ndim < 2) return;
  // FluxStokes<F1,E3[,E2]>, <F2,E3[,E1]>, <F3,E1,E2>: dB/dt for every face element
  struct Term {
    int edge, axis;
    double sign;
  };
  // (edge, axis of the upper edge, sign) per face, from AxisFromFaceEdge and the
  // cyclic-permutation sign rule, grid_update.hpp:59-78,115-125
  const Term f1[2] = {{2, 1, 1.0}, {1, 2, -1.0}};
  const Term f2[2] = {{2, 0, -1.0}, {0, 2, 1.0}};
  const Term f3[2] = {{0, 1, -1.0}, {1, 0, 1.0}};
  const Term *terms[3] = {f1, f2, f3};
  const int nterms = ndim > 2 ? 2 : 1;
  const long soff[3] = {1, sh.S[0], (long)sh.S[0] * sh.S[1]};
  for (int face = 0; face < (ndim > 2 ? 3 : 2); ++face) {
    const int di = face == 0, dj = face == 1, dk = face == 2 && ndim > 2;
#pragma omp parallel for schedule(static)
    for (int b = 0; b < u.nblocks; ++b) {
      const km_block_t &blk = u.blocks[b];
      double *out = facep(dudt, face, b);
      // cylindrical r, z, phi is an odd permutation (:122-125); E3 edges have length
      // |r| 2 pi (geometry.hpp:143-150), the rest their Dx
      const double gsign = Cyl(cfg) ? -1.0 : 1.0;
      auto edge_len = [&](int e, int ii) {
        return (Cyl(cfg) && e == 2) ? std::abs(RFace(blk, sh, ii)) * kPhiSpan : blk.dx[e];
      };
      for (int k = sh.ks; k <= sh.ke + dk; ++k)
        for (int j = sh.js; j <= sh.je + dj; ++j)
          for (int i = sh.is; i <= sh.ie + di; ++i) {
            const long s = sh.s(k, j, i);
            const double A = Area(cfg, blk, sh, face, i);
            out[s] = 0.;
            for (int t = 0; t < nterms; ++t) {
              const Term &tm = terms[face][t];
              const double *E = edgep(u, tm.edge, b);
              const double Lm = edge_len(tm.edge, i);
              const double Lp = edge_len(tm.edge, i + (tm.axis == 0));
              out[s] += (gsign * tm.sign) * (Lm * E[s] - Lp * E[s + soff[tm.axis]]);
            }
            out[s] *= 1. / (A + 1.e-36);
          }
    }
  }
}

// ---------------------------------------------------------------------------------
// ApplyDuDt_impl, grid.cpp:213-283
// ---------------------------------------------------------------------------------
// hydro::AddSourceTerms, hydro.cpp:199-234 (no-op in Cartesian geometry), with
// AddGeometricSource, primconsflux.hpp:144-159: primitives of u, into dudt
void AddSourceTerms(const km_cfg_t &cfg, const km_pack_t &u, const km_pack_t &dudt) {
  if (!Cyl(cfg)) return;
  const Shape sh(u);
  const int mhd = cfg.mhd;
#pragma omp parallel for collapse(2) schedule(static)
  for (int b = 0; b < u.nblocks; ++b)
    for (int k = sh.ks; k <= sh.ke; ++k)
      for (int j = sh.js; j <= sh.je; ++j)
        for (int i = sh.is; i <= sh.ie; ++i) {
          const long c = sh.c(k, j, i);
          const double r = RCentre(u.blocks[b], sh, i);
          const double rho = cellp(u, b, KM_DENS)[c];
          const double vphi = cellp(u, b, KM_VEL0(mhd) + 2)[c];
          double &d = cellp(dudt, b, KM_MOM0)[c];
          d += (rho * vphi * vphi + cellp(u, b, KM_PRES(mhd))[c]) / r;
          if (mhd != KM_MHD_OFF) {
            const double b0 = cellp(u, b, KM_MAGC0)[c], b1 = cellp(u, b, KM_MAGC0 + 1)[c],
                         b2 = cellp(u, b, KM_MAGC0 + 2)[c];
            const double pmag = 0.5 * (b0 * b0 + b1 * b1 + b2 * b2);
            d += (pmag - b2 * b2) / r;
          }
        }
}

// hydro::PrepareConserved, the per-stage task before ApplyDuDt (hydro.cpp:244-286): with CT
// in cylindrical geometry the cell B_phi of md0 becomes B_phi / r_centroid
void SpeciesToConserved(const km_cfg_t &cfg, const km_pack_t &u);

void StagePrepareConserved(const km_cfg_t &cfg, const km_pack_t &u) {
  SpeciesToConserved(cfg, u);
  if (!Cyl(cfg) || cfg.mhd == KM_MHD_OFF) return;
  const Shape sh(u);
#pragma omp parallel for collapse(2) schedule(static)
  for (int b = 0; b < u.nblocks; ++b)
    for (int k = sh.ks; k <= sh.ke; ++k)
      for (int j = sh.js; j <= sh.je; ++j)
        for (int i = sh.is; i <= sh.ie; ++i)
          cellp(u, b, KM_MAGC0 + 2)[sh.c(k, j, i)] *= 1.0 / RCentroid(u.blocks[b], sh, i);
}

// material::PrepareConserved, material.cpp:132-155: X_s -> rho X_s (more than one species)
void SpeciesToConserved(const km_cfg_t &cfg, const km_pack_t &u) {
  if (cfg.nspecies < 2) return;
  const Shape sh(u);
#pragma omp parallel for collapse(2) schedule(static)
  for (int b = 0; b < u.nblocks; ++b)
    for (int k = sh.ks; k <= sh.ke; ++k)
      for (int j = sh.js; j <= sh.je; ++j)
        for (int i = sh.is; i <= sh.ie; ++i) {
          const long c = sh.c(k, j, i);
          for (int sp = 0; sp < cfg.nspecies; ++sp)
            cellp(u, b, KM_MFRAC(cfg.mhd, sp))[c] *= cellp(u, b, KM_DENS)[c];
        }
}

// material::PreparePrimitive (material.cpp:157-197) and PostMeshInitialization (:104-130):
// each mass fraction divided by the sum over species, summed in species order. The
// reference's PreparePrimitive indexes its per-row scratch sum by the interior i (which
// starts at nghost) in a scratch of nx entries; the division by the sum is restated here
// as PostMeshInitialization writes it.
void SpeciesNormalize(const km_cfg_t &cfg, const km_pack_t &u) {
  if (cfg.nspecies < 2) return;
  const Shape sh(u);
#pragma omp parallel for collapse(2) schedule(static)
  for (int b = 0; b < u.nblocks; ++b)
    for (int k = sh.ks; k <= sh.ke; ++k)
      for (int j = sh.js; j <= sh.je; ++j)
        for (int i = sh.is; i <= sh.ie; ++i) {
          const long c = sh.c(k, j, i);
          double sum = 0.0;
          for (int sp = 0; sp < cfg.nspecies; ++sp) sum += cellp(u, b, KM_MFRAC(cfg.mhd, sp))[c];
          for (int sp = 0; sp < cfg.nspecies; ++sp)
            cellp(u, b, KM_MFRAC(cfg.mhd, sp))[c] *= 1. / sum;
        }
}

void ApplyDuDt(const km_cfg_t &cfg, const km_pack_t &base, const km_pack_t &u0,
               const km_pack_t &u1, const km_pack_t &dudt, double beta, double dt) {
  const Shape sh(u0);
  const int nc = ncons(cfg.mhd);
  const int nsp = cfg.nspecies > 1 ? cfg.nspecies : 0;
#pragma omp parallel for collapse(2) schedule(static)
  for (int b = 0; b < u0.nblocks; ++b)
    for (int k = sh.ks; k <= sh.ke; ++k)
      for (int j = sh.js; j <= sh.je; ++j)
        for (int i = sh.is; i <= sh.ie; ++i) {
          const long c = sh.c(k, j, i);
          for (int fv = 0; fv < nc + nsp; ++fv) {
            const int var = fv < nc ? fv : KM_MFRAC(cfg.mhd, fv - nc);
            double &p0 = cellp(u0, b, var)[c];
            p0 = beta * cellp(base, b, var)[c] + (1.0 - beta) * p0;
            cellp(u1, b, var)[c] = p0 + beta * dt * cellp(dudt, b, var)[c];
          }
        }
  if (cfg.mhd != KM_MHD_CT || sh.ndim < 2) return;
  for (int face = 0; face < (sh.ndim > 2 ? 3 : 2); ++face) {
    const int di = face == 0, dj = face == 1, dk = face == 2;
#pragma omp parallel for schedule(static)
    for (int b = 0; b < u0.nblocks; ++b) {
      double *p0 = facep(u0, face, b), *p1 = facep(u1, face, b);
      const double *pb = facep(base, face, b), *pd = facep(dudt, face, b);
      for (int k = sh.ks; k <= sh.ke + dk; ++k)
        for (int j = sh.js; j <= sh.je + dj; ++j)
          for (int i = sh.is; i <= sh.ie + di; ++i) {
            const long s = sh.s(k, j, i);
            p0[s] = beta * pb[s] + (1.0 - beta) * p0[s];
            p1[s] = p0[s] + beta * dt * pd[s];
          }
    }
  }
}

// CT cell-centred B from face B, primconsflux.cpp:102-128
inline void FaceAverageB(const km_cfg_t &cfg, const Shape &sh, const km_pack_t &u, int b, int k,
                         int j, int i) {
  const km_block_t &blk = u.blocks[b];
  const int ndim = sh.ndim;
  const long c = sh.c(k, j, i), s = sh.s(k, j, i);
  if (ndim > 1) {
    const double *B1 = facep(u, 0, b), *B2 = facep(u, 1, b);
    const double V = Volume(cfg, blk, sh, i);
    cellp(u, b, KM_MAGC0 + 0)[c] =
        0.5 * blk.dx[0] *
        (Area(cfg, blk, sh, 0, i + 1) * B1[s + 1] + Area(cfg, blk, sh, 0, i) * B1[s]) / V;
    cellp(u, b, KM_MAGC0 + 1)[c] =
        0.5 * blk.dx[1] *
        (Area(cfg, blk, sh, 1, i) * B2[s + sh.S[0]] + Area(cfg, blk, sh, 1, i) * B2[s]) / V;
  }
  if (ndim > 2) {
    const double *B3 = facep(u, 2, b);
    cellp(u, b, KM_MAGC0 + 2)[c] =
        0.5 * blk.dx[2] *
        (blk.area[2] * B3[s + (long)sh.S[0] * sh.S[1]] + blk.area[2] * B3[s]) / blk.vol;
  }
}

// PreparePrimitive: hydro PreparePrimitive_impl (primconsflux.cpp:82-142) then eos
// EosWrapped(ener) (eos.cpp:71-114)
void PreparePrimitive(const km_cfg_t &cfg, const km_pack_t &u) {
  const Shape sh(u);
  const int mhd = cfg.mhd;
  const IdealGas eos(cfg.gamma, cfg.abar);
#pragma omp parallel for collapse(2) schedule(static)
  for (int b = 0; b < u.nblocks; ++b)
    for (int k = sh.ks; k <= sh.ke; ++k)
      for (int j = sh.js; j <= sh.je; ++j)
        for (int i = sh.is; i <= sh.ie; ++i) {
          const long c = sh.c(k, j, i);
          if (mhd == KM_MHD_CT) FaceAverageB(cfg, sh, u, b, k, j, i);
          // Cons2Prim, primconsflux.hpp:76-98
          const double dens = cellp(u, b, KM_DENS)[c];
          const double idens = 1.0 / dens;
          double emag = 0., ekin = 0.;
          for (int dir = 0; dir < 3; dir++) {
            const double v = idens * cellp(u, b, KM_MOM0 + dir)[c];
            cellp(u, b, KM_VEL0(mhd) + dir)[c] = v;
            ekin += v * v;
            if (mhd != KM_MHD_OFF) {
              const double B = cellp(u, b, KM_MAGC0 + dir)[c];
              emag += B * B;
            }
          }
          ekin *= 0.5 * dens;
          emag *= 0.5;
          const double eint = cellp(u, b, KM_ENER)[c] - ekin - emag;
          const double sie = idens * eint;
          cellp(u, b, KM_EINT(mhd))[c] = sie;
          if (Cyl(cfg)) {  // primconsflux.cpp:129-137: v_phi from the angular momentum
            double &vphi = cellp(u, b, KM_VEL0(mhd) + 2)[c];
            vphi *= Ratio(1.0, RCentroid(u.blocks[b], sh, i));
            if (mhd != KM_MHD_OFF) cellp(u, b, KM_MAGC0 + 2)[c] *= RCentroid(u.blocks[b], sh, i);
          }
          // EosMode::ener -> temperature | pressure | bulk_modulus
          cellp(u, b, KM_PRES(mhd))[c] = eos.P(dens, sie);
          cellp(u, b, KM_TEMP(mhd))[c] = eos.T(dens, sie);
          cellp(u, b, KM_BMOD(mhd))[c] = eos.B(dens, sie);
        }
  SpeciesNormalize(cfg, u);  // material PreparePrimitive, after hydro and eos
}

// PostMeshInitialization: eos PrepareConserved(mode_init) (eos.cpp:165-168; which outputs
// each mode fills: eos_singularity.hpp:17-36)
// then hydro ConvertToConserved_impl (primconsflux.cpp:19-79)
void PrepareConserved(const km_cfg_t &cfg, const km_pack_t &u) {
  const Shape sh(u);
  const int mhd = cfg.mhd;
  const IdealGas eos(cfg.gamma, cfg.abar);
#pragma omp parallel for collapse(2) schedule(static)
  for (int b = 0; b < u.nblocks; ++b)
    for (int k = sh.ks; k <= sh.ke; ++k)
      for (int j = sh.js; j <= sh.je; ++j)
        for (int i = sh.is; i <= sh.ie; ++i) {
          const long c = sh.c(k, j, i);
          const double dens = cellp(u, b, KM_DENS)[c];
          if (cfg.mode_init == KM_EOS_INIT_DENS_ENER) {
            // EosMode::ener -> temperature | pressure | bulk_modulus
            const double sie = cellp(u, b, KM_EINT(mhd))[c];
            cellp(u, b, KM_PRES(mhd))[c] = eos.P(dens, sie);
            cellp(u, b, KM_TEMP(mhd))[c] = eos.T(dens, sie);
            cellp(u, b, KM_BMOD(mhd))[c] = eos.B(dens, sie);
          } else if (cfg.mode_init == KM_EOS_INIT_DENS_TEMP) {
            // EosMode::temp -> specific_internal_energy | pressure | bulk_modulus
            const double sie = eos.sie_from_T(cellp(u, b, KM_TEMP(mhd))[c]);
            cellp(u, b, KM_EINT(mhd))[c] = sie;
            cellp(u, b, KM_PRES(mhd))[c] = eos.P(dens, sie);
            cellp(u, b, KM_BMOD(mhd))[c] = eos.B(dens, sie);
          } else {
            // EosMode::pres -> temperature | specific_internal_energy | bulk_modulus
            const double sie = eos.sie_from_PT(dens, cellp(u, b, KM_PRES(mhd))[c]);
            cellp(u, b, KM_EINT(mhd))[c] = sie;
            cellp(u, b, KM_TEMP(mhd))[c] = eos.T(dens, sie);
            cellp(u, b, KM_BMOD(mhd))[c] = eos.B(dens, sie);
          }
          if (mhd == KM_MHD_CT) FaceAverageB(cfg, sh, u, b, k, j, i);
          double V[10], U[8];
          for (int r = 0; r < nrecon(mhd); ++r) V[r] = cellp(u, b, recon_slot(r, mhd))[c];
          Prim2Cons(mhd, V, U);
          // cylindrical: conserve angular momentum, primconsflux.cpp:70-74
          if (Cyl(cfg)) U[KM_MOM0 + 2] *= RCentroid(u.blocks[b], sh, i);
          for (int v = 0; v < ncons(mhd); ++v) cellp(u, b, v)[c] = U[v];
        }
  SpeciesNormalize(cfg, u);  // material PostMeshInitialization
}

// EstimateTimeStep::dispatch, hydro_time_step.cpp:19-65
double EstimateDt(const km_cfg_t &cfg, const km_pack_t &u) {
  const Shape sh(u);
  const int mhd = cfg.mhd;
  const int ndim = sh.ndim;
  const double cfl = cfg.cfl / static_cast<double>(ndim);
  double dt_min = kBig;
#pragma omp parallel for collapse(2) reduction(min : dt_min) schedule(static)
  for (int b = 0; b < u.nblocks; ++b)
    for (int k = sh.ks; k <= sh.ke; ++k) {
      const km_block_t &blk = u.blocks[b];
      for (int j = sh.js; j <= sh.je; ++j)
        for (int i = sh.is; i <= sh.ie; ++i) {
          const long c = sh.c(k, j, i);
          double V[10];
          for (int r = 0; r < nrecon(mhd); ++r) V[r] = cellp(u, b, recon_slot(r, mhd))[c];
          for (int dir = 0; dir < ndim; dir++) {
            const double cfast = FastSpeed(mhd, dir, V);
            dt_min = kmin(dt_min, blk.dx[dir] / (std::fabs(V[R_VEL + dir]) + cfast));
          }
          if (Cyl(cfg))  // :56-60
            dt_min = kmin(dt_min, RCentre(blk, sh, i) / std::fabs(V[R_VEL + 2]));
        }
    }
  return dt_min * cfl;
}

// A zone the run cannot continue from: density or pressure not positive, or a
// non-finite density, pressure or velocity (checked on the state the CFL estimate reads)
bool AnyNonPhysical(const km_cfg_t &cfg, const km_pack_t &u) {
  const Shape sh(u);
  const int mhd = cfg.mhd;
  int bad = 0;
#pragma omp parallel for collapse(2) reduction(| : bad) schedule(static)
  for (int b = 0; b < u.nblocks; ++b)
    for (int k = sh.ks; k <= sh.ke; ++k)
      for (int j = sh.js; j <= sh.je; ++j)
        for (int i = sh.is; i <= sh.ie; ++i) {
          const long c = sh.c(k, j, i);
          const double rho = cellp(u, b, KM_DENS)[c], p = cellp(u, b, KM_PRES(mhd))[c];
          double vsum = 0.;
          for (int d = 0; d < 3; ++d) vsum += std::fabs(cellp(u, b, KM_VEL0(mhd) + d)[c]);
          if (!std::isfinite(rho) || rho <= 0. || !std::isfinite(p) || p <= 0. ||
              !std::isfinite(vsum))
            bad = 1;
        }
  return bad != 0;
}

// parthenon/time step control (kamayan_driver.cpp:28-53), Parthenon's documented rule:
// a new dt is capped by dt_user, clamped to [dt_floor, dt_ceil] and replaced by dt_force
// when that is set (positive; the defaults, -DBL_MAX and the deck writer's -1e300
// (python/kamayan/code_units/driver.py:10-21), leave it off); the first one is the
// estimate capped by dt_init, or dt_init itself with dt_init_force when dt_init is given
// (positive, below 1e300); later ones grow by at most dt_factor.
double DtLimits(const km_cfg_t &c, double x) {
  x = kmin(x, c.dt_user);
  x = kmin(kmax(x, c.dt_floor), c.dt_ceil);
  if (c.dt_force > 0.0) x = c.dt_force;
  return x;
}
double FirstDt(const km_cfg_t &c, double est) {
  double x = kmin(est, c.dt_init);
  if (c.dt_init_force && c.dt_init > 0.0 && c.dt_init < 1e300) x = c.dt_init;
  return DtLimits(c, x);
}

// FillDerived_impl, hydro.cpp:150-194
void FillDivB(const km_cfg_t &cfg, const km_pack_t &u) {
  if (cfg.mhd != KM_MHD_CT) return;
  const Shape sh(u);
  const long soff[3] = {1, sh.S[0], (long)sh.S[0] * sh.S[1]};
#pragma omp parallel for collapse(2) schedule(static)
  for (int b = 0; b < u.nblocks; ++b)
    for (int k = sh.ks; k <= sh.ke; ++k) {
      const km_block_t &blk = u.blocks[b];
      for (int j = sh.js; j <= sh.je; ++j)
        for (int i = sh.is; i <= sh.ie; ++i) {
          const long s = sh.s(k, j, i);
          double divb = 0.;
          for (int dir = 0; dir < sh.ndim; dir++) {
            const double *B = facep(u, dir, b);
            divb += Area(cfg, blk, sh, dir, i + (dir == 0)) * B[s + soff[dir]] -
                    Area(cfg, blk, sh, dir, i) * B[s];
          }
          cellp(u, b, KM_DIVB(cfg.mhd))[sh.c(k, j, i)] = divb / Volume(cfg, blk, sh, i);
        }
    }
}

// ---------------------------------------------------------------------------------
// Same-level ghost fill for a uniform block grid. Parthenon's boundary exchange
// (kamayan_driver.cpp:150-151) copies neighbour interiors into every ghost zone (faces,
// edges, corners) and applies the physical boundary conditions; for a single-level
// grid that is the index map below, applied per dimension.
// ---------------------------------------------------------------------------------
struct Grid {
  int nb[3];
  int bc[6];
};

// Map a global index g along a dimension of N zones (cells) to [0, N) or -1 (reflect
// flips sign). Returns the mapped index and sets *flip for reflecting boundaries.
inline int map_cell(int g, int N, int bclo, int bchi, bool *flip) {
  if (g < 0) {
    if (bclo == KM_BC_PERIODIC) return ((g % N) + N) % N;
    if (bclo == KM_BC_OUTFLOW) return 0;
    *flip = !*flip;
    return -g - 1;
  }
  if (g >= N) {
    if (bchi == KM_BC_PERIODIC) return g % N;
    if (bchi == KM_BC_OUTFLOW) return N - 1;
    *flip = !*flip;
    return 2 * N - g - 1;
  }
  return g;
}

// faces along their staggered dimension: N+1 face positions, 0 and N coincide for
// periodic grids.
inline int map_face(int g, int N, int bclo, int bchi, bool *flip) {
  if (g < 0) {
    if (bclo == KM_BC_PERIODIC) return ((g % N) + N) % N;
    if (bclo == KM_BC_OUTFLOW) return 0;
    *flip = !*flip;
    return -g;
  }
  if (g > N) {
    if (bchi == KM_BC_PERIODIC) return g % N;
    if (bchi == KM_BC_OUTFLOW) return N;
    *flip = !*flip;
    return 2 * N - g;
  }
  return g;
}

void GhostFill(const km_cfg_t &cfg, const Grid &grid, const km_pack_t &u, int v0, int v1,
               bool with_faces) {
  const Shape sh(u);
  const int ndim = sh.ndim;
  const int mhd = cfg.mhd;
  const int N[3] = {grid.nb[0] * sh.nx[0], grid.nb[1] * sh.nx[1], grid.nb[2] * sh.nx[2]};
  const int st[3] = {sh.is, sh.js, sh.ks};
  // vector components flip sign under reflection of their own axis
  auto vec_axis = [&](int v) -> int {
    if (v >= KM_MOM0 && v < KM_MOM0 + 3) return v - KM_MOM0;
    if (v >= KM_VEL0(mhd) && v < KM_VEL0(mhd) + 3) return v - KM_VEL0(mhd);
    if (mhd == KM_MHD_CT && v >= KM_MAGC0 && v < KM_MAGC0 + 3) return v - KM_MAGC0;
    return -1;
  };
#pragma omp parallel for schedule(static)
  for (int b = 0; b < u.nblocks; ++b) {
    const int bl[3] = {b % grid.nb[0], (b / grid.nb[0]) % grid.nb[1],
                       b / (grid.nb[0] * grid.nb[1])};
    for (int k = 0; k < sh.E[2]; ++k)
      for (int j = 0; j < sh.E[1]; ++j)
        for (int i = 0; i < sh.E[0]; ++i) {
          const int idx[3] = {i, j, k};
          bool ghost = false;
          for (int d = 0; d < ndim; ++d)
            if (idx[d] < st[d] || idx[d] >= st[d] + sh.nx[d]) ghost = true;
          if (!ghost) continue;
          int src[3];
          bool flip[3] = {false, false, false};
          int sb[3];
          for (int d = 0; d < 3; ++d) {
            if (d >= ndim) {
              sb[d] = 0;
              src[d] = idx[d];
              continue;
            }
            const int g = bl[d] * sh.nx[d] + (idx[d] - st[d]);
            // Axisymmetric (boundary_conditions.cpp:23-60): a mirror across r = 0 that
            // also negates vector component 3 (phi)
            if (d == 0 && g < 0 && grid.bc[0] == KM_BC_AXISYMMETRIC) flip[2] = !flip[2];
            const int m = map_cell(g, N[d], grid.bc[2 * d], grid.bc[2 * d + 1], &flip[d]);
            sb[d] = m / sh.nx[d];
            src[d] = m % sh.nx[d] + st[d];
          }
          const int srcb = (sb[2] * grid.nb[1] + sb[1]) * grid.nb[0] + sb[0];
          const long dc = sh.c(k, j, i), scx = sh.c(src[2], src[1], src[0]);
          for (int v = v0; v < v1; ++v) {
            double val = cellp(u, srcb, v)[scx];
            const int ax = vec_axis(v);
            if (ax >= 0 && flip[ax]) val = -val;
            cellp(u, b, v)[dc] = val;
          }
        }
    if (!with_faces || mhd != KM_MHD_CT) continue;
    for (int face = 0; face < 3; ++face) {
      for (int k = 0; k < sh.S[2]; ++k)
        for (int j = 0; j < sh.S[1]; ++j)
          for (int i = 0; i < sh.S[0]; ++i) {
            const int idx[3] = {i, j, k};
            bool ghost = false, valid = true;
            for (int d = 0; d < 3; ++d) {
              const int hi = st[d] + sh.nx[d] + (d == face ? 1 : 0);  // exclusive
              const int ext = d < ndim ? sh.E[d] + (d == face ? 1 : 0) : 1 + (d == face);
              if (idx[d] >= ext) valid = false;
              if (d < ndim && (idx[d] < st[d] || idx[d] >= hi)) ghost = true;
            }
            if (!valid || !ghost) continue;
            int src[3], sb[3];
            bool flip[3] = {false, false, false};
            for (int d = 0; d < 3; ++d) {
              if (d >= ndim) {
                sb[d] = 0;
                src[d] = idx[d];
                continue;
              }
              const int g = bl[d] * sh.nx[d] + (idx[d] - st[d]);
              int m;
              if (d == face) {
                if (d == 0 && g < 0 && grid.bc[0] == KM_BC_AXISYMMETRIC) {
                  // Axisymmetric on F1 (boundary_conditions.cpp:50-58): the ghost face at
                  // index i takes -(face 2 is - 1 - i), i.e. global -g - 1
                  m = -g - 1;
                  flip[d] = !flip[d];
                } else {
                  m = map_face(g, N[d], grid.bc[2 * d], grid.bc[2 * d + 1], &flip[d]);
                }
                sb[d] = std::min(m / sh.nx[d], grid.nb[d] - 1);
                src[d] = m - sb[d] * sh.nx[d] + st[d];
              } else {
                m = map_cell(g, N[d], grid.bc[2 * d], grid.bc[2 * d + 1], &flip[d]);
                sb[d] = m / sh.nx[d];
                src[d] = m % sh.nx[d] + st[d];
              }
            }
            const int srcb = (sb[2] * grid.nb[1] + sb[1]) * grid.nb[0] + sb[0];
            double val = facep(u, face, srcb)[sh.s(src[2], src[1], src[0])];
            if (flip[face]) val = -val;
            facep(u, face, b)[sh.s(k, j, i)] = val;
          }
    }
  }
}

// ---------------------------------------------------------------------------------
// Block-tree (AMR) meshes. Parthenon's boundary exchange on a multilevel mesh
// (kamayan_driver.cpp:145-151) fills a ghost zone from the leaf covering it: a copy from
// a same-level leaf, RestrictAverage (refinement_operations.hpp:52-121) of a finer leaf,
// or ProlongateSharedMinMod (:123-268) from the coarse buffer when the leaf is coarser,
// the coarse buffer holding coarse-leaf values and restricted same-level values.
//
// The restatement builds, per level L, the composite field R_L over the whole domain:
// leaf values where a level-L leaf lies, the restriction of R_{L+1} where finer leaves
// lie, undefined under coarser leaves. Ghost zones then read R_L, or prolongate from
// R_{L-1}. A read of a twice-restricted value (depth >= 2) is reported as an error:
// the 2:1-balanced meshes accepted here never need one.
// Flux correction (AddFluxCorrectionTasks, kamayan_driver.cpp:198-207) replaces a
// block's face flux by the area-weighted RestrictAverage of the fine faces behind it.
// Parthenon internals are not vendored: parity unpinned beyond the invariants in
// tests/test_oracle_amr.py (conservation, linear exactness, uniform equivalence).
// ---------------------------------------------------------------------------------
struct Tree {
  int ndim = 1, nx[3] = {1, 1, 1}, nb0[3] = {1, 1, 1}, lmax = 0, bc[6] = {0, 0, 0, 0, 0, 0};
  bool cyl = false;   // cylindrical (r, z): geometry-templated refinement operators
  double x1min = 0.;  // the domain's lower x1 edge
  std::vector<int> level, loc;  // per leaf: level, loc[3*b + d]
  // leaf id by (level, block location) key
  std::vector<std::vector<int>> at;  // at[L][flat block index at level L] (-1: none)
  int nbl(int L, int d) const { return d < ndim ? (nb0[d] << L) : 1; }
  long flat(int L, const int l[3]) const {
    return ((long)l[2] * nbl(L, 1) + l[1]) * nbl(L, 0) + l[0];
  }
  int zones(int L, int d) const { return d < ndim ? nbl(L, d) * nx[d] : 1; }
};

// returns an error text, or empty
std::string BuildTree(const km_mesh_desc_t &m, const int32_t *leaves, int n, Tree &t) {
  t.ndim = m.ndim;
  for (int d = 0; d < 3; ++d) {
    t.nx[d] = d < m.ndim ? m.nx[d] : 1;
    t.nb0[d] = d < m.ndim ? m.nb[d] : 1;
  }
  for (int d = 0; d < 6; ++d) t.bc[d] = m.bc[d];
  t.x1min = m.xmin[0];
  t.lmax = 0;
  for (int b = 0; b < n; ++b) t.lmax = std::max(t.lmax, (int)leaves[4 * b]);
  t.at.assign(t.lmax + 1, {});
  for (int L = 0; L <= t.lmax; ++L)
    t.at[L].assign((size_t)t.nbl(L, 0) * t.nbl(L, 1) * t.nbl(L, 2), -1);
  t.level.resize(n);
  t.loc.resize(3 * (size_t)n);
  long covered = 0;  // in units of finest blocks
  for (int b = 0; b < n; ++b) {
    const int L = leaves[4 * b];
    int l[3];
    for (int d = 0; d < 3; ++d) {
      l[d] = d < m.ndim ? leaves[4 * b + 1 + d] : 0;
      if (l[d] < 0 || l[d] >= t.nbl(L, d)) return "leaf outside the root grid";
      t.loc[3 * b + d] = l[d];
    }
    t.level[b] = L;
    int &e = t.at[L][t.flat(L, l)];
    if (e >= 0) return "duplicate leaf";
    e = b;
    covered += 1L << (m.ndim * (t.lmax - L));
  }
  const long want = (long)t.nbl(t.lmax, 0) * t.nbl(t.lmax, 1) * t.nbl(t.lmax, 2);
  if (covered != want) return "leaves do not tile the domain";
  return "";
}

// Parthenon's uniform mesh generator applied to the root grid refined L times
void LeafMetric(const km_mesh_desc_t &m, int L, const int l[3], km_block_t &blk) {
  std::memset(&blk, 0, sizeof(blk));
  for (int d = 0; d < 3; ++d) {
    double lo, hi;
    if (d < m.ndim) {
      const long nb = (long)m.nb[d] << L;
      const double xl = static_cast<double>(2 * (long)l[d] - nb) / (2 * nb);
      const double xr = static_cast<double>(2 * ((long)l[d] + 1) - nb) / (2 * nb);
      lo = 0.5 * (m.xmin[d] + m.xmax[d]) + (xl * m.xmax[d] - xl * m.xmin[d]);
      hi = 0.5 * (m.xmin[d] + m.xmax[d]) + (xr * m.xmax[d] - xr * m.xmin[d]);
      if (l[d] == 0) lo = m.xmin[d];
      if (l[d] == nb - 1) hi = m.xmax[d];
    } else {
      lo = m.xmin[d];
      hi = m.xmax[d];
    }
    const int n = d < m.ndim ? m.nx[d] : 1;
    blk.xmin[d] = lo;
    blk.dx[d] = (hi - lo) / static_cast<double>(n);
    blk.lx[d] = d < m.ndim ? l[d] : 0;
  }
  blk.area[0] = blk.dx[1] * blk.dx[2];
  blk.area[1] = blk.dx[0] * blk.dx[2];
  blk.area[2] = blk.dx[0] * blk.dx[1];
  blk.vol = blk.dx[0] * blk.dx[1] * blk.dx[2];
  blk.level = L;
}

// Coordinates<geom> element measures of leaf blk at local index i along x1
// (geometry.hpp:124-191): cylindrical cell volumes 0.5 2pi (r+^2 - r-^2) dz, r-face areas
// r dz 2pi, z-face areas 0.5 (r+^2 - r-^2) 2pi; Cartesian ones from the block metric
double CellVol(bool cyl, const km_block_t &blk, const Shape &sh, int i) {
  if (!cyl) return blk.vol;
  const double rp = RFace(blk, sh, i + 1), rm = RFace(blk, sh, i);
  return 0.5 * kPhiSpan * (rp * rp - rm * rm) * blk.dx[1];
}
double FaceArea(bool cyl, const km_block_t &blk, const Shape &sh, int d, int i) {
  if (!cyl || d == 2) return blk.area[d];
  if (d == 0) return RFace(blk, sh, i) * blk.dx[1] * kPhiSpan;
  const double rp = RFace(blk, sh, i + 1), rm = RFace(blk, sh, i);
  return 0.5 * (rp * rp - rm * rm) * kPhiSpan;
}
// radius of the element at global index G (a face, or a cell centre) with spacing h, for
// the r-weighted face operators: measured from the domain's x1 edge, negative past the
// axis (the coarse buffer's own coordinates)
double RadiusAt(const Tree &t, int G, bool centre, double h) {
  return centre ? t.x1min + (G + 0.5) * h : t.x1min + G * h;
}
constexpr double kDblEps = 2.220446049250313e-16;

// composite per-level data of one ghost fill (variable independent part)
struct Composite {
  std::vector<std::vector<signed char>> depth;  // -1 undefined, 0 leaf value, 1 restricted ..
  std::vector<std::vector<int>> leaf;           // covering leaf where depth == 0
  std::vector<std::vector<double>> val;         // current variable
};

long ZoneIndex(const Tree &t, int L, const int z[3]) {
  return ((long)z[2] * t.zones(L, 1) + z[1]) * t.zones(L, 0) + z[0];
}

void BuildComposite(const Tree &t, const km_pack_t &u, Composite &c) {
  const int nl = t.lmax + 1;
  c.depth.assign(nl, {});
  c.leaf.assign(nl, {});
  c.val.assign(nl, {});
  for (int L = t.lmax; L >= 0; --L) {
    const long nz = (long)t.zones(L, 0) * t.zones(L, 1) * t.zones(L, 2);
    c.depth[L].assign(nz, -1);
    c.leaf[L].assign(nz, -1);
    c.val[L].assign(nz, 0.0);
    for (size_t b = 0; b < t.level.size(); ++b) {
      if (t.level[b] != L) continue;
      for (int k = 0; k < t.nx[2]; ++k)
        for (int j = 0; j < t.nx[1]; ++j)
          for (int i = 0; i < t.nx[0]; ++i) {
            const int z[3] = {t.loc[3 * b] * t.nx[0] + i, t.loc[3 * b + 1] * t.nx[1] + j,
                              t.loc[3 * b + 2] * t.nx[2] + k};
            const long zi = ZoneIndex(t, L, z);
            c.depth[L][zi] = 0;
            c.leaf[L][zi] = (int)b;
          }
    }
    if (L == t.lmax) continue;
    for (int k = 0; k < t.zones(L, 2); ++k)
      for (int j = 0; j < t.zones(L, 1); ++j)
        for (int i = 0; i < t.zones(L, 0); ++i) {
          const int z[3] = {i, j, k};
          const long zi = ZoneIndex(t, L, z);
          if (c.depth[L][zi] >= 0) continue;
          const int ch[3] = {2 * i, t.ndim > 1 ? 2 * j : 0, t.ndim > 2 ? 2 * k : 0};
          const signed char dch = c.depth[L + 1][ZoneIndex(t, L + 1, ch)];
          if (dch >= 0) c.depth[L][zi] = (signed char)(dch + 1);
        }
  }
  (void)u;
}

// RestrictAverage of the 2^d children of zone z of level L from R_{L+1}
double RestrictFrom(const Tree &t, const km_pack_t &u, const Composite &c, int L,
                    const int z[3]) {
  const bool x2 = t.ndim > 1, x3 = t.ndim > 2;
  double vol[2][2][2], terms[2][2][2];
  for (int ok = 0; ok < 2; ++ok)
    for (int oj = 0; oj < 2; ++oj)
      for (int oi = 0; oi < 2; ++oi) vol[ok][oj][oi] = terms[ok][oj][oi] = 0;
  for (int ok = 0; ok < 1 + x3; ++ok)
    for (int oj = 0; oj < 1 + x2; ++oj)
      for (int oi = 0; oi < 2; ++oi) {
        const int ch[3] = {2 * z[0] + oi, x2 ? 2 * z[1] + oj : 0, x3 ? 2 * z[2] + ok : 0};
        const long ci = ZoneIndex(t, L + 1, ch);
        const int lf = c.leaf[L + 1][ci];
        const Shape sh(u);
        vol[ok][oj][oi] =
            CellVol(t.cyl, u.blocks[lf], sh, ch[0] - t.loc[3 * lf] * t.nx[0] + sh.is);
        terms[ok][oj][oi] = vol[ok][oj][oi] * c.val[L + 1][ci];
      }
  const double tvol = ((vol[0][0][0] + vol[0][1][0]) + (vol[0][0][1] + vol[0][1][1])) +
                      ((vol[1][0][0] + vol[1][1][0]) + (vol[1][0][1] + vol[1][1][1]));
  return tvol > 0.0
             ? (((terms[0][0][0] + terms[0][1][0]) + (terms[0][0][1] + terms[0][1][1])) +
                ((terms[1][0][0] + terms[1][1][0]) + (terms[1][0][1] + terms[1][1][1]))) /
                   tvol
             : 0.0;
}

// value of R_L at an in-domain zone; an undefined or twice-restricted read sets *bad
double ReadR(const Tree &t, const km_pack_t &u, const Composite &c, int L, const int z[3],
             int *bad) {
  const long zi = ZoneIndex(t, L, z);
  const int dep = c.depth[L][zi];
  if (dep == 0) return c.val[L][zi];
  if (dep == 1) return RestrictFrom(t, u, c, L, z);
  *bad = 1;
  return 0.0;
}

void FillLevelValues(const Tree &t, const km_pack_t &u, Composite &c, int v) {
  const Shape sh(u);
  for (int L = 0; L <= t.lmax; ++L)
    for (size_t b = 0; b < t.level.size(); ++b) {
      if (t.level[b] != L) continue;
      const double *q = cellp(u, (int)b, v);
      for (int k = 0; k < t.nx[2]; ++k)
        for (int j = 0; j < t.nx[1]; ++j)
          for (int i = 0; i < t.nx[0]; ++i) {
            const int z[3] = {t.loc[3 * b] * t.nx[0] + i, t.loc[3 * b + 1] * t.nx[1] + j,
                              t.loc[3 * b + 2] * t.nx[2] + k};
            c.val[L][ZoneIndex(t, L, z)] = q[sh.c(k + sh.ks, j + sh.js, i + sh.is)];
          }
    }
}

inline double Sign(double x) { return std::copysign(1.0, x); }

// GradMinMod, refinement_operations.hpp:42-48
double GradMinMod(double fc, double fm, double fp, double dxm, double dxp) {
  const double gxm = (fc - fm) / dxm;
  const double gxp = (fp - fc) / dxp;
  return 0.5 * (Sign(gxm) + Sign(gxp)) * kmin(std::fabs(gxm), std::fabs(gxp));
}


void AmrGhostFill(const km_cfg_t &cfg, const Tree &t, const km_pack_t &u, int v0, int v1) {
  const Shape sh(u);
  const int mhd = cfg.mhd;
  const int st[3] = {sh.is, sh.js, sh.ks};
  auto vec_axis = [&](int v) -> int {
    if (v >= KM_MOM0 && v < KM_MOM0 + 3) return v - KM_MOM0;
    if (v >= KM_VEL0(mhd) && v < KM_VEL0(mhd) + 3) return v - KM_VEL0(mhd);
    if (mhd == KM_MHD_CT && v >= KM_MAGC0 && v < KM_MAGC0 + 3) return v - KM_MAGC0;
    return -1;
  };
  Composite c;
  BuildComposite(t, u, c);
  int bad = 0;
  for (int v = v0; v < v1; ++v) {
    FillLevelValues(t, u, c, v);
    const int ax = vec_axis(v);
    const int nleaf = (int)t.level.size();
#pragma omp parallel for schedule(dynamic, 4) reduction(| : bad)
    for (int b = 0; b < nleaf; ++b) {
      const int L = t.level[b];
      const km_block_t &blk = u.blocks[b];
      double *dst = cellp(u, (int)b, v);
      for (int k = 0; k < sh.E[2]; ++k)
        for (int j = 0; j < sh.E[1]; ++j)
          for (int i = 0; i < sh.E[0]; ++i) {
            const int idx[3] = {i, j, k};
            bool ghost = false;
            for (int d = 0; d < t.ndim; ++d)
              if (idx[d] < st[d] || idx[d] >= st[d] + sh.nx[d]) ghost = true;
            if (!ghost) continue;
            int z[3] = {0, 0, 0};
            bool flip[3] = {false, false, false};
            for (int d = 0; d < t.ndim; ++d) {
              const int g0 = t.loc[3 * b + d] * t.nx[d] + (idx[d] - st[d]);
              z[d] = map_cell(g0, t.zones(L, d), t.bc[2 * d], t.bc[2 * d + 1], &flip[d]);
              // the axisymmetric mirror also negates a vector's phi component
              if (d == 0 && g0 < 0 && t.bc[0] == KM_BC_AXISYMMETRIC) flip[2] = !flip[2];
            }
            double x;
            if (c.depth[L][ZoneIndex(t, L, z)] >= 0) {
              x = ReadR(t, u, c, L, z, &bad);
            } else {
              // prolongation from the coarse buffer (level L-1)
              if (L == 0) {
                bad |= 1;
                continue;
              }
              const int Lc = L - 1;
              const int cz[3] = {z[0] >> 1, t.ndim > 1 ? z[1] >> 1 : 0,
                                 t.ndim > 2 ? z[2] >> 1 : 0};
              const long czi = ZoneIndex(t, Lc, cz);
              if (c.depth[Lc][czi] != 0) {
                bad |= 1;
                continue;
              }
              const double fc = c.val[Lc][czi];
              double g[3] = {0., 0., 0.}, dxf[3] = {0., 0., 0.};
              for (int d = 0; d < t.ndim; ++d) {
                double f[2];
                for (int side = 0; side < 2; ++side) {
                  int n[3] = {cz[0], cz[1], cz[2]};
                  bool nf[3] = {false, false, false};
                  n[d] += side ? 1 : -1;
                  if (d == 0 && n[0] < 0 && t.bc[0] == KM_BC_AXISYMMETRIC) nf[2] = !nf[2];
                  n[d] = map_cell(n[d], t.zones(Lc, d), t.bc[2 * d], t.bc[2 * d + 1], &nf[d]);
                  f[side] = ReadR(t, u, c, Lc, n, &bad);
                  if (ax >= 0 && nf[ax]) f[side] = -f[side];
                }
                // coarse-buffer centre spacing and child offset of this block
                const double dxc = 2.0 * blk.dx[d];
                g[d] = GradMinMod(fc, f[0], f[1], dxc, dxc);
                dxf[d] = 0.5 * blk.dx[d];
              }
              const bool oi = z[0] & 1, oj = t.ndim > 1 && (z[1] & 1),
                         ok = t.ndim > 2 && (z[2] & 1);
              // refinement_operations.hpp:241-259, minmod slopes on both sides
              const double a = g[0] * dxf[0], bb = g[1] * dxf[1], cc = g[2] * dxf[2];
              if (!oi && !oj && !ok) x = fc - (a + bb + cc);
              else if (oi && !oj && !ok) x = fc + (a - bb - cc);
              else if (!oi && oj && !ok) x = fc - (a - bb + cc);
              else if (oi && oj && !ok) x = fc + (a + bb - cc);
              else if (!oi && !oj && ok) x = fc - (a + bb - cc);
              else if (oi && !oj && ok) x = fc + (a - bb + cc);
              else if (!oi && oj && ok) x = fc - (a - bb - cc);
              else x = fc + (a + bb + cc);
            }
            if (ax >= 0 && flip[ax]) x = -x;
            dst[sh.c(k, j, i)] = x;
          }
    }
  }
  if (bad) throw std::runtime_error("block-tree ghost fill met an unsupported level jump");
}

// ---- CT face fields on block-tree meshes ------------------------------------------
// A ghost face of a level-L block is taken from the leaf holding the cell beside it on
// the block's side (above a face below the block, below a face above it): a copy, the
// RestrictAverage of the finer leaf's faces (face elements average over their two
// transverse axes, weighted by face area), or, under a coarser leaf,
// ProlongateSharedMinMod for fine faces on coarse faces (transverse minmod slopes from the
// neighbouring coarse faces) and, once every shared face is set,
// ProlongateInternalTothAndRoe (refinement_operations.hpp:270-398) for the fine faces
// inside coarse cells, components in x1, x2, x3 order. A block's own faces are never
// written. The leaf holding a cell is found through the tree's per-level tables.

// leaf holding level-L cell z (in the domain), searching levels L+1, L, L-1
int LeafOfCell(const Tree &t, int L, const int z[3], int *lev) {
  for (int Lq = std::min(L + 1, t.lmax); Lq >= std::max(L - 1, 0); --Lq) {
    int l[3] = {0, 0, 0};
    for (int d = 0; d < t.ndim; ++d) {
      const int pos = Lq >= L ? (z[d] << (Lq - L)) : (z[d] >> (L - Lq));
      l[d] = pos / t.nx[d];
    }
    const int b = t.at[Lq][t.flat(Lq, l)];
    if (b >= 0) {
      *lev = Lq;
      return b;
    }
  }
  return -1;
}

struct FaceRef {
  int lev, leaf;
  int li[3];  // staggered element in the leaf (finer: lower corner of the fine faces)
};

// face d at level-L global index G (face index along d, cells across), seen from side `up`
bool FindFace(const Tree &t, int L, int d, const int G[3], bool up, FaceRef *f) {
  int z[3] = {G[0], G[1], G[2]};
  if (d < t.ndim) {
    const int N = t.zones(L, d);
    z[d] = up ? G[d] : G[d] - 1;
    z[d] = std::max(0, std::min(N - 1, z[d]));
  }
  f->leaf = LeafOfCell(t, L, z, &f->lev);
  if (f->leaf < 0) return false;
  const int sh = f->lev - L;  // +1 finer, 0, -1 coarser
  for (int e = 0; e < 3; ++e) {
    if (e >= t.ndim) {
      f->li[e] = 0;
      continue;
    }
    const int Ge = sh > 0 ? 2 * G[e] : (sh < 0 ? G[e] / 2 : G[e]);
    f->li[e] = Ge - t.loc[3 * f->leaf + e] * t.nx[e] + t.nx[e] * 0;
  }
  return true;
}

double FaceAt(const Shape &sh, const km_pack_t &u, int d, int leaf, const int li[3]) {
  const int lo[3] = {sh.is, sh.js, sh.ks};
  return facep(u, d, leaf)[sh.s(li[2] + lo[2], li[1] + lo[1], li[0] + lo[0])];
}

// value of a face found at level L: copy, or RestrictAverage of the finer faces
double FaceValue(const Tree &t, const Shape &sh, const km_pack_t &u, int L, int d,
                 const FaceRef &f) {
  if (f.lev == L) return FaceAt(sh, u, d, f.leaf, f.li);
  const bool inc[3] = {d != 0, d != 1 && t.ndim > 1, d != 2 && t.ndim > 2};
  double vol[2][2][2], terms[2][2][2];
  for (int ok = 0; ok < 2; ++ok)
    for (int oj = 0; oj < 2; ++oj)
      for (int oi = 0; oi < 2; ++oi) vol[ok][oj][oi] = terms[ok][oj][oi] = 0;
  for (int ok = 0; ok < 1 + inc[2]; ++ok)
    for (int oj = 0; oj < 1 + inc[1]; ++oj)
      for (int oi = 0; oi < 1 + inc[0]; ++oi) {
        const int li[3] = {f.li[0] + oi, f.li[1] + oj, f.li[2] + ok};
        vol[ok][oj][oi] = FaceArea(t.cyl, u.blocks[f.leaf], sh, d, li[0] + sh.is);
        terms[ok][oj][oi] = vol[ok][oj][oi] * FaceAt(sh, u, d, f.leaf, li);
      }
  const double tvol = ((vol[0][0][0] + vol[0][1][0]) + (vol[0][0][1] + vol[0][1][1])) +
                      ((vol[1][0][0] + vol[1][1][0]) + (vol[1][0][1] + vol[1][1][1]));
  return tvol > 0.0
             ? (((terms[0][0][0] + terms[0][1][0]) + (terms[0][0][1] + terms[0][1][1])) +
                ((terms[1][0][0] + terms[1][1][0]) + (terms[1][0][1] + terms[1][1][1]))) /
                   tvol
             : 0.0;
}

// is (k, j, i) of face dim d a ghost element of the block (not an own face, not padding)
bool GhostFaceElem(const Tree &t, const Shape &sh, int d, const int idx[3]) {
  const int lo[3] = {sh.is, sh.js, sh.ks};
  bool ghost = false;
  for (int e = 0; e < 3; ++e) {
    const int ext = e < t.ndim ? sh.E[e] + (e == d) : 1;
    if (idx[e] >= ext) return false;
    const int hi = lo[e] + sh.nx[e] - 1 + (e == d);
    if (e < t.ndim && (idx[e] < lo[e] || idx[e] > hi)) ghost = true;
  }
  return ghost;
}

// ProlongateInternalTothAndRoe (refinement_operations.hpp:278-398), Cartesian: the fine
// face (component d) at idx = (i, j, k) inside the coarse cell whose fine lower corner is
// c, from the coarse cell's shared fine faces of all three components in block b. The
// reference's (kk, jj, ii) loop order is reached by permuting the offsets per component;
// the coarse spacings enter unpermuted, as the reference writes them.
double TothRoe(const Shape &sh, const km_pack_t &u, int b, const km_block_t &blk, int d,
               const int c[3], const int idx[3], bool cyl, double x1min) {
  const int g2 = sh.ndim > 1, g3 = sh.ndim > 2;
  const int fi = c[0], fj = c[1], fk = c[2];
  auto elem = [&](int ok, int oj, int oi, int out[3]) {  // {k, j, i}
    if (d == 0) {
      out[0] = fk + ok * g3, out[1] = fj + oj * g2, out[2] = fi + oi;
    } else if (d == 1) {
      out[0] = fk + oj * g3, out[1] = fj + oi * g2, out[2] = fi + ok;
    } else {
      out[0] = fk + oi * g3, out[1] = fj + ok * g2, out[2] = fi + oj;
    }
  };
  // cylindrical (:336-356): the constraint holds for r B; r faces enter times their radius,
  // z faces times their cell-centre radius, phi faces (1-D and 2-D meshes) times 1
  auto radius = [&](int comp, int ii) -> double {
    if (!cyl || comp == 2) return 1.0;
    const int G = blk.lx[0] * sh.nx[0] + (ii - sh.is);
    return comp == 1 ? x1min + (G + 0.5) * blk.dx[0] : x1min + G * blk.dx[0];
  };
  auto fine = [&](int eidx, int ok, int oj, int oi) {
    int q[3];
    elem(ok, oj, oi, q);
    const int comp = (d + eidx) % 3;
    return facep(u, comp, b)[sh.s(q[0], q[1], q[2])] * radius(comp, q[2]);
  };
  int ok = 0, oj = 0;
  const int o[3] = {idx[0] - fi, idx[1] - fj, idx[2] - fk};
  if (d == 0) {
    ok = o[2];
    oj = o[1];
  } else if (d == 1) {
    ok = o[0];
    oj = o[2];
  } else {
    ok = o[1];
    oj = o[0];
  }
  auto sg = [](int off) -> double { return off == 0 ? -1.0 : 1.0; };
  double Uxx = 0.0, Vxyz = 0.0, Wxyz = 0.0;
  for (const int v : {0, 1})
    for (const int uu : {0, 2})
      for (const int tt : {0, 1}) {
        const double fine2 = fine(1, v, uu, tt) * 1.0;
        const double fine3 = fine(2, uu, v, tt) * 1.0;
        Uxx += sg(tt) * sg(uu) * (fine2 + fine3);
        Vxyz += sg(tt) * sg(uu) * sg(v) * fine2;
        Wxyz += sg(tt) * sg(uu) * sg(v) * fine3;
      }
  Uxx *= 0.125;
  // coarse Dx per axis; Dx along phi is 2 pi in cylindrical geometry (geometry.hpp:35-43)
  const double cdx = 2.0 * blk.dx[0], cdy = 2.0 * blk.dx[1];
  const double cdz = cyl ? kPhiSpan : 2.0 * blk.dx[2];
  const double dx2 = cdx * cdx, dy2 = cdy * cdy, dz2 = cdz * cdz;
  Vxyz *= 0.125 * dz2 / (dx2 + dz2);
  Wxyz *= 0.125 * dy2 / (dx2 + dy2);
  double x =
      0.5 * (fine(0, ok, oj, 0) * 1.0 + fine(0, ok, oj, 2) * 1.0) + Uxx + sg(ok) * Vxyz +
      sg(oj) * Wxyz;
  x *= cyl ? 1.0 / (radius(d, idx[0]) + kDblEps) : 1.0;
  return x;
}

void AmrGhostFaces(const km_cfg_t &cfg, const Tree &t, const km_pack_t &u) {
  (void)cfg;
  const Shape sh(u);
  const int lo[3] = {sh.is, sh.js, sh.ks};
  const int nleaf = (int)t.level.size();
  int bad = 0;
  // phase 1: copies, restrictions, shared prolongation
  for (int d = 0; d < 3; ++d) {
#pragma omp parallel for schedule(dynamic, 4) reduction(| : bad)
    for (int b = 0; b < nleaf; ++b) {
      const int L = t.level[b];
      const km_block_t &blk = u.blocks[b];
      for (int k = 0; k < sh.S[2]; ++k)
        for (int j = 0; j < sh.S[1]; ++j)
          for (int i = 0; i < sh.S[0]; ++i) {
            const int idx[3] = {i, j, k};
            if (!GhostFaceElem(t, sh, d, idx)) continue;
            int G[3] = {0, 0, 0};
            bool fl[3] = {false, false, false};
            for (int e = 0; e < t.ndim; ++e) {
              const int G0 = t.loc[3 * b + e] * t.nx[e] + (idx[e] - lo[e]);
              if (e == d && e == 0 && G0 < 0 && t.bc[0] == KM_BC_AXISYMMETRIC) {
                // axisymmetric: x1 faces mirror like cells (boundary_conditions.cpp:60)
                G[e] = -G0 - 1;
                fl[e] = !fl[e];
                continue;
              }
              G[e] = e == d ? map_face(G0, t.zones(L, e), t.bc[2 * e], t.bc[2 * e + 1], &fl[e])
                            : map_cell(G0, t.zones(L, e), t.bc[2 * e], t.bc[2 * e + 1], &fl[e]);
            }
            const bool up = d >= t.ndim || idx[d] <= lo[d] + sh.nx[d] - 1;
            FaceRef f;
            if (!FindFace(t, L, d, G, up, &f)) {
              bad |= 1;
              continue;
            }
            // cylindrical r and z faces prolongate r B: radii of the coarse elements
            const bool sr = t.cyl && d < 2;
            auto rc = [&](int C0) { return sr ? RadiusAt(t, C0, d == 1, 2.0 * blk.dx[0]) : 1.0; };
            double x;
            if (f.lev >= L) {
              x = FaceValue(t, sh, u, L, d, f);
            } else {
              if (d < t.ndim && (G[d] & 1)) continue;  // inside a coarse cell
              const double fc = FaceAt(sh, u, d, f.leaf, f.li) * rc(G[0] / 2);
              double a[3] = {0., 0., 0.};
              for (int e = 0; e < t.ndim; ++e) {
                if (e == d) continue;
                double fv[2];
                for (int side = 0; side < 2; ++side) {
                  int C[3] = {0, 0, 0};
                  bool nf[3] = {false, false, false};
                  for (int q = 0; q < t.ndim; ++q) C[q] = G[q] / 2;
                  C[e] += side ? 1 : -1;
                  const double rn = rc(C[0]);  // the neighbour's logical radius
                  C[e] = map_cell(C[e], t.zones(L - 1, e), t.bc[2 * e], t.bc[2 * e + 1], &nf[e]);
                  FaceRef s2;
                  if (!FindFace(t, L - 1, d, C, up, &s2) || s2.lev < L - 1) {
                    bad |= 1;
                    fv[side] = 0.;
                    continue;
                  }
                  fv[side] = FaceValue(t, sh, u, L - 1, d, s2) * rn;
                }
                const double dxc = 2.0 * blk.dx[e];
                a[e] = GradMinMod(fc, fv[0], fv[1], dxc, dxc) * (0.5 * blk.dx[e]);
              }
              const bool oi = d != 0 && (G[0] & 1), oj = d != 1 && t.ndim > 1 && (G[1] & 1),
                         ok = d != 2 && t.ndim > 2 && (G[2] & 1);
              if (!oi && !oj && !ok) x = fc - (a[0] + a[1] + a[2]);
              else if (oi && !oj && !ok) x = fc + (a[0] - a[1] - a[2]);
              else if (!oi && oj && !ok) x = fc - (a[0] - a[1] + a[2]);
              else if (oi && oj && !ok) x = fc + (a[0] + a[1] - a[2]);
              else if (!oi && !oj && ok) x = fc - (a[0] + a[1] - a[2]);
              else if (oi && !oj && ok) x = fc + (a[0] - a[1] + a[2]);
              else if (!oi && oj && ok) x = fc - (a[0] - a[1] - a[2]);
              else x = fc + (a[0] + a[1] + a[2]);
              if (sr) x *= 1.0 / (RadiusAt(t, G[0], d == 1, blk.dx[0]) + kDblEps);
            }
            if (d < t.ndim && fl[d]) x = -x;
            facep(u, d, b)[sh.s(k, j, i)] = x;
          }
    }
  }
  // phase 2: Toth & Roe for fine faces inside coarse cells
  for (int d = 0; d < 3; ++d) {
#pragma omp parallel for schedule(dynamic, 4) reduction(| : bad)
    for (int b = 0; b < nleaf; ++b) {
      const int L = t.level[b];
      const km_block_t &blk = u.blocks[b];
      for (int k = 0; k < sh.S[2]; ++k)
        for (int j = 0; j < sh.S[1]; ++j)
          for (int i = 0; i < sh.S[0]; ++i) {
            const int idx[3] = {i, j, k};
            if (!GhostFaceElem(t, sh, d, idx)) continue;
            if (d < t.ndim && ((idx[d] - lo[d]) % 2 == 0)) continue;
            int G[3] = {0, 0, 0};
            bool fl[3] = {false, false, false};
            for (int e = 0; e < t.ndim; ++e) {
              const int G0 = t.loc[3 * b + e] * t.nx[e] + (idx[e] - lo[e]);
              if (e == d && e == 0 && G0 < 0 && t.bc[0] == KM_BC_AXISYMMETRIC) {
                G[e] = -G0 - 1;
                fl[e] = !fl[e];
                continue;
              }
              G[e] = e == d ? map_face(G0, t.zones(L, e), t.bc[2 * e], t.bc[2 * e + 1], &fl[e])
                            : map_cell(G0, t.zones(L, e), t.bc[2 * e], t.bc[2 * e + 1], &fl[e]);
            }
            const bool up = d >= t.ndim || idx[d] <= lo[d] + sh.nx[d] - 1;
            FaceRef f;
            if (!FindFace(t, L, d, G, up, &f)) {
              bad |= 1;
              continue;
            }
            if (f.lev >= L) continue;
            // lower corner (fine) of the coarse cell around this element
            int c[3] = {0, 0, 0};
            for (int e = 0; e < t.ndim; ++e) {
              const int off = idx[e] - lo[e];
              c[e] = lo[e] + off - (((off % 2) + 2) % 2);
            }
            facep(u, d, b)[sh.s(k, j, i)] = TothRoe(sh, u, b, blk, d, c, idx, t.cyl, t.x1min);
          }
    }
  }
  if (bad) throw std::runtime_error("block-tree face fill met an unsupported level jump");
}

// EMF correction: an edge on a block's boundary that a finer leaf also has on its
// boundary (across a face, an edge or a corner of the block) takes the length-weighted
// RestrictAverage of the fine edges on it, so every block sharing the edge agrees.
void AmrEmfCorrect(const Tree &t, const km_pack_t &u) {
  const Shape sh(u);
  const int lo[3] = {sh.is, sh.js, sh.ks};
  const int nleaf = (int)t.level.size();
#pragma omp parallel for schedule(dynamic, 4)
  for (int b = 0; b < nleaf; ++b) {
    const int L = t.level[b];
    if (L >= t.lmax) continue;
    for (int te = (t.ndim < 3 ? 2 : 0); te < 3; ++te) {
      if (!u.edge[te]) continue;
      const int a0 = te == 0 ? 1 : 0, a1 = te == 2 ? 1 : 2;  // the two axes across
      const int ncell_t = te < t.ndim ? sh.nx[te] : 1;
      const int n0 = sh.nx[a0] + 1, n1 = a1 < t.ndim ? sh.nx[a1] + 1 : 1;
      for (int m1 = 0; m1 < n1; ++m1)
        for (int m0 = 0; m0 < n0; ++m0) {
          const bool onb = m0 == 0 || m0 == sh.nx[a0] ||
                           (a1 < t.ndim && (m1 == 0 || m1 == sh.nx[a1]));
          if (!onb) continue;
          for (int ct = 0; ct < ncell_t; ++ct) {
            int gn[3] = {0, 0, 0};  // level-L global node across, cell along te
            gn[a0] = t.loc[3 * b + a0] * t.nx[a0] + m0;
            if (a1 < t.ndim) gn[a1] = t.loc[3 * b + a1] * t.nx[a1] + m1;
            if (te < t.ndim) gn[te] = t.loc[3 * b + te] * t.nx[te] + ct;
            for (int q = 0; q < 4; ++q) {
              int c[3] = {gn[0], gn[1], gn[2]}, fnode[3] = {0, 0, 0};
              bool skip = false;
              const int axes[2] = {a0, a1};
              for (int m = 0; m < 2; ++m) {
                const int e = axes[m];
                if (e >= t.ndim) continue;
                const int below = (q >> m) & 1;
                const int N = t.zones(L, e);
                int ce = gn[e] - below;
                if (ce < 0 || ce >= N) {
                  if (t.bc[2 * e + (ce < 0 ? 0 : 1)] != KM_BC_PERIODIC) {
                    skip = true;
                    break;
                  }
                  ce = (ce + N) % N;
                }
                c[e] = ce;
                fnode[e] = 2 * (ce + below);
              }
              if (skip) continue;
              int zf[3] = {0, 0, 0};
              for (int e = 0; e < t.ndim; ++e) zf[e] = 2 * c[e];
              int lev = -1;
              const int fb = LeafOfCell(t, L + 1, zf, &lev);
              if (fb < 0 || lev != L + 1) continue;
              int fe[3] = {0, 0, 0};
              for (int e = 0; e < t.ndim; ++e) {
                if (e == te)
                  fe[e] = 2 * gn[e] - t.loc[3 * fb + e] * t.nx[e] + lo[e];
                else
                  fe[e] = fnode[e] - t.loc[3 * fb + e] * t.nx[e] + lo[e];
              }
              const bool inc[3] = {te == 0, te == 1 && t.ndim > 1, te == 2 && t.ndim > 2};
              // edge length: dx along the edge, |r| 2 pi for phi edges in cylindrical
              // geometry (geometry.hpp:143-150)
              const double len =
                  (t.cyl && te == 2) ? std::abs(RFace(u.blocks[fb], sh, fe[0])) * kPhiSpan
                                     : u.blocks[fb].dx[te];
              const double *E = edgep(u, te, fb);
              double vol[2][2][2], terms[2][2][2];
              for (int ok = 0; ok < 2; ++ok)
                for (int oj = 0; oj < 2; ++oj)
                  for (int oi = 0; oi < 2; ++oi) vol[ok][oj][oi] = terms[ok][oj][oi] = 0;
              for (int ok = 0; ok < 1 + inc[2]; ++ok)
                for (int oj = 0; oj < 1 + inc[1]; ++oj)
                  for (int oi = 0; oi < 1 + inc[0]; ++oi) {
                    vol[ok][oj][oi] = len;
                    terms[ok][oj][oi] =
                        vol[ok][oj][oi] * E[sh.s(fe[2] + ok, fe[1] + oj, fe[0] + oi)];
                  }
              const double tvol =
                  ((vol[0][0][0] + vol[0][1][0]) + (vol[0][0][1] + vol[0][1][1])) +
                  ((vol[1][0][0] + vol[1][1][0]) + (vol[1][0][1] + vol[1][1][1]));
              int ce3[3] = {0, 0, 0};
              ce3[a0] = lo[a0] + m0;
              if (a1 < t.ndim) ce3[a1] = lo[a1] + m1;
              if (te < t.ndim) ce3[te] = lo[te] + ct;
              edgep(u, te, b)[sh.s(ce3[2], ce3[1], ce3[0])] =
                  tvol > 0.0 ? (((terms[0][0][0] + terms[0][1][0]) +
                                 (terms[0][0][1] + terms[0][1][1])) +
                                ((terms[1][0][0] + terms[1][1][0]) +
                                 (terms[1][0][1] + terms[1][1][1]))) /
                                   tvol
                             : 0.0;
              break;
            }
          }
        }
    }
  }
}

void AmrFluxCorrect(const km_cfg_t &cfg, const Tree &t, const km_pack_t &u) {
  const Shape sh(u);
  const int nc = ncons(cfg.mhd) + KM_NSPEC_SLOTS(cfg.nspecies);  // every WithFluxes field
  const int st[3] = {sh.is, sh.js, sh.ks};
  for (size_t b = 0; b < t.level.size(); ++b) {
    const int L = t.level[b];
    if (L == t.lmax) continue;
    for (int dir = 0; dir < t.ndim; ++dir)
      for (int side = 0; side < 2; ++side) {
        // neighbour block position across this face at level L + 1 (finer candidates)
        const int Lf = L + 1;
        int nbpos = side ? 2 * (t.loc[3 * b + dir] + 1) : 2 * t.loc[3 * b + dir] - 1;
        const int nbmax = t.nbl(Lf, dir);
        if (nbpos < 0 || nbpos >= nbmax) {
          if (t.bc[2 * dir + side] != KM_BC_PERIODIC) continue;
          nbpos = (nbpos + nbmax) % nbmax;
        }
        // the 2^(ndim-1) fine blocks covering this face
        int ta = -1, tb = -1;
        for (int d = 0; d < t.ndim; ++d)
          if (d != dir) (ta < 0 ? ta : tb) = d;
        for (int q2 = 0; q2 < (tb >= 0 ? 2 : 1); ++q2)
          for (int q1 = 0; q1 < (ta >= 0 ? 2 : 1); ++q1) {
            int l[3] = {0, 0, 0};
            l[dir] = nbpos;
            if (ta >= 0) l[ta] = 2 * t.loc[3 * b + ta] + q1;
            if (tb >= 0) l[tb] = 2 * t.loc[3 * b + tb] + q2;
            const int fb = t.at[Lf][t.flat(Lf, l)];
            if (fb < 0) continue;  // not a leaf: same level or coarser behind this face
            const int fnorm = side ? st[dir] : st[dir] + sh.nx[dir];  // fine face index
            const int cnorm = side ? st[dir] + sh.nx[dir] : st[dir];  // coarse face index
            const int na = ta >= 0 ? sh.nx[ta] / 2 : 1, nbn = tb >= 0 ? sh.nx[tb] / 2 : 1;
            for (int jb = 0; jb < nbn; ++jb)
              for (int ja = 0; ja < na; ++ja) {
                int ci[3] = {0, 0, 0}, fi[3] = {0, 0, 0};
                ci[dir] = cnorm;
                fi[dir] = fnorm;
                if (ta >= 0) {
                  ci[ta] = st[ta] + q1 * (sh.nx[ta] / 2) + ja;
                  fi[ta] = st[ta] + 2 * ja;
                }
                if (tb >= 0) {
                  ci[tb] = st[tb] + q2 * (sh.nx[tb] / 2) + jb;
                  fi[tb] = st[tb] + 2 * jb;
                }
                const bool inc[3] = {dir != 0, dir != 1 && t.ndim > 1, dir != 2 && t.ndim > 2};
                const km_block_t &fbk = u.blocks[fb];
                for (int v = 0; v < nc; ++v) {
                  const double *F = fluxp(u, dir, fb, v);
                  double vol[2][2][2], terms[2][2][2];
                  for (int ok = 0; ok < 2; ++ok)
                    for (int oj = 0; oj < 2; ++oj)
                      for (int oi = 0; oi < 2; ++oi) vol[ok][oj][oi] = terms[ok][oj][oi] = 0;
                  for (int ok = 0; ok < 1 + inc[2]; ++ok)
                    for (int oj = 0; oj < 1 + inc[1]; ++oj)
                      for (int oi = 0; oi < 1 + inc[0]; ++oi) {
                        vol[ok][oj][oi] = FaceArea(t.cyl, fbk, sh, dir, fi[0] + oi);
                        terms[ok][oj][oi] =
                            vol[ok][oj][oi] * F[sh.c(fi[2] + ok, fi[1] + oj, fi[0] + oi)];
                      }
                  const double tvol =
                      ((vol[0][0][0] + vol[0][1][0]) + (vol[0][0][1] + vol[0][1][1])) +
                      ((vol[1][0][0] + vol[1][1][0]) + (vol[1][0][1] + vol[1][1][1]));
                  fluxp(u, dir, (int)b, v)[sh.c(ci[2], ci[1], ci[0])] =
                      tvol > 0.0 ? (((terms[0][0][0] + terms[0][1][0]) +
                                     (terms[0][0][1] + terms[0][1][1])) +
                                    ((terms[1][0][0] + terms[1][1][0]) +
                                     (terms[1][0][1] + terms[1][1][1]))) /
                                       tvol
                                 : 0.0;
                }
              }
          }
      }
  }
}

// ---------------------------------------------------------------------------------
// Uniform mesh metric: Parthenon's uniform mesh generator and UniformCartesian
// ---------------------------------------------------------------------------------
void MakeBlocks(const km_mesh_desc_t &m, std::vector<km_block_t> &blocks) {
  const int nbt = m.nb[0] * m.nb[1] * m.nb[2];
  blocks.resize(nbt);
  for (int b = 0; b < nbt; ++b) {
    const int l[3] = {b % m.nb[0], (b / m.nb[0]) % m.nb[1], b / (m.nb[0] * m.nb[1])};
    km_block_t &blk = blocks[b];
    for (int d = 0; d < 3; ++d) {
      double lo, hi;
      if (d < m.ndim) {
        const double xl = static_cast<double>(2 * l[d] - m.nb[d]) / (2 * m.nb[d]);
        const double xr = static_cast<double>(2 * (l[d] + 1) - m.nb[d]) / (2 * m.nb[d]);
        lo = 0.5 * (m.xmin[d] + m.xmax[d]) + (xl * m.xmax[d] - xl * m.xmin[d]);
        hi = 0.5 * (m.xmin[d] + m.xmax[d]) + (xr * m.xmax[d] - xr * m.xmin[d]);
        if (l[d] == 0) lo = m.xmin[d];
        if (l[d] == m.nb[d] - 1) hi = m.xmax[d];
      } else {
        lo = m.xmin[d];
        hi = m.xmax[d];
      }
      const int n = d < m.ndim ? m.nx[d] : 1;
      blk.xmin[d] = lo;
      blk.dx[d] = (hi - lo) / static_cast<double>(n);
      blk.lx[d] = l[d];
    }
    blk.area[0] = blk.dx[1] * blk.dx[2];
    blk.area[1] = blk.dx[0] * blk.dx[2];
    blk.area[2] = blk.dx[0] * blk.dx[1];
    blk.vol = blk.dx[0] * blk.dx[1] * blk.dx[2];
    blk.level = 0;
  }
}

// ---------------------------------------------------------------------------------
// Adaptive refinement. Tagging is kamayan's grid::AMRLoehner::operator()
// (grid_refinement.cpp:30-152), run by the RefinementTagging task of the last stage
// (kamayan_driver.cpp:178-181). The tag bookkeeping, the tree update and the data
// movement of a remesh belong to Parthenon (not vendored); they are restated here from
// Parthenon's published algorithm (MeshRefinement::SetRefinement,
// Mesh::UpdateMeshBlockTree with MeshBlockTree::Refine/Derefine,
// RedistributeAndRefineMeshBlocks, EvolutionDriver::InitializeBlockTimeStepsAfterAMR):
// parity unpinned against Parthenon itself; pinned by the invariants and hand-worked
// cases in tests/test_oracle_adapt.py.
// ---------------------------------------------------------------------------------

// largest num/den of the Loehner estimator over the interior of block b (:61-140):
//   d_p(z) = 0.5 (u(z+e_p) - u(z-e_p)) / dx_p   on the interior and one ghost layer
//   num    = sum_q sum_p [0.5 (d_p(z+e_q) - d_p(z-e_q)) / dx_q]^2
//   den    = 1e-12 + sum_q sum_p [0.5 (|d_p(z+e_q)| + |d_p(z-e_q)|) / dx_p
//            + filter (|u(z+e_q)| + |u(z-e_q)| + |u(z-e_p)| + |u(z+e_p)|) / (dx_q dx_p)]^2
// squares are x*x (std::pow(x, 2) rounds the same product once)
double LoehnerMaxErr2(const km_pack_t &u, int b, int var, double filter) {
  const Shape sh(u);
  const km_block_t &blk = u.blocks[b];
  const double *q = cellp(u, b, var);
  const long so[3] = {1, sh.E[0], (long)sh.E[0] * sh.E[1]};
  const int nd = sh.ndim;
  // first derivatives over interior +- 1 (the scratch field FirstDer)
  const long nc = sh.ncell();
  std::vector<double> der(3 * (size_t)nc, 0.0);
  const int k3 = nd > 2, k2 = nd > 1;
  for (int k = sh.ks - k3; k <= sh.ke + k3; ++k)
    for (int j = sh.js - k2; j <= sh.je + k2; ++j)
      for (int i = sh.is - 1; i <= sh.ie + 1; ++i) {
        const long c = sh.c(k, j, i);
        for (int p = 0; p < nd; ++p)
          der[p * nc + c] = 0.5 * (q[c + so[p]] - q[c - so[p]]) / blk.dx[p];
      }
  double mx = 0.0;
  for (int k = sh.ks; k <= sh.ke; ++k)
    for (int j = sh.js; j <= sh.je; ++j) {
      double row = 0.0;
      for (int i = sh.is; i <= sh.ie; ++i) {
        const long c = sh.c(k, j, i);
        double num = 0.0, den = 1.e-12;
        for (int qd = 0; qd < nd; ++qd)
          for (int p = 0; p < nd; ++p) {
            const double dp = der[p * nc + c + so[qd]], dm = der[p * nc + c - so[qd]];
            const double a = 0.5 * (dp - dm) / blk.dx[qd];
            num += a * a;
            const double e =
                0.5 * (std::fabs(dp) + std::fabs(dm)) / blk.dx[p] +
                filter *
                    (std::fabs(q[c + so[qd]]) + std::fabs(q[c - so[qd]]) +
                     std::fabs(q[c - so[p]]) + std::fabs(q[c + so[p]])) /
                    (blk.dx[qd] * blk.dx[p]);
            den += e * e;
          }
        row = den == 0.0 ? row : std::max(row, num / den);
      }
      mx = std::max(mx, row);
    }
  return mx;
}

void TagLoehner(const km_pack_t &u, const km_amr_criterion_t &cr, int32_t *tags, bool reset) {
#pragma omp parallel for schedule(dynamic, 1)
  for (int b = 0; b < u.nblocks; ++b) {
    if (reset) tags[b] = KM_AMR_DEREFINE;
    const double err = std::sqrt(LoehnerMaxErr2(u, b, cr.var, cr.filter));
    int flag = KM_AMR_SAME;
    if (err > cr.refine_tol && u.blocks[b].level < cr.max_level) flag = KM_AMR_REFINE;
    if (err < cr.derefine_tol) flag = KM_AMR_DEREFINE;
    tags[b] = std::max(tags[b], flag);
  }
}

// ---- the leaf set ----
struct LeafSet {
  int ndim, nb0[3], bc[6];
  std::vector<std::vector<char>> leaf;  // leaf[L][flat] at each level 0..Lcap
  int nbl(int L, int d) const { return d < ndim ? (nb0[d] << L) : 1; }
  long flat(int L, const int l[3]) const {
    return ((long)l[2] * nbl(L, 1) + l[1]) * nbl(L, 0) + l[0];
  }
  bool is_leaf(int L, const int l[3]) const { return leaf[L][flat(L, l)] != 0; }
  void set(int L, const int l[3], bool on) { leaf[L][flat(L, l)] = on ? 1 : 0; }
  // position l + o at level L through periodic boundaries; false outside the domain
  bool shift(int L, const int l[3], const int o[3], int out[3]) const {
    for (int d = 0; d < 3; ++d) {
      out[d] = l[d] + o[d];
      if (d >= ndim) {
        out[d] = 0;
        continue;
      }
      const int n = nbl(L, d);
      if (out[d] < 0 || out[d] >= n) {
        if (bc[2 * d] != KM_BC_PERIODIC) return false;
        out[d] = (out[d] + n) % n;
      }
    }
    return true;
  }
  // level of the leaf holding node (L, l), or L + 1 if the node is split
  int cover(int L, const int l[3]) const {
    for (int Lq = L; Lq >= 0; --Lq) {
      int a[3];
      for (int d = 0; d < 3; ++d) a[d] = l[d] >> (L - Lq);
      if (is_leaf(Lq, a)) return Lq;
    }
    return L + 1;
  }
};

void ForEachDir(int ndim, const std::function<void(const int o[3])> &f) {
  for (int oz = (ndim > 2 ? -1 : 0); oz <= (ndim > 2 ? 1 : 0); ++oz)
    for (int oy = (ndim > 1 ? -1 : 0); oy <= (ndim > 1 ? 1 : 0); ++oy)
      for (int ox = -1; ox <= 1; ++ox) {
        if (!ox && !oy && !oz) continue;
        const int o[3] = {ox, oy, oz};
        f(o);
      }
}

// MeshBlockTree::Refine: coarser leaves touching (L, l) split first
void RefineLeaf(LeafSet &s, int L, const int l[3], int &applied) {
  if (!s.is_leaf(L, l)) return;
  ForEachDir(s.ndim, [&](const int o[3]) {
    int n[3];
    if (!s.shift(L, l, o, n)) return;
    const int cl = s.cover(L, n);
    if (cl < L) {
      int a[3];
      for (int d = 0; d < 3; ++d) a[d] = n[d] >> (L - cl);
      RefineLeaf(s, cl, a, applied);
    }
  });
  ++applied;
  s.set(L, l, false);
  for (int c = 0; c < (1 << s.ndim); ++c) {
    int ch[3];
    for (int d = 0; d < 3; ++d) ch[d] = d < s.ndim ? 2 * l[d] + ((c >> d) & 1) : 0;
    s.set(L + 1, ch, true);
  }
}

// MeshBlockTree::Derefine of parent (L, l): its children must all be leaves and no
// split child of a neighbouring node may touch it
void DerefineParent(LeafSet &s, int L, const int l[3], int &applied) {
  const int nch = 1 << s.ndim;
  for (int c = 0; c < nch; ++c) {
    int ch[3];
    for (int d = 0; d < 3; ++d) ch[d] = d < s.ndim ? 2 * l[d] + ((c >> d) & 1) : 0;
    if (!s.is_leaf(L + 1, ch)) return;
  }
  bool ok = true;
  ForEachDir(s.ndim, [&](const int o[3]) {
    int n[3];
    if (!ok || !s.shift(L, l, o, n)) return;
    if (s.cover(L, n) <= L) return;  // a leaf at level <= L
    // node n is split: its children on the side facing (L, l)
    for (int c = 0; c < nch; ++c) {
      int ch[3];
      bool touch = true;
      for (int d = 0; d < 3; ++d) {
        const int bit = (c >> d) & 1;
        if (d >= s.ndim) {
          ch[d] = 0;
          continue;
        }
        if (o[d] == -1 && bit != 1) touch = false;
        if (o[d] == 1 && bit != 0) touch = false;
        ch[d] = 2 * n[d] + bit;
      }
      if (!touch) continue;
      if (s.cover(L + 1, ch) > L + 1) ok = false;
    }
  });
  if (!ok) return;
  for (int c = 0; c < nch; ++c) {
    int ch[3];
    for (int d = 0; d < 3; ++d) ch[d] = d < s.ndim ? 2 * l[d] + ((c >> d) & 1) : 0;
    s.set(L + 1, ch, false);
  }
  s.set(L, l, true);
  ++applied;
}

// Morton key of a leaf's lower corner on level `lev` (x in the lowest bit)
uint64_t MortonKey(int ndim, int L, const int l[3], int lev) {
  uint64_t key = 0;
  int c[3];
  for (int d = 0; d < 3; ++d) c[d] = d < ndim ? (l[d] << (lev - L)) : 0;
  for (int bit = 20; bit >= 0; --bit)
    for (int d = ndim - 1; d >= 0; --d) key = (key << 1) | ((c[d] >> bit) & 1);
  return key;
}

// SetRefinement + UpdateMeshBlockTree; returns true when the tree changed
bool UpdateTree(const km_mesh_desc_t &m, int numlevel, int threshold,
                const std::vector<int32_t> &leaves, const int32_t *tags,
                std::vector<int32_t> &deref, std::vector<int32_t> &out_leaves,
                std::vector<int32_t> &out_deref, std::vector<int32_t> &origin) {
  const int n = (int)leaves.size() / 4;
  const int maxlev = numlevel - 1;
  LeafSet s;
  s.ndim = m.ndim;
  for (int d = 0; d < 3; ++d) s.nb0[d] = d < m.ndim ? m.nb[d] : 1;
  for (int d = 0; d < 6; ++d) s.bc[d] = m.bc[d];
  s.leaf.assign(maxlev + 2, {});
  for (int L = 0; L <= maxlev + 1; ++L)
    s.leaf[L].assign((size_t)s.nbl(L, 0) * s.nbl(L, 1) * s.nbl(L, 2), 0);
  for (int b = 0; b < n; ++b) {
    const int l[3] = {leaves[4 * b + 1], leaves[4 * b + 2], leaves[4 * b + 3]};
    s.set(leaves[4 * b], l, true);
  }
  // MeshRefinement::SetRefinement per leaf
  std::vector<int> flag(n, 0);
  for (int b = 0; b < n; ++b) {
    const int L = leaves[4 * b];
    const int l[3] = {leaves[4 * b + 1], leaves[4 * b + 2], leaves[4 * b + 3]};
    const int aret = std::max(-1, (int)tags[b]);
    if (aret >= 0) deref[b] = 0;
    if (aret > 0) {
      flag[b] = L == maxlev ? 0 : 1;
    } else if (aret < 0) {
      if (L == 0) {
        deref[b] = 0;
      } else {
        deref[b] += 1;
        bool finer = false;
        ForEachDir(s.ndim, [&](const int o[3]) {
          int nn[3];
          if (s.shift(L, l, o, nn) && s.cover(L, nn) > L) finer = true;
        });
        flag[b] = (!finer && deref[b] >= threshold) ? -1 : 0;
      }
    }
  }
  // UpdateMeshBlockTree
  const int nleaf = 1 << m.ndim;
  int nref = 0, nderef = 0;
  for (int b = 0; b < n; ++b) {
    nref += flag[b] == 1;
    nderef += flag[b] == -1;
  }
  std::vector<std::array<int, 4>> ref, par;
  for (int b = 0; b < n; ++b)
    if (flag[b] == 1) ref.push_back({leaves[4 * b], leaves[4 * b + 1], leaves[4 * b + 2], leaves[4 * b + 3]});
  if (nderef >= nleaf) {
    // parents all of whose children are leaves flagged -1
    std::vector<std::vector<char>> fl(maxlev + 2);
    for (int L = 0; L <= maxlev + 1; ++L) fl[L].assign(s.leaf[L].size(), 0);
    for (int b = 0; b < n; ++b)
      if (flag[b] == -1) {
        const int l[3] = {leaves[4 * b + 1], leaves[4 * b + 2], leaves[4 * b + 3]};
        fl[leaves[4 * b]][s.flat(leaves[4 * b], l)] = 1;
      }
    for (int b = 0; b < n; ++b) {
      if (flag[b] != -1) continue;
      const int L = leaves[4 * b];
      const int l[3] = {leaves[4 * b + 1], leaves[4 * b + 2], leaves[4 * b + 3]};
      bool first = true;
      for (int d = 0; d < m.ndim; ++d)
        if (l[d] & 1) first = false;
      if (!first) continue;
      bool all = true;
      for (int c = 0; c < nleaf; ++c) {
        int ch[3];
        for (int d = 0; d < 3; ++d) ch[d] = d < m.ndim ? l[d] + ((c >> d) & 1) : 0;
        if (!fl[L][s.flat(L, ch)]) all = false;
      }
      if (all) par.push_back({L - 1, l[0] >> 1, m.ndim > 1 ? l[1] >> 1 : 0, m.ndim > 2 ? l[2] >> 1 : 0});
    }
  }
  int applied = 0;
  std::stable_sort(ref.begin(), ref.end(), [](const auto &a, const auto &b) { return a[0] < b[0]; });
  std::stable_sort(par.begin(), par.end(), [](const auto &a, const auto &b) { return a[0] > b[0]; });
  for (const auto &r : ref) {
    const int l[3] = {r[1], r[2], r[3]};
    RefineLeaf(s, r[0], l, applied);
  }
  for (const auto &p : par) {
    const int l[3] = {p[1], p[2], p[3]};
    DerefineParent(s, p[0], l, applied);
  }
  if (applied == 0) {
    out_leaves = leaves;
    out_deref = deref;
    origin.assign(9 * (size_t)n, -1);
    for (int b = 0; b < n; ++b) {
      origin[9 * b] = 0;
      origin[9 * b + 1] = b;
    }
    return false;
  }
  // new leaves in Morton order; origins against the old leaves
  struct NL {
    uint64_t key;
    int L, l[3];
  };
  std::vector<NL> nl;
  for (int L = 0; L <= maxlev + 1; ++L)
    for (int k = 0; k < s.nbl(L, 2); ++k)
      for (int j = 0; j < s.nbl(L, 1); ++j)
        for (int i = 0; i < s.nbl(L, 0); ++i) {
          const int l[3] = {i, j, k};
          if (s.is_leaf(L, l)) nl.push_back({MortonKey(m.ndim, L, l, maxlev + 1), L, {i, j, k}});
        }
  std::sort(nl.begin(), nl.end(), [](const NL &a, const NL &b) { return a.key < b.key; });
  // old leaf index by (level, position)
  std::vector<std::vector<int>> old(maxlev + 2);
  for (int L = 0; L <= maxlev + 1; ++L) old[L].assign(s.leaf[L].size(), -1);
  for (int b = 0; b < n; ++b) {
    const int l[3] = {leaves[4 * b + 1], leaves[4 * b + 2], leaves[4 * b + 3]};
    old[leaves[4 * b]][s.flat(leaves[4 * b], l)] = b;
  }
  out_leaves.clear();
  out_deref.clear();
  origin.assign(9 * nl.size(), -1);
  for (size_t i = 0; i < nl.size(); ++i) {
    const NL &e = nl[i];
    out_leaves.insert(out_leaves.end(), {e.L, e.l[0], e.l[1], e.l[2]});
    const int ob = old[e.L][s.flat(e.L, e.l)];
    int *o = &origin[9 * i];
    if (ob >= 0) {
      o[0] = 0;
      o[1] = ob;
      out_deref.push_back(deref[ob]);
      continue;
    }
    out_deref.push_back(0);
    if (e.L > 0) {
      const int pl[3] = {e.l[0] >> 1, m.ndim > 1 ? e.l[1] >> 1 : 0, m.ndim > 2 ? e.l[2] >> 1 : 0};
      const int pb = old[e.L - 1][s.flat(e.L - 1, pl)];
      if (pb >= 0) {
        o[0] = 1;
        o[1] = pb;
        continue;
      }
    }
    o[0] = 2;
    for (int c = 0; c < nleaf; ++c) {
      int ch[3];
      for (int d = 0; d < 3; ++d) ch[d] = d < m.ndim ? 2 * e.l[d] + ((c >> d) & 1) : 0;
      o[1 + c] = old[e.L + 1][s.flat(e.L + 1, ch)];
      if (o[1 + c] < 0) throw std::runtime_error("remesh: new leaf with no source");
    }
  }
  return true;
}

// RedistributeAndRefineMeshBlocks data movement for cell slots [v0, v1)
void Remesh(const km_pack_t &o, const km_pack_t &nw, const int32_t *origin, int v0, int v1,
            bool cyl) {
  const Shape sh(nw);
  const int st[3] = {sh.is, sh.js, sh.ks};
  const int nd = sh.ndim;
  const long so[3] = {1, sh.E[0], (long)sh.E[0] * sh.E[1]};
#pragma omp parallel for schedule(dynamic, 1)
  for (int b = 0; b < nw.nblocks; ++b) {
    const int32_t *org = origin + 9 * b;
    const km_block_t &blk = nw.blocks[b];
    for (int v = v0; v < v1; ++v) {
      double *dst = cellp(nw, b, v);
      if (org[0] == 0) {
        std::memcpy(dst, cellp(o, org[1], v), sh.ncell() * sizeof(double));
        continue;
      }
      for (int k = sh.ks; k <= sh.ke; ++k)
        for (int j = sh.js; j <= sh.je; ++j)
          for (int i = sh.is; i <= sh.ie; ++i) {
            const int idx[3] = {i, j, k};
            double x;
            if (org[0] == 1) {
              // ProlongateSharedMinMod from the parent's zones
              const km_block_t &pb = o.blocks[org[1]];
              const double *q = cellp(o, org[1], v);
              int pc[3] = {0, 0, 0}, par[3] = {0, 0, 0};
              for (int d = 0; d < nd; ++d) {
                const int G = blk.lx[d] * sh.nx[d] + (idx[d] - st[d]);
                par[d] = G & 1;
                pc[d] = (G >> 1) - pb.lx[d] * sh.nx[d] + st[d];
              }
              const long cc = sh.c(pc[2], pc[1], pc[0]);
              const double fc = q[cc];
              double g[3] = {0., 0., 0.}, dxf[3] = {0., 0., 0.};
              for (int d = 0; d < nd; ++d) {
                const double dxc = 2.0 * blk.dx[d];
                g[d] = GradMinMod(fc, q[cc - so[d]], q[cc + so[d]], dxc, dxc);
                dxf[d] = 0.5 * blk.dx[d];
              }
              const bool oi = par[0], oj = par[1], ok = par[2];
              const double a = g[0] * dxf[0], bb = g[1] * dxf[1], c3 = g[2] * dxf[2];
              if (!oi && !oj && !ok) x = fc - (a + bb + c3);
              else if (oi && !oj && !ok) x = fc + (a - bb - c3);
              else if (!oi && oj && !ok) x = fc - (a - bb + c3);
              else if (oi && oj && !ok) x = fc + (a + bb - c3);
              else if (!oi && !oj && ok) x = fc - (a + bb - c3);
              else if (oi && !oj && ok) x = fc + (a - bb + c3);
              else if (!oi && oj && ok) x = fc - (a - bb - c3);
              else x = fc + (a + bb + c3);
            } else {
              // RestrictAverage of the child holding the 2^d fine zones
              int c = 0, f[3] = {0, 0, 0};
              for (int d = 0; d < nd; ++d) {
                const int lc = idx[d] - st[d];
                const int half = (2 * lc) / sh.nx[d];
                c |= half << d;
                f[d] = 2 * lc - half * sh.nx[d] + st[d];
              }
              const int cb = org[1 + c];
              const double *q = cellp(o, cb, v);
              double V[2][2][2], T[2][2][2];
              for (int ok = 0; ok < 2; ++ok)
                for (int oj = 0; oj < 2; ++oj)
                  for (int oi = 0; oi < 2; ++oi) {
                    V[ok][oj][oi] = T[ok][oj][oi] = 0.0;
                    if ((ok == 0 || nd > 2) && (oj == 0 || nd > 1)) {
                      const double vol = CellVol(cyl, o.blocks[cb], sh, f[0] + oi);
                      V[ok][oj][oi] = vol;
                      T[ok][oj][oi] = vol * q[sh.c(f[2] + ok, f[1] + oj, f[0] + oi)];
                    }
                  }
              const double tvol = ((V[0][0][0] + V[0][1][0]) + (V[0][0][1] + V[0][1][1])) +
                                  ((V[1][0][0] + V[1][1][0]) + (V[1][0][1] + V[1][1][1]));
              x = tvol > 0.0 ? (((T[0][0][0] + T[0][1][0]) + (T[0][0][1] + T[0][1][1])) +
                                ((T[1][0][0] + T[1][1][0]) + (T[1][0][1] + T[1][1][1]))) /
                                   tvol
                             : 0.0;
            }
            dst[sh.c(k, j, i)] = x;
          }
    }
  }
}

// The CT face field's share of RedistributeAndRefineMeshBlocks, written as Parthenon's
// refinement operators loop: a kept leaf copies its face arrays; a split leaf's child
// runs ProlongateSharedMinMod (refinement_operations.hpp:123-268) over the parent's faces
// in its quadrant -- each coarse face writes the fine faces on it, with minmod slopes
// across the face from the parent's neighbouring faces -- then
// ProlongateInternalTothAndRoe over each coarse cell of the quadrant; a merged parent
// takes RestrictAverage (:52-121) of each child's faces over the child's quadrant, the
// children in x1-fastest order, so the upper child writes a face two children share.
void RemeshFaces(const km_pack_t &o, const km_pack_t &nw, const int32_t *origin, bool cyl,
                 double x1min) {
  Tree tr;  // the radii of the r-weighted operators
  tr.cyl = cyl;
  tr.x1min = x1min;
  const Shape sh(nw);
  const int nd = sh.ndim;
  const int lo[3] = {sh.is, sh.js, sh.ks};
  const int n[3] = {sh.nx[0], nd > 1 ? sh.nx[1] : 1, nd > 2 ? sh.nx[2] : 1};
  const int hn[3] = {n[0] / 2, nd > 1 ? n[1] / 2 : 1, nd > 2 ? n[2] / 2 : 1};
  const long ss[3] = {1, sh.S[0], (long)sh.S[0] * sh.S[1]};
#pragma omp parallel for schedule(dynamic, 1)
  for (int b = 0; b < nw.nblocks; ++b) {
    const int32_t *org = origin + 9 * b;
    const km_block_t &blk = nw.blocks[b];
    if (org[0] == 0) {
      for (int d = 0; d < 3; ++d)
        std::memcpy(facep(nw, d, b), facep(o, d, org[1]), sh.nstag() * sizeof(double));
      continue;
    }
    if (org[0] == 1) {
      const int pb = org[1];
      int q0[3] = {0, 0, 0};  // the child's quadrant: first parent cell, parent-local
      for (int e = 0; e < nd; ++e) q0[e] = lo[e] + (blk.lx[e] & 1) * hn[e];
      for (int d = 0; d < 3; ++d) {
        const double *q = facep(o, d, pb);
        double *f = facep(nw, d, b);
        const int cn[3] = {hn[0] + (d == 0), hn[1] + (d == 1 && nd > 1),
                           hn[2] + (d == 2 && nd > 2)};
        for (int ck = 0; ck < cn[2]; ++ck)
          for (int cj = 0; cj < cn[1]; ++cj)
            for (int ci = 0; ci < cn[0]; ++ci) {
              const int cc[3] = {ci, cj, ck};
              const long ps = sh.s(q0[2] + ck, q0[1] + cj, q0[0] + ci);
              // cylindrical r and z faces prolongate r B (global coarse index along x1)
              const bool sr = cyl && d < 2;
              const int C0 = blk.lx[0] * hn[0] + ci;
              auto rc = [&](int Cq) { return sr ? RadiusAt(tr, Cq, d == 1, 2.0 * blk.dx[0]) : 1.0; };
              const double fc = q[ps] * rc(C0);
              double a[3] = {0., 0., 0.};
              for (int e = 0; e < nd; ++e) {
                if (e == d) continue;
                const double dxc = 2.0 * blk.dx[e];
                const double fm = q[ps - ss[e]] * rc(C0 - (e == 0)),
                             fp = q[ps + ss[e]] * rc(C0 + (e == 0));
                a[e] = GradMinMod(fc, fm, fp, dxc, dxc) * (0.5 * blk.dx[e]);
              }
              const int ni = (d != 0) ? 2 : 1, nj = (d != 1 && nd > 1) ? 2 : 1,
                        nk = (d != 2 && nd > 2) ? 2 : 1;
              for (int ok = 0; ok < nk; ++ok)
                for (int oj = 0; oj < nj; ++oj)
                  for (int oi = 0; oi < ni; ++oi) {
                    double x;
                    if (!oi && !oj && !ok) x = fc - (a[0] + a[1] + a[2]);
                    else if (oi && !oj && !ok) x = fc + (a[0] - a[1] - a[2]);
                    else if (!oi && oj && !ok) x = fc - (a[0] - a[1] + a[2]);
                    else if (oi && oj && !ok) x = fc + (a[0] + a[1] - a[2]);
                    else if (!oi && !oj && ok) x = fc - (a[0] + a[1] - a[2]);
                    else if (oi && !oj && ok) x = fc + (a[0] - a[1] + a[2]);
                    else if (!oi && oj && ok) x = fc - (a[0] - a[1] - a[2]);
                    else x = fc + (a[0] + a[1] + a[2]);
                    int fi[3] = {0, 0, 0};
                    const int off[3] = {oi, oj, ok};
                    for (int e = 0; e < nd; ++e) fi[e] = lo[e] + 2 * cc[e] + off[e];
                    if (sr) {
                      const int Gf = blk.lx[0] * n[0] + (fi[0] - lo[0]);
                      x *= 1.0 / (RadiusAt(tr, Gf, d == 1, blk.dx[0]) + kDblEps);
                    }
                    f[sh.s(fi[2], fi[1], fi[0])] = x;
                  }
            }
      }
      // Toth & Roe over every coarse cell of the quadrant (all shared faces are in place)
      for (int d = 0; d < nd; ++d)
        for (int ck = 0; ck < hn[2]; ++ck)
          for (int cj = 0; cj < hn[1]; ++cj)
            for (int ci = 0; ci < hn[0]; ++ci) {
              const int c[3] = {lo[0] + 2 * ci, nd > 1 ? lo[1] + 2 * cj : 0,
                                nd > 2 ? lo[2] + 2 * ck : 0};
              const int na = (d != 0 ? 2 : 1), nb2 = (d != 1 && nd > 1 ? 2 : 1),
                        nc2 = (d != 2 && nd > 2 ? 2 : 1);
              for (int ok = 0; ok < nc2; ++ok)
                for (int oj = 0; oj < nb2; ++oj)
                  for (int oi = 0; oi < na; ++oi) {
                    int idx[3] = {c[0] + oi, c[1] + oj, c[2] + ok};
                    idx[d] = c[d] + 1;  // the face inside the coarse cell
                    facep(nw, d, b)[sh.s(idx[2], idx[1], idx[0])] =
                        TothRoe(sh, nw, b, blk, d, c, idx, cyl, x1min);
                  }
            }
      continue;
    }
    // merged parent: each child restricts its faces into the parent's quadrant
    for (int ch = 0; ch < (1 << nd); ++ch) {
      const int cb = org[1 + ch];
      int q0[3] = {0, 0, 0};
      for (int e = 0; e < nd; ++e) q0[e] = lo[e] + ((ch >> e) & 1) * hn[e];
      for (int d = 0; d < 3; ++d) {
        const double *q = facep(o, d, cb);
        const bool inc[3] = {d != 0, d != 1 && nd > 1, d != 2 && nd > 2};
        const int cn[3] = {hn[0] + (d == 0), hn[1] + (d == 1 && nd > 1),
                           hn[2] + (d == 2 && nd > 2)};
        for (int ck = 0; ck < cn[2]; ++ck)
          for (int cj = 0; cj < cn[1]; ++cj)
            for (int ci = 0; ci < cn[0]; ++ci) {
              const int cc[3] = {ci, cj, ck};
              int fl[3] = {0, 0, 0};
              for (int e = 0; e < nd; ++e) fl[e] = lo[e] + 2 * cc[e];
              double V[2][2][2], T[2][2][2];
              for (int ok = 0; ok < 2; ++ok)
                for (int oj = 0; oj < 2; ++oj)
                  for (int oi = 0; oi < 2; ++oi) {
                    V[ok][oj][oi] = T[ok][oj][oi] = 0.0;
                    if ((ok == 0 || inc[2]) && (oj == 0 || inc[1]) && (oi == 0 || inc[0])) {
                      const double A = FaceArea(cyl, o.blocks[cb], sh, d, fl[0] + oi);
                      V[ok][oj][oi] = A;
                      T[ok][oj][oi] = A * q[sh.s(fl[2] + ok, fl[1] + oj, fl[0] + oi)];
                    }
                  }
              const double tvol = ((V[0][0][0] + V[0][1][0]) + (V[0][0][1] + V[0][1][1])) +
                                  ((V[1][0][0] + V[1][1][0]) + (V[1][0][1] + V[1][1][1]));
              const double x =
                  tvol > 0.0 ? (((T[0][0][0] + T[0][1][0]) + (T[0][0][1] + T[0][1][1])) +
                                ((T[1][0][0] + T[1][1][0]) + (T[1][0][1] + T[1][1][1]))) /
                                   tvol
                             : 0.0;
              facep(nw, d, b)[sh.s(q0[2] + ck, q0[1] + cj, q0[0] + ci)] = x;
            }
      }
    }
  }
}

// ---------------------------------------------------------------------------------
// Whole-mesh driver: KamayanDriver::MakeTaskCollection / BuildTaskListRKStage
// (kamayan_driver.cpp:121-248) over a single MeshData partition.
// ---------------------------------------------------------------------------------
struct Container {
  std::vector<std::vector<double>> cell;  // owned cell arrays (some slots unused)
  std::vector<std::vector<double>> face[3], edge[3], flux[3];
  std::vector<double *> cell_tab, flux_tab[3], face_tab[3], edge_tab[3];
  km_pack_t pack;
};

}  // namespace

struct km_cpu_sim_t {
  km_cfg_t cfg;
  km_mesh_desc_t mesh;
  std::vector<km_block_t> blocks;
  Grid grid;
  bool amr = false;  // block-tree mesh
  Tree tree;
  int nblocks = 0;
  int nvar, nc, nstages;
  int nflux;  // nc plus one per species (more than one species)
  Container base, stage[2], dudt;
  double time = 0., dt = 0., dt_est = 0.;
  int ncycle = 0;
  bool dt_override = false;
  // failure: a non-physical zone seen by a CFL pass (bit 1), dt below dt_min (2) or above
  // dt_max (4) for their cycle limits; below/above count the cycles in a row
  int fail = 0, below = 0, above = 0;
  // adaptive refinement (block-tree meshes)
  bool adaptive = false;
  km_amr_adapt_t adapt{};
  std::vector<int32_t> leaves, deref;
};

namespace {

void MakeContainer(km_cpu_sim_t &s, Container &c, bool full, Container *share_prims,
                   bool with_flux) {
  const km_mesh_desc_t &m = s.mesh;
  const int nbt = s.nblocks;
  km_pack_t &p = c.pack;
  std::memset(&p, 0, sizeof(p));
  p.nblocks = nbt;
  p.ndim = m.ndim;
  p.ng = m.ng;
  for (int d = 0; d < 3; ++d) p.nx[d] = d < m.ndim ? m.nx[d] : 1;
  p.nvar = s.nvar;
  p.nflux = s.nflux;
  p.blocks = s.blocks.data();
  Shape sh(p);
  // stage containers own the conserved fields (hydro conserved slots and mass fractions)
  const int nhydro = KM_NVAR(s.cfg.mhd);
  c.cell.assign((size_t)nbt * s.nvar, {});
  c.cell_tab.assign((size_t)nbt * s.nvar, nullptr);
  for (int b = 0; b < nbt; ++b)
    for (int v = 0; v < s.nvar; ++v) {
      const size_t t = (size_t)b * s.nvar + v;
      if (v < s.nc || v >= nhydro || full) {
        c.cell[t].assign(sh.ncell(), 0.0);
        c.cell_tab[t] = c.cell[t].data();
      } else if (share_prims) {
        c.cell_tab[t] = share_prims->cell_tab[t];
      }
    }
  p.cell = c.cell_tab.data();
  if (with_flux) {
    for (int d = 0; d < 3; ++d) {
      c.flux[d].assign((size_t)nbt * s.nflux, std::vector<double>(sh.ncell(), 0.0));
      c.flux_tab[d].resize((size_t)nbt * s.nflux);
      for (size_t t = 0; t < c.flux_tab[d].size(); ++t) c.flux_tab[d][t] = c.flux[d][t].data();
      p.flux[d] = c.flux_tab[d].data();
    }
  }
  if (s.cfg.mhd == KM_MHD_CT) {
    for (int d = 0; d < 3; ++d) {
      c.face[d].assign(nbt, std::vector<double>(sh.nstag(), 0.0));
      c.face_tab[d].resize(nbt);
      for (int b = 0; b < nbt; ++b) c.face_tab[d][b] = c.face[d][b].data();
      p.face[d] = c.face_tab[d].data();
      if (with_flux) {
        c.edge[d].assign(nbt, std::vector<double>(sh.nstag(), 0.0));
        c.edge_tab[d].resize(nbt);
        for (int b = 0; b < nbt; ++b) c.edge_tab[d][b] = c.edge[d][b].data();
        p.edge[d] = c.edge_tab[d].data();
      }
    }
  }
}

// Ghost fill of md1 after a stage: every FillGhost field (fields.hpp:160-161,
// hydro.cpp:102-105): conserved, VELOCITY, PRES, BMOD, EINT, and face B; TEMP is not
// FillGhost (eos.cpp:67).
void FillGhosts(km_cpu_sim_t &s, const km_pack_t &p) {
  const int nh = KM_NVAR(s.cfg.mhd);  // mass fractions (FillGhost) follow at nh..nvar-1
  if (s.amr) {
    AmrGhostFill(s.cfg, s.tree, p, 0, KM_TEMP(s.cfg.mhd));
    if (s.nvar > nh) AmrGhostFill(s.cfg, s.tree, p, nh, s.nvar);
    if (s.cfg.mhd == KM_MHD_CT) AmrGhostFaces(s.cfg, s.tree, p);
  } else {
    GhostFill(s.cfg, s.grid, p, 0, KM_TEMP(s.cfg.mhd), true);
    if (s.nvar > nh) GhostFill(s.cfg, s.grid, p, nh, s.nvar, false);
  }
}

const double kBetaRK1[1] = {1.0};
const double kBetaRK2[2] = {1.0, 0.5};
const double kBetaRK3[3] = {1.0, 0.25, 2.0 / 3.0};

void Cycle(km_cpu_sim_t &s) {
  const double *betas = s.nstages == 1 ? kBetaRK1 : s.nstages == 2 ? kBetaRK2 : kBetaRK3;
  // stage_name: base, 1, .., n-1, base
  auto container = [&](int idx) -> Container & {
    if (idx == 0 || idx == s.nstages) return s.base;
    return s.stage[idx - 1];
  };
  const double dt = s.dt;
  for (int stage = 1; stage <= s.nstages; ++stage) {
    Container &md0 = container(stage - 1);
    Container &md1 = container(stage);
    // md0 needs flux arrays: point md0's flux/edge tables at the shared scratch
    km_pack_t p0 = md0.pack;
    for (int d = 0; d < 3; ++d) {
      p0.flux[d] = s.dudt.flux_tab[d].data();
      p0.edge[d] = s.cfg.mhd == KM_MHD_CT ? s.dudt.edge_tab[d].data() : nullptr;
    }
    CalculateFluxes(s.cfg, p0);
    CalculateEMF(s.cfg, p0);
    if (s.amr) {
      AmrFluxCorrect(s.cfg, s.tree, p0);
      if (s.cfg.mhd == KM_MHD_CT) AmrEmfCorrect(s.tree, p0);
    }
    FluxesToDuDt(s.cfg, p0, s.dudt.pack);
    AddSourceTerms(s.cfg, p0, s.dudt.pack);  // AddTasksOneStep
    StagePrepareConserved(s.cfg, p0);         // PrepareConserved on md0
    ApplyDuDt(s.cfg, s.base.pack, p0, md1.pack, s.dudt.pack, betas[stage - 1], dt);
    PreparePrimitive(s.cfg, md1.pack);
    if (stage == s.nstages) {
      s.dt_est = EstimateDt(s.cfg, s.base.pack);
      if (AnyNonPhysical(s.cfg, s.base.pack)) s.fail |= 1;
      FillDivB(s.cfg, s.base.pack);
    }
    FillGhosts(s, md1.pack);
  }
  s.time += dt;
  s.ncycle += 1;
}

// Tagging of the filled base state, tree update and, when the tree changed, the remesh,
// ghost fill and dt estimate on the new mesh (EvolutionDriver::Execute order: Step,
// LoadBalancingAndAdaptiveRefinement, InitializeBlockTimeStepsAfterAMR, SetGlobalTimeStep)
void AdaptStep(km_cpu_sim_t &s) {
  std::vector<int32_t> tags(s.nblocks, KM_AMR_DEREFINE);
  for (int c = 0; c < s.adapt.ncrit; ++c) TagLoehner(s.base.pack, s.adapt.crit[c], tags.data(), false);
  std::vector<int32_t> nl, nd, org;
  if (!UpdateTree(s.mesh, s.adapt.numlevel, s.adapt.derefine_count, s.leaves, tags.data(),
                  s.deref, nl, nd, org))
    return;
  const int n = (int)nl.size() / 4;
  Tree t;
  const std::string err = BuildTree(s.mesh, nl.data(), n, t);
  if (!err.empty()) throw std::runtime_error("remesh produced an invalid tree: " + err);
  t.cyl = Cyl(s.cfg);
  std::vector<km_block_t> old_blocks = std::move(s.blocks);
  Container old_base = std::move(s.base);
  s.blocks.assign(n, km_block_t{});
  for (int b = 0; b < n; ++b) LeafMetric(s.mesh, t.level[b], &t.loc[3 * b], s.blocks[b]);
  s.tree = std::move(t);
  s.nblocks = n;
  MakeContainer(s, s.base, true, nullptr, false);
  Remesh(old_base.pack, s.base.pack, org.data(), 0, s.nvar, Cyl(s.cfg));
  if (s.cfg.mhd == KM_MHD_CT)
    RemeshFaces(old_base.pack, s.base.pack, org.data(), Cyl(s.cfg), s.mesh.xmin[0]);
  for (int i = 0; i < 2; ++i) MakeContainer(s, s.stage[i], false, &s.base, false);
  MakeContainer(s, s.dudt, false, nullptr, true);
  s.leaves = std::move(nl);
  s.deref = std::move(nd);
  FillGhosts(s, s.base.pack);
  FillDivB(s.cfg, s.base.pack);
  s.dt_est = EstimateDt(s.cfg, s.base.pack);
  if (AnyNonPhysical(s.cfg, s.base.pack)) s.fail |= 1;
}

}  // namespace

extern "C" {

// ---- single-stencil / single-face entry points used by the known-answer tests ----
int km_cpu_reconstruct(int32_t recon, int32_t limiter, const double *s5, double *vM,
                       double *vP) {
  Stencil st{s5 + 2, 1};
  Reconstruct(recon, limiter, st, *vM, *vP);
  return KM_OK;
}

int km_cpu_riemann(int32_t riemann, int32_t mhd, int32_t dir, const double *vL,
                   const double *vR, double *flux) {
  RiemannFlux(riemann, mhd, dir, vL, vR, flux);
  return KM_OK;
}

// state = {dens, temp, eint, pres, bmod}; mode 0 = ener, 1 = pres (dens_pres)
int km_cpu_eos(int32_t mode, double gamma, double abar, double *state) {
  const IdealGas eos(gamma, abar);
  if (mode == 1) state[2] = eos.sie_from_PT(state[0], state[3]);
  if (mode == 0) state[3] = eos.P(state[0], state[2]);
  state[1] = eos.T(state[0], state[2]);
  state[4] = eos.B(state[0], state[2]);
  return KM_OK;
}

int km_cpu_fast_speed(int32_t mhd, int32_t dir, const double *V, double *cf) {
  *cf = FastSpeed(mhd, dir, V);
  return KM_OK;
}

int km_cpu_prim2cons(int32_t mhd, const double *V, double *U) {
  Prim2Cons(mhd, V, U);
  return KM_OK;
}

int km_cpu_prim2flux(int32_t mhd, int32_t dir, const double *V, double *F) {
  Prim2Flux(mhd, dir, V, F);
  return KM_OK;
}

// ---- pack-level twins of the km_* tasks (host pointers, stream ignored) ----
int km_cpu_calculate_fluxes(const km_cfg_t *cfg, const km_pack_t *u, void *) {
  CalculateFluxes(*cfg, *u);
  return KM_OK;
}
int km_cpu_calculate_emf(const km_cfg_t *cfg, const km_pack_t *u, void *) {
  CalculateEMF(*cfg, *u);
  return KM_OK;
}
int km_cpu_fluxes_to_dudt(const km_cfg_t *cfg, const km_pack_t *u, const km_pack_t *dudt,
                          void *) {
  FluxesToDuDt(*cfg, *u, *dudt);
  return KM_OK;
}
int km_cpu_add_source_terms(const km_cfg_t *cfg, const km_pack_t *u, const km_pack_t *dudt,
                            void *) {
  AddSourceTerms(*cfg, *u, *dudt);
  return KM_OK;
}
int km_cpu_prepare_conserved_stage(const km_cfg_t *cfg, const km_pack_t *u, void *) {
  StagePrepareConserved(*cfg, *u);
  return KM_OK;
}
int km_cpu_apply_dudt(const km_cfg_t *cfg, const km_pack_t *base, const km_pack_t *u0,
                      const km_pack_t *u1, const km_pack_t *dudt, double beta, double dt,
                      void *) {
  ApplyDuDt(*cfg, *base, *u0, *u1, *dudt, beta, dt);
  return KM_OK;
}
int km_cpu_prepare_primitive(const km_cfg_t *cfg, const km_pack_t *u, void *) {
  PreparePrimitive(*cfg, *u);
  return KM_OK;
}
int km_cpu_prepare_conserved(const km_cfg_t *cfg, const km_pack_t *u, void *) {
  PrepareConserved(*cfg, *u);
  return KM_OK;
}
int km_cpu_estimate_dt(const km_cfg_t *cfg, const km_pack_t *u, double *dt, void *) {
  *dt = EstimateDt(*cfg, *u);
  return KM_OK;
}
int km_cpu_fill_divb(const km_cfg_t *cfg, const km_pack_t *u, void *) {
  FillDivB(*cfg, *u);
  return KM_OK;
}
// isentropic_vortex history variables: ErrorHistory (isentropic_vortex.hpp:150-200) of
// DENS, VELOCITY(0..2), PRES (+ MAGC(0..2)) against VortexData::State / StateMHD
// (:31-135) at the cell centroids, divided by the x1-x2 extent; params as the problem's
// (density, pressure, velx, vely, strength, mhd_strength)
int km_cpu_history_vortex_errors(const km_cfg_t *cfg, const km_pack_t *u,
                                 const km_mesh_desc_t *mesh, const double *params,
                                 int32_t nparams, double *errors) {
  auto P = [&](int i, double d) { return (params && i < nparams) ? params[i] : d; };
  const double density = P(0, 1.0), pressure = P(1, 1.0), velx = P(2, 1.0), vely = P(3, 1.0),
               strength = P(4, 5.0), mstr = P(5, 1.0), gamma = cfg->gamma;
  const double pi = 3.141592653589793;
  const int mhd = cfg->mhd;
  const int nv = mhd == KM_MHD_CT ? 8 : 5;
  const int slot[8] = {KM_DENS,       KM_VEL0(mhd) + 0, KM_VEL0(mhd) + 1, KM_VEL0(mhd) + 2,
                       KM_PRES(mhd),  KM_MAGC0,         KM_MAGC0 + 1,     KM_MAGC0 + 2};
  const Shape sh(*u);
  double err[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int b = 0; b < u->nblocks; ++b) {
    const km_block_t &blk = u->blocks[b];
    for (int k = sh.ks; k <= sh.ke; ++k)
      for (int j = sh.js; j <= sh.je; ++j)
        for (int i = sh.is; i <= sh.ie; ++i) {
          const double x = Cyl(*cfg) ? RCentroid(blk, sh, i)
                                     : (blk.xmin[0] - sh.is * blk.dx[0]) + (i + 0.5) * blk.dx[0];
          const double y = (blk.xmin[1] - sh.js * blk.dx[1]) + (j + 0.5) * blk.dx[1];
          const double r2 = Cyl(*cfg) ? x * x : x * x + y * y;
          const double dv = strength * std::exp(0.5 * (1.0 - r2)) / (2.0 * pi);
          double st[8] = {0, 0, 0, 0, 0, 0, 0, 0};
          if (mhd == KM_MHD_OFF) {
            const double T = pressure / density - (gamma - 1.0) * strength * strength *
                                                      std::exp(1.0 - r2) /
                                                      (8.0 * gamma * pi * pi);
            st[0] = std::pow(T, 1.0 / (gamma - 1.0));
            st[4] = T * st[0];
          } else {
            const double dB = mstr * std::exp(0.5 * (1.0 - r2)) / (2.0 * pi);
            const double ex = std::exp(1.0 - r2);
            st[0] = 1.0;
            st[4] = pressure + (mstr * mstr / (8.0 * pi * pi)) * (1.0 - r2) * ex -
                    (strength * strength / (8.0 * pi * pi)) * ex;
            if (Cyl(*cfg)) {
              st[7] = x * dB;
            } else {
              st[5] = -y * dB;
              st[6] = x * dB;
            }
          }
          if (Cyl(*cfg)) {
            st[3] = x * dv;
          } else {
            st[1] = velx - y * dv;
            st[2] = vely + x * dv;
          }
          const long c = sh.c(k, j, i);
          const double vol = Volume(*cfg, blk, sh, i);
          for (int v = 0; v < nv; ++v) err[v] += std::abs(st[v] - cellp(*u, b, slot[v])[c]) * vol;
        }
  }
  const double area = (mesh->xmax[0] - mesh->xmin[0]) * (mesh->xmax[1] - mesh->xmin[1]);
  for (int v = 0; v < nv; ++v) errors[v] = err[v] / area;
  return KM_OK;
}
int km_cpu_ghost_fill_uniform(const km_cfg_t *cfg, const km_pack_t *u, const int32_t nb[3],
                              const int32_t bc[6], int32_t v0, int32_t v1,
                              int32_t with_faces) {
  Grid g;
  for (int d = 0; d < 3; ++d) g.nb[d] = nb[d];
  for (int d = 0; d < 6; ++d) g.bc[d] = bc[d];
  GhostFill(*cfg, g, *u, v0, v1, with_faces != 0);
  return KM_OK;
}
int km_cpu_mesh_blocks(const km_mesh_desc_t *mesh, km_block_t *out) {
  std::vector<km_block_t> blocks;
  MakeBlocks(*mesh, blocks);
  std::memcpy(out, blocks.data(), blocks.size() * sizeof(km_block_t));
  return KM_OK;
}

// ---- whole-mesh twin of km_sim_* ----
int km_cpu_sim_create(const km_cfg_t *cfg, const km_mesh_desc_t *mesh, km_cpu_sim_t **out) {
  auto *s = new km_cpu_sim_t();
  s->cfg = *cfg;
  s->mesh = *mesh;
  s->cfg.ndim = mesh->ndim;
  s->nvar = KM_NVAR(cfg->mhd) + KM_NSPEC_SLOTS(cfg->nspecies);
  s->nc = KM_NCONS(cfg->mhd);
  s->nflux = s->nc + KM_NSPEC_SLOTS(cfg->nspecies);
  s->nstages = cfg->nstages;
  for (int d = 0; d < 3; ++d) s->grid.nb[d] = mesh->nb[d];
  for (int d = 0; d < 6; ++d) s->grid.bc[d] = mesh->bc[d];
  MakeBlocks(*mesh, s->blocks);
  s->nblocks = (int)s->blocks.size();
  MakeContainer(*s, s->base, true, nullptr, false);
  for (int i = 0; i < 2; ++i) MakeContainer(*s, s->stage[i], false, &s->base, false);
  MakeContainer(*s, s->dudt, false, nullptr, true);
  *out = s;
  return KM_OK;
}

int km_cpu_amr_blocks(const km_mesh_desc_t *mesh, const int32_t *leaves, int32_t n,
                      km_block_t *out) {
  Tree t;
  const std::string err = BuildTree(*mesh, leaves, n, t);
  if (!err.empty()) {
    std::fprintf(stderr, "km_cpu_amr_blocks: %s\n", err.c_str());
    return KM_EINVAL;
  }
  for (int b = 0; b < n; ++b) LeafMetric(*mesh, t.level[b], &t.loc[3 * b], out[b]);
  return KM_OK;
}

int km_cpu_amr_ghost_fill(const km_cfg_t *cfg, const km_pack_t *u, const km_mesh_desc_t *mesh,
                          const int32_t *leaves, int32_t n, int32_t v0, int32_t v1) {
  Tree t;
  const std::string err = BuildTree(*mesh, leaves, n, t);
  t.cyl = Cyl(*cfg);
  if (!err.empty() || n != u->nblocks) {
    std::fprintf(stderr, "km_cpu_amr_ghost_fill: %s\n", err.empty() ? "leaf count" : err.c_str());
    return KM_EINVAL;
  }
  try {
    AmrGhostFill(*cfg, t, *u, v0, v1);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "km_cpu_amr_ghost_fill: %s\n", e.what());
    return KM_EINVAL;
  }
  return KM_OK;
}

int km_cpu_flux_correct(const km_cfg_t *cfg, const km_pack_t *u, const km_mesh_desc_t *mesh,
                        const int32_t *leaves, int32_t n) {
  Tree t;
  const std::string err = BuildTree(*mesh, leaves, n, t);
  t.cyl = Cyl(*cfg);
  if (!err.empty() || n != u->nblocks) return KM_EINVAL;
  AmrFluxCorrect(*cfg, t, *u);
  if (u->edge[2] && (u->ndim < 3 || (u->edge[0] && u->edge[1]))) AmrEmfCorrect(t, *u);
  return KM_OK;
}

int km_cpu_amr_ghost_faces(const km_cfg_t *cfg, const km_pack_t *u, const km_mesh_desc_t *mesh,
                           const int32_t *leaves, int32_t n) {
  Tree t;
  const std::string err = BuildTree(*mesh, leaves, n, t);
  t.cyl = Cyl(*cfg);
  if (!err.empty() || n != u->nblocks || !u->face[0] || !u->face[1] || !u->face[2]) return KM_EINVAL;
  try {
    AmrGhostFaces(*cfg, t, *u);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "km_cpu_amr_ghost_faces: %s\n", e.what());
    return KM_EINVAL;
  }
  return KM_OK;
}

int km_cpu_sim_create_amr(const km_cfg_t *cfg, const km_mesh_desc_t *mesh,
                          const int32_t *leaves, int32_t n, km_cpu_sim_t **out) {
  auto *s = new km_cpu_sim_t();
  s->cfg = *cfg;
  s->mesh = *mesh;
  s->cfg.ndim = mesh->ndim;
  s->nvar = KM_NVAR(cfg->mhd) + KM_NSPEC_SLOTS(cfg->nspecies);
  s->nc = KM_NCONS(cfg->mhd);
  s->nflux = s->nc + KM_NSPEC_SLOTS(cfg->nspecies);
  s->nstages = cfg->nstages;
  const std::string err = BuildTree(*mesh, leaves, n, s->tree);
  s->tree.cyl = Cyl(*cfg);
  if (!err.empty() || (cfg->mhd == KM_MHD_CT && mesh->ng % 2)) {
    std::fprintf(stderr, "km_cpu_sim_create_amr: %s\n", err.empty() ? "CT needs even nghost" : err.c_str());
    delete s;
    return KM_EINVAL;
  }
  s->amr = true;
  s->leaves.assign(leaves, leaves + 4 * (size_t)n);
  s->deref.assign(n, 0);
  s->blocks.resize(n);
  for (int b = 0; b < n; ++b) LeafMetric(*mesh, s->tree.level[b], &s->tree.loc[3 * b], s->blocks[b]);
  s->nblocks = n;
  MakeContainer(*s, s->base, true, nullptr, false);
  for (int i = 0; i < 2; ++i) MakeContainer(*s, s->stage[i], false, &s->base, false);
  MakeContainer(*s, s->dudt, false, nullptr, true);
  *out = s;
  return KM_OK;
}

int km_cpu_sim_destroy(km_cpu_sim_t *s) {
  delete s;
  return KM_OK;
}

static void copy_state(km_cpu_sim_t *s, const double *cell, const double *const face[3],
                       bool in, double *cell_out, double *const face_out[3]) {
  const km_pack_t &p = s->base.pack;
  Shape sh(p);
  const long nce = sh.ncell(), nst = sh.nstag();
  for (int b = 0; b < p.nblocks; ++b) {
    for (int v = 0; v < s->nvar; ++v) {
      double *dst = p.cell[(long)b * s->nvar + v];
      if (in)
        std::memcpy(dst, cell + ((long)b * s->nvar + v) * nce, nce * sizeof(double));
      else
        std::memcpy(cell_out + ((long)b * s->nvar + v) * nce, dst, nce * sizeof(double));
    }
    if (s->cfg.mhd == KM_MHD_CT)
      for (int d = 0; d < 3; ++d) {
        if (in && face && face[d])
          std::memcpy(p.face[d][b], face[d] + (long)b * nst, nst * sizeof(double));
        if (!in && face_out && face_out[d])
          std::memcpy(face_out[d] + (long)b * nst, p.face[d][b], nst * sizeof(double));
      }
  }
}

int km_cpu_sim_set_primitive(km_cpu_sim_t *s, const double *cell, const double *const face[3]) {
  copy_state(s, cell, face, true, nullptr, nullptr);
  PrepareConserved(s->cfg, s->base.pack);
  try {
    FillGhosts(*s, s->base.pack);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "km_cpu_sim_set_primitive: %s\n", e.what());
    return KM_EINVAL;
  }
  s->dt_est = EstimateDt(s->cfg, s->base.pack);
  s->fail = AnyNonPhysical(s->cfg, s->base.pack) ? 1 : 0;
  s->below = s->above = 0;
  s->dt = FirstDt(s->cfg, s->dt_est);
  s->time = 0.;
  s->ncycle = 0;
  return KM_OK;
}

int km_cpu_sim_get_state(km_cpu_sim_t *s, double *cell, double *const face[3]) {
  copy_state(s, nullptr, nullptr, false, cell, face);
  return KM_OK;
}

int km_cpu_sim_set_dt(km_cpu_sim_t *s, double dt) {
  s->dt = dt;
  s->dt_override = true;
  return KM_OK;
}

// Parthenon EvolutionDriver time-step rule: the new dt is min(estimate, dt_factor * old dt)
// within the limits of DtLimits, clipped so the run ends at tlim. A run whose CFL pass saw a
// non-physical zone, or whose dt stayed below dt_min (above dt_max) for dt_min_cycle_limit
// (dt_max_cycle_limit) cycles in a row, fails with KM_ESTATE once its cycles are done.
int km_cpu_sim_advance(km_cpu_sim_t *s, int32_t ncycles, double tlim) {
  for (int n = 0; n < ncycles; ++n) {
    if (tlim > 0.) {
      if (s->time >= tlim) break;
      if (s->time + s->dt > tlim) s->dt = tlim - s->time;
    }
    try {
      Cycle(*s);
      if (s->adaptive) AdaptStep(*s);
    } catch (const std::exception &e) {
      std::fprintf(stderr, "km_cpu_sim_advance: %s\n", e.what());
      return KM_EINVAL;
    }
    const double used = s->dt;
    s->below = used < s->cfg.dt_min ? s->below + 1 : 0;
    s->above = used > s->cfg.dt_max ? s->above + 1 : 0;
    if (s->below > 0 && s->below >= s->cfg.dt_min_cycle_limit) s->fail |= 2;
    if (s->above > 0 && s->above >= s->cfg.dt_max_cycle_limit) s->fail |= 4;
    if (!s->dt_override) s->dt = DtLimits(s->cfg, kmin(s->cfg.dt_factor * used, s->dt_est));
    s->dt_override = false;
  }
  return s->fail ? KM_ESTATE : KM_OK;
}

// ---- adaptive refinement twins (host pointers) ----
int km_cpu_amr_tag_loehner(const km_pack_t *u, const km_amr_criterion_t *crit, int32_t *tags,
                           int32_t reset) {
  TagLoehner(*u, *crit, tags, reset != 0);
  return KM_OK;
}

int km_cpu_amr_update_tree(const km_mesh_desc_t *mesh, int32_t numlevel,
                           int32_t derefine_count, const int32_t *leaves, int32_t nleaves,
                           const int32_t *tags, int32_t *deref_count, int32_t *new_leaves,
                           int32_t *new_deref, int32_t *origin, int32_t *nnew,
                           int32_t *changed) {
  std::vector<int32_t> lv(leaves, leaves + 4 * (size_t)nleaves);
  std::vector<int32_t> dc(deref_count, deref_count + nleaves), nl, nd, org;
  bool ch;
  try {
    ch = UpdateTree(*mesh, numlevel, derefine_count, lv, tags, dc, nl, nd, org);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "km_cpu_amr_update_tree: %s\n", e.what());
    return KM_EINVAL;
  }
  const int n = (int)nl.size() / 4;
  if (n > *nnew) {
    *nnew = n;
    return KM_EINVAL;
  }
  std::copy(dc.begin(), dc.end(), deref_count);
  std::copy(nl.begin(), nl.end(), new_leaves);
  std::copy(nd.begin(), nd.end(), new_deref);
  std::copy(org.begin(), org.end(), origin);
  *nnew = n;
  *changed = ch ? 1 : 0;
  return KM_OK;
}

int km_cpu_amr_remesh_geom(const km_pack_t *old_u, const km_pack_t *new_u, const int32_t *origin,
                           int32_t v0, int32_t v1, const km_mesh_desc_t *mesh, int32_t geometry);
int km_cpu_amr_remesh(const km_pack_t *old_u, const km_pack_t *new_u, const int32_t *origin,
                      int32_t v0, int32_t v1) {
  return km_cpu_amr_remesh_geom(old_u, new_u, origin, v0, v1, nullptr, KM_GEOM_CARTESIAN);
}

int km_cpu_amr_remesh_geom(const km_pack_t *old_u, const km_pack_t *new_u, const int32_t *origin,
                           int32_t v0, int32_t v1, const km_mesh_desc_t *mesh, int32_t geometry) {
  const bool cyl = geometry == KM_GEOM_CYLINDRICAL;
  if (cyl && !mesh) return KM_EINVAL;
  Remesh(*old_u, *new_u, origin, v0, v1, cyl);
  bool faces = true;
  for (int d = 0; d < 3; ++d) faces = faces && old_u->face[d] && new_u->face[d];
  if (faces) RemeshFaces(*old_u, *new_u, origin, cyl, cyl ? mesh->xmin[0] : 0.0);
  return KM_OK;
}

int km_cpu_sim_set_adaptive(km_cpu_sim_t *s, const km_amr_adapt_t *a) {
  if (!s->amr || a->ncrit < 0 || a->ncrit > KM_AMR_MAX_CRIT ||
      a->numlevel < 1)
    return KM_EINVAL;
  s->adaptive = true;
  s->adapt = *a;
  s->deref.assign(s->nblocks, 0);
  return KM_OK;
}

int km_cpu_sim_amr_leaves(const km_cpu_sim_t *s, int32_t *leaves, int32_t *nleaves) {
  const int n = (int)s->leaves.size() / 4;
  if (*nleaves < n) {
    *nleaves = n;
    return KM_EINVAL;
  }
  std::copy(s->leaves.begin(), s->leaves.end(), leaves);
  *nleaves = n;
  return KM_OK;
}

int km_cpu_sim_info(const km_cpu_sim_t *s, double *time, double *dt, int32_t *ncycle) {
  if (time) *time = s->time;
  if (dt) *dt = s->dt;
  if (ncycle) *ncycle = s->ncycle;
  return KM_OK;
}

}  // extern "C"
