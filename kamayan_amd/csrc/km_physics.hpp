// km_physics.hpp -- per-cell / per-face arithmetic of the hydro/MHD path for the gfx950
// kernels: slope limiters, FOG/PLM/PPM/WENO-Z reconstruction, state conversions,
// fast speeds and the HLL/HLLC Riemann solvers.
//
// Templated on the arithmetic type T (double for the FP64 form, float for the
// reconstruction/Riemann half of the FP32-mixed form). For T = double every expression
// keeps the operation order of the reference source lines cited beside it, so that a
// build with -ffp-contract=off rounds each operation as the reference does.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

#define KM_HD __host__ __device__ __forceinline__

namespace km {

template <class T>
struct Num;
template <>
struct Num<double> {
  static constexpr double tiny = DBL_MIN;  // std::numeric_limits<Real>::min()
};
template <>
struct Num<float> {
  static constexpr float tiny = FLT_MIN;
};

// Kokkos::min / Kokkos::max (== std::min / std::max), including their +-0 behaviour
template <class T>
KM_HD T kmin(T a, T b) {
  return (b < a) ? b : a;
}
template <class T>
KM_HD T kmax(T a, T b) {
  return (a < b) ? b : a;
}
KM_HD double kabs(double x) { return __builtin_fabs(x); }
KM_HD float kabs(float x) { return __builtin_fabsf(x); }
KM_HD double ksign(double x) { return __builtin_copysign(1.0, x); }
KM_HD float ksign(float x) { return __builtin_copysignf(1.0f, x); }

// Division and square root: IEEE for double (the reference's rounding); the FP32 form
// uses the hardware reciprocal / square root, whose ~1 ulp error is far below the
// FP32 form's own truncation of the states.
// FAST (the FP64-FMA form): the hardware reciprocal estimate refined by one Newton step,
// which squares its relative error (a few ulps of double remain), and for a quotient one
// residual correction q + (a - b q) / b, within an ulp or so of the IEEE result. Every
// divisor here is a normal number (densities, wave-speed differences, WENO weights
// >= 1e-36); the form's contract is 1e-12 relative L1 agreement with the IEEE form.
template <bool FAST = false>
KM_HD double krcp(double b) {
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (FAST) {
    double r = __builtin_amdgcn_rcp(b);
    r = __builtin_fma(__builtin_fma(-b, r, 1.0), r, r);
    return r;
  }
#endif
  return 1.0 / b;
}
template <bool FAST = false>
KM_HD double kdiv(double a, double b) {
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (FAST) {
    const double r = krcp<true>(b);
    const double q = a * r;
    return __builtin_fma(__builtin_fma(-b, q, a), r, q);
  }
#endif
  return a / b;
}
KM_HD double ksqrt(double x) { return __builtin_sqrt(x); }
#if defined(__HIP_DEVICE_COMPILE__)
template <bool FAST = false>
KM_HD float kdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
template <bool FAST = false>
KM_HD float krcp(float b) { return __builtin_amdgcn_rcpf(b); }
KM_HD float ksqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
#else
template <bool FAST = false>
KM_HD float kdiv(float a, float b) { return a / b; }
template <bool FAST = false>
KM_HD float krcp(float b) { return 1.0f / b; }
KM_HD float ksqrt(float x) { return __builtin_sqrtf(x); }
#endif

enum { RECON_FOG = 0, RECON_PLM = 1, RECON_PPM = 2, RECON_WENOZ = 3 };
enum { LIM_MINMOD = 0, LIM_VAN_LEER = 1, LIM_MC = 2 };
enum { RIEMANN_HLL = 0, RIEMANN_HLLC = 1 };

// number of cells on each side of a reconstructed cell that the method reads
template <int RECON>
struct ReconWidth {
  static constexpr int value = RECON == RECON_FOG ? 0 : RECON == RECON_PLM ? 1 : 2;
};

// LimitedSlope, reconstruction.hpp:24-39
template <int LIM, bool FAST = false, class T>
KM_HD T limited_slope(T a, T b) {
  if constexpr (LIM == LIM_MC) {
    return (ksign(a) + ksign(b)) * kmin(kabs(a), kmin(T(.25) * kabs(a + b), kabs(b)));
  } else if constexpr (LIM == LIM_VAN_LEER) {
    if (a * b > T(0.)) return kdiv<FAST>(T(2.) * a * b, a + b);
    return T(0.);
  } else {
    return T(0.5) * (ksign(a) + ksign(b)) * kmin(kabs(a), kabs(b));
  }
}

// Reconstruct (reconstruction.hpp:16-142) on a window w[0..4] = cells -2..+2 around the
// reconstructed cell (w[2]); vM at the cell's lower face, vP at its upper face.
template <int RECON, int LIM, bool FAST = false, class T>
KM_HD void reconstruct(const T *w, T &vM, T &vP) {
  const T s0 = w[2];
  if constexpr (RECON == RECON_FOG) {
    vM = s0;
    vP = s0;
  } else if constexpr (RECON == RECON_PLM) {  // :48-59
    const T dvL = s0 - w[1];
    const T dvR = w[3] - s0;
    const T del = limited_slope<LIM, FAST>(dvL, dvR);
    vM = s0 - T(0.5) * del;
    vP = s0 + T(0.5) * del;
  } else if constexpr (RECON == RECON_PPM) {  // :61-88, Slope(idx) = Lim(s(i+1)-s(i), s(i)-s(i-1))
    const T dv_p = limited_slope<LIM, FAST>(w[4] - w[3], w[3] - s0);
    const T dv_0 = limited_slope<LIM, FAST>(w[3] - s0, s0 - w[1]);
    const T dv_m = limited_slope<LIM, FAST>(s0 - w[1], w[1] - w[0]);
    constexpr T sixth = T(1. / 6.);
    T m = T(0.5) * (w[1] + s0) - sixth * (dv_0 - dv_m);
    T p = T(0.5) * (s0 + w[3]) - sixth * (dv_p - dv_0);
    if ((p - s0) * (s0 - m) <= T(0.)) {
      m = s0;
      p = s0;
    }
    if (-(p - m) * (p - m) > T(6.) * (p - m) * (s0 - T(0.5) * (p + m))) {
      p = T(3.0) * s0 - T(2.) * m;
    }
    if ((p - m) * (p - m) < T(6.) * (p - m) * (s0 - T(0.5) * (p + m))) {
      m = T(3.0) * s0 - T(2.) * p;
    }
    vM = m;
    vP = p;
  } else {  // WENO-Z, :90-142
    constexpr T sixth = T(1. / 6.);
    // eno_reconstruction(+1) and (-1)
    const T ep0 = sixth * (-w[4] + T(5.) * w[3] + T(2.) * s0);
    const T ep1 = sixth * (T(2.) * w[3] + T(5.) * s0 - w[1]);
    const T ep2 = sixth * (T(11.) * s0 - T(7.) * w[1] + T(2.) * w[0]);
    const T em0 = sixth * (-w[0] + T(5.) * w[1] + T(2.) * s0);
    const T em1 = sixth * (T(2.) * w[1] + T(5.) * s0 - w[3]);
    const T em2 = sixth * (T(11.) * s0 - T(7.) * w[3] + T(2.) * w[4]);
    constexpr T c1312 = T(13. / 12.);
    const T a0 = w[0] - T(2.) * w[1] + s0;
    const T a1 = w[0] - T(4.) * w[1] + T(3.) * s0;
    const T b0 = c1312 * (a0 * a0) + T(0.25) * (a1 * a1);
    const T a2 = w[1] - T(2.) * s0 + w[3];
    const T a3 = w[1] - w[3];
    const T b1 = c1312 * (a2 * a2) + T(0.25) * (a3 * a3);
    const T a4 = s0 - T(2.) * w[3] + w[4];
    const T a5 = w[4] - T(4.) * w[3] + T(3.) * s0;
    const T b2 = c1312 * (a4 * a4) + T(0.25) * (a5 * a5);
    constexpr T eps = T(1.e-36);
    const T tau = kabs(b2 - b0);
    // tau / (eps + beta_k) is evaluated identically for both sides; computing it once
    // per k gives the same bits as the reference's two evaluations
    // FAST: a reciprocal and a product for each ratio, as for the normalisations below
    T r0 = FAST ? tau * krcp<FAST>(eps + b0) : kdiv(tau, eps + b0);
    T r1 = FAST ? tau * krcp<FAST>(eps + b1) : kdiv(tau, eps + b1);
    T r2 = FAST ? tau * krcp<FAST>(eps + b2) : kdiv(tau, eps + b2);
    if constexpr (sizeof(T) == 4) {
      // FP32: an indicator that rounds to 0 gives tau/eps ~ 1e22, whose square overflows
      // float. Clamping at 1e15 keeps every weight finite; a ratio that large already
      // makes its stencil's weight dominate by > 1e20.
      r0 = kmin(r0, T(1.e15));
      r1 = kmin(r1, T(1.e15));
      r2 = kmin(r2, T(1.e15));
    }
    const T q0 = T(1.0) + r0 * r0;
    const T q1 = T(1.0) + r1 * r1;
    const T q2 = T(1.0) + r2 * r2;
    // vM: weno_weighting(-1): w0 uses beta[0], w2 uses beta[2]
    {
      const T w0 = T(3.) * q0, w1 = T(6.) * q1, w2 = q2;
      const T norm = w0 + w1 + w2;
      vM = FAST ? (w0 * em0 + w1 * em1 + w2 * em2) * krcp<FAST>(norm)
                : kdiv(w0 * em0 + w1 * em1 + w2 * em2, norm);
    }
    // vP: weno_weighting(+1): w0 uses beta[2], w2 uses beta[0]
    {
      const T w0 = T(3.) * q2, w1 = T(6.) * q1, w2 = q0;
      const T norm = w0 + w1 + w2;
      vP = FAST ? (w0 * ep0 + w1 * ep1 + w2 * ep2) * krcp<FAST>(norm)
                : kdiv(w0 * ep0 + w1 * ep1 + w2 * ep2, norm);
    }
  }
}

// ---- state conversions, primconsflux.hpp:21-138 ----
// recon-variable order (hydro_types.hpp:61-74,110): DENS, VELOCITY[3], PRES, BMOD,
// EINT, MAGC[3]; conserved order: DENS, MOMENTUM[3], ENER, MAGC[3]
enum { R_DENS = 0, R_VEL = 1, R_PRES = 4, R_BMOD = 5, R_EINT = 6, R_MAGC = 7 };
enum { C_DENS = 0, C_MOM = 1, C_ENER = 4, C_MAGC = 5 };

template <int MHD>
struct HV {
  static constexpr int NR = MHD ? 10 : 7;
  static constexpr int NC = MHD ? 8 : 5;
};

template <int MHD, class T>
KM_HD T total_pres(const T *V) {  // :21-30
  T pres = V[R_PRES];
  if constexpr (MHD) {
    pres += T(0.5) * (V[R_MAGC] * V[R_MAGC] + V[R_MAGC + 1] * V[R_MAGC + 1] +
                      V[R_MAGC + 2] * V[R_MAGC + 2]);
  }
  return pres;
}

template <int MHD, int D1, bool FAST = false, class T>
KM_HD T fast_speed(const T *V) {  // :31-53
  const T idens = krcp<FAST>(V[R_DENS]);
  const T a2 = idens * V[R_BMOD];
  T cfast2;
  if constexpr (!MHD) {
    cfast2 = a2;
  } else {
    const T bb2 = V[R_MAGC + D1] * V[R_MAGC + D1] * idens;
    T b2 = T(0.0);
    for (int dir = 0; dir < 3; dir++) b2 += V[R_MAGC + dir] * V[R_MAGC + dir];
    b2 *= idens;
    cfast2 = T(0.5) * ((a2 + b2) + ksqrt((a2 - b2) * (a2 - b2) + T(4.) * a2 * (b2 - bb2)));
  }
  return ksqrt(cfast2);
}

template <int MHD, class T>
KM_HD void prim2cons(const T *V, T *U) {  // :55-74
  U[C_DENS] = V[R_DENS];
  T emag = T(0.);
  T ekin = T(0.);
#pragma unroll
  for (int dir = 0; dir < 3; dir++) {
    U[C_MOM + dir] = V[R_DENS] * V[R_VEL + dir];
    ekin += V[R_VEL + dir] * V[R_VEL + dir];
    if constexpr (MHD) {
      U[C_MAGC + dir] = V[R_MAGC + dir];
      emag += V[R_MAGC + dir] * V[R_MAGC + dir];
    }
  }
  const T eint = V[R_EINT] * V[R_DENS];
  ekin *= T(0.5) * V[R_DENS];
  emag *= T(0.5);
  U[C_ENER] = eint + ekin + emag;
}

template <int MHD, int D1, class T>
KM_HD void prim2flux(const T *V, T *F) {  // :100-138
  constexpr int D2 = (D1 + 1) % 3;
  constexpr int D3 = (D1 + 2) % 3;
  F[C_DENS] = V[R_DENS] * V[R_VEL + D1];
  T ptot = V[R_PRES];
  T B2 = T(0.);
  T uB = T(0.);
  T ekin = T(0.);
#pragma unroll
  for (int dir = 0; dir < 3; dir++) {
    F[C_MOM + dir] = F[C_DENS] * V[R_VEL + dir];
    ekin += V[R_VEL + dir] * V[R_VEL + dir];
    if constexpr (MHD) {
      B2 += V[R_MAGC + dir] * V[R_MAGC + dir];
      uB += V[R_VEL + dir] * V[R_MAGC + dir];
      F[C_MOM + dir] -= V[R_MAGC + D1] * V[R_MAGC + dir];
    }
  }
  ekin *= T(0.5) * V[R_DENS];
  const T emag = T(0.5) * B2;
  ptot += T(0.5) * B2;
  const T etot = V[R_EINT] * V[R_DENS] + ekin + emag;
  F[C_MOM + D1] += ptot;
  F[C_ENER] = (etot + ptot) * V[R_VEL + D1];
  if constexpr (MHD) {
    F[C_ENER] -= uB * V[R_MAGC + D1];
    F[C_MAGC + D1] = T(0.);
    F[C_MAGC + D2] = V[R_VEL + D1] * V[R_MAGC + D2] - V[R_VEL + D2] * V[R_MAGC + D1];
    F[C_MAGC + D3] = V[R_VEL + D1] * V[R_MAGC + D3] - V[R_VEL + D3] * V[R_MAGC + D1];
  }
}

// HLL, riemann_solver.hpp:20-57
template <int MHD, int D1, bool FAST = false, class T>
KM_HD void riemann_hll(const T *vL, const T *vR, T *flux) {
  constexpr int NC = HV<MHD>::NC;
  const T cfL = fast_speed<MHD, D1, FAST>(vL);
  const T cfR = fast_speed<MHD, D1, FAST>(vR);
  const T tiny = Num<T>::tiny;
  const T sL = kmin(-tiny, kmin(vL[R_VEL + D1] - cfL, vR[R_VEL + D1] - cfR));
  const T sR = kmax(tiny, kmax(vL[R_VEL + D1] + cfL, vR[R_VEL + D1] + cfR));
  const T sRmsLi = krcp<FAST>(sR - sL);
  T UL[NC], UR[NC], FL[NC], FR[NC];
  prim2cons<MHD>(vL, UL);
  prim2cons<MHD>(vR, UR);
  prim2flux<MHD, D1>(vR, FR);
  prim2flux<MHD, D1>(vL, FL);
#pragma unroll
  for (int v = 0; v < NC; ++v)
    flux[v] = sRmsLi * (sR * FL[v] - sL * FR[v] + sR * sL * (UR[v] - UL[v]));
}

// HLLC, riemann_solver.hpp:59-149
template <int MHD, int D1, bool FAST = false, class T>
KM_HD void riemann_hllc(const T *vL, const T *vR, T *flux) {
  constexpr int NC = HV<MHD>::NC;
  constexpr int D2 = (D1 + 1) % 3;
  constexpr int D3 = (D1 + 2) % 3;
  const T cfL = fast_speed<MHD, D1, FAST>(vL);
  const T cfR = fast_speed<MHD, D1, FAST>(vR);
  const T tiny = Num<T>::tiny;
  const T sL = kmin(-tiny, kmin(vL[R_VEL + D1] - cfL, vR[R_VEL + D1] - cfR));
  const T sR = kmax(tiny, kmax(vL[R_VEL + D1] + cfL, vR[R_VEL + D1] + cfR));
  T UL[NC], UR[NC], FL[NC], FR[NC];
  prim2cons<MHD>(vL, UL);
  prim2cons<MHD>(vR, UR);
  prim2flux<MHD, D1>(vR, FR);
  prim2flux<MHD, D1>(vL, FL);
  const T total_presL = total_pres<MHD>(vL);
  const T total_presR = total_pres<MHD>(vR);
  const T uL = vL[R_VEL + D1], uR = vR[R_VEL + D1];
  T ustar = total_presR - total_presL + UL[C_MOM + D1] * (sL - uL) - UR[C_MOM + D1] * (sR - uR);
  ustar = kdiv<FAST>(ustar * T(1.), vL[R_DENS] * (sL - uL) - vR[R_DENS] * (sR - uR));
  const T pstar = T(0.5) * (total_presL + total_presR + vL[R_DENS] * (sL - uL) * (ustar - uL) +
                            vR[R_DENS] * (sR - uR) * (ustar - uR));
  T UstarL[NC], UstarR[NC];
  auto hllc_state = [&](const T S, const T *U, const T Pu, T *Us) {
    const T susi = krcp<FAST>(S - ustar + tiny);
    Us[C_DENS] = susi * (S * U[C_DENS] - U[C_MOM + D1]);
    Us[C_MOM + D1] = Us[C_DENS] * ustar;
    if constexpr (FAST) {
      // one reciprocal of the state's density for its three scaled components
      const T q = Us[C_DENS] * krcp<FAST>(U[C_DENS]);
      Us[C_MOM + D2] = U[C_MOM + D2] * q;
      Us[C_MOM + D3] = U[C_MOM + D3] * q;
      Us[C_ENER] = U[C_ENER] * q + susi * (pstar * ustar - Pu);
    } else {
      Us[C_MOM + D2] = kdiv(U[C_MOM + D2] * Us[C_DENS], U[C_DENS]);
      Us[C_MOM + D3] = kdiv(U[C_MOM + D3] * Us[C_DENS], U[C_DENS]);
      Us[C_ENER] = kdiv(U[C_ENER] * Us[C_DENS], U[C_DENS]) + susi * (pstar * ustar - Pu);
    }
    if constexpr (MHD) {  // :111-133
      const T sRsLi = krcp(sR - sL);
      auto hll = [&](int v) { return sRsLi * (sR * UR[v] - sL * UL[v] + FL[v] - FR[v]); };
      Us[C_MAGC + D1] = hll(C_MAGC + D1);
      Us[C_MAGC + D2] = hll(C_MAGC + D2);
      Us[C_MAGC + D3] = hll(C_MAGC + D3);
      Us[C_MOM + D2] -=
          (Us[C_MAGC + D1] * Us[C_MAGC + D2] - U[C_MAGC + D1] * U[C_MAGC + D2]) * susi;
      Us[C_MOM + D3] -=
          (Us[C_MAGC + D1] * Us[C_MAGC + D3] - U[C_MAGC + D1] * U[C_MAGC + D3]) * susi;
      Us[C_ENER] -=
          susi * (kdiv(Us[C_MAGC + D1] * (Us[C_MAGC + D1] * hll(C_MOM + D1) +
                                          Us[C_MAGC + D2] * hll(C_MOM + D2) +
                                          Us[C_MAGC + D3] * hll(C_MOM + D3)),
                       hll(C_DENS)) -
                  kdiv(U[C_MAGC + D1] * (U[C_MAGC + D1] * U[C_MOM + D1] +
                                         U[C_MAGC + D2] * U[C_MOM + D2] +
                                         U[C_MAGC + D3] * U[C_MOM + D3]),
                       U[C_DENS]));
    }
  };
  hllc_state(sL, UL, total_presL * uL, UstarL);
  hllc_state(sR, UR, total_presR * uR, UstarR);
  const T biasL = -kmin(-tiny, ksign(ustar));
  const T biasR = kmax(tiny, ksign(ustar));
#pragma unroll
  for (int v = 0; v < NC; ++v)
    flux[v] = biasR * (FL[v] + sL * (UstarL[v] - UL[v])) + biasL * (FR[v] + sR * (UstarR[v] - UR[v]));
}

template <int RIEMANN, int MHD, int D1, bool FAST = false, class T>
KM_HD void riemann(const T *vL, const T *vR, T *flux) {
  if constexpr (RIEMANN == RIEMANN_HLLC)
    riemann_hllc<MHD, D1, FAST>(vL, vR, flux);
  else
    riemann_hll<MHD, D1, FAST>(vL, vR, flux);
}

// Ideal-gas closure of singularity-eos IdealGas (equation_of_state.hpp:40-64):
// P = gm1*rho*e, BMOD = max(0, (gm1+1)*gm1*rho*e), T = max(0, e/cv)
struct Eos {
  double gm1, gp1gm1, cv;
};
KM_HD double eos_pres(const Eos &e, double rho, double sie) { return e.gm1 * rho * sie; }
KM_HD double eos_bmod(const Eos &e, double rho, double sie) {
  return kmax(0.0, (e.gm1 + 1) * e.gm1 * rho * sie);
}
KM_HD double eos_temp(const Eos &e, double rho, double sie) { return kmax(0.0, sie / e.cv); }

}  // namespace km
