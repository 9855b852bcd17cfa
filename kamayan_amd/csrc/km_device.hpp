// km_device.hpp -- device-side pack geometry and the register-window sweep shared by the
// flux kernels.
//
// Sweep design (gfx950): one lane owns one pencil of cells along the sweep direction and
// marches along it, keeping the reconstruction stencil of every reconstructed variable
// in a 5-slot register window. Each step loads one new cell, reconstructs one cell (both
// face values at once) and solves the Riemann problem at the face it shares with the
// previous cell, whose upper state is still in registers. Every cell is reconstructed
// exactly once per direction and every face is solved exactly once, as in the
// reference's pencil loops (hydro_add_flux_tasks.cpp:79-261), but with no scratch
// memory and no barriers. For the x2/x3 sweeps consecutive lanes own consecutive x1
// columns, so each window load is one coalesced 512-byte wavefront access.
#pragma once

#include "../../include/km_hydro.h"
#include "km_physics.hpp"

namespace km {

struct Geo {
  int ndim, ng, nx[3], E[3], S[3];
  int lo[3], hi[3];  // interior cell bounds per dimension (inclusive)
  long sc[3];        // cell-array strides
  long ss[3];        // staggered-array strides
  KM_HD explicit Geo(const km_pack_t &p) {
    ndim = p.ndim;
    ng = p.ng;
    for (int d = 0; d < 3; ++d) nx[d] = p.nx[d];
    E[0] = nx[0] + 2 * ng;
    E[1] = ndim > 1 ? nx[1] + 2 * ng : 1;
    E[2] = ndim > 2 ? nx[2] + 2 * ng : 1;
    S[0] = E[0] + 1;
    S[1] = E[1] + (ndim > 1 ? 1 : 0);
    S[2] = E[2] + (ndim > 2 ? 1 : 0);
    for (int d = 0; d < 3; ++d) {
      lo[d] = d < ndim ? ng : 0;
      hi[d] = d < ndim ? ng + nx[d] - 1 : 0;
    }
    sc[0] = 1;
    sc[1] = E[0];
    sc[2] = (long)E[0] * E[1];
    ss[0] = 1;
    ss[1] = S[0];
    ss[2] = (long)S[0] * S[1];
  }
  KM_HD long c(int k, int j, int i) const { return ((long)k * E[1] + j) * E[0] + i; }
  KM_HD long s(int k, int j, int i) const { return ((long)k * S[1] + j) * S[0] + i; }
  KM_HD long ncell() const { return (long)E[0] * E[1] * E[2]; }
  KM_HD long nstag() const { return (long)S[0] * S[1] * S[2]; }
};

KM_HD double *cell_ptr(const km_pack_t &p, int b, int v) { return p.cell[(long)b * p.nvar + v]; }

// ---- coordinates of a block ------------------------------------------------------
// Parthenon's uniform coordinates: face i of dimension d sits at x0 + i dx with
// x0 = xmin - lo dx (lo = first interior index), cell i's centre at x0 + (i + 1/2) dx.
KM_HD double x_face(const km_block_t &b, const Geo &g, int d, int i) {
  return (b.xmin[d] - g.lo[d] * b.dx[d]) + i * b.dx[d];
}
KM_HD double x_centre(const km_block_t &b, const Geo &g, int d, int i) {
  return (b.xmin[d] - g.lo[d] * b.dx[d]) + (i + 0.5) * b.dx[d];
}
// Cylindrical (r, z, phi) geometry, kamayan's grid::Coordinates<cylindrical>
// (geometry.hpp:35-167): Dx along phi is 2 pi; the r faces have area r dz 2 pi, the z
// faces (r+^2 - r-^2)/2 2 pi, the cells (1/2) 2 pi (r+^2 - r-^2) dz; the radial centroid
// is (2/3)(r+^2 |r+| - r-^2 |r-|)/(r+^2 - r-^2).
constexpr double kTwoPi = 2.0 * 3.141592653589793;
KM_HD double cyl_area_r(const km_block_t &b, const Geo &g, int i) {
  return x_face(b, g, 0, i) * b.dx[1] * kTwoPi;
}
KM_HD double cyl_area_z(const km_block_t &b, const Geo &g, int i) {
  const double rp = x_face(b, g, 0, i + 1), rm = x_face(b, g, 0, i);
  return 0.5 * (rp * rp - rm * rm) * kTwoPi;
}
KM_HD double cyl_volume(const km_block_t &b, const Geo &g, int i) {
  const double rp = x_face(b, g, 0, i + 1), rm = x_face(b, g, 0, i);
  return 0.5 * kTwoPi * (rp * rp - rm * rm) * b.dx[1];
}
KM_HD double cyl_centroid_r(const km_block_t &b, const Geo &g, int i) {
  const double r0 = x_face(b, g, 0, i), r1 = x_face(b, g, 0, i + 1);
  const double r0sq = r0 * r0, r1sq = r1 * r1;
  return (2.0 / 3.0) * (r1sq * __builtin_fabs(r1) - r0sq * __builtin_fabs(r0)) / (r1sq - r0sq);
}
// face area of the dir-face at cell (k, j, i) (the face on the cell's low side) and cell
// volume, for either geometry
KM_HD double face_area(const km_block_t &b, const Geo &g, bool cyl, int dir, int i) {
  if (!cyl || dir == 2) return b.area[dir];
  return dir == 0 ? cyl_area_r(b, g, i) : cyl_area_z(b, g, i);
}
KM_HD double cell_volume(const km_block_t &b, const Geo &g, bool cyl, int i) {
  return cyl ? cyl_volume(b, g, i) : b.vol;
}
// utils::Ratio (robust.hpp:29-32): a / (b + sgn(b) 10 DBL_MIN)
KM_HD double robust_ratio(double a, double b) {
  const int sgn = (0.0 <= b) - (b < 0.0);
  return a / (b + sgn * (10.0 * 2.2250738585072014e-308));
}

// failure bits of a simulation's device flag word
constexpr int KM_FAIL_STATE = 1;   // a zone with density or pressure <= 0, or non-finite
constexpr int KM_FAIL_DT_MIN = 2;  // dt below parthenon/time/dt_min for its cycle limit
constexpr int KM_FAIL_DT_MAX = 4;  // dt above parthenon/time/dt_max for its cycle limit
// A zone the hydro update cannot continue from: rho or P not positive, or a NaN or
// infinity in rho, P or the velocity (comparisons with NaN are false).
KM_HD bool unphysical(double dens, double pres, double v0, double v1, double v2) {
  const double inf = __builtin_huge_val();
  return !(dens > 0.0 && dens < inf) || !(pres > 0.0 && pres < inf) ||
         !(__builtin_fabs(v0) + __builtin_fabs(v1) + __builtin_fabs(v2) < inf);
}

// Minimum of positive doubles into *p by their bit patterns (which order like the values).
// The atomic is skipped when v is not below the value this lane reads first: within a
// launch that value can only be stale towards larger ones (the minimum only falls), so the
// result is still the minimum of every v, while most lanes of a large launch skip the
// contended atomic on one address.
__device__ __forceinline__ void dt_min_push(double *p, double v) {
  if (v < *p)
    atomicMin(reinterpret_cast<unsigned long long *>(p),
              static_cast<unsigned long long>(__double_as_longlong(v)));
}

// ghost zones of one block, enumerated without the interior: the two x3 slabs (full
// x1-x2 planes), then the two x2 slabs (interior x3 rows), then the two x1 slabs
// (interior x2-x3 rows, ng cells at each end); x1 runs fastest within each piece.
// stag = d enumerates the ghost faces of the d-staggered array instead (one more
// interior face and one more element along d); stag = -1 the cells. h in [1, ng)
// enumerates only the ghost zones within h zones of the interior (edges and corners
// included), h <= 0 all of them.
KM_HD long ghost_count(const Geo &g, int stag = -1, int h = 0) {
  if (h <= 0 || h > g.ng) h = g.ng;
  long all = 1, in = 1;
  for (int d = 0; d < 3; ++d) {
    const int s = d == stag ? 1 : 0;
    all *= d < g.ndim ? g.nx[d] + s + 2 * h : g.E[d] + s;
    in *= d < g.ndim ? g.nx[d] + s : 1;
  }
  return all - in;
}
KM_HD void ghost_cell(const Geo &g, long r, int idx[3], int stag = -1, int h = 0) {
  if (h <= 0 || h > g.ng) h = g.ng;
  const int ng = h;
  int E[3], n[3], lo[3];
  for (int d = 0; d < 3; ++d) {
    const int s = d == stag ? 1 : 0;
    n[d] = d < g.ndim ? g.nx[d] + s : 1;
    E[d] = d < g.ndim ? n[d] + 2 * h : g.E[d] + s;
    lo[d] = d < g.ndim ? h : 0;
  }
  // the enumeration over the block shrunk to h ghost layers, then shifted by ng - h
  const int sh = g.ng - h;
  auto done = [&]() {
    for (int d = 0; d < g.ndim; ++d) idx[d] += sh;
  };
  if (g.ndim > 2) {
    const long zs = (long)ng * E[1] * E[0];
    if (r < 2 * zs) {
      const int side = r >= zs;
      long q = r - (side ? zs : 0);
      idx[0] = (int)(q % E[0]);
      q /= E[0];
      idx[1] = (int)(q % E[1]);
      idx[2] = (int)(q / E[1]) + (side ? ng + n[2] : 0);
      done();
      return;
    }
    r -= 2 * zs;
  }
  if (g.ndim > 1) {
    const long ys = (long)ng * E[0] * n[2];
    if (r < 2 * ys) {
      const int side = r >= ys;
      long q = r - (side ? ys : 0);
      idx[0] = (int)(q % E[0]);
      q /= E[0];
      idx[1] = (int)(q % ng) + (side ? ng + n[1] : 0);
      idx[2] = lo[2] + (int)(q / ng);
      done();
      return;
    }
    r -= 2 * ys;
  }
  const int i = (int)(r % ng);
  long q = r / ng;
  const int side = (int)(q % 2);
  q /= 2;
  idx[0] = i + (side ? ng + n[0] : 0);
  idx[1] = lo[1] + (int)(q % n[1]);
  idx[2] = lo[2] + (int)(q / n[1]);
  done();
}

// The ghost zones a directional sweep reads: for each active dimension d, the two slabs
// of h zones next to the d faces over the interior extent of the other dimensions (no
// edges or corners, and nothing deeper than h); x1 runs fastest within a slab.
KM_HD long face_ghost_count(const Geo &g, int h) {
  long n = 0;
  for (int d = 0; d < g.ndim; ++d) {
    long slab = h;
    for (int e = 0; e < g.ndim; ++e)
      if (e != d) slab *= g.nx[e];
    n += 2 * slab;
  }
  return n;
}
KM_HD void face_ghost_cell(const Geo &g, int h, long r, int idx[3]) {
  for (int d = 0; d < g.ndim; ++d) {
    int m[3];
    for (int e = 0; e < 3; ++e) m[e] = e == d ? h : (e < g.ndim ? g.nx[e] : 1);
    const long slab = (long)m[0] * m[1] * m[2];
    if (r < 2 * slab || d == g.ndim - 1) {
      const int side = r >= slab;
      long q = r - (side ? slab : 0);
      int o[3];
      o[0] = (int)(q % m[0]);
      q /= m[0];
      o[1] = (int)(q % m[1]);
      o[2] = (int)(q / m[1]);
      for (int e = 0; e < 3; ++e)
        idx[e] = e == d ? (side ? g.hi[d] + 1 + o[e] : g.lo[d] - h + o[e]) : g.lo[e] + o[e];
      return;
    }
    r -= 2 * slab;
  }
}

// loads/stores through the global address space (the pack tables hold generic
// pointers; telling the compiler they are global gives global_load/store instead of
// flat ones)
#if defined(__HIP_DEVICE_COMPILE__)
typedef __attribute__((address_space(1))) const double gconst_double;
typedef __attribute__((address_space(1))) double gdouble;
__device__ __forceinline__ double ldg(const double *p, long i) {
  return ((gconst_double *)p)[i];
}
__device__ __forceinline__ void stg(double *p, long i, double v) { ((gdouble *)p)[i] = v; }
// v into p[i] and p[i + 1] with one 16-byte store (p + i 16-byte aligned)
typedef __attribute__((address_space(1))) double2 gdouble2;
__device__ __forceinline__ void stg2(double *p, long i, double v) {
  *(gdouble2 *)(p + i) = make_double2(v, v);
}
#else
__host__ __device__ inline double ldg(const double *p, long i) { return p[i]; }
__host__ __device__ inline void stg(double *p, long i, double v) { p[i] = v; }
__host__ __device__ inline void stg2(double *p, long i, double v) { p[i] = p[i + 1] = v; }
#endif
// Writing part of a cache line costs several times a whole line on MI355X (a probe of rows
// of 16 of 24 doubles: 1.3 against 6.6 TB/s), and the interior of a ghosted row starts
// and ends inside lines. A kernel that writes every interior zone of a row uses stg_row:
// v at p[o] and, for a zone at an x1 end of its row (edge bit 1: the first, bit 2: the
// last), also in the gx x1 ghost zones beyond it, so that the row is written whole. Only
// for arrays whose x1 ghost zones are stale until the next ghost exchange or fill.
KM_HD int row_edge(const Geo &g, int i) {
  return (i == g.lo[0] ? 1 : 0) | (i == g.hi[0] ? 2 : 0);
}
__device__ __forceinline__ void stg_row(double *p, long o, double v, int edge, int gx) {
  stg(p, o, v);
  if (!edge) return;
#pragma unroll 1
  for (int side = 1; side <= 2; side <<= 1) {
    if (!(edge & side)) continue;
    double *q0 = p + (side == 1 ? o - gx : o + 1);
    if (gx % 2 == 0 && (reinterpret_cast<unsigned long>(q0) & 15) == 0) {
#pragma unroll 1
      for (int e = 0; e < gx; e += 2) stg2(q0, e, v);
    } else {
#pragma unroll 1
      for (int e = 0; e < gx; ++e) stg(q0, e, v);
    }
  }
}
KM_HD double *flux_ptr(const km_pack_t &p, int d, int b, int v) {
  return p.flux[d][(long)b * p.nflux + v];
}

// slot of recon variable r in a hydro state pack
template <int MHD>
KM_HD int recon_slot(int r) {
  constexpr int nc = MHD ? 8 : 5;
  if (r == R_DENS) return 0;
  if (r < R_PRES) return nc + (r - R_VEL);
  if (r == R_PRES) return nc + 3;
  if (r == R_BMOD) return nc + 4;
  if (r == R_EINT) return nc + 5;
  return 5 + (r - R_MAGC);
}

// Ideal-gas closure: BMOD = (gm1 + 1) gm1 rho e = gamma PRES in every cell, and every
// reconstruction here is positively homogeneous of degree one (the PLM limiters and the PPM
// conditions scale with the data; WENO-Z up to its 1e-36 epsilon), so the reconstructed
// BMOD equals gamma times the reconstructed PRES up to rounding. The forms held to a
// tolerance (FP64-FMA, FP32-mixed) take it from PRES instead of reconstructing it; the
// FP64 form reconstructs it as the reference does.
template <class T, bool FAST>
constexpr bool bmod_from_pres() {
  return FAST || sizeof(T) == 4;
}

// runtime limiter choice (uniform across a wavefront) over a compile-time method
template <int RECON, bool FAST = false, class T>
KM_HD void reconstruct_rt(int lim, const T *w, T &vM, T &vP) {
  if constexpr (RECON == RECON_PLM || RECON == RECON_PPM) {
    if (lim == LIM_MC)
      reconstruct<RECON, LIM_MC, FAST>(w, vM, vP);
    else if (lim == LIM_VAN_LEER)
      reconstruct<RECON, LIM_VAN_LEER, FAST>(w, vM, vP);
    else
      reconstruct<RECON, LIM_MINMOD, FAST>(w, vM, vP);
  } else {
    reconstruct<RECON, LIM_MINMOD, FAST>(w, vM, vP);
  }
}

// March along one pencil: reconstruct cells c0..c1 and call face(c, vL, vR) for every
// face c in (c0, c1], the face between cells c-1 and c (stored at index c).
// load(c, T p[NR]) supplies the recon variables of cell c.
template <class T, int RECON, int NR, class Load, class Face>
__device__ __forceinline__ void march(int lim, int c0, int c1, Load &&load, Face &&face) {
  constexpr int W = ReconWidth<RECON>::value;
  T win[NR][5];
#pragma unroll
  for (int o = -W; o < W; ++o) {
    T p[NR];
    load(c0 + o, p);
#pragma unroll
    for (int r = 0; r < NR; ++r) win[r][2 + o] = p[r];
  }
  T vPprev[NR];
  for (int c = c0; c <= c1; ++c) {
    {
      T p[NR];
      load(c + W, p);
#pragma unroll
      for (int r = 0; r < NR; ++r) win[r][2 + W] = p[r];
    }
    T vM[NR], vP[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) reconstruct_rt<RECON>(lim, win[r], vM[r], vP[r]);
    if (c > c0) face(c, vPprev, vM);
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      vPprev[r] = vP[r];
#pragma unroll
      for (int o = 2 - W; o < 2 + W; ++o) win[r][o] = win[r][o + 1];
    }
  }
}

// March with the next cell's raw data fetched one step ahead, so the load latency
// overlaps the current step's reconstruction and Riemann solve.
// fetch(c) -> R (raw loads, no arithmetic); unpack(const R&, T p[NR]).
// fetch0 serves the window prefill (cells c0-W .. c0+W), fetch the steps after it.
// BP: the reconstructed BMOD is gam times the reconstructed PRES (bmod_from_pres).
template <class T, int RECON, int NR, class R, bool FAST = false, bool BP = false,
          class Fetch0, class Fetch, class Unpack, class Face>
__device__ __forceinline__ void march_pf(int lim, int c0, int c1, Fetch0 &&fetch0, Fetch &&fetch,
                                         Unpack &&unpack, Face &&face, T gam = T(0)) {
  constexpr int W = ReconWidth<RECON>::value;
  T win[NR][5];
#pragma unroll
  for (int o = -W; o < W; ++o) {
    T p[NR];
    unpack(fetch0(c0 + o), p);
#pragma unroll
    for (int r = 0; r < NR; ++r) win[r][2 + o] = p[r];
  }
  R nxt = fetch0(c0 + W);
  T vPprev[NR];
  for (int c = c0; c <= c1; ++c) {
    const R cur = nxt;
    if (c < c1) nxt = fetch(c + 1 + W);
    {
      T p[NR];
      unpack(cur, p);
#pragma unroll
      for (int r = 0; r < NR; ++r) win[r][2 + W] = p[r];
    }
    T vM[NR], vP[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r)
      if (!BP || r != R_BMOD) reconstruct_rt<RECON, FAST>(lim, win[r], vM[r], vP[r]);
    if constexpr (BP) {
      vM[R_BMOD] = gam * vM[R_PRES];
      vP[R_BMOD] = gam * vP[R_PRES];
    }
    if (c > c0) face(c, vPprev, vM);
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      vPprev[r] = vP[r];
#pragma unroll
      for (int o = 2 - W; o < 2 + W; ++o) win[r][o] = win[r][o + 1];
    }
  }
}

}  // namespace km
