// km_tasks.hip -- one gfx950 kernel per task of the reference's RK stage
// (BuildTaskListRKStage, kamayan_driver.cpp:186-248): fluxes per direction, edge EMFs,
// flux divergence / Stokes, stage combine, cons->prim + EOS, CFL dt, divB and the
// same-level ghost fill. These are the drop-in replacements for the individual hydro
// and grid tasks; the fused stage path (km_fused.hip) reuses their arithmetic.
#include <hip/hip_runtime.h>

#include "km_device.hpp"
#include "km_internal.hpp"

namespace km {

// ---------------------------------------------------------------------------------
// hydro::CalculateFluxes (hydro_add_flux_tasks.cpp:36-264): one lane per pencil.
// bx holds the (CT-widened, :59-70) interior bounds; cells bx.lo-1..bx.hi+1 along DIR are
// reconstructed and faces bx.lo..bx.hi+1 solved, as in the reference's loops.
// ---------------------------------------------------------------------------------
template <class T, int RECON, int RIEM, int MHD, int DIR>
__global__ void __launch_bounds__(256) k_flux(km_pack_t u, int lim, Box bx, int cyl) {
  constexpr int NR = HV<MHD>::NR, NC = HV<MHD>::NC;
  constexpr int DA = DIR == 0 ? 1 : 0;  // transverse dim mapped to consecutive lanes
  constexpr int DB = DIR == 2 ? 1 : 2;
  const Geo g(u);
  const int na = bx.hi[DA] - bx.lo[DA] + 1;
  const int nbb = bx.hi[DB] - bx.lo[DB] + 1;
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= (long)na * nbb * u.nblocks) return;
  int idx[3];
  idx[DA] = bx.lo[DA] + (int)(t % na);
  t /= na;
  idx[DB] = bx.lo[DB] + (int)(t % nbb);
  const int b = (int)(t / nbb);
  idx[DIR] = 0;
  const long cb = g.c(idx[2], idx[1], idx[0]);
  const long cs = g.sc[DIR];
  const double *q[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) q[r] = cell_ptr(u, b, recon_slot<MHD>(r));
  double *f[NC];
#pragma unroll
  for (int v = 0; v < NC; ++v) f[v] = flux_ptr(u, DIR, b, v);
  const double *bface = MHD ? u.face[DIR][b] : nullptr;
  const long sb = g.s(idx[2], idx[1], idx[0]);
  const long ssd = g.ss[DIR];
  // T = double: the reference's arithmetic; T = float: the FP32-mixed form (states and
  // Riemann solve in FP32, fluxes stored in FP64)
  march<T, RECON, NR>(
      lim, bx.lo[DIR] - 1, bx.hi[DIR] + 1,
      [&](int c, T *p) {
#pragma unroll
        for (int r = 0; r < NR; ++r) p[r] = T(ldg(q[r], cb + c * cs));
      },
      [&](int c, T *vL, T *vR) {
        if constexpr (MHD) {  // :104-107,167-170,233-236
          const T bf = T(ldg(bface, sb + c * ssd));
          vL[R_MAGC + DIR] = bf;
          vR[R_MAGC + DIR] = bf;
        }
        T F[NC];
        riemann<RIEM, MHD, DIR>(vL, vR, F);
        double out[NC];
#pragma unroll
        for (int v = 0; v < NC; ++v) out[v] = double(F[v]);
        // cylindrical CT: the B_phi flux divided by the face's x2 position, as written
        // (hydro_add_flux_tasks.cpp:171-178)
        if (MHD && DIR == 1 && cyl)
          out[C_MAGC + 2] *= robust_ratio(1.0, x_face(u.blocks[b], g, 1, c));
#pragma unroll
        for (int v = 0; v < NC; ++v) stg(f[v], cb + c * cs, out[v]);
      });
}

// ---------------------------------------------------------------------------------
// hydro::CalculateFluxes along x1 with one lane per face. The march above gives every lane
// its own row; along x1 that makes each load and store touch 64 separate cache lines.
// Here the lanes of a wavefront cover consecutive faces of a row, so every access is
// contiguous. Each lane reconstructs the two cells beside its face from the same
// stencil values the march reads, so the faces are bitwise the march's.
// ---------------------------------------------------------------------------------
template <class T, int RECON, int RIEM, int MHD>
__global__ void __launch_bounds__(256) k_flux_x1(km_pack_t u, int lim, Box bx, int cyl) {
  constexpr int NR = HV<MHD>::NR, NC = HV<MHD>::NC;
  constexpr int W = ReconWidth<RECON>::value;
  const Geo g(u);
  const int nf = bx.hi[0] - bx.lo[0] + 2;  // faces bx.lo .. bx.hi + 1
  const int n1 = bx.hi[1] - bx.lo[1] + 1, n2 = bx.hi[2] - bx.lo[2] + 1;
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= (long)nf * n1 * n2 * u.nblocks) return;
  const int c = bx.lo[0] + (int)(t % nf);
  t /= nf;
  const int j = bx.lo[1] + (int)(t % n1);
  t /= n1;
  const int k = bx.lo[2] + (int)(t % n2);
  const int b = (int)(t / n2);
  const long row = g.c(k, j, 0);
  T vL[NR], vR[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    const double *q = cell_ptr(u, b, recon_slot<MHD>(r));
    T s[2 * W + 2];  // cells c-1-W .. c+W
#pragma unroll
    for (int o = 0; o < 2 * W + 2; ++o) s[o] = T(ldg(q, row + c - 1 - W + o));
    T wl[5], wr[5];
#pragma unroll
    for (int o = -W; o <= W; ++o) {
      wl[2 + o] = s[W + o];
      wr[2 + o] = s[W + 1 + o];
    }
    T m, p;
    reconstruct_rt<RECON>(lim, wl, m, p);
    vL[r] = p;
    reconstruct_rt<RECON>(lim, wr, m, p);
    vR[r] = m;
  }
  if constexpr (MHD) {  // :104-107
    const T bf = T(ldg(u.face[0][b], g.s(k, j, c)));
    vL[R_MAGC + 0] = bf;
    vR[R_MAGC + 0] = bf;
  }
  T F[NC];
  riemann<RIEM, MHD, 0>(vL, vR, F);
  double out[NC];
#pragma unroll
  for (int v = 0; v < NC; ++v) out[v] = double(F[v]);
  if (cyl) {  // :108-117: the phi momentum flux carries r (angular momentum)
    const double rf = x_face(u.blocks[b], g, 0, c);
    out[C_MOM + 2] *= rf;
    if constexpr (MHD) out[C_MAGC + 2] *= robust_ratio(1.0, rf);
  }
#pragma unroll
  for (int v = 0; v < NC; ++v) stg(flux_ptr(u, 0, b, v), row + c, out[v]);
}

// ---------------------------------------------------------------------------------
// hydro::CalculateEMF (hydro_add_flux_tasks.cpp:387-442), arithmetic averaging
// E3(i-1/2,j-1/2) = 1/4 [Fy(Bx)(i-1,j) + Fy(Bx)(i,j) - Fx(By)(i,j-1) - Fx(By)(i,j)] and
// cyclic; each edge family over its own interior range.
// ---------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_emf(km_pack_t u, double esign) {
  const Geo g(u);
  const long n = g.nstag();
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= n * u.nblocks) return;
  const int b = (int)(t / n);
  long r = t % n;
  const int i = (int)(r % g.S[0]);
  r /= g.S[0];
  const int j = (int)(r % g.S[1]);
  const int k = (int)(r / g.S[1]);
  const long s = g.s(k, j, i);
  const long sj = g.sc[1], sk = g.sc[2];
  const bool in_i = i >= g.lo[0] && i <= g.hi[0], in_i1 = i >= g.lo[0] && i <= g.hi[0] + 1;
  const bool in_j = j >= g.lo[1] && j <= g.hi[1], in_j1 = j >= g.lo[1] && j <= g.hi[1] + 1;
  const bool in_k = k >= g.lo[2] && k <= g.hi[2];
  const bool in_k1 = k >= g.lo[2] && k <= g.hi[2] + (g.ndim > 2 ? 1 : 0);
  if (i >= g.E[0] || j >= g.E[1] || k >= g.E[2]) return;
  const long c = g.c(k, j, i);
  if (in_i1 && in_j1 && in_k) {
    const double *fyBx = flux_ptr(u, 1, b, C_MAGC + 0), *fxBy = flux_ptr(u, 0, b, C_MAGC + 1);
    u.edge[2][b][s] = esign * 0.25 * (fyBx[c - 1] + fyBx[c] - fxBy[c - sj] - fxBy[c]);
  }
  if (g.ndim > 2) {
    if (in_i && in_j1 && in_k1) {
      const double *fzBy = flux_ptr(u, 2, b, C_MAGC + 1), *fyBz = flux_ptr(u, 1, b, C_MAGC + 2);
      u.edge[0][b][s] = esign * 0.25 * (fzBy[c - sj] + fzBy[c] - fyBz[c - sk] - fyBz[c]);
    }
    if (in_i1 && in_j && in_k1) {
      const double *fxBz = flux_ptr(u, 0, b, C_MAGC + 2), *fzBx = flux_ptr(u, 2, b, C_MAGC + 0);
      u.edge[1][b][s] = esign * 0.25 * (fxBz[c - sk] + fxBz[c] - fzBx[c - 1] - fzBx[c]);
    }
  }
}

// interior cell (b, k, j, i) of thread t; false past the end
__device__ __forceinline__ bool interior_cell(const Geo &g, int nblocks, int &b, int &k, int &j,
                                              int &i) {
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long n = (long)g.nx[0] * (g.ndim > 1 ? g.nx[1] : 1) * (g.ndim > 2 ? g.nx[2] : 1);
  if (t >= n * nblocks) return false;
  b = (int)(t / n);
  long r = t % n;
  i = g.lo[0] + (int)(r % g.nx[0]);
  r /= g.nx[0];
  const int ny = g.ndim > 1 ? g.nx[1] : 1;
  j = g.lo[1] + (int)(r % ny);
  k = g.lo[2] + (int)(r / ny);
  return true;
}

// ---------------------------------------------------------------------------------
// Mass scalars (MFRAC, material unit). Fluxes (hydro_add_flux_tasks.cpp:119-135,
// 181-197, 240-256): one lane per face along DIR; each species is reconstructed as the
// primitives are, and its flux is the face's mass flux times the upwind side's mass
// fraction, F_s = F_rho > 0 ? F_rho X_s(left) : F_rho X_s(right). Along x2 and x3 the
// reference's loops run s < upper bound, so the last species gets no flux there (its flux
// arrays stay zero). slot0: cell slot of species 0; flux0: its flux slot.
// ---------------------------------------------------------------------------------
template <int RECON, int DIR>
__global__ void __launch_bounds__(256) k_species_flux(km_pack_t u, int lim, Box bx, int nsp,
                                                      int slot0, int flux0) {
  constexpr int W = ReconWidth<RECON>::value;
  constexpr int DA = DIR == 0 ? 1 : 0, DB = DIR == 2 ? 1 : 2;
  const Geo g(u);
  const int nf = bx.hi[DIR] - bx.lo[DIR] + 2;
  const int na = bx.hi[DA] - bx.lo[DA] + 1, nbb = bx.hi[DB] - bx.lo[DB] + 1;
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= (long)nf * na * nbb * u.nblocks) return;
  int idx[3];
  idx[DIR] = bx.lo[DIR] + (int)(t % nf);
  t /= nf;
  idx[DA] = bx.lo[DA] + (int)(t % na);
  t /= na;
  idx[DB] = bx.lo[DB] + (int)(t % nbb);
  const int b = (int)(t / nbb);
  const long c = g.c(idx[2], idx[1], idx[0]);
  const long cs = g.sc[DIR];
  const double frho = flux_ptr(u, DIR, b, 0)[c];
  const int ns = DIR == 0 ? nsp : nsp - 1;
  for (int sp = 0; sp < ns; ++sp) {
    const double *q = cell_ptr(u, b, slot0 + sp);
    double st[2 * W + 2];  // cells c-1-W .. c+W along DIR
#pragma unroll
    for (int o = 0; o < 2 * W + 2; ++o) st[o] = ldg(q, c + (o - 1 - W) * cs);
    double wl[5], wr[5];
#pragma unroll
    for (int o = -W; o <= W; ++o) {
      wl[2 + o] = st[W + o];
      wr[2 + o] = st[W + 1 + o];
    }
    double m, pl, pr;
    reconstruct_rt<RECON>(lim, wl, m, pl);
    reconstruct_rt<RECON>(lim, wr, pr, m);
    stg(flux_ptr(u, DIR, b, flux0 + sp), c, frho > 0.0 ? frho * pl : frho * pr);
  }
}

// FluxesToDuDt for the species: dudt slot slot0 + s from flux slot flux0 + s
__global__ void __launch_bounds__(256) k_species_dudt(km_pack_t u, km_pack_t dudt, int nsp,
                                                      int slot0, int flux0, int cyl) {
  const Geo g(u);
  int b, k, j, i;
  if (!interior_cell(g, u.nblocks, b, k, j, i)) return;
  const km_block_t &blk = u.blocks[b];
  const long c = g.c(k, j, i);
  const double vol = cell_volume(blk, g, cyl, i);
  for (int sp = 0; sp < nsp; ++sp) {
    double du = 0.;
#pragma unroll
    for (int dir = 0; dir < 3; ++dir) {
      if (dir < g.ndim) {
        const double *F = flux_ptr(u, dir, b, flux0 + sp);
        du -= (face_area(blk, g, cyl, dir, i + (dir == 0)) * F[c + g.sc[dir]] -
               face_area(blk, g, cyl, dir, i) * F[c]);
      }
    }
    cell_ptr(dudt, b, slot0 + sp)[c] = du / vol;
  }
}

// material::PrepareConserved, the per-stage task on md0 (material.cpp:132-155):
// X_s -> rho X_s
__global__ void __launch_bounds__(256) k_species_to_conserved(km_pack_t u, int nsp, int slot0) {
  const Geo g(u);
  int b, k, j, i;
  if (!interior_cell(g, u.nblocks, b, k, j, i)) return;
  const long c = g.c(k, j, i);
  const double rho = cell_ptr(u, b, KM_DENS)[c];
  for (int sp = 0; sp < nsp; ++sp) {
    double *x = cell_ptr(u, b, slot0 + sp);
    x[c] = x[c] * rho;
  }
}

// material::PreparePrimitive (material.cpp:157-197) and PostMeshInitialization (:104-130):
// X_s -> X_s / sum_s X_s, the sum taken in species order
__global__ void __launch_bounds__(256) k_species_normalize(km_pack_t u, int nsp, int slot0) {
  const Geo g(u);
  int b, k, j, i;
  if (!interior_cell(g, u.nblocks, b, k, j, i)) return;
  const long c = g.c(k, j, i);
  double sum = 0.0;
  for (int sp = 0; sp < nsp; ++sp) sum += cell_ptr(u, b, slot0 + sp)[c];
  for (int sp = 0; sp < nsp; ++sp) {
    double *x = cell_ptr(u, b, slot0 + sp);
    x[c] = x[c] * (1.0 / sum);
  }
}

// One stage of the species in one pass per zone, in the reference's task order: md0's
// X_s -> rho X_s with md0's density (the per-stage material::PrepareConserved, which runs
// before ApplyDuDt mixes md0), the flux divergence, ApplyDuDt, and md1's normalisation
// (material::PreparePrimitive). Every value rounds as in the separate tasks. It runs before
// the hydro update, which rewrites md0's density. When base and md0 are one container the
// mix reads the converted value back, as the reference's ApplyDuDt does.
__global__ void __launch_bounds__(256) k_species_update(km_pack_t base, km_pack_t u0,
                                                        km_pack_t u1, int nsp, int slot0,
                                                        int flux0, double beta, double dt,
                                                        const double *dtp, const int *inst,
                                                        int cyl) {
  const Geo g(u0);
  int b, k, j, i;
  if (!interior_cell(g, u0.nblocks, b, k, j, i)) return;
  if (dtp) dt = dtp[inst ? inst[b] : 0];
  const km_block_t &blk = u0.blocks[b];
  const long c = g.c(k, j, i);
  const double rho = cell_ptr(u0, b, KM_DENS)[c];
  const double vol = cell_volume(blk, g, cyl, i);
  double sum = 0.0;
  for (int sp = 0; sp < nsp; ++sp) {
    double *p0 = cell_ptr(u0, b, slot0 + sp);
    p0[c] = p0[c] * rho;
    double du = 0.;
#pragma unroll
    for (int dir = 0; dir < 3; ++dir) {
      if (dir < g.ndim) {
        const double *F = flux_ptr(u0, dir, b, flux0 + sp);
        du -= (face_area(blk, g, cyl, dir, i + (dir == 0)) * F[c + g.sc[dir]] -
               face_area(blk, g, cyl, dir, i) * F[c]);
      }
    }
    du = du / vol;
    const double mix = beta * cell_ptr(base, b, slot0 + sp)[c] + (1.0 - beta) * p0[c];
    p0[c] = mix;
    const double x1 = mix + beta * dt * du;
    cell_ptr(u1, b, slot0 + sp)[c] = x1;
    sum += x1;
  }
  for (int sp = 0; sp < nsp; ++sp) {
    double *x = cell_ptr(u1, b, slot0 + sp);
    x[c] = x[c] * (1.0 / sum);
  }
}

// grid::FluxDivergence (grid_update.hpp:15-57)
__global__ void __launch_bounds__(256) k_flux_div(km_pack_t u, km_pack_t dudt, int nc, int cyl) {
  const Geo g(u);
  int b, k, j, i;
  if (!interior_cell(g, u.nblocks, b, k, j, i)) return;
  const km_block_t &blk = u.blocks[b];
  const long c = g.c(k, j, i);
  const long sc[3] = {g.sc[0], g.sc[1], g.sc[2]};
  double am[3], ap[3];
#pragma unroll
  for (int dir = 0; dir < 3; ++dir) {
    am[dir] = face_area(blk, g, cyl, dir, i);
    ap[dir] = face_area(blk, g, cyl, dir, i + (dir == 0));
  }
  const double vol = cell_volume(blk, g, cyl, i);
  for (int var = 0; var < nc; ++var) {
    double du = 0.;
#pragma unroll
    for (int dir = 0; dir < 3; ++dir) {
      if (dir < g.ndim) {
        const double *F = flux_ptr(u, dir, b, var);
        du -= (ap[dir] * F[c + sc[dir]] - am[dir] * F[c]);
      }
    }
    cell_ptr(dudt, b, var)[c] = du / vol;
  }
}

// hydro::AddSourceTerms (hydro.cpp:199-234) with AddGeometricSource
// (primconsflux.hpp:144-159): in cylindrical geometry the r momentum gains
// (rho v_phi^2 + P) / r at the cell centre (+ (pmag - Bphi^2) / r with MHD)
template <int MHD>
__global__ void __launch_bounds__(256) k_geom_source(km_pack_t u, km_pack_t dudt) {
  constexpr int nc = HV<MHD>::NC;
  const Geo g(u);
  int b, k, j, i;
  if (!interior_cell(g, u.nblocks, b, k, j, i)) return;
  const long c = g.c(k, j, i);
  const double xi = x_centre(u.blocks[b], g, 0, i);
  const double rho = cell_ptr(u, b, KM_DENS)[c], vphi = cell_ptr(u, b, nc + 2)[c];
  double s = (rho * vphi * vphi + cell_ptr(u, b, nc + 3)[c]) / xi;
  double *d = cell_ptr(dudt, b, KM_MOM0);
  d[c] = d[c] + s;
  if constexpr (MHD) {
    const double b0 = cell_ptr(u, b, KM_MAGC0)[c], b1 = cell_ptr(u, b, KM_MAGC0 + 1)[c],
                 b2 = cell_ptr(u, b, KM_MAGC0 + 2)[c];
    const double pmag = 0.5 * (b0 * b0 + b1 * b1 + b2 * b2);
    d[c] = d[c] + (pmag - b2 * b2) / xi;
  }
}

// grid::FluxStokes<geom, Face, edges...> (grid_update.hpp:80-137)
template <int FACE>
__global__ void __launch_bounds__(256) k_stokes(km_pack_t u, km_pack_t dudt, int cyl) {
  const Geo g(u);
  const int di = FACE == 0, dj = FACE == 1, dk = FACE == 2;
  const int ni = g.nx[0] + di, nj = (g.ndim > 1 ? g.nx[1] : 1) + dj,
            nk = (g.ndim > 2 ? g.nx[2] : 1) + dk;
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long n = (long)ni * nj * nk;
  if (t >= n * u.nblocks) return;
  const int b = (int)(t / n);
  long r = t % n;
  const int i = g.lo[0] + (int)(r % ni);
  r /= ni;
  const int j = g.lo[1] + (int)(r % nj);
  const int k = g.lo[2] + (int)(r / nj);
  const km_block_t &blk = u.blocks[b];
  const long s = g.s(k, j, i);
  // (edge, axis of the upper edge, sign): AxisFromFaceEdge + permutation sign,
  // grid_update.hpp:59-78,115-125; template edge order F1:E3,E2  F2:E3,E1  F3:E1,E2
  constexpr int e0 = FACE == 0 ? 2 : FACE == 1 ? 2 : 0;
  constexpr int a0 = FACE == 0 ? 1 : FACE == 1 ? 0 : 1;
  constexpr double s0 = FACE == 0 ? 1.0 : -1.0;
  constexpr int e1 = FACE == 0 ? 1 : FACE == 1 ? 0 : 1;
  constexpr int a1 = FACE == 0 ? 2 : FACE == 1 ? 2 : 0;
  constexpr double s1 = FACE == 0 ? -1.0 : 1.0;
  // cylindrical r, z, phi is an odd permutation (:122-125): every sign flips; E3 edges
  // have length |r| 2 pi (geometry.hpp:143-150), the others their Dx
  const double gs = cyl ? -1.0 : 1.0;
  auto len = [&](int e, int ii) {
    return (cyl && e == 2) ? __builtin_fabs(x_face(blk, g, 0, ii)) * kTwoPi : blk.dx[e];
  };
  double acc = 0.;
  {
    const double *E = u.edge[e0][b];
    acc += (gs * s0) * (len(e0, i) * E[s] - len(e0, i + (a0 == 0)) * E[s + g.ss[a0]]);
  }
  if (g.ndim > 2) {
    const double *E = u.edge[e1][b];
    acc += (gs * s1) * (len(e1, i) * E[s] - len(e1, i + (a1 == 0)) * E[s + g.ss[a1]]);
  }
  acc *= 1. / (face_area(blk, g, cyl, FACE, i) + 1.e-36);
  dudt.face[FACE][b][s] = acc;
}

// grid::ApplyDuDt_impl (grid.cpp:213-241), cell part
__global__ void __launch_bounds__(256) k_apply_cell(km_pack_t base, km_pack_t u0, km_pack_t u1,
                                                    km_pack_t dudt, int v0, int v1,
                                                    double beta, double dt, const double *dtp,
                                                    const int *inst) {
  const Geo g(u0);
  int b, k, j, i;
  if (!interior_cell(g, u0.nblocks, b, k, j, i)) return;
  if (dtp) dt = dtp[inst ? inst[b] : 0];
  const long c = g.c(k, j, i);
  for (int var = v0; var < v1; ++var) {
    double *p0 = cell_ptr(u0, b, var);
    const double mix = beta * cell_ptr(base, b, var)[c] + (1.0 - beta) * p0[c];
    p0[c] = mix;
    cell_ptr(u1, b, var)[c] = mix + beta * dt * cell_ptr(dudt, b, var)[c];
  }
}

// grid::ApplyDuDt_impl (grid.cpp:262-282), face part
template <int FACE>
__global__ void __launch_bounds__(256) k_apply_face(km_pack_t base, km_pack_t u0, km_pack_t u1,
                                                    km_pack_t dudt, double beta, double dt,
                                                    const double *dtp, const int *inst) {
  const Geo g(u0);
  const int di = FACE == 0, dj = FACE == 1, dk = FACE == 2;
  const int ni = g.nx[0] + di, nj = (g.ndim > 1 ? g.nx[1] : 1) + dj,
            nk = (g.ndim > 2 ? g.nx[2] : 1) + dk;
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long n = (long)ni * nj * nk;
  if (t >= n * u0.nblocks) return;
  const int b = (int)(t / n);
  long r = t % n;
  const int i = g.lo[0] + (int)(r % ni);
  r /= ni;
  const int j = g.lo[1] + (int)(r % nj);
  const int k = g.lo[2] + (int)(r / nj);
  const long s = g.s(k, j, i);
  if (dtp) dt = dtp[inst ? inst[b] : 0];
  double *p0 = u0.face[FACE][b];
  const double v0 = beta * base.face[FACE][b][s] + (1.0 - beta) * p0[s];
  p0[s] = v0;
  u1.face[FACE][b][s] = v0 + beta * dt * dudt.face[FACE][b][s];
}

// CT cell-centred B from face B (primconsflux.cpp:102-128)
template <int MHD>
__device__ __forceinline__ void face_average_b(const km_pack_t &u, const Geo &g, int b, int k,
                                               int j, int i, int cyl) {
  if constexpr (MHD) {
    const km_block_t &blk = u.blocks[b];
    const long c = g.c(k, j, i), s = g.s(k, j, i);
    if (g.ndim > 1) {
      const double *B1 = u.face[0][b], *B2 = u.face[1][b];
      const double V = cell_volume(blk, g, cyl, i);
      cell_ptr(u, b, 5)[c] = 0.5 * blk.dx[0] *
                             (face_area(blk, g, cyl, 0, i + 1) * B1[s + 1] +
                              face_area(blk, g, cyl, 0, i) * B1[s]) /
                             V;
      cell_ptr(u, b, 6)[c] = 0.5 * blk.dx[1] *
                             (face_area(blk, g, cyl, 1, i) * B2[s + g.ss[1]] +
                              face_area(blk, g, cyl, 1, i) * B2[s]) /
                             V;
    }
    if (g.ndim > 2) {
      const double *B3 = u.face[2][b];
      cell_ptr(u, b, 7)[c] =
          0.5 * blk.dx[2] * (blk.area[2] * B3[s + g.ss[2]] + blk.area[2] * B3[s]) / blk.vol;
    }
  }
}

// primitives of one zone from its conserved values U (cell-centred B in U[5..7] for MHD);
// rinv != 0: cylindrical geometry, where the stored phi momentum is r rho v_phi and
// PreparePrimitive_impl (primconsflux.cpp:129-137) scales v_phi by Ratio(1, r_centroid)
// after Cons2Prim (whose kinetic energy uses the stored phi momentum as is)
template <int MHD>
__device__ __forceinline__ void primitive_zone(const km_pack_t &u, int b, long c, const Eos &eos,
                                               const double *U, double rinv = 0.0,
                                               double rc = 0.0) {
  constexpr int nc = HV<MHD>::NC;
  const double dens = U[0];
  const double idens = 1.0 / dens;
  double emag = 0., ekin = 0.;
#pragma unroll
  for (int dir = 0; dir < 3; dir++) {
    const double v = idens * U[1 + dir];
    cell_ptr(u, b, nc + dir)[c] = (dir == 2 && rinv != 0.0) ? v * rinv : v;
    ekin += v * v;
    if constexpr (MHD) {
      const double B = U[5 + dir];
      emag += B * B;
    }
  }
  if constexpr (MHD)
    if (rinv != 0.0) cell_ptr(u, b, KM_MAGC0 + 2)[c] = U[7] * rc;  // B_phi back from B_phi / r
  ekin *= 0.5 * dens;
  emag *= 0.5;
  const double eint = U[4] - ekin - emag;
  const double sie = idens * eint;
  cell_ptr(u, b, nc + 5)[c] = sie;
  cell_ptr(u, b, nc + 3)[c] = eos_pres(eos, dens, sie);
  cell_ptr(u, b, nc + 6)[c] = eos_temp(eos, dens, sie);
  cell_ptr(u, b, nc + 4)[c] = eos_bmod(eos, dens, sie);
}

// hydro PreparePrimitive_impl (primconsflux.cpp:82-142) + eos EosWrapped(ener)
template <int MHD>
__global__ void __launch_bounds__(256) k_prepare_primitive(km_pack_t u, Eos eos, int cyl) {
  constexpr int nc = HV<MHD>::NC;
  const Geo g(u);
  int b, k, j, i;
  if (!interior_cell(g, u.nblocks, b, k, j, i)) return;
  face_average_b<MHD>(u, g, b, k, j, i, cyl);
  const long c = g.c(k, j, i);
  double U[nc];
#pragma unroll
  for (int v = 0; v < nc; ++v) U[v] = cell_ptr(u, b, v)[c];
  const double rc = cyl ? cyl_centroid_r(u.blocks[b], g, i) : 0.0;
  primitive_zone<MHD>(u, b, c, eos, U, cyl ? robust_ratio(1.0, rc) : 0.0, rc);
}

// hydro::PrepareConserved (hydro.cpp:244-286), the per-stage task before ApplyDuDt: with CT
// in cylindrical geometry the cell B_phi of md0 becomes B_phi / r_centroid
__global__ void __launch_bounds__(256) k_cyl_stage_conserved(km_pack_t u) {
  const Geo g(u);
  int b, k, j, i;
  if (!interior_cell(g, u.nblocks, b, k, j, i)) return;
  double *q = cell_ptr(u, b, KM_MAGC0 + 2);
  const long c = g.c(k, j, i);
  q[c] = q[c] * (1.0 / cyl_centroid_r(u.blocks[b], g, i));
}

// FluxesToDuDt (grid_update.hpp:20-57), the cell part of ApplyDuDt (grid.cpp:213-241)
// and, for hydro, PreparePrimitive of one interior zone in a single pass: the zone's
// dU/dt and new conserved state stay in registers, with the arithmetic of k_flux_div,
// k_apply_cell and k_prepare_primitive unchanged. MHD keeps the primitive pass apart,
// since cell-centred B comes from face B that k_apply_face updates afterwards.
template <int MHD>
__global__ void __launch_bounds__(256) k_update_cell(km_pack_t base, km_pack_t u0, km_pack_t u1,
                                                     double beta, double dt, const double *dtp,
                                                     const int *inst, Eos eos, int cyl) {
  constexpr int nc = HV<MHD>::NC;
  const Geo g(u0);
  int b, k, j, i;
  if (!interior_cell(g, u0.nblocks, b, k, j, i)) return;
  if (dtp) dt = dtp[inst ? inst[b] : 0];
  const km_block_t &blk = u0.blocks[b];
  const long c = g.c(k, j, i);
  const long sc[3] = {g.sc[0], g.sc[1], g.sc[2]};
  double am[3], ap[3];
#pragma unroll
  for (int dir = 0; dir < 3; ++dir) {
    am[dir] = face_area(blk, g, cyl, dir, i);
    ap[dir] = face_area(blk, g, cyl, dir, i + (dir == 0));
  }
  const double vol = cell_volume(blk, g, cyl, i);
  double du[nc], U[nc];
#pragma unroll
  for (int var = 0; var < nc; ++var) {
    double d = 0.;
#pragma unroll
    for (int dir = 0; dir < 3; ++dir) {
      if (dir < g.ndim) {
        const double *F = flux_ptr(u0, dir, b, var);
        d -= (ap[dir] * F[c + sc[dir]] - am[dir] * F[c]);
      }
    }
    du[var] = d / vol;
  }
  if (cyl && !MHD) {  // hydro::AddSourceTerms between FluxesToDuDt and ApplyDuDt
    const double xi = x_centre(blk, g, 0, i);
    const double rho = cell_ptr(u0, b, KM_DENS)[c], vphi = cell_ptr(u0, b, nc + 2)[c];
    du[1] = du[1] + (rho * vphi * vphi + cell_ptr(u0, b, nc + 3)[c]) / xi;
  }
#pragma unroll
  for (int var = 0; var < nc; ++var)
    U[var] = beta * cell_ptr(base, b, var)[c] + (1.0 - beta) * cell_ptr(u0, b, var)[c];
#pragma unroll
  for (int var = 0; var < nc; ++var) {
    cell_ptr(u0, b, var)[c] = U[var];
    U[var] = U[var] + beta * dt * du[var];
    cell_ptr(u1, b, var)[c] = U[var];
  }
  if constexpr (!MHD)
    primitive_zone<0>(u1, b, c, eos, U, cyl ? robust_ratio(1.0, cyl_centroid_r(blk, g, i)) : 0.0);
}

// eos PrepareConserved(mode_init) (eos.cpp:165-168, the outputs of SingularityEosFill,
// eos_singularity.hpp:17-36) then hydro ConvertToConserved_impl (primconsflux.cpp:19-79):
//   dens_pres: EINT = P / (gm1 rho), TEMP, BMOD
//   dens_ener: PRES = gm1 rho e, TEMP, BMOD
//   dens_temp: EINT = max(0, cv T), PRES, BMOD
template <int MHD>
__global__ void __launch_bounds__(256) k_prepare_conserved(km_pack_t u, Eos eos, int cyl,
                                                           int mode) {
  constexpr int nc = HV<MHD>::NC, NR = HV<MHD>::NR;
  const Geo g(u);
  int b, k, j, i;
  if (!interior_cell(g, u.nblocks, b, k, j, i)) return;
  const long c = g.c(k, j, i);
  const double dens = cell_ptr(u, b, 0)[c];
  if (mode == KM_EOS_INIT_DENS_ENER) {
    const double sie = cell_ptr(u, b, nc + 5)[c];
    cell_ptr(u, b, nc + 3)[c] = eos_pres(eos, dens, sie);
    cell_ptr(u, b, nc + 6)[c] = eos_temp(eos, dens, sie);
    cell_ptr(u, b, nc + 4)[c] = eos_bmod(eos, dens, sie);
  } else if (mode == KM_EOS_INIT_DENS_TEMP) {
    const double sie = kmax(0.0, eos.cv * cell_ptr(u, b, nc + 6)[c]);
    cell_ptr(u, b, nc + 5)[c] = sie;
    cell_ptr(u, b, nc + 3)[c] = eos_pres(eos, dens, sie);
    cell_ptr(u, b, nc + 4)[c] = eos_bmod(eos, dens, sie);
  } else {
    const double sie = cell_ptr(u, b, nc + 3)[c] / (eos.gm1 * dens);
    cell_ptr(u, b, nc + 5)[c] = sie;
    cell_ptr(u, b, nc + 6)[c] = eos_temp(eos, dens, sie);
    cell_ptr(u, b, nc + 4)[c] = eos_bmod(eos, dens, sie);
  }
  face_average_b<MHD>(u, g, b, k, j, i, cyl);
  double V[NR], U[nc];
#pragma unroll
  for (int r = 0; r < NR; ++r) V[r] = cell_ptr(u, b, recon_slot<MHD>(r))[c];
  prim2cons<MHD>(V, U);
  // cylindrical: conserve angular momentum (primconsflux.cpp:70-74)
  if (cyl) U[KM_MOM0 + 2] *= cyl_centroid_r(u.blocks[b], g, i);
#pragma unroll
  for (int v = 0; v < nc; ++v) cell_ptr(u, b, v)[c] = U[v];
}

__device__ __forceinline__ void atomic_min_pos(double *out, double v) { dt_min_push(out, v); }

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = kmin(v, __shfl_xor(v, off, 64));
  return v;
}

__global__ void k_fill_value(double *p, double v, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}

// hydro EstimateTimeStep (hydro_time_step.cpp:19-65): min over the interior of
// dx/(|v| + cf); *out must hold DBL_MAX on entry
template <int MHD>
__global__ void __launch_bounds__(256) k_estimate_dt(km_pack_t u, double *out, const int *inst,
                                                     int cyl, int *fail) {
  constexpr int NR = HV<MHD>::NR;
  const Geo g(u);
  int b = 0, k, j, i;
  double dtl = DBL_MAX;
  bool bad = false;
  const bool valid = interior_cell(g, u.nblocks, b, k, j, i);
  if (valid) {
    const km_block_t &blk = u.blocks[b];
    const long c = g.c(k, j, i);
    double V[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) V[r] = cell_ptr(u, b, recon_slot<MHD>(r))[c];
    dtl = kmin(dtl, blk.dx[0] / (kabs(V[R_VEL + 0]) + fast_speed<MHD, 0>(V)));
    if (g.ndim > 1) dtl = kmin(dtl, blk.dx[1] / (kabs(V[R_VEL + 1]) + fast_speed<MHD, 1>(V)));
    if (g.ndim > 2) dtl = kmin(dtl, blk.dx[2] / (kabs(V[R_VEL + 2]) + fast_speed<MHD, 2>(V)));
    // cylindrical: the azimuthal crossing time r / |v_phi| (hydro_time_step.cpp:56-60)
    if (cyl) dtl = kmin(dtl, x_centre(blk, g, 0, i) / kabs(V[R_VEL + 2]));
    bad = unphysical(V[R_DENS], V[R_PRES], V[R_VEL + 0], V[R_VEL + 1], V[R_VEL + 2]);
  }
  if (fail && __any(bad) && (threadIdx.x & 63) == 0) atomicOr(fail, KM_FAIL_STATE);
  if (inst) {
    // instances may change inside a workgroup: reduce per wavefront when it holds one
    // instance, else per lane
    const int my = valid ? inst[b] : -1;
    const int first = __shfl(my, 0, 64);
    if (__all(my == first)) {
      dtl = wave_min(dtl);
      if ((threadIdx.x & 63) == 0 && first >= 0) atomic_min_pos(out + first, dtl);
    } else if (valid) {
      atomic_min_pos(out + my, dtl);
    }
    return;
  }
  dtl = wave_min(dtl);
  __shared__ double part[4];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = dtl;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = part[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = kmin(m, part[w]);
    atomic_min_pos(out, m);
  }
}

__global__ void k_scale_value(double *p, double s, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = p[t] * s;
}

// hydro FillDerived_impl (hydro.cpp:150-194)
__global__ void __launch_bounds__(256) k_fill_divb(km_pack_t u, int slot, int cyl) {
  const Geo g(u);
  int b, k, j, i;
  if (!interior_cell(g, u.nblocks, b, k, j, i)) return;
  const km_block_t &blk = u.blocks[b];
  const long s = g.s(k, j, i);
  double divb = 0.;
  const long ss[3] = {g.ss[0], g.ss[1], g.ss[2]};
#pragma unroll
  for (int dir = 0; dir < 3; dir++) {
    if (dir < g.ndim) {
      const double *B = u.face[dir][b];
      divb += face_area(blk, g, cyl, dir, i + (dir == 0)) * B[s + ss[dir]] -
              face_area(blk, g, cyl, dir, i) * B[s];
    }
  }
  cell_ptr(u, b, slot)[g.c(k, j, i)] = divb / cell_volume(blk, g, cyl, i);
}

// ---------------------------------------------------------------------------------
// Same-level ghost fill of a uniform block grid (Parthenon boundary exchange plus
// periodic / outflow / reflect / axisymmetric boundaries): every ghost zone, including
// edges and corners, maps per dimension to a source zone of the same or a neighbouring
// block. Reflect and axisymmetric mirror the zone and flip the normal vector component;
// axisymmetric (Axisymmetric, boundary_conditions.cpp:23-60) flips the phi component too.
// ---------------------------------------------------------------------------------
__device__ __forceinline__ int map_cell(int g, int N, int lo, int hi, bool &flip) {
  if (g < 0) {
    if (lo == KM_BC_PERIODIC) return ((g % N) + N) % N;
    if (lo == KM_BC_OUTFLOW) return 0;
    flip = !flip;
    return -g - 1;
  }
  if (g >= N) {
    if (hi == KM_BC_PERIODIC) return g % N;
    if (hi == KM_BC_OUTFLOW) return N - 1;
    flip = !flip;
    return 2 * N - g - 1;
  }
  return g;
}
__device__ __forceinline__ int map_face(int g, int N, int lo, int hi, bool &flip) {
  if (g < 0) {
    if (lo == KM_BC_PERIODIC) return ((g % N) + N) % N;
    if (lo == KM_BC_OUTFLOW) return 0;
    flip = !flip;
    return -g;
  }
  if (g > N) {
    if (hi == KM_BC_PERIODIC) return g % N;
    if (hi == KM_BC_OUTFLOW) return N;
    flip = !flip;
    return 2 * N - g;
  }
  return g;
}

__global__ void __launch_bounds__(256) k_ghost_cells(km_pack_t u, UniformGrid gr0, int v0, int v1,
                                                     int nc, Instances in, int depth, int skip) {
  const Geo g(u);
  const long n = ghost_count(g, -1, depth);
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= n * u.nblocks) return;
  const int b = (int)(t / n);
  int idx[3];
  ghost_cell(g, t % n, idx, -1, depth);
  const UniformGrid gr = in.inst ? in.grids[in.inst[b]] : gr0;
  const int lb = b - gr.first;
  const int bl[3] = {lb % gr.nb[0], (lb / gr.nb[0]) % gr.nb[1], lb / (gr.nb[0] * gr.nb[1])};
  int src[3], sb[3];
  bool flip[3] = {false, false, false};
  bool axis = false;  // mirrored across an axisymmetric r = 0 boundary: phi flips too
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (d >= g.ndim) {
      sb[d] = 0;
      src[d] = idx[d];
      continue;
    }
    const int N = gr.nb[d] * g.nx[d];
    const int gi = bl[d] * g.nx[d] + (idx[d] - g.lo[d]);
    if ((gi < 0 && gr.bc[2 * d] == KM_BC_AXISYMMETRIC) ||
        (gi >= N && gr.bc[2 * d + 1] == KM_BC_AXISYMMETRIC))
      axis = !axis;
    const int m = map_cell(gi, N, gr.bc[2 * d], gr.bc[2 * d + 1], flip[d]);
    sb[d] = m / g.nx[d];
    src[d] = m % g.nx[d] + g.lo[d];
  }
  const int srcb = gr.first + (sb[2] * gr.nb[1] + sb[1]) * gr.nb[0] + sb[0];
  const long dc = g.c(idx[2], idx[1], idx[0]), scx = g.c(src[2], src[1], src[0]);
  const bool fl0 = flip[0], fl1 = flip[1], fl2 = flip[2] != axis;
  // variables [v0, v1) without the slot skip go in chunks whose loads are all issued
  // before their stores
  constexpr int kChunk = 4;
  const bool sk = skip >= v0 && skip < v1;
  const int nv = (v1 - v0) - (sk ? 1 : 0);
  auto var = [&](int k) { return v0 + k + (sk && v0 + k >= skip ? 1 : 0); };
  for (int k = 0; k < nv; k += kChunk) {
    double x[kChunk];
#pragma unroll
    for (int q = 0; q < kChunk; ++q) x[q] = k + q < nv ? cell_ptr(u, srcb, var(k + q))[scx] : 0.0;
#pragma unroll
    for (int q = 0; q < kChunk; ++q) {
      if (k + q >= nv) continue;
      const int w = var(k + q);
      int ax = -1;
      if (w >= 1 && w < 4) ax = w - 1;
      else if (w >= nc && w < nc + 3) ax = w - nc;
      else if (nc == 8 && w >= 5 && w < 8) ax = w - 5;
      const bool f = (ax == 0 && fl0) || (ax == 1 && fl1) || (ax == 2 && fl2);
      cell_ptr(u, b, w)[dc] = f ? -x[q] : x[q];
    }
  }
}

__global__ void __launch_bounds__(256) k_ghost_faces(km_pack_t u, UniformGrid gr0, int face,
                                                     Instances in, int depth) {
  const Geo g(u);
  const long n = ghost_count(g, face, depth);
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= n * u.nblocks) return;
  const int b = (int)(t / n);
  int idx[3];
  ghost_cell(g, t % n, idx, face, depth);
  const UniformGrid gr = in.inst ? in.grids[in.inst[b]] : gr0;
  const int lb = b - gr.first;
  const int bl[3] = {lb % gr.nb[0], (lb / gr.nb[0]) % gr.nb[1], lb / (gr.nb[0] * gr.nb[1])};
  int src[3], sb[3];
  bool flip[3] = {false, false, false};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (d >= g.ndim) {
      sb[d] = 0;
      src[d] = idx[d];
      continue;
    }
    const int N = gr.nb[d] * g.nx[d];
    const int gi = bl[d] * g.nx[d] + (idx[d] - g.lo[d]);
    if (d == face) {
      int m;
      if (d == 0 && gi < 0 && gr.bc[0] == KM_BC_AXISYMMETRIC) {
        // Axisymmetric on F1 faces (boundary_conditions.cpp:50-58): global face -g - 1,
        // negated
        m = -gi - 1;
        flip[d] = !flip[d];
      } else {
        m = map_face(gi, N, gr.bc[2 * d], gr.bc[2 * d + 1], flip[d]);
      }
      sb[d] = min(m / g.nx[d], gr.nb[d] - 1);
      src[d] = m - sb[d] * g.nx[d] + g.lo[d];
    } else {
      const int m = map_cell(gi, N, gr.bc[2 * d], gr.bc[2 * d + 1], flip[d]);
      sb[d] = m / g.nx[d];
      src[d] = m % g.nx[d] + g.lo[d];
    }
  }
  const int srcb = gr.first + (sb[2] * gr.nb[1] + sb[1]) * gr.nb[0] + sb[0];
  double val = u.face[face][srcb][g.s(src[2], src[1], src[0])];
  if ((face == 0 && flip[0]) || (face == 1 && flip[1]) || (face == 2 && flip[2])) val = -val;
  u.face[face][b][g.s(idx[2], idx[1], idx[0])] = val;
}

// ---------------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------------
static inline unsigned nblk(long n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

template <class T, int RECON, int RIEM, int MHD>
static void launch_flux_dirs(const km_pack_t &u, int lim, const Box &bx, int cyl,
                             hipStream_t st) {
  const Geo g(u);
  auto cnt = [&](int da, int db) {
    return (long)(bx.hi[da] - bx.lo[da] + 1) * (bx.hi[db] - bx.lo[db] + 1) * u.nblocks;
  };
  const long nface_x1 = (long)(bx.hi[0] - bx.lo[0] + 2) * cnt(1, 2);
  hipLaunchKernelGGL((k_flux_x1<T, RECON, RIEM, MHD>), dim3(nblk(nface_x1)), dim3(256), 0, st, u,
                     lim, bx, cyl);
  if (g.ndim > 1)
    hipLaunchKernelGGL((k_flux<T, RECON, RIEM, MHD, 1>), dim3(nblk(cnt(0, 2))), dim3(256), 0, st, u,
                       lim, bx, cyl);
  if (g.ndim > 2)
    hipLaunchKernelGGL((k_flux<T, RECON, RIEM, MHD, 2>), dim3(nblk(cnt(0, 1))), dim3(256), 0, st, u,
                       lim, bx, cyl);
}

template <int RECON, int RIEM>
static void launch_flux_mhd(const km_pack_t &u, int mhd, bool fp32, int lim, const Box &bx,
                            int cyl, hipStream_t st) {
  if (fp32) {
    if (mhd)
      launch_flux_dirs<float, RECON, RIEM, 1>(u, lim, bx, cyl, st);
    else
      launch_flux_dirs<float, RECON, RIEM, 0>(u, lim, bx, cyl, st);
  } else {
    if (mhd)
      launch_flux_dirs<double, RECON, RIEM, 1>(u, lim, bx, cyl, st);
    else
      launch_flux_dirs<double, RECON, RIEM, 0>(u, lim, bx, cyl, st);
  }
}

template <int RECON>
static void launch_flux_riem(const km_pack_t &u, int riem, int mhd, bool fp32, int lim,
                             const Box &bx, int cyl,
                             hipStream_t st) {
  if (riem == RIEMANN_HLLC)
    launch_flux_mhd<RECON, RIEMANN_HLLC>(u, mhd, fp32, lim, bx, cyl, st);
  else
    launch_flux_mhd<RECON, RIEMANN_HLL>(u, mhd, fp32, lim, bx, cyl, st);
}

template <int RECON>
static void launch_species_flux_recon(const km_pack_t &u, int lim, const Box &bx, int nsp,
                                      int slot0, int flux0, hipStream_t st) {
  const Geo g(u);
  auto cnt = [&](int da, int db) {
    return (long)(bx.hi[da] - bx.lo[da] + 1) * (bx.hi[db] - bx.lo[db] + 1) * u.nblocks;
  };
  hipLaunchKernelGGL((k_species_flux<RECON, 0>),
                     dim3(nblk((long)(bx.hi[0] - bx.lo[0] + 2) * cnt(1, 2))), dim3(256), 0, st, u,
                     lim, bx, nsp, slot0, flux0);
  if (g.ndim > 1)
    hipLaunchKernelGGL((k_species_flux<RECON, 1>),
                       dim3(nblk((long)(bx.hi[1] - bx.lo[1] + 2) * cnt(0, 2))), dim3(256), 0, st,
                       u, lim, bx, nsp, slot0, flux0);
  if (g.ndim > 2)
    hipLaunchKernelGGL((k_species_flux<RECON, 2>),
                       dim3(nblk((long)(bx.hi[2] - bx.lo[2] + 2) * cnt(0, 1))), dim3(256), 0, st,
                       u, lim, bx, nsp, slot0, flux0);
}

static void launch_species_fluxes(const km_cfg_t &cfg, const km_pack_t &u, const Box &bx,
                                  hipStream_t st) {
  const int nsp = cfg.nspecies, s0 = KM_MFRAC(cfg.mhd, 0), f0 = KM_NCONS(cfg.mhd);
  switch (cfg.recon) {
    case KM_RECON_FOG:
      launch_species_flux_recon<RECON_FOG>(u, cfg.limiter, bx, nsp, s0, f0, st);
      break;
    case KM_RECON_PLM:
      launch_species_flux_recon<RECON_PLM>(u, cfg.limiter, bx, nsp, s0, f0, st);
      break;
    case KM_RECON_PPM:
      launch_species_flux_recon<RECON_PPM>(u, cfg.limiter, bx, nsp, s0, f0, st);
      break;
    default: launch_species_flux_recon<RECON_WENOZ>(u, cfg.limiter, bx, nsp, s0, f0, st); break;
  }
}

void launch_calculate_fluxes(const km_cfg_t &cfg, const km_pack_t &u, const Box &bx,
                             hipStream_t st) {
  // the per-task path has two arithmetic forms: FP64 (also used for FP64_FMA, whose
  // contraction only exists in the fused sweeps) and FP32-mixed
  const bool fp32 = cfg.precision == KM_PREC_FP32_MIXED;
  const int r = cfg.riemann, m = cfg.mhd, l = cfg.limiter;
  const int cyl = cfg.geometry == KM_GEOM_CYLINDRICAL;
  switch (cfg.recon) {
    case KM_RECON_FOG: launch_flux_riem<RECON_FOG>(u, r, m, fp32, l, bx, cyl, st); break;
    case KM_RECON_PLM: launch_flux_riem<RECON_PLM>(u, r, m, fp32, l, bx, cyl, st); break;
    case KM_RECON_PPM: launch_flux_riem<RECON_PPM>(u, r, m, fp32, l, bx, cyl, st); break;
    default: launch_flux_riem<RECON_WENOZ>(u, r, m, fp32, l, bx, cyl, st); break;
  }
  if (cfg.nspecies > 1) launch_species_fluxes(cfg, u, bx, st);  // after the mass flux
}

void launch_calculate_emf(const km_cfg_t &cfg, const km_pack_t &u, hipStream_t st) {
  const Geo g(u);
  if (g.ndim < 2) return;  // no edges to evolve in 1-D (hydro_add_flux_tasks.cpp:413)
  // GetEdgeEMF's sign: -1 in cylindrical geometry (hydro_add_flux_tasks.cpp:398)
  const double esign = cfg.geometry == KM_GEOM_CYLINDRICAL ? -1.0 : 1.0;
  hipLaunchKernelGGL(k_emf, dim3(nblk(g.nstag() * u.nblocks)), dim3(256), 0, st, u, esign);
}

static long interior_count(const Geo &g, int nblocks) {
  return (long)g.nx[0] * (g.ndim > 1 ? g.nx[1] : 1) * (g.ndim > 2 ? g.nx[2] : 1) * nblocks;
}

static long face_count(const Geo &g, int face, int nblocks) {
  return (long)(g.nx[0] + (face == 0)) * ((g.ndim > 1 ? g.nx[1] : 1) + (face == 1)) *
         ((g.ndim > 2 ? g.nx[2] : 1) + (face == 2)) * nblocks;
}

void launch_fluxes_to_dudt(const km_cfg_t &cfg, const km_pack_t &u, const km_pack_t &dudt,
                           hipStream_t st, bool species) {
  species = species && cfg.nspecies > 1;
  const Geo g(u);
  const int nc = cfg.mhd ? 8 : 5;
  const int cyl = cfg.geometry == KM_GEOM_CYLINDRICAL;
  hipLaunchKernelGGL(k_flux_div, dim3(nblk(interior_count(g, u.nblocks))), dim3(256), 0, st, u,
                     dudt, nc, cfg.geometry == KM_GEOM_CYLINDRICAL ? 1 : 0);
  if (species)
    hipLaunchKernelGGL(k_species_dudt, dim3(nblk(interior_count(g, u.nblocks))), dim3(256), 0,
                       st, u, dudt, cfg.nspecies, KM_MFRAC(cfg.mhd, 0), nc, cyl);
  if (cfg.mhd && g.ndim >= 2) {
    hipLaunchKernelGGL(k_stokes<0>, dim3(nblk(face_count(g, 0, u.nblocks))), dim3(256), 0, st,
                       u, dudt, cyl);
    hipLaunchKernelGGL(k_stokes<1>, dim3(nblk(face_count(g, 1, u.nblocks))), dim3(256), 0, st,
                       u, dudt, cyl);
    if (g.ndim > 2)
      hipLaunchKernelGGL(k_stokes<2>, dim3(nblk(face_count(g, 2, u.nblocks))), dim3(256), 0, st,
                         u, dudt, cyl);
  }
}

void launch_add_source_terms(const km_cfg_t &cfg, const km_pack_t &u, const km_pack_t &dudt,
                             hipStream_t st) {
  if (cfg.geometry != KM_GEOM_CYLINDRICAL) return;  // Cartesian: no source (hydro.cpp:201-205)
  const Geo g(u);
  const unsigned grid = nblk(interior_count(g, u.nblocks));
  if (cfg.mhd)
    hipLaunchKernelGGL(k_geom_source<1>, dim3(grid), dim3(256), 0, st, u, dudt);
  else
    hipLaunchKernelGGL(k_geom_source<0>, dim3(grid), dim3(256), 0, st, u, dudt);
}

void launch_apply_dudt(const km_cfg_t &cfg, const km_pack_t &base, const km_pack_t &u0,
                       const km_pack_t &u1, const km_pack_t &dudt, double beta, double dt,
                       hipStream_t st, const double *dt_dev, const int *inst) {
  const Geo g(u0);
  const int nc = cfg.mhd ? 8 : 5;
  hipLaunchKernelGGL(k_apply_cell, dim3(nblk(interior_count(g, u0.nblocks))), dim3(256), 0, st,
                     base, u0, u1, dudt, 0, nc, beta, dt, dt_dev, inst);
  if (cfg.nspecies > 1)
    hipLaunchKernelGGL(k_apply_cell, dim3(nblk(interior_count(g, u0.nblocks))), dim3(256), 0,
                       st, base, u0, u1, dudt, KM_MFRAC(cfg.mhd, 0),
                       KM_MFRAC(cfg.mhd, cfg.nspecies), beta, dt, dt_dev, inst);
  if (cfg.mhd && g.ndim >= 2) {
    hipLaunchKernelGGL(k_apply_face<0>, dim3(nblk(face_count(g, 0, u0.nblocks))), dim3(256), 0,
                       st, base, u0, u1, dudt, beta, dt, dt_dev, inst);
    hipLaunchKernelGGL(k_apply_face<1>, dim3(nblk(face_count(g, 1, u0.nblocks))), dim3(256), 0,
                       st, base, u0, u1, dudt, beta, dt, dt_dev, inst);
    if (g.ndim > 2)
      hipLaunchKernelGGL(k_apply_face<2>, dim3(nblk(face_count(g, 2, u0.nblocks))), dim3(256),
                         0, st, base, u0, u1, dudt, beta, dt, dt_dev, inst);
  }
}

void launch_update(const km_cfg_t &cfg, const km_pack_t &base, const km_pack_t &u0,
                   const km_pack_t &u1, const km_pack_t &dudt, double beta, hipStream_t st,
                   const double *dt_dev, const int *inst) {
  const Geo g(u0);
  const Eos eos = make_eos(cfg);
  const unsigned grid = nblk(interior_count(g, u0.nblocks));
  const int cyl = cfg.geometry == KM_GEOM_CYLINDRICAL;
  if (cfg.nspecies > 1)
    hipLaunchKernelGGL(k_species_update, dim3(grid), dim3(256), 0, st, base, u0, u1,
                       cfg.nspecies, KM_MFRAC(cfg.mhd, 0), KM_NCONS(cfg.mhd), beta, 0.0,
                       dt_dev, inst, cyl);
  if (!cfg.mhd) {
    hipLaunchKernelGGL(k_update_cell<0>, dim3(grid), dim3(256), 0, st, base, u0, u1, beta, 0.0,
                       dt_dev, inst, eos, cyl);
    return;
  }
  if (cyl) {
    // the reference's task order, which the per-stage PrepareConserved on md0 splits:
    // FluxesToDuDt, AddSourceTerms (reads md0's B_phi), PrepareConserved (md0's B_phi to
    // B_phi / r), ApplyDuDt
    launch_fluxes_to_dudt(cfg, u0, dudt, st, false);
    launch_add_source_terms(cfg, u0, dudt, st);
    hipLaunchKernelGGL(k_cyl_stage_conserved, dim3(grid), dim3(256), 0, st, u0);
    hipLaunchKernelGGL(k_apply_cell, dim3(grid), dim3(256), 0, st, base, u0, u1, dudt, 0, 8,
                       beta, 0.0, dt_dev, inst);
  } else {
    hipLaunchKernelGGL(k_update_cell<1>, dim3(grid), dim3(256), 0, st, base, u0, u1, beta, 0.0,
                       dt_dev, inst, eos, cyl);
  }
  if (g.ndim >= 2 && !cyl) {
    hipLaunchKernelGGL(k_stokes<0>, dim3(nblk(face_count(g, 0, u0.nblocks))), dim3(256), 0, st,
                       u0, dudt, cyl);
    hipLaunchKernelGGL(k_stokes<1>, dim3(nblk(face_count(g, 1, u0.nblocks))), dim3(256), 0, st,
                       u0, dudt, cyl);
    if (g.ndim > 2)
      hipLaunchKernelGGL(k_stokes<2>, dim3(nblk(face_count(g, 2, u0.nblocks))), dim3(256), 0, st,
                         u0, dudt, cyl);
  }
  if (g.ndim >= 2) {
    hipLaunchKernelGGL(k_apply_face<0>, dim3(nblk(face_count(g, 0, u0.nblocks))), dim3(256), 0,
                       st, base, u0, u1, dudt, beta, 0.0, dt_dev, inst);
    hipLaunchKernelGGL(k_apply_face<1>, dim3(nblk(face_count(g, 1, u0.nblocks))), dim3(256), 0,
                       st, base, u0, u1, dudt, beta, 0.0, dt_dev, inst);
    if (g.ndim > 2)
      hipLaunchKernelGGL(k_apply_face<2>, dim3(nblk(face_count(g, 2, u0.nblocks))), dim3(256),
                         0, st, base, u0, u1, dudt, beta, 0.0, dt_dev, inst);
  }
  hipLaunchKernelGGL(k_prepare_primitive<1>, dim3(grid), dim3(256), 0, st, u1, eos, cyl);
}

// EMF (k_emf), FluxStokes (k_stokes) and the face part of ApplyDuDt (k_apply_face) of one
// face in one lane, Cartesian 3-D CT. The face's four edge EMFs are averaged from u0's
// stored transverse B fluxes with k_emf's expression, Stokes and the stage combine follow
// with k_stokes' and k_apply_face's, so the updated field is bitwise the three kernels'.
// u0's face field is not overwritten by the combine (u0 is not read again in the cycle).
template <int FACE>
__device__ __forceinline__ void ct_face(const km_pack_t &base, const km_pack_t &u0,
                                        const km_pack_t &u1, const Geo &g, int b, int k, int j,
                                        int i, double beta, double dt) {
  const km_block_t &blk = u0.blocks[b];
  const long s = g.s(k, j, i), c = g.c(k, j, i);
  const long sj = g.sc[1], sk = g.sc[2];
  constexpr double esign = 1.0;
  // edge EMF of family e at the edge with the lower corner of cell offset cc
  auto emf = [&](int e, long cc) -> double {
    if (e == 2) {
      const double *fyBx = flux_ptr(u0, 1, b, C_MAGC + 0), *fxBy = flux_ptr(u0, 0, b, C_MAGC + 1);
      return esign * 0.25 * (fyBx[cc - 1] + fyBx[cc] - fxBy[cc - sj] - fxBy[cc]);
    }
    if (e == 0) {
      const double *fzBy = flux_ptr(u0, 2, b, C_MAGC + 1), *fyBz = flux_ptr(u0, 1, b, C_MAGC + 2);
      return esign * 0.25 * (fzBy[cc - sj] + fzBy[cc] - fyBz[cc - sk] - fyBz[cc]);
    }
    const double *fxBz = flux_ptr(u0, 0, b, C_MAGC + 2), *fzBx = flux_ptr(u0, 2, b, C_MAGC + 0);
    return esign * 0.25 * (fxBz[cc - sk] + fxBz[cc] - fzBx[cc - 1] - fzBx[cc]);
  };
  constexpr int e0 = FACE == 0 ? 2 : FACE == 1 ? 2 : 0;
  constexpr int a0 = FACE == 0 ? 1 : FACE == 1 ? 0 : 1;
  constexpr double s0 = FACE == 0 ? 1.0 : -1.0;
  constexpr int e1 = FACE == 0 ? 1 : FACE == 1 ? 0 : 1;
  constexpr int a1 = FACE == 0 ? 2 : FACE == 1 ? 2 : 0;
  constexpr double s1 = FACE == 0 ? -1.0 : 1.0;
  constexpr double gs = 1.0;
  double acc = 0.;
  acc += (gs * s0) * (blk.dx[e0] * emf(e0, c) - blk.dx[e0] * emf(e0, c + g.sc[a0]));
  acc += (gs * s1) * (blk.dx[e1] * emf(e1, c) - blk.dx[e1] * emf(e1, c + g.sc[a1]));
  acc *= 1. / (blk.area[FACE] + 1.e-36);
  const double v0 = beta * base.face[FACE][b][s] + (1.0 - beta) * u0.face[FACE][b][s];
  u1.face[FACE][b][s] = v0 + beta * dt * acc;
}

__global__ void __launch_bounds__(256) k_ct_faces(km_pack_t base, km_pack_t u0, km_pack_t u1,
                                                  double beta, const double *dtp,
                                                  const int *inst) {
  // one lane per (block, k, j, i) of the block's (nx+1)^3 face box: the lane updates the
  // x1, x2 and x3 faces at that index that exist, so the three families' EMFs read the
  // same flux neighbourhood close together in time (L2 reuse across families)
  const Geo g(u0);
  const int n0 = g.nx[0] + 1, n1 = g.nx[1] + 1, n2 = g.nx[2] + 1;
  const long per = (long)n0 * n1 * n2;
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= per * u0.nblocks) return;
  const int b = (int)(t / per);
  long r = t % per;
  const int ii = (int)(r % n0);
  r /= n0;
  const int jj = (int)(r % n1);
  const int kk = (int)(r / n1);
  const int i = g.lo[0] + ii, j = g.lo[1] + jj, k = g.lo[2] + kk;
  const double dt = dtp[inst ? inst[b] : 0];
  const bool in_i = ii < g.nx[0], in_j = jj < g.nx[1], in_k = kk < g.nx[2];
  if (in_j && in_k) ct_face<0>(base, u0, u1, g, b, k, j, i, beta, dt);
  if (in_i && in_k) ct_face<1>(base, u0, u1, g, b, k, j, i, beta, dt);
  if (in_i && in_j) ct_face<2>(base, u0, u1, g, b, k, j, i, beta, dt);
}

void launch_ct_faces(const km_pack_t &base, const km_pack_t &u0, const km_pack_t &u1,
                     double beta, hipStream_t st, const double *dt_dev, const int *inst) {
  const Geo g(u0);
  const long n = (long)(g.nx[0] + 1) * (g.nx[1] + 1) * (g.nx[2] + 1) * u0.nblocks;
  hipLaunchKernelGGL(k_ct_faces, dim3(nblk(n)), dim3(256), 0, st, base, u0, u1, beta, dt_dev,
                     inst);
}

void launch_prepare_primitive(const km_cfg_t &cfg, const km_pack_t &u, hipStream_t st) {
  const Geo g(u);
  const Eos eos = make_eos(cfg);
  const unsigned grid = nblk(interior_count(g, u.nblocks));
  const int cyl = cfg.geometry == KM_GEOM_CYLINDRICAL;
  if (cfg.mhd)
    hipLaunchKernelGGL(k_prepare_primitive<1>, dim3(grid), dim3(256), 0, st, u, eos, cyl);
  else
    hipLaunchKernelGGL(k_prepare_primitive<0>, dim3(grid), dim3(256), 0, st, u, eos, cyl);
  if (cfg.nspecies > 1)
    hipLaunchKernelGGL(k_species_normalize, dim3(grid), dim3(256), 0, st, u, cfg.nspecies,
                       KM_MFRAC(cfg.mhd, 0));
}

void launch_prepare_conserved(const km_cfg_t &cfg, const km_pack_t &u, hipStream_t st) {
  const Geo g(u);
  const Eos eos = make_eos(cfg);
  const unsigned grid = nblk(interior_count(g, u.nblocks));
  const int cyl = cfg.geometry == KM_GEOM_CYLINDRICAL;
  if (cfg.mhd)
    hipLaunchKernelGGL(k_prepare_conserved<1>, dim3(grid), dim3(256), 0, st, u, eos, cyl,
                       cfg.mode_init);
  else
    hipLaunchKernelGGL(k_prepare_conserved<0>, dim3(grid), dim3(256), 0, st, u, eos, cyl,
                       cfg.mode_init);
  // material::PostMeshInitialization: the mass fractions sum to one
  if (cfg.nspecies > 1)
    hipLaunchKernelGGL(k_species_normalize, dim3(grid), dim3(256), 0, st, u, cfg.nspecies,
                       KM_MFRAC(cfg.mhd, 0));
}

void launch_estimate_dt(const km_cfg_t &cfg, const km_pack_t &u, double *dt_dev,
                        hipStream_t st, const int *inst, int ninst, int *fail, int part) {
  const Geo g(u);
  if (part & 1)
    hipLaunchKernelGGL(k_fill_value, dim3(nblk(ninst, 64)), dim3(64), 0, st, dt_dev, DBL_MAX,
                       ninst);
  const unsigned grid = nblk(interior_count(g, u.nblocks));
  const int cyl = cfg.geometry == KM_GEOM_CYLINDRICAL;
  if (cfg.mhd)
    hipLaunchKernelGGL(k_estimate_dt<1>, dim3(grid), dim3(256), 0, st, u, dt_dev, inst, cyl, fail);
  else
    hipLaunchKernelGGL(k_estimate_dt<0>, dim3(grid), dim3(256), 0, st, u, dt_dev, inst, cyl, fail);
  if (part & 2)
    hipLaunchKernelGGL(k_scale_value, dim3(nblk(ninst, 64)), dim3(64), 0, st, dt_dev,
                       cfg.cfl / static_cast<double>(g.ndim), ninst);
}

void launch_fill_divb(const km_cfg_t &cfg, const km_pack_t &u, hipStream_t st) {
  const Geo g(u);
  hipLaunchKernelGGL(k_fill_divb, dim3(nblk(interior_count(g, u.nblocks))), dim3(256), 0, st, u,
                     KM_DIVB(cfg.mhd), cfg.geometry == KM_GEOM_CYLINDRICAL ? 1 : 0);
}

void launch_prepare_conserved_stage(const km_cfg_t &cfg, const km_pack_t &u, hipStream_t st) {
  const Geo g(u);
  const unsigned grid = nblk(interior_count(g, u.nblocks));
  if (cfg.geometry == KM_GEOM_CYLINDRICAL && cfg.mhd != KM_MHD_OFF)
    hipLaunchKernelGGL(k_cyl_stage_conserved, dim3(grid), dim3(256), 0, st, u);
  if (cfg.nspecies > 1)
    hipLaunchKernelGGL(k_species_to_conserved, dim3(grid), dim3(256), 0, st, u, cfg.nspecies,
                       KM_MFRAC(cfg.mhd, 0));
}

void launch_ghost_fill(const UniformGrid &gr, const km_pack_t &u, int v0, int v1, int nc,
                       bool faces, hipStream_t st, Instances in, int depth, int face_depth,
                       int skip) {
  const Geo g(u);
  if (v1 > v0)
    hipLaunchKernelGGL(k_ghost_cells, dim3(nblk(ghost_count(g, -1, depth) * u.nblocks)),
                       dim3(256), 0, st, u, gr, v0, v1, nc, in, depth, skip);
  if (faces)
    for (int f = 0; f < (g.ndim > 2 ? 3 : g.ndim); ++f)
      hipLaunchKernelGGL(k_ghost_faces, dim3(nblk(ghost_count(g, f, face_depth) * u.nblocks)),
                         dim3(256), 0, st, u, gr, f, in, face_depth);
}

}  // namespace km
