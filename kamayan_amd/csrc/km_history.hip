// km_history.hip -- history outputs (Parthenon's "hst" file type) of the isentropic vortex:
// the problem registers one history variable per primitive, each the volume-weighted L1
// distance of the current state from the vortex's initial state divided by the x1-x2
// extent of the domain (isentropic_vortex.cpp:74-101, ErrorHistory isentropic_vortex.hpp:
// 150-200). The reduction runs on the device: one workgroup per mesh block sums its
// interior in a fixed order, and the host adds the block sums in pack order.
#include <hip/hip_runtime.h>

#include <cmath>

#include "km_device.hpp"
#include "km_internal.hpp"

namespace km {
namespace {

constexpr double kPiH = 3.141592653589793;
constexpr int kHistThreads = 256;

// VortexData::State / StateMHD (isentropic_vortex.hpp:31-135) at (x, y), in the order
// DENS, VEL0..2, PRES, MAGC0..2. Cartesian: r^2 = x^2 + y^2 and the swirl is added to the
// mean flow (velx, vely); cylindrical: r^2 = x^2 and the swirl is azimuthal, v_phi = x dv.
__device__ void vortex_state(const HistVortex &p, bool cyl, bool mhd, double x, double y,
                             double out[8]) {
  const double r2 = cyl ? x * x : x * x + y * y;
  const double dv = p.strength * exp(0.5 * (1.0 - r2)) / (2.0 * kPiH);
  for (int v = 0; v < 8; ++v) out[v] = 0.0;
  if (!mhd) {
    const double T = p.pressure / p.density - (p.gamma - 1.0) * p.strength * p.strength *
                                                  exp(1.0 - r2) /
                                                  (8.0 * p.gamma * kPiH * kPiH);
    out[0] = pow(T, 1.0 / (p.gamma - 1.0));
    out[4] = T * out[0];
  } else {
    const double ex = exp(1.0 - r2);
    out[0] = 1.0;
    out[4] = p.pressure +
             (p.mhd_strength * p.mhd_strength / (8.0 * kPiH * kPiH)) * (1.0 - r2) * ex -
             (p.strength * p.strength / (8.0 * kPiH * kPiH)) * ex;
    const double dB = p.mhd_strength * exp(0.5 * (1.0 - r2)) / (2.0 * kPiH);
    if (cyl) {
      out[7] = x * dB;
    } else {
      out[5] = -y * dB;
      out[6] = +x * dB;
    }
  }
  if (cyl) {
    out[3] = x * dv;
  } else {
    out[1] = p.velx - y * dv;
    out[2] = p.vely + x * dv;
  }
}

// part[b * nv + v] = sum over block b's interior of |exact_v - q_v| V, with the exact state
// at the cell centroid (the radial centroid in cylindrical geometry)
__global__ void __launch_bounds__(kHistThreads) k_vortex_error(km_pack_t u, HistVortex p,
                                                               int cyl, int mhd, double *part) {
  const Geo g(u);
  const int b = blockIdx.x;
  const km_block_t &blk = u.blocks[b];
  const int nc = mhd ? 8 : 5;
  const int nv = mhd ? 8 : 5;
  const int slot[8] = {KM_DENS, nc + 0, nc + 1, nc + 2, nc + 3, KM_MAGC0, KM_MAGC0 + 1,
                       KM_MAGC0 + 2};
  const int n0 = g.nx[0], n1 = g.ndim > 1 ? g.nx[1] : 1, n2 = g.ndim > 2 ? g.nx[2] : 1;
  const long n = (long)n0 * n1 * n2;
  double acc[8] = {0., 0., 0., 0., 0., 0., 0., 0.};
  for (long t = threadIdx.x; t < n; t += kHistThreads) {
    const int i = g.lo[0] + (int)(t % n0);
    const int j = g.lo[1] + (int)((t / n0) % n1);
    const int k = g.lo[2] + (int)(t / ((long)n0 * n1));
    const double x0 = cyl ? cyl_centroid_r(blk, g, i) : x_centre(blk, g, 0, i);
    const double y0 = x_centre(blk, g, 1, j);
    double ex[8];
    vortex_state(p, cyl != 0, mhd != 0, x0, y0, ex);
    const double vol = cell_volume(blk, g, cyl != 0, i);
    const long c = g.c(k, j, i);
#pragma unroll
    for (int v = 0; v < 8; ++v)
      if (v < nv) acc[v] += fabs(ex[v] - cell_ptr(u, b, slot[v])[c]) * vol;
  }
  __shared__ double red[8][kHistThreads];
#pragma unroll
  for (int v = 0; v < 8; ++v) red[v][threadIdx.x] = acc[v];
  __syncthreads();
  for (int s = kHistThreads / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
#pragma unroll
      for (int v = 0; v < 8; ++v) red[v][threadIdx.x] += red[v][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int v = 0; v < nv; ++v) part[(long)b * nv + v] = red[v][0];
}

}  // namespace

void launch_history_vortex(const km_cfg_t &cfg, const km_pack_t &u, const HistVortex &p,
                           double *part, hipStream_t st) {
  if (u.nblocks < 1) return;
  hipLaunchKernelGGL(k_vortex_error, dim3(u.nblocks), dim3(kHistThreads), 0, st, u, p,
                     cfg.geometry == KM_GEOM_CYLINDRICAL ? 1 : 0, cfg.mhd != KM_MHD_OFF ? 1 : 0,
                     part);
}

}  // namespace km
