// km_problems.cpp -- block metric of a uniform mesh and the problem generators of the
// benchmark configurations (host-side; they fill primitives on block interiors the way
// a kamayan ProblemGenerator does before PostMeshInitialization).
#include <array>
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/km_hydro.h"
#include "km_internal.hpp"

namespace {

constexpr double kPi = 3.14159265358979323846;

// deterministic uniform doubles in [0, 1) from a 64-bit generator
struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() {  // splitmix64
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return static_cast<double>(next() >> 11) * (1.0 / 9007199254740992.0); }
};

// per-zone noise that does not depend on the block decomposition
double zone_noise(uint64_t seed, uint64_t salt, long gi, long gj, long gk) {
  Rng r(seed ^ (salt * 0xD1B54A32D192ED03ull) ^ ((uint64_t)gi * 0x9E3779B97F4A7C15ull) ^
        ((uint64_t)gj * 0xC2B2AE3D27D4EB4Full) ^ ((uint64_t)gk * 0x165667B19E3779F9ull));
  r.next();
  return 2.0 * r.uniform() - 1.0;
}

struct Layout {
  int ndim, ng, nx[3], E[3], S[3], nvar, nblocks;
  explicit Layout(const km_mesh_desc_t &m, int nv) {
    ndim = m.ndim;
    ng = m.ng;
    for (int d = 0; d < 3; ++d) nx[d] = d < ndim ? m.nx[d] : 1;
    E[0] = nx[0] + 2 * ng;
    E[1] = ndim > 1 ? nx[1] + 2 * ng : 1;
    E[2] = ndim > 2 ? nx[2] + 2 * ng : 1;
    S[0] = E[0] + 1;
    S[1] = E[1] + (ndim > 1);
    S[2] = E[2] + (ndim > 2);
    nvar = nv;
    nblocks = m.nb[0] * m.nb[1] * m.nb[2];
  }
  long ncell() const { return (long)E[0] * E[1] * E[2]; }
  long nstag() const { return (long)S[0] * S[1] * S[2]; }
  long c(int k, int j, int i) const { return ((long)k * E[1] + j) * E[0] + i; }
  long s(int k, int j, int i) const { return ((long)k * S[1] + j) * S[0] + i; }
  int lo(int d) const { return d < ndim ? ng : 0; }
  int hi(int d) const { return d < ndim ? ng + nx[d] - 1 : 0; }
};

// Parthenon UniformCartesian::Xc / Xf: xmin_ = block xmin - istart*dx, then
// Xc(idx) = xmin_ + (idx + 0.5) dx, Xf(idx) = xmin_ + idx dx
double xc(const km_block_t &b, int d, int idx, int lo) {
  const double x0 = b.xmin[d] - lo * b.dx[d];
  return x0 + (idx + 0.5) * b.dx[d];
}
double xf(const km_block_t &b, int d, int idx, int lo) {
  const double x0 = b.xmin[d] - lo * b.dx[d];
  return x0 + idx * b.dx[d];
}
// radial centroid of cylindrical cell idx (geometry.hpp:55-63)
double xc_cyl(const km_block_t &b, int idx, int lo) {
  const double r0 = xf(b, 0, idx, lo), r1 = xf(b, 0, idx + 1, lo);
  const double r0sq = r0 * r0, r1sq = r1 * r1;
  return (2.0 / 3.0) * (r1sq * std::fabs(r1) - r0sq * std::fabs(r0)) / (r1sq - r0sq);
}

struct Mode {
  int k[3];
  double phase;
};

std::vector<Mode> draw_modes(Rng &rng, int n, int kmax, int ndim) {
  // at most as many modes as there are distinct +-k pairs with 0 < |k| <= kmax (four in
  // one dimension for kmax 4)
  int avail = 0;
  for (int a = -kmax; a <= kmax; ++a)
    for (int b = ndim > 1 ? -kmax : 0; b <= (ndim > 1 ? kmax : 0); ++b)
      for (int c = ndim > 2 ? -kmax : 0; c <= (ndim > 2 ? kmax : 0); ++c) {
        const int k2 = a * a + b * b + c * c;
        const int first = a != 0 ? a : b != 0 ? b : c;
        if (k2 > 0 && k2 <= kmax * kmax && first > 0) ++avail;
      }
  n = std::min(n, avail);
  std::vector<Mode> modes;
  while ((int)modes.size() < n) {
    Mode m;
    int k2 = 0;
    for (int d = 0; d < 3; ++d) {
      m.k[d] = d < ndim ? (int)(rng.uniform() * (2 * kmax + 1)) - kmax : 0;
      k2 += m.k[d] * m.k[d];
    }
    m.phase = 2.0 * kPi * rng.uniform();
    if (k2 == 0 || k2 > kmax * kmax) continue;
    // one representative per +-k pair: first nonzero component positive
    int first = 0;
    for (int d = 0; d < 3; ++d)
      if (m.k[d] != 0) {
        first = m.k[d];
        break;
      }
    if (first < 0) continue;
    bool dup = false;
    for (const Mode &o : modes)
      if (o.k[0] == m.k[0] && o.k[1] == m.k[1] && o.k[2] == m.k[2]) dup = true;
    if (!dup) modes.push_back(m);
  }
  return modes;
}

double mode_sum(const std::vector<Mode> &modes, double amp, double x, double y, double z) {
  double s = 0.;
  for (const Mode &m : modes)
    s += amp * std::cos(2.0 * kPi * (m.k[0] * x + m.k[1] * y + m.k[2] * z) + m.phase);
  return s;
}

}  // namespace

namespace km {
void make_uniform_blocks(const km_mesh_desc_t &m, km_block_t *blocks) {
  const int nbt = m.nb[0] * m.nb[1] * m.nb[2];
  for (int b = 0; b < nbt; ++b) {
    const int l[3] = {b % m.nb[0], (b / m.nb[0]) % m.nb[1], b / (m.nb[0] * m.nb[1])};
    km_block_t &blk = blocks[b];
    std::memset(&blk, 0, sizeof(blk));
    for (int d = 0; d < 3; ++d) {
      double lo, hi;
      if (d < m.ndim) {
        // uniform mesh generator: x(r) = 0.5 (xmin + xmax) + (r xmax - r xmin),
        // r = (2 l - nb) / (2 nb); the domain edges are exact
        const double rl = static_cast<double>(2 * l[d] - m.nb[d]) / (2 * m.nb[d]);
        const double rr = static_cast<double>(2 * (l[d] + 1) - m.nb[d]) / (2 * m.nb[d]);
        lo = 0.5 * (m.xmin[d] + m.xmax[d]) + (rl * m.xmax[d] - rl * m.xmin[d]);
        hi = 0.5 * (m.xmin[d] + m.xmax[d]) + (rr * m.xmax[d] - rr * m.xmin[d]);
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
}  // namespace km

extern "C" {

int km_mesh_blocks(const km_mesh_desc_t *mesh, km_block_t *blocks_host) {
  if (!mesh || !blocks_host) return KM_EINVAL;
  km::make_uniform_blocks(*mesh, blocks_host);
  return KM_OK;
}

int km_problem_fill(const char *name, const km_cfg_t *cfg, const km_mesh_desc_t *mesh,
                    uint64_t seed, const double *params, int32_t nparams, double *cell,
                    double *const face[3]) {
  if (!name || !cfg || !mesh || !cell) return KM_EINVAL;
  std::vector<km_block_t> blocks((size_t)mesh->nb[0] * mesh->nb[1] * mesh->nb[2]);
  km::make_uniform_blocks(*mesh, blocks.data());
  return km_problem_fill_blocks(name, cfg, mesh, blocks.data(), (int32_t)blocks.size(), seed,
                                params, nparams, cell, face);
}

int km_problem_fill_blocks(const char *name, const km_cfg_t *cfg, const km_mesh_desc_t *mesh,
                           const km_block_t *blocks, int32_t nblocks, uint64_t seed,
                           const double *params, int32_t nparams, double *cell,
                           double *const face[3]) {
  if (!name || !cfg || !mesh || !blocks || !cell || nblocks < 0) return KM_EINVAL;
  const std::string pname(name);
  const int mhd = cfg->mhd;
  const int nsp = cfg->nspecies;
  Layout L(*mesh, KM_NVAR(mhd) + KM_NSPEC_SLOTS(nsp));
  L.nblocks = nblocks;
  const int nc = KM_NCONS(mhd);
  const int sV = nc, sP = nc + 3;
  auto P = [&](int n, double dflt) { return (params && n < nparams) ? params[n] : dflt; };
  std::memset(cell, 0, sizeof(double) * L.ncell() * L.nvar * L.nblocks);
  if (mhd == KM_MHD_CT && face)
    for (int d = 0; d < 3; ++d)
      if (face[d]) std::memset(face[d], 0, sizeof(double) * L.nstag() * L.nblocks);
  auto at = [&](int b, int v) { return cell + ((long)b * L.nvar + v) * L.ncell(); };
  // with more than one species the material is species 0 throughout, unless the problem
  // sets the mass fractions itself
  if (nsp > 1)
    for (int b = 0; b < L.nblocks; ++b)
      for (int k = L.lo(2); k <= L.hi(2); ++k)
        for (int j = L.lo(1); j <= L.hi(1); ++j)
          for (int i = L.lo(0); i <= L.hi(0); ++i) at(b, KM_MFRAC(mhd, 0))[L.c(k, j, i)] = 1.0;

  if (pname == "isentropic_vortex") {
    // VortexData::State (isentropic_vortex.hpp:31-51) at cell centroids; in cylindrical
    // geometry State<cylindrical> (:53-72): r^2 = x1^2 at the radial centroid and the
    // vortex velocity is azimuthal, v_phi = r dv
    const bool cyl = cfg->geometry == KM_GEOM_CYLINDRICAL;
    const double density = P(0, 1.0), pressure = P(1, 1.0), velx = P(2, 1.0),
                 vely = P(3, 1.0), strength = P(4, 5.0), gamma = cfg->gamma;
    const double mstr = P(5, 1.0);  // mhd_strength (isentropic_vortex.cpp:49)
    if (mhd != KM_MHD_OFF) {
      // StateMHD (isentropic_vortex.hpp:74-124): unit density, magnetic tension and
      // centrifugal force balanced by pressure; B = (-y, x) dB (Cartesian) or B_phi = r dB
      const double k8 = 8.0 * kPi * kPi;
      for (int b = 0; b < L.nblocks; ++b)
        for (int k = L.lo(2); k <= L.hi(2); ++k)
          for (int j = L.lo(1); j <= L.hi(1); ++j)
            for (int i = L.lo(0); i <= L.hi(0); ++i) {
              const double x = cyl ? xc_cyl(blocks[b], i, L.lo(0)) : xc(blocks[b], 0, i, L.lo(0));
              const double y = xc(blocks[b], 1, j, L.lo(1));
              const double r2 = cyl ? x * x : x * x + y * y;
              const double dv = strength * std::exp(0.5 * (1.0 - r2)) / (2.0 * kPi);
              const double dB = mstr * std::exp(0.5 * (1.0 - r2)) / (2.0 * kPi);
              const double ex = std::exp(1.0 - r2);
              const long c = L.c(k, j, i);
              at(b, KM_DENS)[c] = 1.0;
              at(b, sP)[c] = pressure + (mstr * mstr / k8) * (1.0 - r2) * ex -
                             (strength * strength / k8) * ex;
              if (cyl) {
                at(b, sV + 2)[c] = x * dv;
                at(b, KM_MAGC0 + 2)[c] = x * dB;
              } else {
                at(b, sV + 0)[c] = velx - y * dv;
                at(b, sV + 1)[c] = vely + x * dv;
                at(b, KM_MAGC0 + 0)[c] = -y * dB;
                at(b, KM_MAGC0 + 1)[c] = +x * dB;
              }
            }
      if (mhd == KM_MHD_CT && L.ndim > 1 && !cyl && face && face[0] && face[1]) {
        // div-free faces from Az = mhd_strength / (2 pi) exp((1 - r^2) / 2)
        // (isentropic_vortex.cpp:154-174)
        auto Az = [&](double xx, double yy) {
          return mstr / (2. * kPi) * std::exp(0.5 * (1.0 - (xx * xx + yy * yy)));
        };
        const int k3 = L.ndim > 2 ? 1 : 0;
        for (int b = 0; b < L.nblocks; ++b) {
          const km_block_t &bk = blocks[b];
          double *B1 = face[0] + (long)b * L.nstag(), *B2 = face[1] + (long)b * L.nstag();
          for (int k = L.lo(2); k <= L.hi(2) + k3; ++k)
            for (int j = L.lo(1); j <= L.hi(1) + 1; ++j)
              for (int i = L.lo(0); i <= L.hi(0) + 1; ++i) {
                const double xfx = xf(bk, 0, i, L.lo(0)), xfy = xc(bk, 1, j, L.lo(1));
                const double dy = bk.dx[1];
                B1[L.s(k, j, i)] = 1. / dy * (Az(xfx, xfy + 0.5 * dy) - Az(xfx, xfy - 0.5 * dy));
                const double yfx = xc(bk, 0, i, L.lo(0)), yfy = xf(bk, 1, j, L.lo(1));
                const double dx = bk.dx[0];
                B2[L.s(k, j, i)] =
                    -1. / dx * (Az(yfx + 0.5 * dx, yfy) - Az(yfx - 0.5 * dx, yfy));
              }
        }
      }
      return KM_OK;
    }
    for (int b = 0; b < L.nblocks; ++b)
      for (int k = L.lo(2); k <= L.hi(2); ++k)
        for (int j = L.lo(1); j <= L.hi(1); ++j)
          for (int i = L.lo(0); i <= L.hi(0); ++i) {
            const double x = cyl ? xc_cyl(blocks[b], i, L.lo(0)) : xc(blocks[b], 0, i, L.lo(0));
            const double y = xc(blocks[b], 1, j, L.lo(1));
            const double r2 = cyl ? x * x : x * x + y * y;
            const double dv = strength * std::exp(0.5 * (1.0 - r2)) / (2.0 * kPi);
            const double T = pressure / density - (gamma - 1.0) * strength * strength *
                                                      std::exp(1.0 - r2) /
                                                      (8.0 * gamma * kPi * kPi);
            const double rho = std::pow(T, 1.0 / (gamma - 1.0));
            const long c = L.c(k, j, i);
            at(b, KM_DENS)[c] = rho;
            at(b, sP)[c] = T * rho;
            if (cyl) {
              at(b, sV + 0)[c] = 0.0;
              at(b, sV + 1)[c] = 0.0;
              at(b, sV + 2)[c] = x * dv;
            } else {
              at(b, sV + 0)[c] = velx - y * dv;
              at(b, sV + 1)[c] = vely + x * dv;
              at(b, sV + 2)[c] = 0.0;
            }
          }
    return KM_OK;
  }

  if (pname == "smooth3d") {
    // density = 1 + 0.1 * sum of 16 modes (|k| <= 4, amplitude 1/4, random phases);
    // each velocity component the same construction at amplitude 0.1; pressure = 1
    Rng rng(seed);
    const double amp = P(0, 0.25);
    const auto mr = draw_modes(rng, 16, 4, L.ndim);
    std::vector<Mode> mv[3];
    for (int d = 0; d < 3; ++d) mv[d] = draw_modes(rng, 16, 4, L.ndim);
    for (int b = 0; b < L.nblocks; ++b)
      for (int k = L.lo(2); k <= L.hi(2); ++k)
        for (int j = L.lo(1); j <= L.hi(1); ++j)
          for (int i = L.lo(0); i <= L.hi(0); ++i) {
            const double x = xc(blocks[b], 0, i, L.lo(0));
            const double y = L.ndim > 1 ? xc(blocks[b], 1, j, L.lo(1)) : 0.;
            const double z = L.ndim > 2 ? xc(blocks[b], 2, k, L.lo(2)) : 0.;
            const long c = L.c(k, j, i);
            at(b, KM_DENS)[c] = 1.0 + 0.1 * mode_sum(mr, amp, x, y, z);
            for (int d = 0; d < 3; ++d) at(b, sV + d)[c] = 0.1 * mode_sum(mv[d], amp, x, y, z);
            at(b, sP)[c] = 1.0;
          }
    return KM_OK;
  }

  if (pname == "kelvin_helmholtz") {
    // shear layer: rho = 2, vx = +0.5 for |y - 1/2| < 1/4, else rho = 1, vx = -0.5;
    // vy = A sin(2 pi n x); P = 2.5
    const double A = P(0, 0.01), n = P(1, 2.0), pres = P(2, 2.5);
    for (int b = 0; b < L.nblocks; ++b)
      for (int k = L.lo(2); k <= L.hi(2); ++k)
        for (int j = L.lo(1); j <= L.hi(1); ++j)
          for (int i = L.lo(0); i <= L.hi(0); ++i) {
            const double x = xc(blocks[b], 0, i, L.lo(0));
            const double y = xc(blocks[b], 1, j, L.lo(1));
            const bool inner = std::fabs(y - 0.5) < 0.25;
            const long c = L.c(k, j, i);
            at(b, KM_DENS)[c] = inner ? 2.0 : 1.0;
            at(b, sV + 0)[c] = inner ? 0.5 : -0.5;
            at(b, sV + 1)[c] = A * std::sin(2.0 * kPi * n * x);
            at(b, sV + 2)[c] = 0.0;
            at(b, sP)[c] = pres;
          }
    return KM_OK;
  }

  if (pname == "sedov") {
    // sedov.cpp:75-143 in 3-D: P = 3 (g-1) E / (4 pi r^3) inside r <= radius
    const double E = P(0, 1.0), radius = P(1, 0.1), rho0 = P(2, 1.0), p0 = P(3, 1.0e-5);
    const double nu = static_cast<double>(L.ndim);
    const double pexp = 3. * (cfg->gamma - 1) * E / ((nu + 1) * kPi * std::pow(radius, nu));
    for (int b = 0; b < L.nblocks; ++b)
      for (int k = L.lo(2); k <= L.hi(2); ++k)
        for (int j = L.lo(1); j <= L.hi(1); ++j)
          for (int i = L.lo(0); i <= L.hi(0); ++i) {
            const double x = xc(blocks[b], 0, i, L.lo(0));
            const double y = L.ndim > 1 ? xc(blocks[b], 1, j, L.lo(1)) : 0.;
            const double z = L.ndim > 2 ? xc(blocks[b], 2, k, L.lo(2)) : 0.;
            const double r = std::sqrt(x * x + y * y + z * z);
            const long c = L.c(k, j, i);
            at(b, KM_DENS)[c] = rho0;
            at(b, sP)[c] = r <= radius ? pexp : p0;
            if (nsp > 1) {  // sedov.cpp:138-141: species 0 inside the blast, 1 outside
              at(b, KM_MFRAC(mhd, 0))[c] = r <= radius ? 1.0 : 0.0;
              at(b, KM_MFRAC(mhd, 1))[c] = r > radius ? 1.0 : 0.0;
            }
          }
    return KM_OK;
  }

  if (pname == "mhd_blast") {
    // mhd_blast.cpp:47-160, Cartesian: rho = 1, v = 0, P = p_explosion inside
    // r = sqrt(x^2 + y^2) <= radius else p_ambient, cell B = (bx, 0, 0), face B from
    // Az = bx y: B1 = (Az(y_{j+1}) - Az(y_j)) / dy, B2 = 0 (and B3 = 0 in 3-D).
    // Cylindrical (r, z) (:122-152): the distance uses the radial centroid and z, the field
    // lies along the axis: cell B = (0, bx, 0), z faces bx, r faces 0.
    if (mhd != KM_MHD_CT || !face || !face[0] || !face[1] || !face[2] || L.ndim < 2)
      return KM_EINVAL;
    const double rho = P(0, 1.0), pamb = P(1, 1.0e-1), pexp = P(2, 1.0e1), bx = P(3, 1.0),
                 radius = P(4, 0.1);
    if (cfg->geometry == KM_GEOM_CYLINDRICAL) {
      if (L.ndim != 2) return KM_EINVAL;
      for (int b = 0; b < L.nblocks; ++b) {
        const km_block_t &bk = blocks[b];
        double *B1 = face[0] + (long)b * L.nstag(), *B2 = face[1] + (long)b * L.nstag();
        for (int j = L.lo(1); j <= L.hi(1) + 1; ++j)
          for (int i = L.lo(0); i <= L.hi(0) + 1; ++i) {
            const long s = L.s(0, j, i);
            if (i <= L.hi(0) && j <= L.hi(1)) {
              const double r = xc_cyl(bk, i, L.lo(0)), z = xc(bk, 1, j, L.lo(1));
              const long c = L.c(0, j, i);
              at(b, KM_DENS)[c] = rho;
              at(b, sV + 0)[c] = 0.;
              at(b, sV + 1)[c] = 0.;
              at(b, sV + 2)[c] = 0.;
              at(b, sP)[c] = std::sqrt(r * r + z * z) <= radius ? pexp : pamb;
              at(b, KM_MAGC0 + 0)[c] = 0.;
              at(b, KM_MAGC0 + 1)[c] = bx;
              at(b, KM_MAGC0 + 2)[c] = 0.;
            }
            if (i <= L.hi(0)) B2[s] = bx;
            if (j <= L.hi(1)) B1[s] = 0.0;
          }
      }
      return KM_OK;
    }
    for (int b = 0; b < L.nblocks; ++b) {
      const km_block_t &bk = blocks[b];
      double *B1 = face[0] + (long)b * L.nstag(), *B2 = face[1] + (long)b * L.nstag();
      const int kx = L.ndim > 2 ? 1 : 0;
      for (int k = L.lo(2); k <= L.hi(2) + kx; ++k)
        for (int j = L.lo(1); j <= L.hi(1) + 1; ++j)
          for (int i = L.lo(0); i <= L.hi(0) + 1; ++i) {
            const double x = xc(bk, 0, i, L.lo(0)), y = xc(bk, 1, j, L.lo(1));
            const bool cell = i <= L.hi(0) && j <= L.hi(1) && k <= L.hi(2);
            if (cell) {
              const long c = L.c(k, j, i);
              const double r = std::sqrt(x * x + y * y);
              at(b, KM_DENS)[c] = rho;
              at(b, sV + 0)[c] = 0.;
              at(b, sV + 1)[c] = 0.;
              at(b, sV + 2)[c] = 0.;
              at(b, sP)[c] = r <= radius ? pexp : pamb;
              at(b, KM_MAGC0 + 0)[c] = bx;
              at(b, KM_MAGC0 + 1)[c] = 0.;
              at(b, KM_MAGC0 + 2)[c] = 0.;
            }
            const long s = L.s(k, j, i);
            if (j <= L.hi(1) && k <= L.hi(2)) {
              const double yp = xf(bk, 1, j + 1, L.lo(1)), ym = xf(bk, 1, j, L.lo(1));
              B1[s] = (bx * yp - bx * ym) / bk.dx[1];
            }
            if (i <= L.hi(0) && k <= L.hi(2)) B2[s] = (bx * xf(bk, 1, j, L.lo(1)) -
                                                       bx * xf(bk, 1, j, L.lo(1))) /
                                                      bk.dx[0];
          }
    }
    return KM_OK;
  }

  if (pname == "mhd_random") {
    // face B = (B0, 0, 0) + discrete curl of an edge-centred vector potential
    // A = sum_m amp sin(2 pi k.x + phi) e_m, e_m a random unit vector normal to k;
    // rho and P = 1 + 0.01 xi with per-zone white noise xi in [-1, 1]; v = 0
    if (mhd != KM_MHD_CT || !face || !face[0] || !face[1] || !face[2] || L.ndim < 2)
      return KM_EINVAL;
    Rng rng(seed);
    const double amp = P(0, 0.05), B0 = P(1, 1.0), noise = P(2, 0.01);
    const int nmodes = (int)P(3, 8);
    auto modes = draw_modes(rng, nmodes, 4, L.ndim);
    std::vector<std::array<double, 3>> evec(modes.size());
    for (size_t m = 0; m < modes.size(); ++m) {
      double kk[3] = {(double)modes[m].k[0], (double)modes[m].k[1], (double)modes[m].k[2]};
      const double kn = std::sqrt(kk[0] * kk[0] + kk[1] * kk[1] + kk[2] * kk[2]);
      for (;;) {
        double r[3] = {2 * rng.uniform() - 1, 2 * rng.uniform() - 1, 2 * rng.uniform() - 1};
        if (L.ndim == 2) r[0] = r[1] = 0, r[2] = 1;  // 2-D: A along z
        const double rk = (r[0] * kk[0] + r[1] * kk[1] + r[2] * kk[2]) / (kn * kn);
        double e[3] = {r[0] - rk * kk[0], r[1] - rk * kk[1], r[2] - rk * kk[2]};
        const double en = std::sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
        if (en < 1e-3) continue;
        evec[m] = {e[0] / en, e[1] / en, e[2] / en};
        break;
      }
    }
    auto Acomp = [&](int comp, double x, double y, double z) {
      double s = 0.;
      for (size_t m = 0; m < modes.size(); ++m)
        s += amp * evec[m][comp] *
             std::sin(2.0 * kPi * (modes[m].k[0] * x + modes[m].k[1] * y + modes[m].k[2] * z) +
                      modes[m].phase);
      return s;
    };
    const bool d3 = L.ndim > 2;
    for (int b = 0; b < L.nblocks; ++b) {
      const km_block_t &bk = blocks[b];
      const int lo0 = L.lo(0), lo1 = L.lo(1), lo2 = L.lo(2);
      auto X = [&](int d, int idx, bool face_pos) {
        const int lo = L.lo(d);
        return face_pos ? xf(bk, d, idx, lo) : xc(bk, d, idx, lo);
      };
      // edge potentials: Ax at (xc, yf, zf), Ay at (xf, yc, zf), Az at (xf, yf, zc)
      auto Ax = [&](int k, int j, int i) {
        return d3 ? Acomp(0, X(0, i, false), X(1, j, true), X(2, k, true)) : 0.0;
      };
      auto Ay = [&](int k, int j, int i) {
        return d3 ? Acomp(1, X(0, i, true), X(1, j, false), X(2, k, true)) : 0.0;
      };
      auto Az = [&](int k, int j, int i) {
        return Acomp(2, X(0, i, true), X(1, j, true), d3 ? X(2, k, false) : 0.0);
      };
      const double dx = bk.dx[0], dy = bk.dx[1], dz = bk.dx[2];
      double *B1 = face[0] + (long)b * L.nstag(), *B2 = face[1] + (long)b * L.nstag(),
             *B3 = face[2] + (long)b * L.nstag();
      for (int k = lo2; k <= L.hi(2) + (d3 ? 1 : 0); ++k)
        for (int j = lo1; j <= L.hi(1) + 1; ++j)
          for (int i = lo0; i <= L.hi(0) + 1; ++i) {
            const long s = L.s(k, j, i);
            if (j <= L.hi(1) && k <= L.hi(2)) {
              double bx = B0 + (Az(k, j + 1, i) - Az(k, j, i)) / dy;
              if (d3) bx -= (Ay(k + 1, j, i) - Ay(k, j, i)) / dz;
              B1[s] = bx;
            }
            if (i <= L.hi(0) && k <= L.hi(2)) {
              double by = -(Az(k, j, i + 1) - Az(k, j, i)) / dx;
              if (d3) by += (Ax(k + 1, j, i) - Ax(k, j, i)) / dz;
              B2[s] = by;
            }
            if (d3 && i <= L.hi(0) && j <= L.hi(1)) {
              B3[s] = (Ay(k, j, i + 1) - Ay(k, j, i)) / dx - (Ax(k, j + 1, i) - Ax(k, j, i)) / dy;
            }
          }
      for (int k = lo2; k <= L.hi(2); ++k)
        for (int j = lo1; j <= L.hi(1); ++j)
          for (int i = lo0; i <= L.hi(0); ++i) {
            const long gi = (long)bk.lx[0] * L.nx[0] + (i - lo0);
            const long gj = (long)bk.lx[1] * L.nx[1] + (j - lo1);
            const long gk = (long)bk.lx[2] * L.nx[2] + (k - lo2);
            const long c = L.c(k, j, i);
            at(b, KM_DENS)[c] = 1.0 + noise * zone_noise(seed, 1, gi, gj, gk);
            at(b, sP)[c] = 1.0 + noise * zone_noise(seed, 2, gi, gj, gk);
          }
    }
    return KM_OK;
  }
  km::set_error("unknown problem '" + pname + "'");
  return KM_EINVAL;
}

}  // extern "C"
