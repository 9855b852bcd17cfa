// km_amr_tree.cpp -- host side of block-tree meshes: leaf validation, block metrics,
// the finest-level leaf lookup, and a ball-refinement helper with 2:1 balancing.
//
// Parthenon keeps the leaves of its MeshBlockTree; a leaf is (level, logical location)
// and a level-L block covers 2^-L of a root block per dimension, holding the same
// nx cells as every other block. Neighbouring leaves (faces, edges and corners,
// periodic wraps included) differ by at most one level.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/km_hydro.h"
#include "km_amr.hpp"
#include "km_internal.hpp"

namespace km {

namespace {

int active(const km_mesh_desc_t &m, int d) { return d < m.ndim; }

// block bounds along d of level-L block l (the uniform generator of a root grid
// refined L times; the domain edges are exact)
void block_bounds(const km_mesh_desc_t &m, int d, int L, int l, double &lo, double &hi) {
  if (!active(m, d)) {
    lo = m.xmin[d];
    hi = m.xmax[d];
    return;
  }
  const long nb = (long)m.nb[d] << L;
  const double rl = static_cast<double>(2 * (long)l - nb) / (2 * nb);
  const double rr = static_cast<double>(2 * ((long)l + 1) - nb) / (2 * nb);
  lo = 0.5 * (m.xmin[d] + m.xmax[d]) + (rl * m.xmax[d] - rl * m.xmin[d]);
  hi = 0.5 * (m.xmin[d] + m.xmax[d]) + (rr * m.xmax[d] - rr * m.xmin[d]);
  if (l == 0) lo = m.xmin[d];
  if (l == nb - 1) hi = m.xmax[d];
}

}  // namespace

void amr_block_metric(const km_mesh_desc_t &m, int L, const int l[3], km_block_t &blk) {
  std::memset(&blk, 0, sizeof(blk));
  for (int d = 0; d < 3; ++d) {
    double lo, hi;
    block_bounds(m, d, L, active(m, d) ? l[d] : 0, lo, hi);
    const int n = active(m, d) ? m.nx[d] : 1;
    blk.xmin[d] = lo;
    blk.dx[d] = (hi - lo) / static_cast<double>(n);
    blk.lx[d] = active(m, d) ? l[d] : 0;
  }
  blk.area[0] = blk.dx[1] * blk.dx[2];
  blk.area[1] = blk.dx[0] * blk.dx[2];
  blk.area[2] = blk.dx[0] * blk.dx[1];
  blk.vol = blk.dx[0] * blk.dx[1] * blk.dx[2];
  blk.level = L;
}

// Ghost zones reach ng cells past a block and a prolongation stencil 2*(ceil(ng/2) + 1);
// with blocks of at least that many cells (and 2 ng), every leaf those reads land in
// touches the block being filled, so 2:1 balance bounds its level to L-1 .. L+1.
int amr_min_cells(int ng) {
  const int stencil = 2 * ((ng + 1) / 2 + 1);
  return stencil > 2 * ng ? stencil : 2 * ng;
}

std::string build_amr_tree(const km_mesh_desc_t &m, const int32_t *leaves, int n,
                           AmrTreeHost &t) {
  if (m.ndim < 1 || m.ndim > 3) return "bad ndim";
  if (n < 1 || !leaves) return "no leaves";
  int lmax = 0;
  for (int b = 0; b < n; ++b) {
    const int L = leaves[4 * b];
    if (L < 0 || L > 20) return "leaf level out of range";
    lmax = std::max(lmax, L);
  }
  t.lmax = lmax;
  for (int d = 0; d < 3; ++d) {
    t.nb0[d] = active(m, d) ? m.nb[d] : 1;
    t.nbf[d] = active(m, d) ? (m.nb[d] << lmax) : 1;
  }
  for (int d = 0; d < 6; ++d) t.bc[d] = m.bc[d];
  for (int d = 0; d < m.ndim; ++d) {
    if ((m.bc[2 * d] == KM_BC_PERIODIC) != (m.bc[2 * d + 1] == KM_BC_PERIODIC))
      return "periodic boundaries must be paired";
    if (lmax > 0 && (m.nx[d] % 2 != 0 || m.nx[d] < amr_min_cells(m.ng)))
      return "refined meshes need an even block size of at least " +
             std::to_string(amr_min_cells(m.ng)) + " cells per active dimension for nghost " +
             std::to_string(m.ng);
    if (m.ng > m.nx[d]) return "ghost width larger than a block";
  }
  const long nf = (long)t.nbf[0] * t.nbf[1] * t.nbf[2];
  t.lookup.assign(nf, -1);
  t.blocks.resize(n);
  for (int b = 0; b < n; ++b) {
    const int L = leaves[4 * b];
    int l[3];
    for (int d = 0; d < 3; ++d) {
      l[d] = leaves[4 * b + 1 + d];
      const long nbL = active(m, d) ? ((long)m.nb[d] << L) : 1;
      if (l[d] < 0 || l[d] >= nbL) return "leaf location outside the root grid";
    }
    amr_block_metric(m, L, l, t.blocks[b]);
    const int s = lmax - L;
    int f0[3], f1[3];
    for (int d = 0; d < 3; ++d) {
      f0[d] = active(m, d) ? (l[d] << s) : 0;
      f1[d] = active(m, d) ? ((l[d] + 1) << s) : 1;
    }
    for (int k = f0[2]; k < f1[2]; ++k)
      for (int j = f0[1]; j < f1[1]; ++j)
        for (int i = f0[0]; i < f1[0]; ++i) {
          int &e = t.lookup[((long)k * t.nbf[1] + j) * t.nbf[0] + i];
          if (e >= 0) return "leaves overlap";
          e = b;
        }
  }
  for (long f = 0; f < nf; ++f)
    if (t.lookup[f] < 0) return "leaves do not cover the domain";
  // 2:1 balance across faces, edges and corners
  for (int k = 0; k < t.nbf[2]; ++k)
    for (int j = 0; j < t.nbf[1]; ++j)
      for (int i = 0; i < t.nbf[0]; ++i) {
        const int me = t.blocks[t.lookup[((long)k * t.nbf[1] + j) * t.nbf[0] + i]].level;
        for (int ok = -1; ok <= 1; ++ok)
          for (int oj = -1; oj <= 1; ++oj)
            for (int oi = -1; oi <= 1; ++oi) {
              int q[3] = {i + oi, j + oj, k + ok};
              const int o[3] = {oi, oj, ok};
              bool skip = false;
              for (int d = 0; d < 3; ++d) {
                if (o[d] != 0 && !active(m, d)) skip = true;
                if (q[d] < 0 || q[d] >= t.nbf[d]) {
                  if (m.bc[2 * d] != KM_BC_PERIODIC) skip = true;
                  q[d] = (q[d] + t.nbf[d]) % t.nbf[d];
                }
              }
              if (skip) continue;
              const int nb = t.blocks[t.lookup[((long)q[2] * t.nbf[1] + q[1]) * t.nbf[0] + q[0]]]
                                 .level;
              if (nb > me + 1 || nb < me - 1)
                return "leaves that touch differ by more than one level";
            }
      }
  return "";
}

AmrGrid AmrTreeHost::grid(const int *lookup_dev) const {
  AmrGrid g;
  for (int d = 0; d < 3; ++d) {
    g.nb0[d] = nb0[d];
    g.nbf[d] = nbf[d];
  }
  g.lmax = lmax;
  for (int d = 0; d < 6; ++d) g.bc[d] = bc[d];
  g.lookup = lookup_dev;
  return g;
}

namespace {

// leaves of the tree whose finest-level positions want at least des[] levels
void leaves_from_desired(const km_mesh_desc_t &m, int lmax, const int nbf[3],
                         const std::vector<int> &des, std::vector<int32_t> &out) {
  out.clear();
  struct Node {
    int L, l[3];
  };
  std::vector<Node> stack;
  // roots in x-fastest order, children in x-fastest order (depth first)
  const int nr[3] = {m.nb[0], active(m, 1) ? m.nb[1] : 1, active(m, 2) ? m.nb[2] : 1};
  for (int k = nr[2] - 1; k >= 0; --k)
    for (int j = nr[1] - 1; j >= 0; --j)
      for (int i = nr[0] - 1; i >= 0; --i) stack.push_back({0, {i, j, k}});
  while (!stack.empty()) {
    const Node nd = stack.back();
    stack.pop_back();
    const int s = lmax - nd.L;
    int want = 0;
    int f0[3], f1[3];
    for (int d = 0; d < 3; ++d) {
      f0[d] = active(m, d) ? (nd.l[d] << s) : 0;
      f1[d] = active(m, d) ? ((nd.l[d] + 1) << s) : 1;
    }
    for (int k = f0[2]; k < f1[2]; ++k)
      for (int j = f0[1]; j < f1[1]; ++j)
        for (int i = f0[0]; i < f1[0]; ++i)
          want = std::max(want, des[((long)k * nbf[1] + j) * nbf[0] + i]);
    if (want <= nd.L) {
      out.push_back(nd.L);
      for (int d = 0; d < 3; ++d) out.push_back(nd.l[d]);
      continue;
    }
    const int n2 = active(m, 2) ? 2 : 1, n1 = active(m, 1) ? 2 : 1;
    for (int ck = n2 - 1; ck >= 0; --ck)
      for (int cj = n1 - 1; cj >= 0; --cj)
        for (int ci = 1; ci >= 0; --ci) {
          Node c{nd.L + 1, {2 * nd.l[0] + ci, active(m, 1) ? 2 * nd.l[1] + cj : 0,
                            active(m, 2) ? 2 * nd.l[2] + ck : 0}};
          stack.push_back(c);
        }
  }
}

}  // namespace

std::string refine_ball(const km_mesh_desc_t &m, int lmax, const double c[3], double r,
                        std::vector<int32_t> &leaves) {
  int nbf[3];
  for (int d = 0; d < 3; ++d) nbf[d] = active(m, d) ? (m.nb[d] << lmax) : 1;
  const long nf = (long)nbf[0] * nbf[1] * nbf[2];
  std::vector<int> des(nf, 0);
  for (int k = 0; k < nbf[2]; ++k)
    for (int j = 0; j < nbf[1]; ++j)
      for (int i = 0; i < nbf[0]; ++i) {
        const int l[3] = {i, j, k};
        double d2 = 0.;
        for (int d = 0; d < m.ndim; ++d) {
          double lo, hi;
          block_bounds(m, d, lmax, l[d], lo, hi);
          const double q = std::min(std::max(c[d], lo), hi) - c[d];
          d2 += q * q;
        }
        if (d2 <= r * r) des[((long)k * nbf[1] + j) * nbf[0] + i] = lmax;
      }
  for (int iter = 0; iter < 64; ++iter) {
    leaves_from_desired(m, lmax, nbf, des, leaves);
    // level of every finest position
    std::vector<int> lev(nf, 0);
    const int n = (int)leaves.size() / 4;
    for (int b = 0; b < n; ++b) {
      const int L = leaves[4 * b], s = lmax - L;
      int f0[3], f1[3];
      for (int d = 0; d < 3; ++d) {
        f0[d] = active(m, d) ? (leaves[4 * b + 1 + d] << s) : 0;
        f1[d] = active(m, d) ? ((leaves[4 * b + 1 + d] + 1) << s) : 1;
      }
      for (int k = f0[2]; k < f1[2]; ++k)
        for (int j = f0[1]; j < f1[1]; ++j)
          for (int i = f0[0]; i < f1[0]; ++i) lev[((long)k * nbf[1] + j) * nbf[0] + i] = L;
    }
    bool changed = false;
    for (int k = 0; k < nbf[2]; ++k)
      for (int j = 0; j < nbf[1]; ++j)
        for (int i = 0; i < nbf[0]; ++i) {
          const long f = ((long)k * nbf[1] + j) * nbf[0] + i;
          for (int ok = -1; ok <= 1; ++ok)
            for (int oj = -1; oj <= 1; ++oj)
              for (int oi = -1; oi <= 1; ++oi) {
                int q[3] = {i + oi, j + oj, k + ok};
                const int o[3] = {oi, oj, ok};
                bool skip = false;
                for (int d = 0; d < 3; ++d) {
                  if (o[d] != 0 && !active(m, d)) skip = true;
                  if (q[d] < 0 || q[d] >= nbf[d]) {
                    if (m.bc[2 * d] != KM_BC_PERIODIC) skip = true;
                    q[d] = (q[d] + nbf[d]) % nbf[d];
                  }
                }
                if (skip) continue;
                const int nl = lev[((long)q[2] * nbf[1] + q[1]) * nbf[0] + q[0]];
                if (nl > lev[f] + 1 && des[f] < nl - 1) {
                  des[f] = nl - 1;
                  changed = true;
                }
              }
        }
    if (!changed) return "";
  }
  return "refinement did not balance";
}

}  // namespace km

using namespace km;

extern "C" {

int km_amr_blocks(const km_mesh_desc_t *mesh, const int32_t *leaves, int32_t nleaves,
                  km_block_t *blocks_host) {
  if (!mesh || !leaves || !blocks_host) return KM_EINVAL;
  AmrTreeHost t;
  const std::string err = build_amr_tree(*mesh, leaves, nleaves, t);
  if (!err.empty()) {
    set_error(err);
    return KM_EINVAL;
  }
  std::memcpy(blocks_host, t.blocks.data(), t.blocks.size() * sizeof(km_block_t));
  return KM_OK;
}

int km_amr_refine_ball(const km_mesh_desc_t *mesh, int32_t lmax, const double center[3],
                       double radius, int32_t *leaves, int32_t *nleaves) {
  if (!mesh || !center || !nleaves || lmax < 0 || lmax > 12) return KM_EINVAL;
  std::vector<int32_t> out;
  const std::string err = refine_ball(*mesh, lmax, center, radius, out);
  if (!err.empty()) {
    set_error(err);
    return KM_EINVAL;
  }
  const int n = (int)out.size() / 4;
  const int cap = *nleaves;
  *nleaves = n;
  if (!leaves || cap < n) {
    set_error("leaf buffer holds " + std::to_string(cap) + " entries, " + std::to_string(n) +
              " needed");
    return KM_EINVAL;
  }
  std::memcpy(leaves, out.data(), out.size() * sizeof(int32_t));
  return KM_OK;
}

}  // extern "C"
