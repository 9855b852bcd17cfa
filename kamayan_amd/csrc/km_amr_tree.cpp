// km_amr_tree.cpp -- host side of block-tree meshes: leaf validation, block metrics,
// the finest-level leaf lookup, and a ball-refinement helper with 2:1 balancing.
//
// Parthenon keeps the leaves of its MeshBlockTree; a leaf is (level, logical location)
// and a level-L block covers 2^-L of a root block per dimension, holding the same
// nx cells as every other block. Neighbouring leaves (faces, edges and corners,
// periodic wraps included) differ by at most one level.
#include <algorithm>
#include <array>
#include <cstdint>
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
  for (int d = 0; d < 6; ++d) {
    if (m.bc[d] == KM_BC_AXISYMMETRIC && (d != 0 || m.ndim > 2))
      return "the axisymmetric boundary is the inner x1 boundary of a 1-D or 2-D mesh";
    t.bc[d] = m.bc[d];
  }
  t.x1min = m.xmin[0];
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
  g.cyl = cyl;
  g.x1min = x1min;
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

// ---- adaptive refinement: Parthenon's refinement flags and tree update ----------
// The mesh is held as the level of the leaf covering each cell of the finest level the
// run allows (numlevel - 1). A node (L, l) is covered by a leaf of level <= L exactly when
// the finest cell at its lower corner is, since leaves are aligned boxes; otherwise it is
// split.
namespace {

struct LevelMap {
  int ndim = 1, top = 0, nr[3] = {1, 1, 1}, nf[3] = {1, 1, 1}, bc[6] = {0, 0, 0, 0, 0, 0};
  std::vector<signed char> lev;
  long at(const int f[3]) const { return ((long)f[2] * nf[1] + f[1]) * nf[0] + f[0]; }
  int nodes(int L, int d) const { return d < ndim ? (nr[d] << L) : 1; }
  // level of the leaf over the lower corner of node (L, l)
  int corner(int L, const int l[3]) const {
    int f[3];
    for (int d = 0; d < 3; ++d) f[d] = d < ndim ? (l[d] << (top - L)) : 0;
    return lev[at(f)];
  }
  void paint(int L, const int l[3], int value) {
    int f0[3], f1[3];
    for (int d = 0; d < 3; ++d) {
      f0[d] = d < ndim ? (l[d] << (top - L)) : 0;
      f1[d] = d < ndim ? ((l[d] + 1) << (top - L)) : 1;
    }
    for (int k = f0[2]; k < f1[2]; ++k)
      for (int j = f0[1]; j < f1[1]; ++j)
        for (int i = f0[0]; i < f1[0]; ++i) lev[((long)k * nf[1] + j) * nf[0] + i] = (signed char)value;
  }
  // node l + o at level L across periodic boundaries; false outside the domain
  bool step(int L, const int l[3], int o0, int o1, int o2, int out[3]) const {
    const int o[3] = {o0, o1, o2};
    for (int d = 0; d < 3; ++d) {
      if (d >= ndim) {
        out[d] = 0;
        continue;
      }
      const int n = nodes(L, d);
      int q = l[d] + o[d];
      if (q < 0 || q >= n) {
        if (bc[2 * d] != KM_BC_PERIODIC) return false;
        q = (q + n) % n;
      }
      out[d] = q;
    }
    return true;
  }
};

template <class F>
void each_neighbour(int ndim, F &&f) {
  const int r2 = ndim > 2 ? 1 : 0, r1 = ndim > 1 ? 1 : 0;
  for (int c = -r2; c <= r2; ++c)
    for (int b = -r1; b <= r1; ++b)
      for (int a = -1; a <= 1; ++a)
        if (a || b || c) f(a, b, c);
}

// split leaf (L, l) after splitting every coarser leaf that touches it
int split_leaf(LevelMap &m, int L, const int l[3]) {
  if (m.corner(L, l) != L) return 0;  // no longer a leaf
  int done = 0;
  each_neighbour(m.ndim, [&](int a, int b, int c) {
    int n[3];
    if (!m.step(L, l, a, b, c, n)) return;
    const int cl = m.corner(L, n);
    if (cl >= L) return;
    int p[3];
    for (int d = 0; d < 3; ++d) p[d] = n[d] >> (L - cl);
    done += split_leaf(m, cl, p);
  });
  m.paint(L, l, L + 1);
  return done + 1;
}

// merge the children of (L, l) unless a node two levels finer would touch it
int merge_children(LevelMap &m, int L, const int l[3]) {
  const int nch = 1 << m.ndim;
  for (int c = 0; c < nch; ++c) {
    int ch[3];
    for (int d = 0; d < 3; ++d) ch[d] = d < m.ndim ? 2 * l[d] + ((c >> d) & 1) : 0;
    if (m.corner(L + 1, ch) != L + 1) return 0;
  }
  bool blocked = false;
  each_neighbour(m.ndim, [&](int a, int b, int c) {
    int n[3];
    if (blocked || !m.step(L, l, a, b, c, n) || m.corner(L, n) <= L) return;
    const int o[3] = {a, b, c};
    for (int q = 0; q < nch; ++q) {
      int ch[3];
      bool faces = true;
      for (int d = 0; d < 3; ++d) {
        const int bit = d < m.ndim ? (q >> d) & 1 : 0;
        // the child of n on the side facing (L, l)
        if ((o[d] < 0 && bit == 0) || (o[d] > 0 && bit == 1)) faces = false;
        ch[d] = d < m.ndim ? 2 * n[d] + bit : 0;
      }
      if (faces && m.corner(L + 1, ch) > L + 1) blocked = true;
    }
  });
  if (blocked) return 0;
  m.paint(L, l, L);
  return 1;
}

uint64_t morton(int ndim, int L, const int l[3], int top) {
  uint64_t key = 0;
  for (int bit = 20; bit >= 0; --bit)
    for (int d = ndim - 1; d >= 0; --d) key = (key << 1) | (((l[d] << (top - L)) >> bit) & 1);
  return key;
}

}  // namespace

std::string update_amr_tree(const km_mesh_desc_t &mesh, int numlevel, int threshold,
                            const int32_t *leaves, int n, const int32_t *tags,
                            int32_t *deref, std::vector<int32_t> &out_leaves,
                            std::vector<int32_t> &out_deref, std::vector<int32_t> &origin,
                            bool &changed) {
  if (numlevel < 1 || numlevel > 16) return "numlevel must be 1..16";
  LevelMap m;
  m.ndim = mesh.ndim;
  m.top = numlevel - 1;
  long nfc = 1;
  for (int d = 0; d < 3; ++d) {
    m.nr[d] = d < mesh.ndim ? mesh.nb[d] : 1;
    m.nf[d] = d < mesh.ndim ? (mesh.nb[d] << m.top) : 1;
    nfc *= m.nf[d];
  }
  if (nfc > (1L << 28)) return "numlevel too deep for the root grid";
  for (int d = 0; d < 6; ++d) m.bc[d] = mesh.bc[d];
  m.lev.assign(nfc, -1);
  std::vector<long> old_key(n);
  auto key_of = [&](int L, const int l[3]) {
    return (((long)L * m.nf[2] + l[2]) * m.nf[1] + l[1]) * m.nf[0] + l[0];
  };
  for (int b = 0; b < n; ++b) {
    const int L = leaves[4 * b];
    const int l[3] = {leaves[4 * b + 1], leaves[4 * b + 2], leaves[4 * b + 3]};
    if (L < 0 || L > m.top) return "leaf level above numlevel - 1";
    m.paint(L, l, L);
    old_key[b] = key_of(L, l);
  }
  for (long f = 0; f < nfc; ++f)
    if (m.lev[f] < 0) return "leaves do not cover the domain";
  // MeshRefinement::SetRefinement
  std::vector<int> flag(n, 0);
  for (int b = 0; b < n; ++b) {
    const int L = leaves[4 * b];
    const int l[3] = {leaves[4 * b + 1], leaves[4 * b + 2], leaves[4 * b + 3]};
    const int t = tags[b] < -1 ? -1 : tags[b];
    if (t >= 0) deref[b] = 0;
    if (t > 0) {
      flag[b] = L < m.top ? 1 : 0;
    } else if (t < 0) {
      if (L == 0) {
        deref[b] = 0;
        continue;
      }
      ++deref[b];
      bool finer = false;
      each_neighbour(m.ndim, [&](int a, int bb, int c) {
        int q[3];
        if (m.step(L, l, a, bb, c, q) && m.corner(L, q) > L) finer = true;
      });
      flag[b] = !finer && deref[b] >= threshold ? -1 : 0;
    }
  }
  // Mesh::UpdateMeshBlockTree: merge candidates are parents whose children are all
  // leaves flagged -1; splits by increasing level, then merges by decreasing level
  std::vector<std::pair<int, std::array<int, 3>>> splits, merges;
  std::vector<char> want_merge(nfc, 0);  // by finest corner of flagged leaves (unique per leaf)
  for (int b = 0; b < n; ++b) {
    const int L = leaves[4 * b];
    const std::array<int, 3> l = {leaves[4 * b + 1], leaves[4 * b + 2], leaves[4 * b + 3]};
    if (flag[b] == 1) splits.push_back({L, l});
    if (flag[b] == -1) {
      int f[3];
      for (int d = 0; d < 3; ++d) f[d] = d < m.ndim ? (l[d] << (m.top - L)) : 0;
      want_merge[m.at(f)] = 1;
    }
  }
  const int nch = 1 << m.ndim;
  for (int b = 0; b < n; ++b) {
    if (flag[b] != -1) continue;
    const int L = leaves[4 * b];
    const int l[3] = {leaves[4 * b + 1], leaves[4 * b + 2], leaves[4 * b + 3]};
    bool even = true;
    for (int d = 0; d < m.ndim; ++d) even = even && (l[d] % 2 == 0);
    if (!even) continue;
    bool all = true;
    for (int c = 0; c < nch && all; ++c) {
      int f[3];
      for (int d = 0; d < 3; ++d) f[d] = d < m.ndim ? ((l[d] + ((c >> d) & 1)) << (m.top - L)) : 0;
      all = want_merge[m.at(f)] && m.lev[m.at(f)] == L;
    }
    if (all) merges.push_back({L - 1, {l[0] / 2, l[1] / 2, l[2] / 2}});
  }
  std::stable_sort(splits.begin(), splits.end(),
                   [](const auto &a, const auto &b) { return a.first < b.first; });
  std::stable_sort(merges.begin(), merges.end(),
                   [](const auto &a, const auto &b) { return a.first > b.first; });
  int applied = 0;
  for (const auto &sp : splits) applied += split_leaf(m, sp.first, sp.second.data());
  for (const auto &mg : merges) applied += merge_children(m, mg.first, mg.second.data());
  changed = applied > 0;
  out_leaves.clear();
  out_deref.clear();
  origin.clear();
  if (!changed) {
    out_leaves.assign(leaves, leaves + 4 * (size_t)n);
    out_deref.assign(deref, deref + n);
    origin.assign(9 * (size_t)n, -1);
    for (int b = 0; b < n; ++b) {
      origin[9 * (size_t)b] = 0;
      origin[9 * (size_t)b + 1] = b;
    }
    return "";
  }
  // leaves of the new map, Morton order
  struct Leaf {
    uint64_t key;
    int L, l[3];
  };
  std::vector<Leaf> nl;
  for (int k = 0; k < m.nf[2]; ++k)
    for (int j = 0; j < m.nf[1]; ++j)
      for (int i = 0; i < m.nf[0]; ++i) {
        const int L = m.lev[((long)k * m.nf[1] + j) * m.nf[0] + i];
        const int s = m.top - L, f[3] = {i, j, k};
        bool corner = true;
        for (int d = 0; d < m.ndim; ++d) corner = corner && (f[d] % (1 << s) == 0);
        if (!corner) continue;
        Leaf e{0, L, {i >> s, m.ndim > 1 ? j >> s : 0, m.ndim > 2 ? k >> s : 0}};
        e.key = morton(m.ndim, L, e.l, m.top);
        nl.push_back(e);
      }
  std::sort(nl.begin(), nl.end(), [](const Leaf &a, const Leaf &b) { return a.key < b.key; });
  std::vector<std::pair<long, int>> byk(n);
  for (int b = 0; b < n; ++b) byk[b] = {old_key[b], b};
  std::sort(byk.begin(), byk.end());
  auto find_old = [&](int L, const int l[3]) -> int {
    const long k = key_of(L, l);
    auto it = std::lower_bound(byk.begin(), byk.end(), std::make_pair(k, -1));
    return it != byk.end() && it->first == k ? it->second : -1;
  };
  origin.assign(9 * nl.size(), -1);
  for (size_t i = 0; i < nl.size(); ++i) {
    const Leaf &e = nl[i];
    out_leaves.insert(out_leaves.end(), {e.L, e.l[0], e.l[1], e.l[2]});
    int32_t *o = &origin[9 * i];
    const int same = find_old(e.L, e.l);
    if (same >= 0) {
      o[0] = 0;
      o[1] = same;
      out_deref.push_back(deref[same]);
      continue;
    }
    out_deref.push_back(0);
    if (e.L > 0) {
      int p[3];
      for (int d = 0; d < 3; ++d) p[d] = e.l[d] >> 1;
      const int parent = find_old(e.L - 1, p);
      if (parent >= 0) {
        o[0] = 1;
        o[1] = parent;
        continue;
      }
    }
    o[0] = 2;
    for (int c = 0; c < nch; ++c) {
      int ch[3];
      for (int d = 0; d < 3; ++d) ch[d] = d < m.ndim ? 2 * e.l[d] + ((c >> d) & 1) : 0;
      o[1 + c] = find_old(e.L + 1, ch);
      if (o[1 + c] < 0) return "remesh produced a leaf with no source";
    }
  }
  return "";
}

}  // namespace km

using namespace km;

extern "C" {

int km_amr_update_tree(const km_mesh_desc_t *mesh, int32_t numlevel, int32_t derefine_count,
                       const int32_t *leaves, int32_t nleaves, const int32_t *tags,
                       int32_t *deref_count, int32_t *new_leaves, int32_t *new_deref,
                       int32_t *origin, int32_t *nnew, int32_t *changed) {
  if (!mesh || !leaves || !tags || !deref_count || !nnew || !changed || nleaves < 1)
    return KM_EINVAL;
  std::vector<int32_t> dc(deref_count, deref_count + nleaves), nl, nd, org;
  bool ch = false;
  const std::string err = update_amr_tree(*mesh, numlevel, derefine_count, leaves, nleaves,
                                          tags, dc.data(), nl, nd, org, ch);
  if (!err.empty()) {
    set_error(err);
    return KM_EINVAL;
  }
  const int n = (int)nl.size() / 4;
  const int cap = *nnew;
  *nnew = n;
  if (!new_leaves || !new_deref || !origin || cap < n) {
    set_error("leaf buffers hold " + std::to_string(cap) + " entries, " + std::to_string(n) +
              " needed");
    return KM_EINVAL;
  }
  std::copy(dc.begin(), dc.end(), deref_count);
  std::copy(nl.begin(), nl.end(), new_leaves);
  std::copy(nd.begin(), nd.end(), new_deref);
  std::copy(org.begin(), org.end(), origin);
  *changed = ch ? 1 : 0;
  return KM_OK;
}


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
