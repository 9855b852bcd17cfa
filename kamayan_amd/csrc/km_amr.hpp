// km_amr.hpp -- block-tree geometry shared by the AMR host code and kernels.
//
// A mesh is a set of leaf blocks (level, logical location) over a root grid of nb0 blocks,
// every block holding the same number of cells (as in Parthenon). lookup[] maps each
// finest-level block position to the leaf covering it, so the leaf that contains any
// zone at any level is found in O(1).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "km_device.hpp"

namespace km {

struct AmrGrid {
  int nb0[3];         // root blocks per dimension
  int lmax;           // finest level present
  int bc[6];          // ix1, ox1, ix2, ox2, ix3, ox3
  int nbf[3];         // lookup extent: nb0 << lmax along active dimensions
  const int *lookup;  // [nbf2][nbf1][nbf0] -> leaf block id
  int cyl;            // cylindrical (r, z) geometry: r-weighted refinement operators
  double x1min;       // domain's lower x1 edge (the radius of global face 0)
};

// Radius of a level-L element along x1 for the r-weighted face operators: global face
// index G (centre = 0) or the centre of cell G (centre = 1), h the level's spacing. Ghost
// positions past the axis have negative radii.
KM_HD double amr_radius(double x1min, int G, int centre, double h) {
  return centre ? x1min + (G + 0.5) * h : x1min + G * h;
}

// global zone count along dimension d at level L
KM_HD int amr_zones(const AmrGrid &t, const Geo &g, int d, int L) {
  return d < g.ndim ? (t.nb0[d] << L) * g.nx[d] : 1;
}

// Physical boundary map of a level-L global zone index along d: periodic wrap, outflow
// clamp, reflect mirror (flip toggles for the normal vector component).
KM_HD int amr_map(const AmrGrid &t, int d, int G, int N, bool &flip) {
  if (G < 0) {
    const int bc = t.bc[2 * d];
    if (bc == KM_BC_PERIODIC) return ((G % N) + N) % N;
    if (bc == KM_BC_OUTFLOW) return 0;
    flip = !flip;  // reflect and axisymmetric mirror the zones
    return -G - 1;
  }
  if (G >= N) {
    const int bc = t.bc[2 * d + 1];
    if (bc == KM_BC_PERIODIC) return G % N;
    if (bc == KM_BC_OUTFLOW) return N - 1;
    flip = !flip;
    return 2 * N - G - 1;
  }
  return G;
}

// the same map for a face index along d (N + 1 faces; 0 and N coincide when periodic)
KM_HD int amr_map_face(const AmrGrid &t, int d, int G, int N, bool &flip) {
  if (G < 0) {
    const int bc = t.bc[2 * d];
    if (bc == KM_BC_PERIODIC) return ((G % N) + N) % N;
    if (bc == KM_BC_OUTFLOW) return 0;
    flip = !flip;
    // the axisymmetric boundary maps x1 faces as it maps cells (boundary_conditions.cpp:
    // 40-58): ghost face -g takes face g - 1
    return bc == KM_BC_AXISYMMETRIC ? -G - 1 : -G;
  }
  if (G > N) {
    const int bc = t.bc[2 * d + 1];
    if (bc == KM_BC_PERIODIC) return G % N;
    if (bc == KM_BC_OUTFLOW) return N;
    flip = !flip;
    return 2 * N - G;
  }
  return G;
}

// leaf containing the (in-domain) level-L zone G[3]
KM_HD int amr_leaf(const AmrGrid &t, const Geo &g, int L, const int G[3]) {
  int f[3];
  for (int d = 0; d < 3; ++d)
    f[d] = d < g.ndim ? ((G[d] << (t.lmax - L)) / g.nx[d]) : 0;
  return t.lookup[((long)f[2] * t.nbf[1] + f[1]) * t.nbf[0] + f[0]];
}

// local (ghost-inclusive) index in leaf `blk` of its own-level global zone G along d,
// clamped to the interior (the mesh checks make the clamp a no-op; it keeps every read
// inside the block whatever the leaves)
KM_HD int amr_local(const km_block_t &blk, const Geo &g, int d, int G) {
  if (d >= g.ndim) return 0;
  const int i = G - blk.lx[d] * g.nx[d] + g.lo[d];
  return i < g.lo[d] ? g.lo[d] : (i > g.hi[d] ? g.hi[d] : i);
}

// local face index in leaf `blk` of its own-level global face G along d (clamped to the
// block's own faces lo .. hi+1)
KM_HD int amr_local_face(const km_block_t &blk, const Geo &g, int d, int G) {
  if (d >= g.ndim) return 0;
  const int i = G - blk.lx[d] * g.nx[d] + g.lo[d];
  return i < g.lo[d] ? g.lo[d] : (i > g.hi[d] + 1 ? g.hi[d] + 1 : i);
}

// the corrected value of edge idx (family te) of block b, if a finer leaf shares it
KM_HD bool amr_emf_edge(const km_pack_t &u, const AmrGrid &t, int te, int b, const int idx[3],
                        double &val) {
  const Geo g(u);
  // edge family te: cell index along te, node indices across
  int ax[2], n = 0;
#pragma unroll
  for (int e = 0; e < 3; ++e)
    if (e != te) ax[n++] = e;
  bool onb = false;
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    if (e >= g.ndim) {
      if (idx[e] != 0) return false;
      continue;
    }
    const int hi = g.hi[e] + (e == te ? 0 : 1);
    if (idx[e] < g.lo[e] || idx[e] > hi) return false;
    if (e != te && (idx[e] == g.lo[e] || idx[e] == g.hi[e] + 1)) onb = true;
  }
  if (!onb) return false;
  const km_block_t &blk = u.blocks[b];
  const int L = blk.level;
  if (L >= t.lmax) return false;
  int node[3];  // level-L global node (across axes) / cell (along te)
#pragma unroll
  for (int e = 0; e < 3; ++e) node[e] = e < g.ndim ? blk.lx[e] * g.nx[e] + (idx[e] - g.lo[e]) : 0;
  // the four cells around the edge: look for a finer leaf
  for (int q = 0; q < 4; ++q) {
    int c[3] = {node[0], node[1], node[2]};
    int fnode[3];
    bool skip = false;
#pragma unroll
    for (int m = 0; m < 2; ++m) {
      const int e = ax[m];
      if (e >= g.ndim) {
        c[e] = 0;
        fnode[e] = 0;
        continue;
      }
      const int below = (q >> m) & 1;  // cell below the node along e
      const int N = amr_zones(t, g, e, L);
      int ce = node[e] - below;
      if (ce < 0 || ce >= N) {
        if (t.bc[2 * e + (ce < 0 ? 0 : 1)] != KM_BC_PERIODIC) {
          skip = true;
          continue;
        }
        ce = (ce + N) % N;
      }
      c[e] = ce;
      fnode[e] = 2 * (ce + below);  // the node in the wrapped frame, at level L+1
    }
    if (skip) continue;
    int cf[3];
#pragma unroll
    for (int e = 0; e < 3; ++e) cf[e] = e < g.ndim ? 2 * c[e] : 0;
    const int fl = amr_leaf(t, g, L + 1, cf);
    const km_block_t &fb = u.blocks[fl];
    if (fb.level != L + 1) continue;
    int fe[3];
#pragma unroll
    for (int e = 0; e < 3; ++e) {
      if (e >= g.ndim) {
        fe[e] = 0;
      } else if (e == te) {
        fe[e] = amr_local(fb, g, e, 2 * node[e]);
      } else {
        int li = fnode[e] - fb.lx[e] * g.nx[e] + g.lo[e];
        fe[e] = li < g.lo[e] ? g.lo[e] : (li > g.hi[e] + 1 ? g.hi[e] + 1 : li);
      }
    }
    const bool inc[3] = {te == 0, te == 1 && g.ndim > 1, te == 2 && g.ndim > 2};
    // edge length: dx along the edge, |r| 2 pi for phi edges in cylindrical geometry
    const double len =
        (t.cyl && te == 2) ? __builtin_fabs(x_face(fb, g, 0, fe[0])) * kTwoPi : fb.dx[te];
    const double *E = u.edge[te][fl];
    const long fidx = g.s(fe[2], fe[1], fe[0]);
    double V[2][2][2], Tm[2][2][2];
#pragma unroll
    for (int ok = 0; ok < 2; ++ok)
#pragma unroll
      for (int oj = 0; oj < 2; ++oj)
#pragma unroll
        for (int oi = 0; oi < 2; ++oi) {
          V[ok][oj][oi] = 0.;
          Tm[ok][oj][oi] = 0.;
          if ((ok == 0 || inc[2]) && (oj == 0 || inc[1]) && (oi == 0 || inc[0])) {
            V[ok][oj][oi] = len;
            Tm[ok][oj][oi] = len * E[fidx + ok * g.ss[2] + oj * g.ss[1] + oi];
          }
        }
    const double tvol = ((V[0][0][0] + V[0][1][0]) + (V[0][0][1] + V[0][1][1])) +
                        ((V[1][0][0] + V[1][1][0]) + (V[1][0][1] + V[1][1][1]));
    val = tvol > 0.0 ? (((Tm[0][0][0] + Tm[0][1][0]) + (Tm[0][0][1] + Tm[0][1][1])) +
                        ((Tm[1][0][0] + Tm[1][1][0]) + (Tm[1][0][1] + Tm[1][1][1]))) /
                           tvol
                     : 0.0;
    return true;
  }
  return false;
}

// host-side tree: leaf metrics in the caller's leaf order and the finest-level lookup
struct AmrTreeHost {
  int nb0[3] = {1, 1, 1}, nbf[3] = {1, 1, 1}, lmax = 0, bc[6] = {0, 0, 0, 0, 0, 0};
  int cyl = 0;        // set by the caller from the configuration's geometry
  double x1min = 0.;
  std::vector<km_block_t> blocks;
  std::vector<int> lookup;
  AmrGrid grid(const int *lookup_dev) const;
};

// validates leaves ({level, lx1, lx2, lx3} each) and fills `t`; returns an error text
// or an empty string
std::string build_amr_tree(const km_mesh_desc_t &m, const int32_t *leaves, int n,
                           AmrTreeHost &t);
// leaves refined to lmax where blocks meet the ball |x - c| <= r, 2:1 balanced
std::string refine_ball(const km_mesh_desc_t &m, int lmax, const double c[3], double r,
                        std::vector<int32_t> &leaves);
// smallest block size (cells per active dimension) a refined mesh accepts
int amr_min_cells(int ng);
// refinement flags and tree update of an adaptive mesh (km_amr_update_tree); deref is
// updated in place; origin rows of 9 as documented at km_amr_update_tree
std::string update_amr_tree(const km_mesh_desc_t &mesh, int numlevel, int threshold,
                            const int32_t *leaves, int n, const int32_t *tags,
                            int32_t *deref, std::vector<int32_t> &out_leaves,
                            std::vector<int32_t> &out_deref, std::vector<int32_t> &origin,
                            bool &changed);
void launch_amr_tag_loehner(const km_pack_t &u, const km_amr_criterion_t &crit, int *tags,
                            bool reset, double *err2_scratch, hipStream_t st);
void launch_amr_remesh(const km_pack_t &old_u, const km_pack_t &new_u, const int32_t *origin,
                       int v0, int v1, bool cyl, hipStream_t st);
// the CT face field's remesh (all three components; new ghost faces are left unset);
// x1min: the domain's lower x1 edge (radii of the cylindrical face operators)
void launch_amr_remesh_faces(const km_pack_t &old_u, const km_pack_t &new_u,
                             const int32_t *origin, bool cyl, double x1min, hipStream_t st);

// cell variables [v0, v1) except the slots [s0, s1) (none when s1 <= s0); depth > 0 fills
// only the face slabs of that depth (what directional sweeps read), else every ghost zone
void launch_amr_ghost_fill(const AmrGrid &t, const km_pack_t &u, int v0, int v1, int nc,
                           hipStream_t st, int s0 = 0, int s1 = 0, int depth = 0);
// Gather list of a tree's ghost fill: one entry per ghost zone of the enumeration
// (face slabs of `depth`, or every ghost zone), the destination block and cell, the
// source leaf (-1: a coarser leaf covers the zone) and, in src, the source cell (bits
// 0-27), restriction of its children (bit 28) and the flips of the x1-x3 vector
// components (bits 29-31)
struct alignas(16) AmrGhostEntry {
  int db, dc, sb;
  unsigned src;
};
// the lists of one tree and depth: ghost entries, and the prolonged parents (as
// b * per_block + r of k_amr_prolong's enumeration)
struct AmrFillList {
  int depth = 0;
  AmrGhostEntry *list = nullptr;
  long n = 0;
  const long *plist = nullptr;
  long np = 0;
};
long amr_ghost_list_size(const km_pack_t &u, int depth);
long amr_prolong_parents(const km_pack_t &u, int depth);
void launch_amr_ghost_list(const AmrGrid &t, const km_pack_t &u, int depth, AmrGhostEntry *out,
                           hipStream_t st);
// flag[p] = 1 for each parent p of the prolongation enumeration that a coarser leaf covers
void launch_amr_prolong_flags(const AmrGrid &t, const km_pack_t &u, int depth,
                              unsigned char *flag, hipStream_t st);
// launch_amr_ghost_fill over the enumeration the lists were built for; s2 (>= s1): one more
// slot to leave out (-1: none)
void launch_amr_ghost_fill_list(const AmrFillList &f, const AmrGrid &t, const km_pack_t &u,
                                int v0, int v1, int nc, hipStream_t st, int s0 = 0, int s1 = 0,
                                int s2 = -1);
// The same with the prolongation on side stream st2, concurrent with the gather on st
// (both read interior zones only and write disjoint ghost zones): fork is recorded on st
// and waited for on st2, join recorded on st2 and waited for on st, so the fill is whole
// on st afterwards (also inside a graph capture of st, as two branches). Sequential on st
// when st2 is null or one of the two lists is empty.
void launch_amr_ghost_fill_list(const AmrFillList &f, const AmrGrid &t, const km_pack_t &u,
                                int v0, int v1, int nc, hipStream_t st, int s0, int s1, int s2,
                                hipStream_t st2, hipEvent_t fork, hipEvent_t join);
void launch_amr_flux_correct(const AmrGrid &t, const km_pack_t &u, int nc, hipStream_t st);
// CT: face B ghost fill across levels and EMF correction at level jumps
void launch_amr_ghost_faces(const AmrGrid &t, const km_pack_t &u, hipStream_t st);
void launch_amr_emf_correct(const AmrGrid &t, const km_pack_t &u, hipStream_t st);

}  // namespace km
