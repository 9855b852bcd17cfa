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
};

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
    flip = !flip;
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

// host-side tree: leaf metrics in the caller's leaf order and the finest-level lookup
struct AmrTreeHost {
  int nb0[3] = {1, 1, 1}, nbf[3] = {1, 1, 1}, lmax = 0, bc[6] = {0, 0, 0, 0, 0, 0};
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

void launch_amr_ghost_fill(const AmrGrid &t, const km_pack_t &u, int v0, int v1, int nc,
                           hipStream_t st);
void launch_amr_flux_correct(const AmrGrid &t, const km_pack_t &u, int nc, hipStream_t st);

}  // namespace km
