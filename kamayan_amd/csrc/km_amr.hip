// km_amr.hip -- ghost fill across refinement levels and flux correction for block-tree
// meshes: Parthenon's boundary exchange and flux correction (kamayan_driver.cpp:145-151,
// 198-207) with kamayan's refinement operators (refinement_operations.hpp:20-268).
//
// Every ghost zone is computed from leaf interiors only, so one launch fills all blocks.
// For a ghost zone of a level-L block, mapped through the physical boundaries:
//   covered by a level-L leaf   -> copy
//   covered by a level-L+1 leaf -> RestrictAverage of its 2^d children (:52-121)
//   covered by a level-L-1 leaf -> ProlongateSharedMinMod (:123-268) of the coarse zone C
//                                  holding it: minmod slopes from C's neighbours at level
//                                  L-1, each a coarse leaf value or the restriction of a
//                                  level-L leaf (what Parthenon's coarse buffer holds there)
// Slopes use the spacings of the fine block's coarse buffer: 2 dx between coarse centres
// and dx/2 from a coarse centre to a child centre (uniform Cartesian coordinates).
// The mesh builder guarantees that every zone read here lies in a leaf that touches the
// block being filled, so the levels met are L-1, L and L+1 only.
#include <hip/hip_runtime.h>

#include "km_amr.hpp"
#include "km_device.hpp"
#include "km_internal.hpp"

namespace km {

namespace {

// a level-L zone (global index, in the domain) as stored in the leaf covering it
struct Src {
  int kind;  // 0: leaf at level L; 1: leaf at L+1 (restrict); 2: leaf at L-1
  int leaf;
  long idx;  // kind 0/2: the zone holding it; kind 1: lower corner of its children
};

__device__ __forceinline__ Src locate(const km_pack_t &u, const AmrGrid &t, const Geo &g, int L,
                                      const int G[3]) {
  Src s;
  s.leaf = amr_leaf(t, g, L, G);
  const km_block_t &nb = u.blocks[s.leaf];
  int li[3];
  if (nb.level == L) {
    s.kind = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) li[d] = amr_local(nb, g, d, G[d]);
  } else if (nb.level > L) {
    s.kind = 1;
#pragma unroll
    for (int d = 0; d < 3; ++d) li[d] = amr_local(nb, g, d, 2 * G[d]);
  } else {
    s.kind = 2;
#pragma unroll
    for (int d = 0; d < 3; ++d) li[d] = amr_local(nb, g, d, G[d] >> 1);
  }
  s.idx = g.c(li[2], li[1], li[0]);
  return s;
}

// map a level-L global index through the physical boundaries (flip[d] toggles when a
// reflecting boundary mirrors along d)
__device__ __forceinline__ void map_zone(const AmrGrid &t, const Geo &g, int L, const int Gin[3],
                                         int G[3], bool flip[3]) {
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    flip[d] = false;
    G[d] = d < g.ndim ? amr_map(t, d, Gin[d], amr_zones(t, g, d, L), flip[d]) : 0;
  }
  // the axisymmetric mirror also negates the phi (x3) component of a vector
  // (boundary_conditions.cpp:66-69); x3 is inactive in the geometries that use it
  if (Gin[0] < 0 && t.bc[0] == KM_BC_AXISYMMETRIC) flip[2] = !flip[2];
}

// RestrictAverage (refinement_operations.hpp:52-121) of the children of a zone (lower
// corner idx in fine block fb): weights are the children's cell volumes (equal in a
// Cartesian block, r-dependent in cylindrical geometry); the pairwise sum order is the
// reference's
__device__ __forceinline__ double restrict_cell(const double *q, const km_block_t &fb,
                                                const Geo &g, long idx, bool cyl) {
  const int i0 = (int)(idx % g.E[0]);
  const double w[2] = {cell_volume(fb, g, cyl, i0), cell_volume(fb, g, cyl, i0 + 1)};
  double V[2][2][2], Tm[2][2][2];
#pragma unroll
  for (int ok = 0; ok < 2; ++ok)
#pragma unroll
    for (int oj = 0; oj < 2; ++oj)
#pragma unroll
      for (int oi = 0; oi < 2; ++oi) {
        V[ok][oj][oi] = 0.;
        Tm[ok][oj][oi] = 0.;
        if ((ok == 0 || g.ndim > 2) && (oj == 0 || g.ndim > 1)) {
          V[ok][oj][oi] = w[oi];
          Tm[ok][oj][oi] = w[oi] * q[idx + ok * g.sc[2] + oj * g.sc[1] + oi];
        }
      }
  const double tvol = ((V[0][0][0] + V[0][1][0]) + (V[0][0][1] + V[0][1][1])) +
                      ((V[1][0][0] + V[1][1][0]) + (V[1][0][1] + V[1][1][1]));
  return tvol > 0.0 ? (((Tm[0][0][0] + Tm[0][1][0]) + (Tm[0][0][1] + Tm[0][1][1])) +
                       ((Tm[1][0][0] + Tm[1][1][0]) + (Tm[1][0][1] + Tm[1][1][1]))) /
                          tvol
                    : 0.0;
}

// value of a located zone (coarse leaf zone or restriction of finer children)
__device__ __forceinline__ double value_at(const km_pack_t &u, const Geo &g, const Src &s, int v,
                                           bool cyl) {
  const double *q = cell_ptr(u, s.leaf, v);
  return s.kind == 1 ? restrict_cell(q, u.blocks[s.leaf], g, s.idx, cyl) : q[s.idx];
}

// axis a vector component points along (momentum, velocity, cell-centred B), else -1
__device__ __forceinline__ int vector_axis(int v, int nc) {
  if (v >= KM_MOM0 && v < KM_MOM0 + 3) return v - KM_MOM0;
  if (v >= nc && v < nc + 3) return v - nc;
  if (nc == 8 && v >= KM_MAGC0 && v < KM_MAGC0 + 3) return v - KM_MAGC0;
  return -1;
}

__device__ __forceinline__ double sgn(double x) { return __builtin_copysign(1.0, x); }

// GradMinMod (refinement_operations.hpp:42-48)
__device__ __forceinline__ double grad_minmod(double fc, double fm, double fp, double dxm,
                                              double dxp) {
  const double gxm = (fc - fm) / dxm;
  const double gxp = (fp - fc) / dxp;
  return 0.5 * (sgn(gxm) + sgn(gxp)) * kmin(kabs(gxm), kabs(gxp));
}

// the child formulas of ProlongateSharedGeneral (refinement_operations.hpp:241-259): the
// fine element at offsets (oi, oj, ok) in its coarse element, from the coarse value and
// the slope terms a = gradient x half the fine spacing along each included axis
__device__ __forceinline__ double child_value(double fc, double a, double bb, double c3, int oi,
                                              int oj, int ok) {
  if (!ok && !oj && !oi) return fc - (a + bb + c3);
  if (!ok && !oj) return fc + (a - bb - c3);
  if (!ok && !oi) return fc - (a - bb + c3);
  if (!ok) return fc + (a + bb - c3);
  if (!oj && !oi) return fc - (a + bb - c3);
  if (!oj) return fc + (a - bb + c3);
  if (!oi) return fc - (a - bb - c3);
  return fc + (a + bb + c3);
}

}  // namespace

// Zones a level-L or a level-L+1 leaf covers: copy or RestrictAverage. Zones a coarser
// leaf covers are left to k_amr_prolong.
// variable k of the range [v0, v1) with the slots [s0, s1) (v0 <= s0 <= s1 <= v1) and the
// slot s2 (s1 <= s2 < v1; -1: none) left out
__device__ __forceinline__ int fill_var(int k, int v0, int s0, int s1, int s2 = -1) {
  int v = v0 + k;
  if (v >= s0) v += s1 - s0;
  if (s2 >= 0 && v >= s2) ++v;
  return v;
}
// how many variables that leaves
__host__ __device__ __forceinline__ int fill_count(int v0, int v1, int s0, int s1, int s2) {
  return (v1 - v0) - (s1 - s0) - (s2 >= 0 ? 1 : 0);
}

// Ghost zone r of block b (the face slabs of depth > 0, else the whole layer): its cell,
// the zone that covers it once mapped through the physical boundaries, and the flips
struct GhostSrc {
  long dc;
  Src s;
  bool fl[3];
};
__device__ __forceinline__ GhostSrc ghost_source(const km_pack_t &u, const AmrGrid &t,
                                                 const Geo &g, int b, long r, int depth) {
  int idx[3];
  if (depth > 0)
    face_ghost_cell(g, depth, r, idx);
  else
    ghost_cell(g, r, idx);
  const km_block_t &blk = u.blocks[b];
  const int L = blk.level;
  int G0[3], G[3];
  GhostSrc o;
#pragma unroll
  for (int d = 0; d < 3; ++d) G0[d] = d < g.ndim ? blk.lx[d] * g.nx[d] + (idx[d] - g.lo[d]) : 0;
  map_zone(t, g, L, G0, G, o.fl);
  o.s = locate(u, t, g, L, G);
  o.dc = g.c(idx[2], idx[1], idx[0]);
  return o;
}

// Copies or restrictions of the variables [v0, v1) without [s0, s1) into zone dc of
// block b from zone sidx of leaf sl (restr: the lower corner of the children).
// Variables go in chunks: every load of a chunk is issued before its stores, so the loads
// of one chunk overlap instead of waiting on the previous variable's store.
__device__ __forceinline__ void ghost_store(const km_pack_t &u, const Geo &g, int b, long dc,
                                            int sl, long sidx, bool restr, const bool fl[3],
                                            int v0, int v1, int nc, int s0, int s1, bool cyl,
                                            int s2 = -1) {
  constexpr int kChunk = 4;
  const km_block_t &sb = u.blocks[sl];
  const int nv = fill_count(v0, v1, s0, s1, s2);
  for (int k = 0; k < nv; k += kChunk) {
    double x[kChunk];
#pragma unroll
    for (int q = 0; q < kChunk; ++q) {
      x[q] = 0.0;
      if (k + q < nv) {
        const double *src = cell_ptr(u, sl, fill_var(k + q, v0, s0, s1, s2));
        x[q] = restr ? restrict_cell(src, sb, g, sidx, cyl) : src[sidx];
      }
    }
#pragma unroll
    for (int q = 0; q < kChunk; ++q) {
      if (k + q < nv) {
        const int v = fill_var(k + q, v0, s0, s1, s2);
        const int ax = vector_axis(v, nc);
        cell_ptr(u, b, v)[dc] = (ax >= 0 && fl[ax]) ? -x[q] : x[q];
      }
    }
  }
}

// depth > 0: only the face slabs of that depth (face_ghost_cell), else every ghost zone
__global__ void __launch_bounds__(256) k_amr_ghost(km_pack_t u, AmrGrid t, int v0, int v1,
                                                   int nc, int s0, int s1, int depth) {
  const Geo g(u);
  const long n = depth > 0 ? face_ghost_count(g, depth) : ghost_count(g);
  const long tt = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (tt >= n * u.nblocks) return;
  const int b = (int)(tt / n);
  const GhostSrc z = ghost_source(u, t, g, b, tt % n, depth);
  if (z.s.kind == 2) return;
  ghost_store(u, g, b, z.dc, z.s.leaf, z.s.idx, z.s.kind == 1, z.fl, v0, v1, nc, s0, s1,
              t.cyl != 0);
}

// The gather list of a tree: ghost_source of every zone k_amr_ghost enumerates, stored
// once per tree (zones a coarser leaf covers get sb = -1)
__global__ void __launch_bounds__(256) k_amr_ghost_list(km_pack_t u, AmrGrid t, int depth,
                                                        AmrGhostEntry *out) {
  const Geo g(u);
  const long n = depth > 0 ? face_ghost_count(g, depth) : ghost_count(g);
  const long tt = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (tt >= n * u.nblocks) return;
  const int b = (int)(tt / n);
  const GhostSrc z = ghost_source(u, t, g, b, tt % n, depth);
  AmrGhostEntry e;
  e.db = b;
  e.dc = (int)z.dc;
  e.sb = z.s.kind == 2 ? -1 : z.s.leaf;
  e.src = (unsigned)z.s.idx | (z.s.kind == 1 ? 1u << 28 : 0u) | (z.fl[0] ? 1u << 29 : 0u) |
          (z.fl[1] ? 1u << 30 : 0u) | (z.fl[2] ? 1u << 31 : 0u);
  out[tt] = e;
}

// k_amr_ghost from the gather list: one entry per lane, no tree walk
__global__ void __launch_bounds__(256) k_amr_ghost_gather(km_pack_t u, const AmrGhostEntry *list,
                                                          long n, int v0, int v1, int nc, int s0,
                                                          int s1, int cyl, int s2) {
  const long tt = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (tt >= n) return;
  const AmrGhostEntry e = list[tt];
  if (e.sb < 0) return;
  const Geo g(u);
  const bool fl[3] = {((e.src >> 29) & 1u) != 0, ((e.src >> 30) & 1u) != 0,
                      ((e.src >> 31) & 1u) != 0};
  ghost_store(u, g, e.db, e.dc, e.sb, (long)(e.src & 0x0fffffffu), ((e.src >> 28) & 1u) != 0,
              fl, v0, v1, nc, s0, s1, cyl != 0, s2);
}

// Zones a level-(L-1) leaf covers: ProlongateSharedMinMod. One thread takes one coarse
// ghost zone of a block, i.e. the up to 2^d fine ghost zones that share a coarse parent
// C: C's value and limited slopes are formed once and each child applies its own
// offsets. The block's ghost layer seen at half resolution has nx/2 interior zones and
// ceil(ng/2) ghosts per side; the physical boundary maps (periodic shift by an even
// count, outflow clamp, reflect mirror) send both children of an unmapped coarse zone
// to children of one mapped C, and a level-(L-1) leaf covers either both or neither.
// depth > 0: the coarse zones of the face slabs of depth ceil(depth / 2) at half
// resolution, whose children cover the fine face slabs of that depth.
// plist (np entries): the parents to fill, as b * n + r; flag: write whether each parent of
// the enumeration is filled here (a coarser leaf covers it) and return.
__global__ void __launch_bounds__(256) k_amr_prolong(km_pack_t u, AmrGrid t, int v0, int v1,
                                                     int nc, int s0, int s1, int depth,
                                                     const long *plist, long np,
                                                     unsigned char *flag, int s2) {
  const Geo g(u);
  km_pack_t half = u;
  half.ng = (u.ng + 1) / 2;
#pragma unroll
  for (int d = 0; d < 3; ++d) half.nx[d] = d < u.ndim ? u.nx[d] / 2 : 1;
  const Geo gc(half);
  const int hc = (depth + 1) / 2;
  const long n = depth > 0 ? face_ghost_count(gc, hc) : ghost_count(gc);
  const long t0 = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t0 >= (plist ? np : n * u.nblocks)) return;
  const long tt = plist ? plist[t0] : t0;
  if (flag) flag[tt] = 0;
  const int b = (int)(tt / n);
  int cz[3];
  if (depth > 0)
    face_ghost_cell(gc, hc, tt % n, cz);
  else
    ghost_cell(gc, tt % n, cz);
  const km_block_t &blk = u.blocks[b];
  const int L = blk.level;
  if (L == 0) return;
  const int nch = 1 << g.ndim;
  // children: padded fine index, their mapped parities and flips (bit 3 c + d of cfl: child
  // c's physical boundary map flips axis d; bit masks rather than bool arrays, because
  // the vector axis that reads them is a runtime value and an array indexed by one would
  // live in scratch memory)
  long dc[8];
  int par[8];
  bool valid[8];
  unsigned cfl = 0u;
  int first = -1;
  int G[3];
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    valid[c] = false;
    dc[c] = 0;
    par[c] = 0;
    if (c >= nch) continue;
    int fi[3], G0[3], Gc[3];
    bool in = true;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      fi[d] = d < g.ndim ? g.ng + 2 * (cz[d] - gc.ng) + ((c >> d) & 1) : 0;
      if (fi[d] < 0 || fi[d] >= g.E[d]) in = false;
      G0[d] = d < g.ndim ? blk.lx[d] * g.nx[d] + (fi[d] - g.lo[d]) : 0;
    }
    if (!in) continue;
    valid[c] = true;
    dc[c] = g.c(fi[2], fi[1], fi[0]);
    bool f3[3];
    map_zone(t, g, L, G0, Gc, f3);
    cfl |= (unsigned)((f3[0] ? 1 : 0) | (f3[1] ? 2 : 0) | (f3[2] ? 4 : 0)) << (3 * c);
    par[c] = (Gc[0] & 1) | (g.ndim > 1 ? (Gc[1] & 1) << 1 : 0) | (g.ndim > 2 ? (Gc[2] & 1) << 2 : 0);
    if (first < 0) {
      first = c;
#pragma unroll
      for (int d = 0; d < 3; ++d) G[d] = Gc[d];
    }
  }
  if (first < 0) return;
  const Src s = locate(u, t, g, L, G);
  if (s.kind != 2) return;
  if (flag) {
    flag[tt] = 1;
    return;
  }
  // the coarse zone C = G/2 and its neighbours at level L-1
  int C[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) C[d] = d < g.ndim ? (G[d] >> 1) : 0;
  Src nbm[3], nbp[3];
  unsigned flm = 0u, flp = 0u;  // bit 3 d + e: the neighbour along -d / +d flips axis e
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (d < g.ndim) {
      int Cm[3] = {C[0], C[1], C[2]}, Cp[3] = {C[0], C[1], C[2]}, M[3];
      Cm[d] -= 1;
      Cp[d] += 1;
      bool f3[3];
      map_zone(t, g, L - 1, Cm, M, f3);
      flm |= (unsigned)((f3[0] ? 1 : 0) | (f3[1] ? 2 : 0) | (f3[2] ? 4 : 0)) << (3 * d);
      nbm[d] = locate(u, t, g, L - 1, M);
      map_zone(t, g, L - 1, Cp, M, f3);
      flp |= (unsigned)((f3[0] ? 1 : 0) | (f3[1] ? 2 : 0) | (f3[2] ? 4 : 0)) << (3 * d);
      nbp[d] = locate(u, t, g, L - 1, M);
    }
  }
  double df[3] = {0., 0., 0.};
#pragma unroll
  for (int d = 0; d < 3; ++d)
    if (d < g.ndim) df[d] = 0.5 * blk.dx[d];
  const int nv = fill_count(v0, v1, s0, s1, s2);
  for (int k = 0; k < nv; ++k) {
    const int v = fill_var(k, v0, s0, s1, s2);
    const int ax = vector_axis(v, nc);
    const double fc = cell_ptr(u, s.leaf, v)[s.idx];
    double gr[3] = {0., 0., 0.};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (d < g.ndim) {
        double fm = value_at(u, g, nbm[d], v, t.cyl != 0);
        double fp = value_at(u, g, nbp[d], v, t.cyl != 0);
        if (ax >= 0 && ((flm >> (3 * d + ax)) & 1u)) fm = -fm;
        if (ax >= 0 && ((flp >> (3 * d + ax)) & 1u)) fp = -fp;
        const double dxc = 2.0 * blk.dx[d];
        gr[d] = grad_minmod(fc, fm, fp, dxc, dxc);
      }
    }
    // the child formulas of ProlongateSharedGeneral (refinement_operations.hpp:241-259)
    // with equal minus/plus slopes and offsets
    const double a = gr[0] * df[0], bb = gr[1] * df[1], cc = gr[2] * df[2];
    double *dst = cell_ptr(u, b, v);
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      if (!valid[c]) continue;
      const int oi = par[c] & 1, oj = (par[c] >> 1) & 1, ok = (par[c] >> 2) & 1;
      double y;
      if (!ok && !oj && !oi) y = fc - (a + bb + cc);
      else if (!ok && !oj) y = fc + (a - bb - cc);
      else if (!ok && !oi) y = fc - (a - bb + cc);
      else if (!ok) y = fc + (a + bb - cc);
      else if (!oj && !oi) y = fc - (a + bb - cc);
      else if (!oj) y = fc + (a - bb + cc);
      else if (!oi) y = fc - (a - bb - cc);
      else y = fc + (a + bb + cc);
      dst[dc[c]] = (ax >= 0 && ((cfl >> (3 * c + ax)) & 1u)) ? -y : y;
    }
  }
}

// Flux correction: a face a block shares with finer leaves takes the area-weighted
// RestrictAverage of the fine face fluxes (face elements average over their two
// transverse axes).
__global__ void __launch_bounds__(256) k_amr_flux_correct(km_pack_t u, AmrGrid t, int nc,
                                                          int ntmax) {
  const Geo g(u);
  const long tt = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long per_block = (long)g.ndim * 2 * ntmax;
  if (tt >= per_block * u.nblocks) return;
  const int b = (int)(tt / per_block);
  long r = tt % per_block;
  const int dir = (int)(r / (2L * ntmax));
  r %= 2L * ntmax;
  const int side = (int)(r / ntmax);
  const int f = (int)(r % ntmax);
  // transverse extents of this face
  int ta = -1, tb = -1;  // the (up to two) transverse active axes
#pragma unroll
  for (int d = 0; d < 3; ++d)
    if (d != dir && d < g.ndim) {
      if (ta < 0) ta = d;
      else tb = d;
    }
  const int na = ta >= 0 ? g.nx[ta] : 1, nbt = tb >= 0 ? g.nx[tb] : 1;
  if (f >= na * nbt) return;
  int li[3] = {0, 0, 0};  // interior offsets
  if (ta >= 0) li[ta] = f % na;
  if (tb >= 0) li[tb] = f / na;
  li[dir] = side ? g.nx[dir] - 1 : 0;
  const km_block_t &blk = u.blocks[b];
  const int L = blk.level;
  int Go[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) Go[d] = d < g.ndim ? blk.lx[d] * g.nx[d] + li[d] : 0;
  Go[dir] += side ? 1 : -1;
  const int N = amr_zones(t, g, dir, L);
  if (Go[dir] < 0 || Go[dir] >= N) {
    if (t.bc[2 * dir + side] != KM_BC_PERIODIC) return;
    Go[dir] = (Go[dir] + N) % N;
  }
  const int fl = amr_leaf(t, g, L, Go);
  const km_block_t &fb = u.blocks[fl];
  if (fb.level <= L) return;
  // the fine faces: children of the neighbour zone on the face this block touches
  int fi[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) fi[d] = amr_local(fb, g, d, 2 * Go[d]);
  fi[dir] = side ? g.lo[dir] : g.hi[dir] + 1;
  const long fidx = g.c(fi[2], fi[1], fi[0]);
  int ci[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) ci[d] = li[d] + g.lo[d];
  ci[dir] = side ? g.hi[dir] + 1 : g.lo[dir];
  const long cidx = g.c(ci[2], ci[1], ci[0]);
  const bool inc[3] = {dir != 0, dir != 1 && g.ndim > 1, dir != 2 && g.ndim > 2};
  // fine face areas (r-dependent for the r and z faces in cylindrical geometry)
  const double A[2] = {face_area(fb, g, t.cyl != 0, dir, fi[0]),
                       face_area(fb, g, t.cyl != 0, dir, fi[0] + (dir != 0))};
  for (int v = 0; v < nc; ++v) {
    const double *F = flux_ptr(u, dir, fl, v);
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
            V[ok][oj][oi] = A[oi];
            Tm[ok][oj][oi] = A[oi] * F[fidx + ok * g.sc[2] + oj * g.sc[1] + oi];
          }
        }
    const double tvol = ((V[0][0][0] + V[0][1][0]) + (V[0][0][1] + V[0][1][1])) +
                        ((V[1][0][0] + V[1][1][0]) + (V[1][0][1] + V[1][1][1]));
    flux_ptr(u, dir, b, v)[cidx] =
        tvol > 0.0 ? (((Tm[0][0][0] + Tm[0][1][0]) + (Tm[0][0][1] + Tm[0][1][1])) +
                      ((Tm[1][0][0] + Tm[1][1][0]) + (Tm[1][0][1] + Tm[1][1][1]))) /
                         tvol
                   : 0.0;
  }
}

// ---------------------------------------------------------------------------------
// Face-centred B (CT) across levels. A ghost face of a level-L block takes its value from
// the leaf holding the cell beside it on the block's side (the cell above a face below
// the block, the cell below a face above it): a same-level copy, RestrictAverage of the
// finer leaf's faces over the two transverse axes (refinement_operations.hpp:52-121 on
// face elements), or, under a coarser leaf, ProlongateSharedMinMod (:123-268) for the
// fine faces lying on coarse faces and ProlongateInternalTothAndRoe (:270-398) for the
// fine faces inside coarse cells. Shared faces are filled first (all components), then
// the internal ones from them and from the block's own faces. A block's own faces are
// never written.
// ---------------------------------------------------------------------------------
namespace {

struct Face {
  int kind;  // 0 copy, 1 restrict, 2 coarser leaf
  int leaf;
  long idx;  // staggered index (kind 1: lower corner of the fine faces)
};

// ghost face of block `blk` at local staggered index idx along face dim d: mapped level-L
// global indices G (face index along d, cell indices across) and flips
__device__ __forceinline__ void face_global(const AmrGrid &t, const Geo &g, const km_block_t &blk,
                                            int d, const int idx[3], int G[3], bool fl[3]) {
  const int L = blk.level;
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    fl[e] = false;
    if (e >= g.ndim) {
      G[e] = 0;
      continue;
    }
    const int G0 = blk.lx[e] * g.nx[e] + (idx[e] - g.lo[e]);
    if (e == d)
      G[e] = amr_map_face(t, e, G0, amr_zones(t, g, e, L), fl[e]);
    else
      G[e] = amr_map(t, e, G0, amr_zones(t, g, e, L), fl[e]);
  }
}

// The leaf holding face G (level L) seen from side `up` along d (up: the cell above the
// face, else the cell below), and where the face sits in it.
__device__ __forceinline__ Face locate_face(const km_pack_t &u, const AmrGrid &t, const Geo &g,
                                            int L, int d, const int G[3], bool up) {
  Face f;
  int Gc[3] = {G[0], G[1], G[2]};
  if (d < g.ndim) {
    const int N = amr_zones(t, g, d, L);
    int c = up ? G[d] : G[d] - 1;
    c = c < 0 ? 0 : (c > N - 1 ? N - 1 : c);
    Gc[d] = c;
  }
  f.leaf = amr_leaf(t, g, L, Gc);
  const km_block_t &nb = u.blocks[f.leaf];
  int li[3];
  if (nb.level == L) {
    f.kind = 0;
#pragma unroll
    for (int e = 0; e < 3; ++e)
      li[e] = e == d ? amr_local_face(nb, g, e, G[e]) : amr_local(nb, g, e, G[e]);
  } else if (nb.level > L) {
    f.kind = 1;
#pragma unroll
    for (int e = 0; e < 3; ++e)
      li[e] = e == d ? amr_local_face(nb, g, e, 2 * G[e]) : amr_local(nb, g, e, 2 * G[e]);
  } else {
    f.kind = 2;
#pragma unroll
    for (int e = 0; e < 3; ++e)
      li[e] = e == d ? amr_local_face(nb, g, e, G[e] >> 1) : amr_local(nb, g, e, G[e] >> 1);
  }
  f.idx = g.s(li[2], li[1], li[0]);
  return f;
}

// RestrictAverage of the fine faces of component d behind a located face (kind 1): the
// transverse active axes are averaged with the fine face areas as weights
__device__ __forceinline__ double face_value(const km_pack_t &u, const Geo &g, int d,
                                             const Face &f, bool cyl) {
  const double *q = u.face[d][f.leaf];
  if (f.kind != 1) return q[f.idx];
  const bool inc[3] = {d != 0, d != 1 && g.ndim > 1, d != 2 && g.ndim > 2};
  // fine face areas: r-dependent for the r and z faces in cylindrical geometry
  const km_block_t &fb = u.blocks[f.leaf];
  const int i0 = (int)(f.idx % g.S[0]);
  const double A[2] = {face_area(fb, g, cyl, d, i0), face_area(fb, g, cyl, d, i0 + (d != 0))};
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
          V[ok][oj][oi] = A[oi];
          Tm[ok][oj][oi] = A[oi] * q[f.idx + ok * g.ss[2] + oj * g.ss[1] + oi];
        }
      }
  const double tvol = ((V[0][0][0] + V[0][1][0]) + (V[0][0][1] + V[0][1][1])) +
                      ((V[1][0][0] + V[1][1][0]) + (V[1][0][1] + V[1][1][1]));
  return tvol > 0.0 ? (((Tm[0][0][0] + Tm[0][1][0]) + (Tm[0][0][1] + Tm[0][1][1])) +
                       ((Tm[1][0][0] + Tm[1][1][0]) + (Tm[1][0][1] + Tm[1][1][1]))) /
                          tvol
                    : 0.0;
}

// a ghost face element of face dim d (false for own faces and for padding)
__device__ __forceinline__ bool ghost_face(const Geo &g, int d, const int idx[3]) {
  bool ghost = false;
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    const int ext = e < g.ndim ? g.E[e] + (e == d ? 1 : 0) : 1;
    if (idx[e] >= ext) return false;
    const int hi = g.hi[e] + (e == d ? 1 : 0);
    if (e < g.ndim && (idx[e] < g.lo[e] || idx[e] > hi)) ghost = true;
  }
  return ghost;
}

// which side of a ghost face faces the block: the cell above for faces below the block
__device__ __forceinline__ bool face_up(const Geo &g, int d, const int idx[3]) {
  return d >= g.ndim || idx[d] <= g.hi[d];
}

__device__ __forceinline__ bool decode_stag(const Geo &g, int nblocks, int &b, int idx[3]) {
  const long n = g.nstag();
  const long tt = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (tt >= n * nblocks) return false;
  b = (int)(tt / n);
  long r = tt % n;
  idx[0] = (int)(r % g.S[0]);
  r /= g.S[0];
  idx[1] = (int)(r % g.S[1]);
  idx[2] = (int)(r / g.S[1]);
  return true;
}

}  // namespace

// phase 1: copies, restrictions and shared-face prolongation
__global__ void __launch_bounds__(256) k_amr_faces_shared(km_pack_t u, AmrGrid t, int d) {
  const Geo g(u);
  int b, idx[3];
  if (!decode_stag(g, u.nblocks, b, idx)) return;
  if (!ghost_face(g, d, idx)) return;
  const km_block_t &blk = u.blocks[b];
  const int L = blk.level;
  int G[3];
  bool fl[3];
  face_global(t, g, blk, d, idx, G, fl);
  const bool up = face_up(g, d, idx);
  const Face f = locate_face(u, t, g, L, d, G, up);
  const long dst = g.s(idx[2], idx[1], idx[0]);
  if (f.kind != 2) {
    double x = face_value(u, g, d, f, t.cyl != 0);
    if (d < g.ndim && fl[d]) x = -x;
    u.face[d][b][dst] = x;
    return;
  }
  if (d < g.ndim && (G[d] & 1)) return;  // inside a coarse cell: phase 2
  // shared fine face on coarse face (G[d]/2, C): transverse minmod slopes at level L-1.
  // Cylindrical r and z faces prolongate r B (SCALE_BY_R, refinement_operations.hpp:
  // 165-173,228-233): each coarse value is scaled by its element's radius (logical
  // index, negative past the axis) and the child divided by its own radius + DBL_EPSILON.
  const bool sr = t.cyl && d < 2;
  const double dxc0 = 2.0 * blk.dx[0];
  auto rc = [&](int C0) { return sr ? amr_radius(t.x1min, C0, d == 1, dxc0) : 1.0; };
  const double fc = u.face[d][f.leaf][f.idx] * rc(G[0] >> 1);
  double a[3] = {0., 0., 0.};
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    if (e == d || e >= g.ndim) continue;
    double fv[2];
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      int C[3], M[3];
      bool nf[3] = {false, false, false};
#pragma unroll
      for (int q = 0; q < 3; ++q) C[q] = q < g.ndim ? G[q] >> 1 : 0;
      C[e] += side ? 1 : -1;
#pragma unroll
      for (int q = 0; q < 3; ++q)
        M[q] = (q == e) ? amr_map(t, q, C[q], amr_zones(t, g, q, L - 1), nf[q]) : C[q];
      const Face s = locate_face(u, t, g, L - 1, d, M, up);
      fv[side] = face_value(u, g, d, s, t.cyl != 0) * rc(C[0]);
    }
    const double dxc = 2.0 * blk.dx[e];
    a[e] = grad_minmod(fc, fv[0], fv[1], dxc, dxc) * (0.5 * blk.dx[e]);
  }
  const int oi = d != 0 ? (G[0] & 1) : 0;
  const int oj = (d != 1 && g.ndim > 1) ? (G[1] & 1) : 0;
  const int ok = (d != 2 && g.ndim > 2) ? (G[2] & 1) : 0;
  double x = child_value(fc, a[0], a[1], a[2], oi, oj, ok);
  if (sr) x *= 1.0 / (amr_radius(t.x1min, G[0], d == 1, blk.dx[0]) + 2.220446049250313e-16);
  if (d < g.ndim && fl[d]) x = -x;
  u.face[d][b][dst] = x;
}

namespace {
// ProlongateInternalTothAndRoe (refinement_operations.hpp:278-398), Cartesian: the value of
// the fine face of component d at idx, inside a coarse cell of block b whose shared fine
// faces (all three components) are already in place. The fine element (kk, jj, ii) of the
// reference's loop is reached through its permuted offsets; the coarse cell spacings
// enter unpermuted, as the reference writes them.
__device__ double toth_roe(const km_pack_t &u, const Geo &g, int b, int d, const int idx[3],
                           const km_block_t &blk, bool cyl, double x1min) {
  const int g2 = g.ndim > 1, g3 = g.ndim > 2;
  // fine lower corner of the coarse cell
  int fc[3];
#pragma unroll
  for (int e = 0; e < 3; ++e) fc[e] = e < g.ndim ? g.lo[e] + ((idx[e] - g.lo[e]) & ~1) : 0;
  const int fi = fc[0], fj = fc[1], fk = fc[2];
  // permuted element index of (ok, oj, oi) for component d
  auto kji = [&](int ok, int oj, int oi, int out[3]) {
    if (d == 0) {
      out[0] = fk + ok * g3, out[1] = fj + oj * g2, out[2] = fi + oi;
    } else if (d == 1) {
      out[0] = fk + oj * g3, out[1] = fj + oi * g2, out[2] = fi + ok;
    } else {
      out[0] = fk + oi * g3, out[1] = fj + ok * g2, out[2] = fi + oj;
    }
  };
  // cylindrical: the divergence constraint holds for r B, so every fine face enters
  // multiplied by its radius (an r face's own, a z face's cell-centre radius, 1 for the
  // phi faces of a 1-D or 2-D mesh) and the result is divided by the target's radius
  auto radius = [&](int comp, int ii) {
    if (!cyl || comp == 2) return 1.0;
    return amr_radius(x1min, blk.lx[0] * g.nx[0] + (ii - g.lo[0]), comp == 1, blk.dx[0]);
  };
  auto fine = [&](int eidx, int ok, int oj, int oi) {
    int p[3];
    kji(ok, oj, oi, p);
    const int comp = (d + eidx) % 3;
    return u.face[comp][b][g.s(p[0], p[1], p[2])] * radius(comp, p[2]);
  };
  // this element's (ok, oj) in the reference's write loop
  int ok = 0, oj = 0;
  {
    const int o[3] = {idx[0] - fi, idx[1] - fj, idx[2] - fk};
    if (d == 0) {
      ok = o[2], oj = o[1];
    } else if (d == 1) {
      oj = o[2], ok = o[0];
    } else {
      ok = o[1], oj = o[0];
    }
  }
  auto sg = [](int off) { return off == 0 ? -1.0 : 1.0; };
  double Uxx = 0.0, Vxyz = 0.0, Wxyz = 0.0;
  for (int v = 0; v < 2; ++v)
    for (int uu = 0; uu <= 2; uu += 2)
      for (int tt = 0; tt < 2; ++tt) {
        const double fine2 = fine(1, v, uu, tt) * 1.0;
        const double fine3 = fine(2, uu, v, tt) * 1.0;
        Uxx += sg(tt) * sg(uu) * (fine2 + fine3);
        Vxyz += sg(tt) * sg(uu) * sg(v) * fine2;
        Wxyz += sg(tt) * sg(uu) * sg(v) * fine3;
      }
  Uxx *= 0.125;
  // coarse spacings; the cylindrical Dx along phi is 2 pi (geometry.hpp:35-43)
  const double cdx = 2.0 * blk.dx[0], cdy = 2.0 * blk.dx[1];
  const double cdz = cyl ? kTwoPi : 2.0 * blk.dx[2];
  const double dx2 = cdx * cdx, dy2 = cdy * cdy, dz2 = cdz * cdz;
  Vxyz *= 0.125 * dz2 / (dx2 + dz2);
  Wxyz *= 0.125 * dy2 / (dx2 + dy2);
  double x = 0.5 * (fine(0, ok, oj, 0) * 1.0 + fine(0, ok, oj, 2) * 1.0) + Uxx + sg(ok) * Vxyz +
             sg(oj) * Wxyz;
  x *= cyl ? 1.0 / (radius(d, idx[0]) + 2.220446049250313e-16) : 1.0;
  return x;
}
}  // namespace

// phase 2: ProlongateInternalTothAndRoe for the ghost fine faces of component d inside a
// coarse cell under a coarser leaf
__global__ void __launch_bounds__(256) k_amr_faces_internal(km_pack_t u, AmrGrid t, int d) {
  const Geo g(u);
  int b, idx[3];
  if (!decode_stag(g, u.nblocks, b, idx)) return;
  if (!ghost_face(g, d, idx)) return;
  if (d < g.ndim && !((idx[d] - g.lo[d]) & 1)) return;  // shared position
  const km_block_t &blk = u.blocks[b];
  int G[3];
  bool fl[3];
  face_global(t, g, blk, d, idx, G, fl);
  const Face f = locate_face(u, t, g, blk.level, d, G, face_up(g, d, idx));
  if (f.kind != 2) return;
  u.face[d][b][g.s(idx[2], idx[1], idx[0])] = toth_roe(u, g, b, d, idx, blk, t.cyl != 0, t.x1min);
}

// EMF correction at level jumps: every edge on a block's boundary that a finer leaf also
// has on its boundary takes the length-weighted RestrictAverage of the fine edges on it
// (edge elements average along their own axis). The finer leaf may sit across a face, an
// edge or a corner of the block; blocks that share the edge then all carry the same value,
// which keeps their shared faces equal. 2:1 balance keeps the fine edges read here out of
// the fine leaves' own corrections.
__global__ void __launch_bounds__(256) k_amr_emf_correct(km_pack_t u, AmrGrid t, int te) {
  const Geo g(u);
  int b, idx[3];
  if (!decode_stag(g, u.nblocks, b, idx)) return;
  double val;
  if (amr_emf_edge(u, t, te, b, idx, val)) u.edge[te][b][g.s(idx[2], idx[1], idx[0])] = val;
}

void launch_amr_ghost_faces(const AmrGrid &t, const km_pack_t &u, hipStream_t st) {
  const Geo g(u);
  const long n = g.nstag() * u.nblocks;
  if (n == 0) return;
  const dim3 grid((unsigned)((n + 255) / 256));
  for (int d = 0; d < 3; ++d) hipLaunchKernelGGL(k_amr_faces_shared, grid, dim3(256), 0, st, u, t, d);
  for (int d = 0; d < 3; ++d)
    hipLaunchKernelGGL(k_amr_faces_internal, grid, dim3(256), 0, st, u, t, d);
}

void launch_amr_emf_correct(const AmrGrid &t, const km_pack_t &u, hipStream_t st) {
  const Geo g(u);
  if (t.lmax == 0 || u.nblocks == 0 || g.ndim < 2) return;
  const long n = g.nstag() * u.nblocks;
  const dim3 grid((unsigned)((n + 255) / 256));
  for (int te = g.ndim < 3 ? 2 : 0; te < 3; ++te)
    hipLaunchKernelGGL(k_amr_emf_correct, grid, dim3(256), 0, st, u, t, te);
}

void launch_amr_ghost_fill(const AmrGrid &t, const km_pack_t &u, int v0, int v1, int nc,
                           hipStream_t st, int s0, int s1, int depth) {
  const Geo g(u);
  if (depth >= g.ng) depth = 0;  // the whole layer
  const long n = (depth > 0 ? face_ghost_count(g, depth) : ghost_count(g)) * u.nblocks;
  if (s0 < v0 || s1 > v1 || s1 <= s0) s0 = s1 = v1;  // nothing left out
  if (n == 0 || v0 >= v1 || (v1 - v0) == (s1 - s0)) return;
  hipLaunchKernelGGL(k_amr_ghost, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, u, t, v0,
                     v1, nc, s0, s1, depth);
  if (t.lmax > 0) {
    km_pack_t half = u;
    half.ng = (u.ng + 1) / 2;
    for (int d = 0; d < 3; ++d) half.nx[d] = d < u.ndim ? u.nx[d] / 2 : 1;
    const Geo gh(half);
    const long nc2 =
        (depth > 0 ? face_ghost_count(gh, (depth + 1) / 2) : ghost_count(gh)) * u.nblocks;
    hipLaunchKernelGGL(k_amr_prolong, dim3((unsigned)((nc2 + 255) / 256)), dim3(256), 0, st, u, t,
                       v0, v1, nc, s0, s1, depth, nullptr, 0L, nullptr, -1);
  }
}

long amr_ghost_list_size(const km_pack_t &u, int depth) {
  const Geo g(u);
  if (depth >= g.ng) depth = 0;
  return (depth > 0 ? face_ghost_count(g, depth) : ghost_count(g)) * u.nblocks;
}

long amr_prolong_parents(const km_pack_t &u, int depth) {
  const Geo g(u);
  if (depth >= g.ng) depth = 0;
  km_pack_t half = u;
  half.ng = (u.ng + 1) / 2;
  for (int d = 0; d < 3; ++d) half.nx[d] = d < u.ndim ? u.nx[d] / 2 : 1;
  const Geo gh(half);
  return (depth > 0 ? face_ghost_count(gh, (depth + 1) / 2) : ghost_count(gh)) * u.nblocks;
}

void launch_amr_ghost_list(const AmrGrid &t, const km_pack_t &u, int depth, AmrGhostEntry *out,
                           hipStream_t st) {
  const Geo g(u);
  if (depth >= g.ng) depth = 0;
  const long n = amr_ghost_list_size(u, depth);
  if (n > 0)
    hipLaunchKernelGGL(k_amr_ghost_list, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, u,
                       t, depth, out);
}

void launch_amr_prolong_flags(const AmrGrid &t, const km_pack_t &u, int depth,
                              unsigned char *flag, hipStream_t st) {
  const Geo g(u);
  if (depth >= g.ng) depth = 0;
  const long n = amr_prolong_parents(u, depth);
  if (n > 0 && t.lmax > 0)
    hipLaunchKernelGGL(k_amr_prolong, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, u, t, 0,
                       0, 0, 0, 0, depth, nullptr, 0L, flag, -1);
}

void launch_amr_ghost_fill_list(const AmrFillList &f, const AmrGrid &t, const km_pack_t &u,
                                int v0, int v1, int nc, hipStream_t st, int s0, int s1, int s2) {
  launch_amr_ghost_fill_list(f, t, u, v0, v1, nc, st, s0, s1, s2, nullptr, nullptr, nullptr);
}

void launch_amr_ghost_fill_list(const AmrFillList &f, const AmrGrid &t, const km_pack_t &u,
                                int v0, int v1, int nc, hipStream_t st, int s0, int s1, int s2,
                                hipStream_t st2, hipEvent_t fork, hipEvent_t join) {
  if (s0 < v0 || s1 > v1 || s1 <= s0) s0 = s1 = v1;  // nothing left out
  if (s2 < s1 || s2 >= v1) s2 = -1;
  if (v0 >= v1 || fill_count(v0, v1, s0, s1, s2) <= 0) return;
  const bool side = st2 && fork && join && f.n > 0 && f.np > 0;
  hipStream_t sp = st;
  if (side) {
    (void)hipEventRecord(fork, st);
    (void)hipStreamWaitEvent(st2, fork, 0);
    sp = st2;
  }
  if (f.np > 0)
    hipLaunchKernelGGL(k_amr_prolong, dim3((unsigned)((f.np + 255) / 256)), dim3(256), 0, sp, u,
                       t, v0, v1, nc, s0, s1, f.depth, f.plist, f.np, nullptr, s2);
  if (f.n > 0)
    hipLaunchKernelGGL(k_amr_ghost_gather, dim3((unsigned)((f.n + 255) / 256)), dim3(256), 0, st,
                       u, f.list, f.n, v0, v1, nc, s0, s1, t.cyl, s2);
  if (side) {
    (void)hipEventRecord(join, st2);
    (void)hipStreamWaitEvent(st, join, 0);
  }
}

void launch_amr_flux_correct(const AmrGrid &t, const km_pack_t &u, int nc, hipStream_t st) {
  const Geo g(u);
  if (t.lmax == 0 || u.nblocks == 0) return;
  int ntmax = 1;
  for (int dir = 0; dir < g.ndim; ++dir) {
    int nt = 1;
    for (int d = 0; d < g.ndim; ++d)
      if (d != dir) nt *= g.nx[d];
    ntmax = nt > ntmax ? nt : ntmax;
  }
  const long n = (long)ntmax * g.ndim * 2 * u.nblocks;
  hipLaunchKernelGGL(k_amr_flux_correct, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, u,
                     t, nc, ntmax);
}

// ---- adaptive refinement ----------------------------------------------------------
// grid::AMRLoehner::operator() (grid_refinement.cpp:61-151), one lane per interior zone:
// first derivatives d_p = 0.5 (u(+e_p) - u(-e_p)) / dx_p are formed where the second
// differences need them (the same values the reference stores in FirstDer), the zone's
// num/den is folded into its block's maximum with a 64-bit atomicMax on the bit pattern
// (non-negative doubles order like their patterns; a NaN ratio is skipped, as Kokkos::max
// skips it), and k_amr_tag turns each block maximum into its tag.
__global__ void __launch_bounds__(256) k_amr_loehner(km_pack_t u, int var, double filter,
                                                     unsigned long long *err2) {
  const Geo g(u);
  const long per = (long)g.nx[0] * (g.ndim > 1 ? g.nx[1] : 1) * (g.ndim > 2 ? g.nx[2] : 1);
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= per * u.nblocks) return;
  const int b = (int)(t / per);
  long r = t % per;
  const int i = g.lo[0] + (int)(r % g.nx[0]);
  r /= g.nx[0];
  const int j = g.lo[1] + (int)(r % (g.ndim > 1 ? g.nx[1] : 1));
  const int k = g.lo[2] + (int)(r / (g.ndim > 1 ? g.nx[1] : 1));
  const km_block_t &blk = u.blocks[b];
  const double *q = cell_ptr(u, b, var);
  const long c = g.c(k, j, i);
  const int nd = g.ndim;
  auto der = [&](int p, long z) { return 0.5 * (q[z + g.sc[p]] - q[z - g.sc[p]]) / blk.dx[p]; };
  double num = 0.0, den = 1.e-12;
  for (int qd = 0; qd < nd; ++qd)
    for (int p = 0; p < nd; ++p) {
      const double dp = der(p, c + g.sc[qd]), dm = der(p, c - g.sc[qd]);
      const double a = 0.5 * (dp - dm) / blk.dx[qd];
      num += a * a;
      const double e = 0.5 * (kabs(dp) + kabs(dm)) / blk.dx[p] +
                       filter *
                           (kabs(q[c + g.sc[qd]]) + kabs(q[c - g.sc[qd]]) +
                            kabs(q[c - g.sc[p]]) + kabs(q[c + g.sc[p]])) /
                           (blk.dx[qd] * blk.dx[p]);
      den += e * e;
    }
  if (den == 0.0) return;
  const double v = num / den;
  if (!(v > 0.0)) return;  // zero adds nothing to a maximum that starts at zero; NaN skipped
  atomicMax(err2 + b, (unsigned long long)__double_as_longlong(v));
}

__global__ void k_amr_tag(const km_block_t *blocks, int nblocks, const unsigned long long *err2,
                          km_amr_criterion_t cr, int *tags, int reset) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nblocks) return;
  const double err = sqrt(__longlong_as_double((long long)err2[b]));
  int flag = KM_AMR_SAME;
  if (err > cr.refine_tol && blocks[b].level < cr.max_level) flag = KM_AMR_REFINE;
  if (err < cr.derefine_tol) flag = KM_AMR_DEREFINE;
  const int cur = reset ? KM_AMR_DEREFINE : tags[b];
  tags[b] = cur > flag ? cur : flag;
}

__global__ void k_zero_u64(unsigned long long *p, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0ull;
}

void launch_amr_tag_loehner(const km_pack_t &u, const km_amr_criterion_t &crit, int *tags,
                            bool reset, double *err2_scratch, hipStream_t st) {
  const Geo g(u);
  auto *e2 = reinterpret_cast<unsigned long long *>(err2_scratch);
  const unsigned nb = (unsigned)((u.nblocks + 255) / 256);
  hipLaunchKernelGGL(k_zero_u64, dim3(nb), dim3(256), 0, st, e2, u.nblocks);
  const long n = (long)g.nx[0] * (g.ndim > 1 ? g.nx[1] : 1) * (g.ndim > 2 ? g.nx[2] : 1) *
                 u.nblocks;
  hipLaunchKernelGGL(k_amr_loehner, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, u,
                     crit.var, crit.filter, e2);
  hipLaunchKernelGGL(k_amr_tag, dim3(nb), dim3(256), 0, st, u.blocks, u.nblocks, e2, crit, tags,
                     reset ? 1 : 0);
}

// RedistributeAndRefineMeshBlocks data movement: one lane per zone of a new block. A
// kept leaf copies the zone; a split leaf's child takes ProlongateSharedMinMod of its
// parent's zone C = G/2 with the parent's C -+ 1 along each axis (interior or ghost
// zones of the parent); a merged parent takes the RestrictAverage of the 2^d zones of
// the child holding 2C. Ghost zones of new leaves are left for the ghost fill.
__global__ void __launch_bounds__(256) k_amr_remesh(km_pack_t o, km_pack_t nw,
                                                    const int32_t *origin, int v0, int v1,
                                                    int cyl) {
  const Geo g(nw);
  const long n = g.ncell();
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= n * nw.nblocks) return;
  const int b = (int)(t / n);
  const long z = t % n;
  int idx[3];
  idx[0] = (int)(z % g.E[0]);
  idx[1] = (int)((z / g.E[0]) % g.E[1]);
  idx[2] = (int)(z / ((long)g.E[0] * g.E[1]));
  const int32_t *org = origin + 9 * (long)b;
  const int kind = org[0];
  if (kind == 0) {
    const int ob = org[1];
    for (int v = v0; v < v1; ++v) cell_ptr(nw, b, v)[z] = cell_ptr(o, ob, v)[z];
    return;
  }
#pragma unroll
  for (int d = 0; d < 3; ++d)
    if (d < g.ndim && (idx[d] < g.lo[d] || idx[d] > g.hi[d])) return;
  const km_block_t &blk = nw.blocks[b];
  if (kind == 1) {
    const int pb = org[1];
    const km_block_t &par = o.blocks[pb];
    int pc[3] = {0, 0, 0}, bit[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (d >= g.ndim) continue;
      const int G = blk.lx[d] * g.nx[d] + (idx[d] - g.lo[d]);
      bit[d] = G & 1;
      pc[d] = (G >> 1) - par.lx[d] * g.nx[d] + g.lo[d];
    }
    const long cc = g.c(pc[2], pc[1], pc[0]);
    double df[3] = {0., 0., 0.};
#pragma unroll
    for (int d = 0; d < 3; ++d)
      if (d < g.ndim) df[d] = 0.5 * blk.dx[d];
    for (int v = v0; v < v1; ++v) {
      const double *q = cell_ptr(o, pb, v);
      const double fc = q[cc];
      double gr[3] = {0., 0., 0.};
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        if (d >= g.ndim) continue;
        const double dxc = 2.0 * blk.dx[d];
        gr[d] = grad_minmod(fc, q[cc - g.sc[d]], q[cc + g.sc[d]], dxc, dxc);
      }
      cell_ptr(nw, b, v)[z] =
          child_value(fc, gr[0] * df[0], gr[1] * df[1], gr[2] * df[2], bit[0], bit[1], bit[2]);
    }
    return;
  }
  int ch = 0, f[3] = {0, 0, 0};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (d >= g.ndim) continue;
    const int lc = idx[d] - g.lo[d];
    const int half = (2 * lc) / g.nx[d];
    ch |= half << d;
    f[d] = 2 * lc - half * g.nx[d] + g.lo[d];
  }
  const int cb = org[1 + ch];
  const long fz = g.c(f[2], f[1], f[0]);
  for (int v = v0; v < v1; ++v)
    cell_ptr(nw, b, v)[z] = restrict_cell(cell_ptr(o, cb, v), o.blocks[cb], g, fz, cyl != 0);
}

// The same data movement for the CT face field, component d, one lane per face element
// of a new block. A kept leaf copies the element (ghost faces included). A split leaf's
// child takes, at every own face that lies on a parent face, ProlongateSharedMinMod of
// that parent face with minmod slopes across it from the parent's neighbouring faces
// (interior or ghost); its faces inside parent cells are left to
// k_amr_remesh_faces_internal. A merged parent takes the area-weighted RestrictAverage of
// the 2^(d-1) child faces on each of its own faces. Ghost faces of new leaves are left to
// the ghost fill.
__global__ void __launch_bounds__(256) k_amr_remesh_faces_shared(km_pack_t o, km_pack_t nw,
                                                                 const int32_t *origin, int d,
                                                                 int cyl, double x1min) {
  const Geo g(nw);
  int b, idx[3];
  if (!decode_stag(g, nw.nblocks, b, idx)) return;
  const int32_t *org = origin + 9 * (long)b;
  const long dst = g.s(idx[2], idx[1], idx[0]);
  if (org[0] == 0) {
    nw.face[d][b][dst] = o.face[d][org[1]][dst];
    return;
  }
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    const int hi = e < g.ndim ? g.hi[e] + (e == d ? 1 : 0) : 0;
    if (idx[e] < g.lo[e] || idx[e] > hi) return;  // ghost face or padding
  }
  const km_block_t &blk = nw.blocks[b];
  if (org[0] == 1) {
    const int pb = org[1];
    const km_block_t &par = o.blocks[pb];
    int G[3] = {0, 0, 0}, pc[3] = {0, 0, 0};
#pragma unroll
    for (int e = 0; e < 3; ++e) {
      if (e >= g.ndim) continue;
      G[e] = blk.lx[e] * g.nx[e] + (idx[e] - g.lo[e]);
      pc[e] = (G[e] >> 1) - par.lx[e] * g.nx[e] + g.lo[e];
    }
    if (d < g.ndim && (G[d] & 1)) return;  // inside a parent cell: phase 2
    const double *q = o.face[d][pb];
    const long pcs = g.s(pc[2], pc[1], pc[0]);
    // cylindrical r and z faces prolongate r B, as in the ghost fill
    const bool sr = cyl && d < 2;
    const double dxc0 = 2.0 * blk.dx[0];
    auto rc = [&](int C0) { return sr ? amr_radius(x1min, C0, d == 1, dxc0) : 1.0; };
    const int C0 = G[0] >> 1;
    const double fc = q[pcs] * rc(C0);
    double a[3] = {0., 0., 0.};
#pragma unroll
    for (int e = 0; e < 3; ++e) {
      if (e == d || e >= g.ndim) continue;
      const double dxc = 2.0 * blk.dx[e];
      const double fm = q[pcs - g.ss[e]] * rc(C0 - (e == 0)), fp = q[pcs + g.ss[e]] * rc(C0 + (e == 0));
      a[e] = grad_minmod(fc, fm, fp, dxc, dxc) * (0.5 * blk.dx[e]);
    }
    const int oi = d != 0 ? (G[0] & 1) : 0;
    const int oj = (d != 1 && g.ndim > 1) ? (G[1] & 1) : 0;
    const int ok = (d != 2 && g.ndim > 2) ? (G[2] & 1) : 0;
    double x = child_value(fc, a[0], a[1], a[2], oi, oj, ok);
    if (sr) x *= 1.0 / (amr_radius(x1min, G[0], d == 1, blk.dx[0]) + 2.220446049250313e-16);
    nw.face[d][b][dst] = x;
    return;
  }
  int ch = 0, f[3] = {0, 0, 0};
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    if (e >= g.ndim) continue;
    const int lc = idx[e] - g.lo[e];
    int half = (2 * lc) / g.nx[e];
    half = half > 1 ? 1 : half;  // the upper face of the block lies in the upper child
    ch |= half << e;
    f[e] = 2 * lc - half * g.nx[e] + g.lo[e];
  }
  Face src;
  src.kind = 1;
  src.leaf = org[1 + ch];
  src.idx = g.s(f[2], f[1], f[0]);
  nw.face[d][b][dst] = face_value(o, g, d, src, cyl != 0);
}

// phase 2 of the face remesh: ProlongateInternalTothAndRoe for the faces of a split leaf's
// child that lie inside parent cells, from the child's shared faces of phase 1
__global__ void __launch_bounds__(256) k_amr_remesh_faces_internal(km_pack_t nw,
                                                                   const int32_t *origin, int d,
                                                                   int cyl, double x1min) {
  const Geo g(nw);
  int b, idx[3];
  if (!decode_stag(g, nw.nblocks, b, idx)) return;
  if (origin[9 * (long)b] != 1 || d >= g.ndim) return;
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    const int hi = e < g.ndim ? g.hi[e] + (e == d ? 1 : 0) : 0;
    if (idx[e] < g.lo[e] || idx[e] > hi) return;
  }
  if (!((idx[d] - g.lo[d]) & 1)) return;  // shared position (block offsets are even)
  nw.face[d][b][g.s(idx[2], idx[1], idx[0])] =
      toth_roe(nw, g, b, d, idx, nw.blocks[b], cyl != 0, x1min);
}

void launch_amr_remesh_faces(const km_pack_t &old_u, const km_pack_t &new_u,
                             const int32_t *origin, bool cyl, double x1min, hipStream_t st) {
  const Geo g(new_u);
  const long n = g.nstag() * new_u.nblocks;
  if (n == 0) return;
  const unsigned grid = (unsigned)((n + 255) / 256);
  for (int d = 0; d < 3; ++d)
    hipLaunchKernelGGL(k_amr_remesh_faces_shared, dim3(grid), dim3(256), 0, st, old_u, new_u,
                       origin, d, cyl ? 1 : 0, x1min);
  for (int d = 0; d < g.ndim; ++d)
    hipLaunchKernelGGL(k_amr_remesh_faces_internal, dim3(grid), dim3(256), 0, st, new_u, origin,
                       d, cyl ? 1 : 0, x1min);
}

void launch_amr_remesh(const km_pack_t &old_u, const km_pack_t &new_u, const int32_t *origin,
                       int v0, int v1, bool cyl, hipStream_t st) {
  const Geo g(new_u);
  const long n = g.ncell() * new_u.nblocks;
  if (n == 0 || v0 >= v1) return;
  hipLaunchKernelGGL(k_amr_remesh, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, old_u,
                     new_u, origin, v0, v1, cyl ? 1 : 0);
}

}  // namespace km
