// km_internal.hpp -- declarations shared by the host runtime and the kernel translation
// units of libkm_hydro.so (not part of the public C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/km_hydro.h"
#include "km_amr.hpp"
#include "km_physics.hpp"

namespace km {

// inclusive cell-index box per dimension (x1, x2, x3)
struct Box {
  int lo[3], hi[3];
};

// block grid of a single-level mesh, blocks stored x-fastest starting at block `first`
struct UniformGrid {
  int nb[3] = {1, 1, 1};
  int bc[6] = {0, 0, 0, 0, 0, 0};
  int first = 0;
};

// Independent instances sharing one block pack (an ensemble): instance of each block and
// the grid of each instance, both device arrays. A single mesh has inst == nullptr and
// uses the grid passed by value.
struct Instances {
  const int *inst = nullptr;
  const UniformGrid *grids = nullptr;
};

inline Eos make_eos(const km_cfg_t &cfg) {
  constexpr double kboltz = 1.380649e-16;  // equation_of_state.hpp:44
  Eos e;
  e.gm1 = cfg.gamma - 1.0;
  e.gp1gm1 = (e.gm1 + 1) * e.gm1;
  e.cv = cfg.gamma * kboltz / cfg.abar;
  return e;
}

// stage arguments of the fused sweeps
struct FusedArgs {
  km_pack_t u0;        // stage input: conserved slots 0..4, ghost zones filled
  km_pack_t base;      // RK base state (conserved)
  km_pack_t u1;        // stage output (conserved, interior written)
  double *du;          // [nblocks][5][nx3][nx2][nx1] accumulator (interior only)
  const double *dt;    // device: step dt per instance
  double *dt_est;      // device: CFL minimum of u1 per instance (bit-ordered), or null
  const int *inst;     // device: instance of each block, or null (one instance)
  const int *nbr;      // [nblocks][3][2] face neighbours: block index, or
                       // KM_NBR_OUTFLOW / KM_NBR_REFLECT at a physical boundary
  double beta;
  int lim;
  Eos eos;
  // block-tree mesh: cells past a block edge are the block's own (filled) ghost zones,
  // primitives are read from the pack, and a face bordering a finer leaf takes the
  // restricted fine flux from the pack's flux arrays
  int amr;
  AmrGrid agrid;
  int *fail;  // device flag word (KM_FAIL_*), or null
  // block trees: the final sweep also writes u1's primitives (PreparePrimitive); only when
  // u1's primitive arrays are not u0's (the kernels read u0's while they write)
  int write_prims;
};

constexpr int KM_NBR_OUTFLOW = -1;
constexpr int KM_NBR_REFLECT = -2;

// task launchers (km_tasks.hip)
void launch_calculate_fluxes(const km_cfg_t &cfg, const km_pack_t &u, const Box &bx,
                             hipStream_t st);
void launch_calculate_emf(const km_cfg_t &cfg, const km_pack_t &u, hipStream_t st);
// species: also the mass fractions' dU/dt (when there is more than one species)
void launch_fluxes_to_dudt(const km_cfg_t &cfg, const km_pack_t &u, const km_pack_t &dudt,
                           hipStream_t st, bool species = true);
// FluxesToDuDt, ApplyDuDt and PreparePrimitive of one stage, with the stage's dt on the
// device (dt_dev, indexed by inst when given)
void launch_update(const km_cfg_t &cfg, const km_pack_t &base, const km_pack_t &u0,
                   const km_pack_t &u1, const km_pack_t &dudt, double beta, hipStream_t st,
                   const double *dt_dev, const int *inst);
// hydro::AddSourceTerms: cylindrical geometric source into dudt (no-op for Cartesian)
void launch_add_source_terms(const km_cfg_t &cfg, const km_pack_t &u, const km_pack_t &dudt,
                             hipStream_t st);
void launch_apply_dudt(const km_cfg_t &cfg, const km_pack_t &base, const km_pack_t &u0,
                       const km_pack_t &u1, const km_pack_t &dudt, double beta, double dt,
                       hipStream_t st, const double *dt_dev = nullptr, const int *inst = nullptr);
void launch_prepare_primitive(const km_cfg_t &cfg, const km_pack_t &u, hipStream_t st);
// EMF, FluxStokes and the face part of ApplyDuDt into u1 in one kernel (Cartesian 3-D CT;
// the EMF from u0's stored transverse B fluxes), with the stage's dt on the device
void launch_ct_faces(const km_pack_t &base, const km_pack_t &u0, const km_pack_t &u1,
                     double beta, hipStream_t st, const double *dt_dev, const int *inst);
void launch_prepare_conserved(const km_cfg_t &cfg, const km_pack_t &u, hipStream_t st);
// fail: device flag word; KM_FAIL_STATE is or-ed in when a zone is non-physical
// part: 1 sets dt_dev to DBL_MAX first, 2 scales the minimum by cfl / ndim after (3: both;
// packs of one mesh reduce into one minimum with 1, 0 ..., 2)
void launch_estimate_dt(const km_cfg_t &cfg, const km_pack_t &u, double *dt_dev, hipStream_t st,
                        const int *inst = nullptr, int ninst = 1, int *fail = nullptr,
                        int part = 3);
void launch_fill_divb(const km_cfg_t &cfg, const km_pack_t &u, hipStream_t st);
// hydro::PrepareConserved per stage (cylindrical CT: md0's B_phi to B_phi / r)
void launch_prepare_conserved_stage(const km_cfg_t &cfg, const km_pack_t &u, hipStream_t st);
// isentropic_vortex parameters of the history reduction (km_history.hip)
struct HistVortex {
  double density, pressure, velx, vely, strength, mhd_strength, gamma;
};
// per-block sums of |exact - q| V of DENS, VEL0..2, PRES (+ MAGC0..2): part[b * nv + v]
void launch_history_vortex(const km_cfg_t &cfg, const km_pack_t &u, const HistVortex &p,
                           double *part, hipStream_t st);
// depth / face_depth in [1, ng): only the ghost cells / faces within that many zones of
// the interior (edges and corners included); <= 0: all of them. skip: a slot of [v0, v1)
// to leave out (-1: none)
void launch_ghost_fill(const UniformGrid &gr, const km_pack_t &u, int v0, int v1, int nc,
                       bool faces, hipStream_t st, Instances in = Instances(), int depth = 0,
                       int face_depth = 0, int skip = -1);

// fused stage sweeps: one translation unit per arithmetic form
//   exact : FP64, -ffp-contract=off (bitwise the reference's operation order)
//   fma   : FP64, -ffp-contract=fast
//   fp32  : FP32 reconstruction/Riemann, FP64 conversion, accumulation and stage combine
namespace exact {
int launch_fused_sweep(const km_cfg_t &cfg, const FusedArgs &a, int ndim, int dir, bool with_dt,
                       int kern, hipStream_t st);
int tile_shapes();
int launch_fine_face_fluxes(const km_cfg_t &cfg, const FusedArgs &a, hipStream_t st);
// the whole stage in one launch (2-D and 3-D, single-level meshes), kern in
// 1..stage3_kernels() on 3-D meshes and 1..stage2_kernels() on 2-D ones
int launch_fused_stage3(const km_cfg_t &cfg, const FusedArgs &a, bool with_dt, int kern,
                        hipStream_t st);
int stage3_kernels();
int stage2_kernels();
// the CT-MHD stage (3-D, Cartesian, single-level meshes; km_fused_mhd.inc): one sweep
// along x(dir+1) with tile shape kern in 1..mhd_tile_shapes(), and the zone pass after
// the face update (dU/dt, stage combine, cell B from faces, PreparePrimitive of u1)
int launch_fused_mhd_sweep(const km_cfg_t &cfg, const FusedArgs &a, int dir, int kern,
                           hipStream_t st);
int mhd_tile_shapes();
int launch_ct_zone(const FusedArgs &a, hipStream_t st);
}
namespace fma {
int launch_fused_sweep(const km_cfg_t &cfg, const FusedArgs &a, int ndim, int dir, bool with_dt,
                       int kern, hipStream_t st);
int tile_shapes();
int launch_fine_face_fluxes(const km_cfg_t &cfg, const FusedArgs &a, hipStream_t st);
// the whole stage in one launch (2-D and 3-D, single-level meshes), kern in
// 1..stage3_kernels() on 3-D meshes and 1..stage2_kernels() on 2-D ones
int launch_fused_stage3(const km_cfg_t &cfg, const FusedArgs &a, bool with_dt, int kern,
                        hipStream_t st);
int stage3_kernels();
int stage2_kernels();
// the CT-MHD stage (3-D, Cartesian, single-level meshes; km_fused_mhd.inc): one sweep
// along x(dir+1) with tile shape kern in 1..mhd_tile_shapes(), and the zone pass after
// the face update (dU/dt, stage combine, cell B from faces, PreparePrimitive of u1)
int launch_fused_mhd_sweep(const km_cfg_t &cfg, const FusedArgs &a, int dir, int kern,
                           hipStream_t st);
int mhd_tile_shapes();
int launch_ct_zone(const FusedArgs &a, hipStream_t st);
}
namespace fp32 {
int launch_fused_sweep(const km_cfg_t &cfg, const FusedArgs &a, int ndim, int dir, bool with_dt,
                       int kern, hipStream_t st);
int tile_shapes();
int launch_fine_face_fluxes(const km_cfg_t &cfg, const FusedArgs &a, hipStream_t st);
// the whole stage in one launch (2-D and 3-D, single-level meshes), kern in
// 1..stage3_kernels() on 3-D meshes and 1..stage2_kernels() on 2-D ones
int launch_fused_stage3(const km_cfg_t &cfg, const FusedArgs &a, bool with_dt, int kern,
                        hipStream_t st);
int stage3_kernels();
int stage2_kernels();
// the CT-MHD stage (3-D, Cartesian, single-level meshes; km_fused_mhd.inc): one sweep
// along x(dir+1) with tile shape kern in 1..mhd_tile_shapes(), and the zone pass after
// the face update (dU/dt, stage combine, cell B from faces, PreparePrimitive of u1)
int launch_fused_mhd_sweep(const km_cfg_t &cfg, const FusedArgs &a, int dir, int kern,
                           hipStream_t st);
int mhd_tile_shapes();
int launch_ct_zone(const FusedArgs &a, hipStream_t st);
}

// thread-local error message for km_last_error()
void set_error(const std::string &msg);

}  // namespace km
