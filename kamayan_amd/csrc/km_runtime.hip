// km_runtime.hip -- host runtime of libkm_hydro.so: the C-ABI task entry points, the
// option dispatcher, the device mesh and the RK-stage driver.
//
// The driver mirrors KamayanDriver::MakeTaskCollection / BuildTaskListRKStage
// (kamayan_driver.cpp:121-248) for a single MeshData partition on one GPU:
//   stage s: fluxes(md0) -> EMF(md0) -> FluxesToDuDt -> ApplyDuDt(base, md0, md1)
//            -> PreparePrimitive(md1) [-> EstimateTimestep, FillDerived on the last stage]
//            -> ghost fill(md1)
// with stage containers stage_name = {base, 1, .., n-1, base} and the SSP-RK beta table
// used by Parthenon's multi-stage integrators. Conserved variables (and CT face B) are
// independent per container; primitives are shared between containers, as the hydro
// unit's field metadata makes them (hydro.cpp:96-105).
//
// The fused path (KM_PATH_FUSED, hydro) replaces the per-task kernels of a stage by one
// sweep per direction (km_fused.inc) and keeps dt on the device, so a whole cycle can be
// captured once in a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/km_hydro.h"
#include "km_amr.hpp"
#include "km_device.hpp"
#include "km_internal.hpp"

namespace km {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }

#define KM_HIP(call)                                                                  \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess) {                                                           \
      set_error(std::string(#call) + ": " + hipGetErrorString(e_));                   \
      return KM_EHIP;                                                                 \
    }                                                                                 \
  } while (0)

static int check_launch() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error(std::string("kernel launch: ") + hipGetErrorString(e));
    return KM_EHIP;
  }
  return KM_OK;
}

// Dispatcher<...>(cfg): the option combinations the kernels are instantiated for. An
// unsupported value fails the way PolymorphicDispatch does (dispatcher.hpp:122-130).
static int check_cfg(const km_cfg_t *cfg, const km_pack_t *u) {
  if (!cfg || !u) {
    set_error("null argument");
    return KM_EINVAL;
  }
  if (cfg->recon < 0 || cfg->recon > 3) {
    set_error("dispatch parm [Reconstruction] not handled");
    return KM_EINVAL;
  }
  if (cfg->limiter < 0 || cfg->limiter > 2) {
    set_error("dispatch parm [SlopeLimiter] not handled");
    return KM_EINVAL;
  }
  if (cfg->riemann != KM_RIEMANN_HLL && cfg->riemann != KM_RIEMANN_HLLC) {
    set_error("dispatch parm [RiemannSolver] not handled");
    return KM_EINVAL;
  }
  if (cfg->fluid != KM_FLUID_ONET) {  // physics_types.hpp:10 instantiates oneT only
    set_error(cfg->fluid == KM_FLUID_THREET ? "ThreeT eos not implemented (physics/fluid = 3t)"
                                            : "dispatch parm [Fluid] not handled");
    return KM_EINVAL;
  }
  if (cfg->mode_init < KM_EOS_INIT_DENS_PRES || cfg->mode_init > KM_EOS_INIT_DENS_TEMP) {
    set_error("eos/mode_init must be dens_pres, dens_ener or dens_temp");
    return KM_EINVAL;
  }
  if (cfg->geometry != KM_GEOM_CARTESIAN && cfg->geometry != KM_GEOM_CYLINDRICAL) {
    set_error("dispatch parm [Geometry] not handled");
    return KM_EINVAL;
  }
  if (cfg->geometry == KM_GEOM_CYLINDRICAL && u->ndim > 2) {
    set_error("cylindrical geometry is provided in 1-D and 2-D (r, z)");
    return KM_EINVAL;
  }
  if (cfg->mhd != KM_MHD_OFF && cfg->mhd != KM_MHD_CT) {
    set_error("dispatch parm [Mhd] not handled");
    return KM_EINVAL;
  }
  if (u->ndim < 1 || u->ndim > 3 || u->nblocks < 0) {
    set_error("bad pack shape");
    return KM_EINVAL;
  }
  if (u->ng < km_cfg_min_ghost(cfg)) {
    set_error("nghost " + std::to_string(u->ng) + " below the " +
              std::to_string(km_cfg_min_ghost(cfg)) + " the options need");
    return KM_EINVAL;
  }
  if (cfg->nspecies < 1 || cfg->nspecies > 64) {
    set_error("material/species must name 1 to 64 species");
    return KM_EINVAL;
  }
  if (u->nvar < KM_NVAR(cfg->mhd) + KM_NSPEC_SLOTS(cfg->nspecies)) {
    set_error("pack has fewer cell variables than the hydro state needs");
    return KM_EINVAL;
  }
  return KM_OK;
}

// flux bounds of CalculateFluxesNested (hydro_add_flux_tasks.cpp:56-70)
static Box flux_box(const km_cfg_t &cfg, const km_pack_t &u) {
  const Geo g(u);
  Box bx;
  for (int d = 0; d < 3; ++d) {
    bx.lo[d] = g.lo[d];
    bx.hi[d] = g.hi[d];
  }
  if (cfg.mhd == KM_MHD_CT) {
    const int w[3] = {g.ndim > 1, g.ndim > 1, g.ndim > 2};
    for (int d = 0; d < 3; ++d) {
      bx.lo[d] -= w[d];
      bx.hi[d] += w[d];
    }
  }
  return bx;
}

static hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

__global__ void k_set(double *p, double v, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}
__global__ void k_scale(double *p, double s, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = p[t] * s;
}
// parthenon/time step control (kamayan_driver.cpp:28-53), restated from Parthenon's
// documented rule (its driver source is not part of the reference tree)
struct DtRule {
  double factor, init, ceil, floor, min, max, user, force;
  int init_force, min_limit, max_limit;
};
DtRule dt_rule(const km_cfg_t &c) {
  return {c.dt_factor, c.dt_init, c.dt_ceil, c.dt_floor, c.dt_min, c.dt_max, c.dt_user,
          c.dt_force, c.dt_init_force, c.dt_min_cycle_limit, c.dt_max_cycle_limit};
}
// the limits every new dt obeys: dt_user, then dt_floor / dt_ceil, then dt_force when
// it is set (positive; the reference's default and its deck writer's -1e300 leave it off)
__device__ double dt_limits(const DtRule &r, double x) {
  x = kmin(x, r.user);
  x = kmin(kmax(x, r.floor), r.ceil);
  if (r.force > 0.0) x = r.force;
  return x;
}
// dt_min / dt_max: the run fails once dt has stayed below dt_min (above dt_max) for
// dt_min_cycle_limit (dt_max_cycle_limit) cycles in a row; cnt[2 * t] / cnt[2 * t + 1]
// count those cycles
__device__ void dt_watch(const DtRule &r, double x, int *cnt, int *fail) {
  if (x < r.min) {
    if (++cnt[0] >= r.min_limit) atomicOr(fail, KM_FAIL_DT_MIN);
  } else {
    cnt[0] = 0;
  }
  if (x > r.max) {
    if (++cnt[1] >= r.max_limit) atomicOr(fail, KM_FAIL_DT_MAX);
  } else {
    cnt[1] = 0;
  }
}
// first dt of a run: the estimate capped by dt_init, or dt_init itself with dt_init_force
// when dt_init is given (positive and below 1e300, the deck writer's "unset"), then the
// limits
__global__ void k_dt_init(double *dt, const double *est, DtRule r, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) {
    double x = kmin(est[t], r.init);
    if (r.init_force && r.init > 0.0 && r.init < 1e300) x = r.init;
    dt[t] = dt_limits(r, x);
  }
}
// after a cycle, per instance: time += dt; dt = min(dt_factor dt, estimate) within the
// limits, unless the caller fixed the step's dt
__global__ void k_dt_update(double *dt, double *time, const double *est, int keep, int n,
                            DtRule r, int *cnt, int *fail) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) {
    const double d = dt[t];
    time[t] = time[t] + d;
    dt_watch(r, d, cnt + 2 * t, fail);
    if (!keep) dt[t] = dt_limits(r, kmin(r.factor * d, est[t]));
  }
}
__global__ void k_set_int(int *p, int v, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}
static inline unsigned nb64(int n) { return (unsigned)((n + 63) / 64); }
// plain device copy, 16 B per lane per iteration, grid-stride
__global__ void __launch_bounds__(256) k_copy16(double2 *dst, const double2 *src, long n) {
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n;
       t += (long)gridDim.x * blockDim.x)
    dst[t] = src[t];
}

}  // namespace km

using namespace km;

struct km_ghost_plan_t {
  UniformGrid grid;
  bool amr = false;  // block-tree plan: agrid over a device lookup table
  AmrGrid agrid{};
  int *lookup_d = nullptr;
  ~km_ghost_plan_t() {
    if (lookup_d) (void)hipFree(lookup_d);
  }
};

// ---------------------------------------------------------------------------------
// device mesh
// ---------------------------------------------------------------------------------
namespace {

struct DevBuf {
  void *p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

struct Cont {
  std::vector<double *> cell_h, face_h[3], flux_h[3], edge_h[3];
  double **cell_d = nullptr, **face_d[3] = {nullptr, nullptr, nullptr};
  double **flux_d[3] = {nullptr, nullptr, nullptr}, **edge_d[3] = {nullptr, nullptr, nullptr};
  km_pack_t pack;
};

}  // namespace

struct km_sim_t {
  km_cfg_t cfg;
  km_mesh_desc_t mesh;
  int path = KM_PATH_TASKS;
  int nb = 0, nvar = 0, nc = 0;
  int nflux = 0;  // conserved fluxes plus one per species (more than one species)
  long ncell = 0, nstag = 0;
  std::vector<std::unique_ptr<DevBuf>> bufs;
  km_block_t *blocks_d = nullptr;
  Cont base, stage[2], dudt, scratch;
  double *du = nullptr;
  int *nbr_d = nullptr;  // [nb][3][2] face neighbours for the fused sweeps
  UniformGrid grid;      // grid of instance 0
  bool amr = false;      // block-tree mesh (task path): ghost fill across levels and
  AmrGrid agrid{};       // flux correction use agrid
  int ninst = 1;         // independent instances sharing the pack (ensemble)
  std::vector<km_mesh_desc_t> meshes;
  std::vector<UniformGrid> grids_h;
  int *inst_d = nullptr;           // instance of each block (null for one instance)
  UniformGrid *grids_d = nullptr;  // grid of each instance
  hipStream_t stream = nullptr;
  // fused path on a block tree: a second stream (default priority, like stream) and the
  // fork / join events that let the stage's ghost-fill prolongation run beside its gather
  hipStream_t side = nullptr;
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  double *dt_d = nullptr, *est_d = nullptr, *time_d = nullptr;
  // failure flag word (KM_FAIL_*) then the dt_min / dt_max cycle counters per instance
  int *fail_d = nullptr;
  int ncycle = 0;
  bool prims_from_ic = false;  // first stage must read the initial primitives
  bool dt_keep = false;
  bool use_graph = true;
  hipGraphExec_t graph_exec = nullptr;
  hipGraph_t graph = nullptr;
  // kernel of each fused sweep: -1 picked by the autotuner at the first cycle that starts
  // from conserved data, 0 the march, k >= 1 pencil-tile shape k
  int kern[3] = {-1, -1, -1};
  // kernel of a whole fused stage on a 3-D single-level mesh: -1 picked by the autotuner,
  // 0 the three sweeps above, k >= 1 the one-launch stage kernel variant k (km_stage3.inc)
  int stage_kern = -1;
  // fused path on a block tree: the stages exchange only the ghost zones their sweeps read
  // (density and the primitives); momentum and energy ghosts are filled from the base
  // interiors when something reads them (state copies, tagging and remesh)
  bool lazy_ghosts = false;
  // fused path on a block tree with two or more stages: every container owns its primitive
  // arrays, so the final sweep writes u1's primitives (and, on the last stage, the CFL
  // minimum) while other tiles still read u0's (RK1 updates u0 in place and keeps the
  // separate kernels)
  bool own_prims = false;
  // fused path on a block tree: the gather lists of the stages' ghost fill (rebuilt with
  // the tree)
  AmrFillList fill;
  // fused path on a block tree with own_prims, or the fused CT stage: the kernels leave
  // TEMP (and, on block trees in the forms that take the reconstructed BMOD from PRES,
  // BMOD) of the base state to be derived when something reads the state
  bool lazy_prims = false;
  void drop_graph() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    graph_exec = nullptr;
    graph = nullptr;
  }
  // optional HIP-event timing of the fused-stage kernels (kinds: sweep x1, x2, x3,
  // ghost fill, whole stage)
  bool profile = false;
  std::vector<hipEvent_t> ev_free;
  struct Rec {
    int kind;
    hipEvent_t a, b;
  };
  std::vector<Rec> recs;
  hipEvent_t ev() {
    hipEvent_t e;
    if (!ev_free.empty()) {
      e = ev_free.back();
      ev_free.pop_back();
    } else {
      (void)hipEventCreate(&e);
    }
    return e;
  }

  // block-tree meshes keep every buffer sized by the leaf count in mesh_bufs, which a
  // remesh replaces as a whole; alloc() files new buffers under *target
  std::vector<std::unique_ptr<DevBuf>> mesh_bufs;
  std::vector<std::unique_ptr<DevBuf>> *target = &bufs;
  // adaptive refinement
  bool adaptive = false;
  km_amr_adapt_t adapt{};
  std::vector<int32_t> leaves, deref;
  int *tags_d = nullptr;
  double *err2_d = nullptr;
  // history output file
  std::FILE *hist = nullptr;
  double hist_dt = 0., hist_next = 0.;
  bool hist_vortex = false;
  HistVortex hist_p{};
  int hist_last_cycle = -1;

  template <class T>
  int alloc(T **out, size_t count) {
    auto b = std::make_unique<DevBuf>();
    KM_HIP(hipMalloc(&b->p, count * sizeof(T) + 256));
    // the fill runs on the null stream, which the simulation's non-blocking stream does
    // not wait for: finish it before a kernel on that stream writes the buffer
    KM_HIP(hipMemset(b->p, 0, count * sizeof(T) + 256));
    KM_HIP(hipStreamSynchronize(nullptr));
    *out = static_cast<T *>(b->p);
    target->push_back(std::move(b));
    return KM_OK;
  }
  int upload_table(double ***dst, const std::vector<double *> &h) {
    int rc = alloc(dst, h.size());
    if (rc) return rc;
    KM_HIP(hipMemcpy(*dst, h.data(), h.size() * sizeof(double *), hipMemcpyHostToDevice));
    return KM_OK;
  }
  ~km_sim_t() {
    for (auto &r : recs) {
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
    }
    for (auto e : ev_free) (void)hipEventDestroy(e);
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (fork_ev) (void)hipEventDestroy(fork_ev);
    if (join_ev) (void)hipEventDestroy(join_ev);
    if (side) (void)hipStreamDestroy(side);
    if (stream) (void)hipStreamDestroy(stream);
    if (hist) std::fclose(hist);
  }
};

// Worklist of km_stage_fused: the mesh topology the fused kernels need beyond the caller's
// packs (face-neighbour table and instance map of single-level meshes or ensembles, the
// leaf lookup of a block tree), the FP64 du scratch of the sweeps and the kernel choices.
struct km_stage_plan_t {
  km_cfg_t cfg{};            // options of the calls the kernel choices were tuned for
  km_mesh_desc_t mesh{};     // instance 0 (ndim, nghost, cells per block)
  int path = KM_PATH_FUSED;  // the fused kernels only
  int nb = 0, ninst = 1;
  bool amr = false;
  AmrGrid agrid{};
  int *lookup_d = nullptr, *nbr_d = nullptr, *inst_d = nullptr;
  double *du = nullptr;
  int kern[3] = {-1, -1, -1};
  int stage_kern = -1;
  hipStream_t stream = nullptr;  // the current call's
  std::vector<hipEvent_t> ev_free;
  std::mutex mu;  // one call at a time per plan (the kernel choices and du are shared)
  hipEvent_t ev() {
    hipEvent_t e;
    if (!ev_free.empty()) {
      e = ev_free.back();
      ev_free.pop_back();
    } else {
      (void)hipEventCreate(&e);
    }
    return e;
  }
  ~km_stage_plan_t() {
    for (auto e : ev_free) (void)hipEventDestroy(e);
    for (void *p : {(void *)lookup_d, (void *)nbr_d, (void *)inst_d, (void *)du})
      if (p) (void)hipFree(p);
  }
};

namespace {

int make_cont(km_sim_t &s, Cont &c, bool full_cells, const Cont *share, bool faces,
              bool fluxes) {
  std::memset(&c.pack, 0, sizeof(c.pack));
  c.pack.nblocks = s.nb;
  c.pack.ndim = s.mesh.ndim;
  c.pack.ng = s.mesh.ng;
  for (int d = 0; d < 3; ++d) c.pack.nx[d] = d < s.mesh.ndim ? s.mesh.nx[d] : 1;
  c.pack.nvar = s.nvar;
  c.pack.nflux = s.nflux;
  c.pack.blocks = s.blocks_d;
  // a stage container owns the conserved fields (the hydro conserved slots and the mass
  // fractions) and shares the derived ones with `share`
  const int nhydro = KM_NVAR(s.cfg.mhd);
  auto owned = [&](int v) { return full_cells || v < s.nc || v >= nhydro; };
  const int own = full_cells ? s.nvar : s.nc + (s.nvar - nhydro);
  double *arena = nullptr;
  int rc = s.alloc(&arena, (size_t)s.nb * own * s.ncell);
  if (rc) return rc;
  c.cell_h.assign((size_t)s.nb * s.nvar, nullptr);
  for (int b = 0; b < s.nb; ++b)
    for (int v = 0, o = 0; v < s.nvar; ++v) {
      const size_t t = (size_t)b * s.nvar + v;
      if (owned(v))
        c.cell_h[t] = arena + ((size_t)b * own + o++) * s.ncell;
      else if (share)
        c.cell_h[t] = share->cell_h[t];
    }
  if ((rc = s.upload_table(&c.cell_d, c.cell_h))) return rc;
  c.pack.cell = c.cell_d;
  const bool ct = s.cfg.mhd == KM_MHD_CT;
  if (faces && ct) {
    for (int d = 0; d < 3; ++d) {
      double *fa = nullptr;
      if ((rc = s.alloc(&fa, (size_t)s.nb * s.nstag))) return rc;
      c.face_h[d].resize(s.nb);
      for (int b = 0; b < s.nb; ++b) c.face_h[d][b] = fa + (size_t)b * s.nstag;
      if ((rc = s.upload_table(&c.face_d[d], c.face_h[d]))) return rc;
      c.pack.face[d] = c.face_d[d];
    }
  }
  if (fluxes) {
    for (int d = 0; d < 3; ++d) {
      if (d >= s.mesh.ndim) continue;
      double *fa = nullptr;
      if ((rc = s.alloc(&fa, (size_t)s.nb * s.nflux * s.ncell))) return rc;
      c.flux_h[d].resize((size_t)s.nb * s.nflux);
      for (size_t t = 0; t < c.flux_h[d].size(); ++t) c.flux_h[d][t] = fa + t * s.ncell;
      if ((rc = s.upload_table(&c.flux_d[d], c.flux_h[d]))) return rc;
      c.pack.flux[d] = c.flux_d[d];
      if (ct) {
        double *ea = nullptr;
        if ((rc = s.alloc(&ea, (size_t)s.nb * s.nstag))) return rc;
        c.edge_h[d].resize(s.nb);
        for (int b = 0; b < s.nb; ++b) c.edge_h[d][b] = ea + (size_t)b * s.nstag;
        if ((rc = s.upload_table(&c.edge_d[d], c.edge_h[d]))) return rc;
        c.pack.edge[d] = c.edge_d[d];
      }
    }
    if (ct && s.mesh.ndim < 3) {  // E3 is needed in 2-D too
      double *ea = nullptr;
      if ((rc = s.alloc(&ea, (size_t)s.nb * s.nstag))) return rc;
      c.edge_h[2].resize(s.nb);
      for (int b = 0; b < s.nb; ++b) c.edge_h[2][b] = ea + (size_t)b * s.nstag;
      if ((rc = s.upload_table(&c.edge_d[2], c.edge_h[2]))) return rc;
      c.pack.edge[2] = c.edge_d[2];
    }
  }
  return KM_OK;
}

const double kBetaRK1[1] = {1.0};
const double kBetaRK2[2] = {1.0, 0.5};
const double kBetaRK3[3] = {1.0, 0.25, 2.0 / 3.0};

Cont &container(km_sim_t &s, int idx) {
  if (idx == 0 || idx == s.cfg.nstages) return s.base;
  return s.stage[idx - 1];
}

// boundary exchange of cell vars [v0, v1) (+ CT face B) on one container
void ghost_fill(km_sim_t &s, const km_pack_t &p, int v0, int v1, bool faces) {
  if (s.amr) {
    launch_amr_ghost_fill(s.agrid, p, v0, v1, s.nc, s.stream);
    if (faces) launch_amr_ghost_faces(s.agrid, p, s.stream);
  } else
    launch_ghost_fill(s.grid, p, v0, v1, s.nc, faces, s.stream, Instances{s.inst_d, s.grids_d});
}

// the blocks [b0, b0 + n) of a pack
km_pack_t sub_pack(const km_pack_t &p, int b0, int n) {
  km_pack_t q = p;
  q.nblocks = n;
  q.blocks = p.blocks + b0;
  q.cell = p.cell + (long)b0 * p.nvar;
  for (int d = 0; d < 3; ++d) {
    if (p.flux[d]) q.flux[d] = p.flux[d] + (long)b0 * p.nflux;
    if (p.face[d]) q.face[d] = p.face[d] + b0;
    if (p.edge[d]) q.edge[d] = p.edge[d] + b0;
  }
  return q;
}

// one RK stage as the reference's task list; with parthenon/mesh/pack_size the per-block
// tasks run pack by pack (MeshData partitions), the boundary exchange and the flux
// correction over the whole mesh
int task_stage(km_sim_t &s, int stage, double beta, double dt) {
  const km_cfg_t &cfg = s.cfg;
  Cont &md0 = container(s, stage - 1), &md1 = container(s, stage);
  km_pack_t p0 = md0.pack;
  for (int d = 0; d < 3; ++d) {
    p0.flux[d] = s.scratch.pack.flux[d];
    p0.edge[d] = s.scratch.pack.edge[d];
  }
  const bool ct = cfg.mhd == KM_MHD_CT;
  const int ps = cfg.pack_size > 0 && cfg.pack_size < s.nb ? cfg.pack_size : s.nb;
  auto packs = [&](auto &&f) {
    for (int b0 = 0; b0 < s.nb; b0 += ps) f(b0, std::min(ps, s.nb - b0));
  };
  packs([&](int b0, int n) {
    const km_pack_t q0 = sub_pack(p0, b0, n);
    launch_calculate_fluxes(cfg, q0, flux_box(cfg, q0), s.stream);
    if (ct) launch_calculate_emf(cfg, q0, s.stream);
  });
  // flux correction at level jumps (AddFluxCorrectionTasks, kamayan_driver.cpp:198-207):
  // cell fluxes and, for CT, the edge EMFs
  if (s.amr) {
    launch_amr_flux_correct(s.agrid, p0, s.nflux, s.stream);
    if (ct) launch_amr_emf_correct(s.agrid, p0, s.stream);
  }
  // FluxesToDuDt, ApplyDuDt, PreparePrimitive
  (void)dt;
  packs([&](int b0, int n) {
    launch_update(cfg, sub_pack(s.base.pack, b0, n), sub_pack(p0, b0, n),
                  sub_pack(md1.pack, b0, n), sub_pack(s.dudt.pack, b0, n), beta, s.stream,
                  s.dt_d, s.inst_d ? s.inst_d + b0 : nullptr);
  });
  if (stage == cfg.nstages) {
    packs([&](int b0, int n) {
      const int part = (b0 == 0 ? 1 : 0) | (b0 + n == s.nb ? 2 : 0);
      launch_estimate_dt(cfg, sub_pack(s.base.pack, b0, n), s.est_d, s.stream,
                         s.inst_d ? s.inst_d + b0 : nullptr, s.ninst, s.fail_d, part);
      if (ct) launch_fill_divb(cfg, sub_pack(s.base.pack, b0, n), s.stream);
    });
  }
  // FillGhost fields: conserved, VELOCITY, PRES, BMOD, EINT, face B (TEMP is not), MFRAC
  ghost_fill(s, md1.pack, 0, KM_TEMP(cfg.mhd), ct);
  if (s.nvar > KM_NVAR(cfg.mhd)) ghost_fill(s, md1.pack, KM_NVAR(cfg.mhd), s.nvar, false);
  return check_launch();
}

// With one RK stage on a single-level mesh the stage input and output would both be the
// base state, and a sweep that writes u1 while other workgroups still read u0 across
// block faces would race; that stage writes into the first stage container, which is
// then copied back (rk1_copy_back).
bool rk1_fused(const km_sim_t &s) { return s.cfg.nstages == 1 && !s.amr; }

__global__ void k_copy_vars(km_pack_t dst, km_pack_t src, int nv, long ncell) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long per = (long)nv * ncell;
  if (t >= per * dst.nblocks) return;
  const int b = (int)(t / per), v = (int)((t % per) / ncell);
  const long c = t % ncell;
  cell_ptr(dst, b, v)[c] = cell_ptr(src, b, v)[c];
}

void rk1_copy_back(km_sim_t &s) {
  const long n = (long)s.nb * s.nc * s.ncell;
  hipLaunchKernelGGL(k_copy_vars, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s.stream,
                     s.base.pack, s.stage[0].pack, s.nc, s.ncell);
}

FusedArgs fused_args(km_sim_t &s, int stage, double beta) {
  FusedArgs a;
  a.u0 = container(s, stage - 1).pack;
  a.base = s.base.pack;
  a.u1 = rk1_fused(s) ? s.stage[0].pack : container(s, stage).pack;
  a.du = s.du;
  a.dt = s.dt_d;
  a.dt_est = s.est_d;
  a.inst = s.inst_d;
  a.nbr = s.nbr_d;
  a.beta = beta;
  a.lim = s.cfg.limiter;
  a.eos = make_eos(s.cfg);
  a.amr = s.amr ? 1 : 0;
  a.agrid = s.agrid;
  a.fail = s.fail_d;
  a.write_prims = s.own_prims ? 1 : 0;
  // block trees: fluxes at level jumps live in the task path's flux arrays; CT: the
  // sweeps store the transverse B fluxes there for the EMF
  if (s.amr || s.cfg.mhd == KM_MHD_CT)
    for (int d = 0; d < 3; ++d) a.u0.flux[d] = s.scratch.pack.flux[d];
  return a;
}

template <class S>
int launch_sweep(const S &s, const FusedArgs &a, int dir, bool last, int kern) {
  const km_cfg_t &cfg = s.cfg;
  const int ndim = s.mesh.ndim;
  if (cfg.mhd == KM_MHD_CT) {  // CT stage sweeps: pencil tiles only (no march)
    if (cfg.precision == KM_PREC_FP32_MIXED)
      return fp32::launch_fused_mhd_sweep(cfg, a, dir, kern, s.stream);
    if (cfg.precision == KM_PREC_FP64_FMA)
      return fma::launch_fused_mhd_sweep(cfg, a, dir, kern, s.stream);
    return exact::launch_fused_mhd_sweep(cfg, a, dir, kern, s.stream);
  }
  if (cfg.precision == KM_PREC_FP32_MIXED)
    return fp32::launch_fused_sweep(cfg, a, ndim, dir, last, kern, s.stream);
  if (cfg.precision == KM_PREC_FP64_FMA)
    return fma::launch_fused_sweep(cfg, a, ndim, dir, last, kern, s.stream);
  return exact::launch_fused_sweep(cfg, a, ndim, dir, last, kern, s.stream);
}

template <class S>
int launch_fine_faces(const S &s, const FusedArgs &a) {
  if (s.cfg.precision == KM_PREC_FP32_MIXED) return fp32::launch_fine_face_fluxes(s.cfg, a, s.stream);
  if (s.cfg.precision == KM_PREC_FP64_FMA) return fma::launch_fine_face_fluxes(s.cfg, a, s.stream);
  return exact::launch_fine_face_fluxes(s.cfg, a, s.stream);
}

template <class S>
int launch_stage3(const S &s, const FusedArgs &a, bool last, int kern) {
  const km_cfg_t &cfg = s.cfg;
  if (cfg.precision == KM_PREC_FP32_MIXED)
    return fp32::launch_fused_stage3(cfg, a, last, kern, s.stream);
  if (cfg.precision == KM_PREC_FP64_FMA) return fma::launch_fused_stage3(cfg, a, last, kern, s.stream);
  return exact::launch_fused_stage3(cfg, a, last, kern, s.stream);
}

// one-launch stage variants of a mesh of ndim dimensions (the same in every form)
int stage3_kernels(const km_cfg_t &cfg, int ndim) {
  (void)cfg;
  return ndim == 2 ? exact::stage2_kernels() : exact::stage3_kernels();
}

// whether the one-launch stage kernel applies: fused path, 2-D or 3-D hydro, single level
template <class S>
bool stage3_applies(const S &s) {
  return s.path == KM_PATH_FUSED && !s.amr && s.mesh.ndim >= 2 && s.cfg.mhd == KM_MHD_OFF;
}

int tile_shapes(const km_cfg_t &cfg) {
  if (cfg.mhd == KM_MHD_CT) return exact::mhd_tile_shapes();  // the same shapes in every form
  if (cfg.precision == KM_PREC_FP32_MIXED) return fp32::tile_shapes();
  if (cfg.precision == KM_PREC_FP64_FMA) return fma::tile_shapes();
  return exact::tile_shapes();
}

// Autotuner of the fused sweeps. For each direction every candidate kernel (the march
// and each pencil-tile shape) runs that direction's sweep of the first stage on the
// simulation's own data; the candidates take turns over five timed rounds and the
// fastest best time wins. The sweeps only write scratch the next stage rewrites before
// reading it: the du accumulator, and u1 of stage 1, which is always a stage container
// (with one stage too, see rk1_fused). Choices are kept
// per (form, reconstruction, limiter, Riemann solver, block shape, block count,
// direction) in a process-wide table, so later simulations of the same shape reuse them.
std::mutex g_tune_mu;
std::map<std::vector<int>, int> g_tune;

// Time candidates fn(0) .. fn(n - 1) (each one launch sequence on the simulation's
// stream) into ms[0 .. n - 1]: each runs once untimed (ms[k] = -1 when it refuses with
// KM_EINVAL, i.e. does not fit the blocks, or the device refuses the launch), then the
// candidates take turns for kRounds HIP-event-timed runs and each keeps its fastest. Taking
// turns exposes every candidate to the same clock and power state, so a drift during the
// tuning does not favour the candidates timed last.
constexpr int kRounds = 5;

template <class S, class Fn>
int time_candidates(S &s, int n, Fn &&fn, std::vector<float> &ms) {
  ms.assign(n, -1.f);
  std::vector<char> ok(n, 0);
  for (int k = 0; k < n; ++k) {
    if (fn(k) != KM_OK || hipGetLastError() != hipSuccess) {
      (void)hipGetLastError();
      continue;
    }
    ok[k] = 1;
  }
  if (int rc = check_launch()) return rc;
  hipEvent_t e0 = s.ev(), e1 = s.ev();
  for (int r = 0; r < kRounds; ++r)
    for (int k = 0; k < n; ++k) {
      if (!ok[k]) continue;
      KM_HIP(hipEventRecord(e0, s.stream));
      if (int rc = fn(k)) return rc;
      KM_HIP(hipEventRecord(e1, s.stream));
      KM_HIP(hipEventSynchronize(e1));
      float t = 0.f;
      KM_HIP(hipEventElapsedTime(&t, e0, e1));
      ms[k] = ms[k] < 0.f ? t : std::min(ms[k], t);
    }
  s.ev_free.push_back(e0);
  s.ev_free.push_back(e1);
  return check_launch();
}

// s: a simulation or a stage plan; a: the arguments of a stage on its data
template <class S>
int autotune(S &s, const FusedArgs &a) {
  const km_cfg_t &cfg = s.cfg;
  const int ndim = s.mesh.ndim;
  const int nshape = tile_shapes(cfg);
  std::lock_guard<std::mutex> lock(g_tune_mu);
  std::vector<float> ms;
  for (int dir = 0; dir < 3; ++dir) {
    if (s.kern[dir] >= 0) continue;
    if (s.path != KM_PATH_FUSED || dir >= ndim) {
      s.kern[dir] = s.amr && s.path == KM_PATH_FUSED ? 1 : 0;
      continue;
    }
    const std::vector<int> key = {cfg.precision, cfg.recon, cfg.limiter, cfg.riemann, ndim,
                                  s.mesh.nx[0], s.mesh.nx[1], s.mesh.nx[2], s.nb, dir,
                                  s.amr ? 1 : 0, cfg.mhd};
    auto it = g_tune.find(key);
    if (it != g_tune.end()) {
      s.kern[dir] = it->second;
      continue;
    }
    // candidates: the march (0) and each tile shape (1..nshape); block trees and CT
    // stages run tiles only
    const int k0 = s.amr || cfg.mhd == KM_MHD_CT ? 1 : 0;
    if (int rc = time_candidates(
            s, nshape + 1 - k0, [&](int i) { return launch_sweep(s, a, dir, false, k0 + i); },
            ms))
      return rc;
    int best = -1;
    for (int i = 0; i < (int)ms.size(); ++i)
      if (ms[i] >= 0.f && (best < 0 || ms[i] < ms[best - k0])) best = k0 + i;
    if (best < 0) {
      set_error("no fused sweep kernel fits these blocks");
      return KM_EINVAL;
    }
    g_tune[key] = best;
    s.kern[dir] = best;
  }
  // the whole stage: the three tuned sweeps (candidate 0) against each one-launch stage
  // variant (1..)
  if (s.stage_kern < 0) {
    if (!stage3_applies(s)) {
      s.stage_kern = 0;
      return KM_OK;
    }
    const std::vector<int> key = {cfg.precision, cfg.recon, cfg.limiter, cfg.riemann, ndim,
                                  s.mesh.nx[0], s.mesh.nx[1], s.mesh.nx[2], s.nb, 3, 0};
    auto it = g_tune.find(key);
    if (it != g_tune.end()) {
      s.stage_kern = it->second;
      return KM_OK;
    }
    if (int rc = time_candidates(
            s, stage3_kernels(cfg, ndim) + 1,
            [&](int k) {
              if (k > 0) return launch_stage3(s, a, false, k);
              for (int dir = 0; dir < ndim; ++dir)
                if (int r = launch_sweep(s, a, dir, false, s.kern[dir])) return r;
              return (int)KM_OK;
            },
            ms))
      return rc;
    int best = 0;
    for (int k = 1; k < (int)ms.size(); ++k)
      if (ms[k] >= 0.f && (ms[best] < 0.f || ms[k] < ms[best])) best = k;
    g_tune[key] = best;
    s.stage_kern = best;
  }
  return KM_OK;
}

// HIP-event records of each sweep of a simulation that profiles (none otherwise)
auto sweep_recorder(km_sim_t &s) {
  return [&s](int dir, bool begin) {
    if (!s.profile) return;
    if (begin) {
      s.recs.push_back({dir, s.ev(), s.ev()});
      (void)hipEventRecord(s.recs.back().a, s.stream);
    } else {
      (void)hipEventRecord(s.recs.back().b, s.stream);
    }
  };
}

// The kernels of one fused stage on a single-level mesh (shared by the simulation driver
// and km_stage_fused): the one-launch stage kernel or the three sweeps, and on the last
// stage the CFL minimum of u1 into est (per instance, scaled by cfl / ndim).
// rec(dir, begin) brackets each sweep (HIP-event profiling).
template <class S, class Rec>
int stage_launch(S &s, const FusedArgs &a, bool last, int ninst, double *est, Rec &&rec) {
  const int ndim = s.mesh.ndim;
  if (last)
    hipLaunchKernelGGL(k_set, dim3(nb64(ninst)), dim3(64), 0, s.stream, est, DBL_MAX, ninst);
  if (s.stage_kern > 0 && stage3_applies(s)) {
    if (int rc = launch_stage3(s, a, last, s.stage_kern)) {
      set_error("the one-launch stage kernel " + std::to_string(s.stage_kern) +
                " does not fit these blocks");
      return rc;
    }
  } else {
    for (int dir = 0; dir < ndim; ++dir) {
      rec(dir, true);
      if (int rc = launch_sweep(s, a, dir, last, s.kern[dir] > 0 ? s.kern[dir] : 0)) return rc;
      rec(dir, false);
    }
  }
  if (last)
    hipLaunchKernelGGL(k_scale, dim3(nb64(ninst)), dim3(64), 0, s.stream, est,
                       s.cfg.cfl / static_cast<double>(ndim), ninst);
  return KM_OK;
}

// The fused kernels of one stage on a block-tree mesh up to u1 (shared by the simulation
// driver and km_stage_fused): the fluxes of faces where a block borders a coarser leaf,
// their restriction onto the coarse face, then the pencil-tile sweeps, which take that flux
// at their end faces.
// With est, the final sweep also reduces the CFL minimum of u1 into est (one instance,
// scaled by cfl / ndim), which a.write_prims requires to be meaningful only on the last
// stage.
template <class S, class Rec>
int amr_stage_launch(S &s, const FusedArgs &a, int nflux, Rec &&rec, double *est = nullptr) {
  const int ndim = s.mesh.ndim;
  if (int rc = launch_fine_faces(s, a)) return rc;
  launch_amr_flux_correct(s.agrid, a.u0, nflux, s.stream);
  if (est) hipLaunchKernelGGL(k_set, dim3(1), dim3(64), 0, s.stream, est, DBL_MAX, 1);
  for (int dir = 0; dir < ndim; ++dir) {
    rec(dir, true);
    if (int rc = launch_sweep(s, a, dir, est != nullptr, s.kern[dir] > 0 ? s.kern[dir] : 1))
      return rc;
    rec(dir, false);
  }
  if (est)
    hipLaunchKernelGGL(k_scale, dim3(1), dim3(64), 0, s.stream, est,
                       s.cfg.cfl / static_cast<double>(ndim), 1);
  return KM_OK;
}

// One RK stage of the fused path on a block-tree mesh. The stage input's ghost zones and
// primitives are current (the previous stage's PreparePrimitive and ghost exchange, as on
// the task path), so the sweeps read the stored primitives of every cell they need. The
// fluxes of faces where a block borders a coarser leaf are computed first and restricted
// onto the coarse face (AddFluxCorrectionTasks), and the coarse block's sweep takes that
// flux at its end face: every face flux, and so du and u1, is the task path's. With
// own_prims the final sweep also writes PreparePrimitive's interior values of u1 and, on
// the last stage, reduces the CFL estimate (the task path's kernels, folded into the
// sweep); otherwise those are the task path's kernels. The ghost exchange of density and
// primitives follows, over the face slabs the sweeps read; the rest of the ghost zones
// are filled when something reads them.
int amr_fused_stage(km_sim_t &s, int stage, double beta) {
  const km_cfg_t &cfg = s.cfg;
  Cont &md1 = container(s, stage);
  const FusedArgs a = fused_args(s, stage, beta);
  hipEvent_t st0 = nullptr, st1 = nullptr;
  if (s.profile) {
    st0 = s.ev();
    st1 = s.ev();
    (void)hipEventRecord(st0, s.stream);
  }
  const bool last = stage == cfg.nstages;
  if (int rc = amr_stage_launch(s, a, s.nflux, sweep_recorder(s),
                                s.own_prims && last ? s.est_d : nullptr))
    return rc;
  if (!s.own_prims) {
    launch_prepare_primitive(cfg, md1.pack, s.stream);
    if (last)
      launch_estimate_dt(cfg, s.base.pack, s.est_d, s.stream, s.inst_d, s.ninst, s.fail_d);
  }
  hipEvent_t g0 = nullptr, g1 = nullptr;
  if (s.profile) {
    g0 = s.ev();
    g1 = s.ev();
    (void)hipEventRecord(g0, s.stream);
  }
  // density and the primitives in the face slabs the next sweeps read (the
  // reconstruction's reach plus one: the cells of the end faces' stencils); momentum,
  // energy, edges and corners are left out, and BMOD too in the forms that take the
  // reconstructed BMOD from PRES (their sweeps never read it)
  const int bmod = cfg.precision == KM_PREC_FP64 ? -1 : KM_VEL0(cfg.mhd) + 4;
  launch_amr_ghost_fill_list(s.fill, s.agrid, md1.pack, 0, KM_TEMP(cfg.mhd), s.nc, s.stream, 1,
                             KM_NCONS(cfg.mhd), bmod, s.side, s.fork_ev, s.join_ev);
  s.lazy_ghosts = true;
  if (s.own_prims) s.lazy_prims = true;
  if (s.profile) {
    (void)hipEventRecord(g1, s.stream);
    s.recs.push_back({3, g0, g1});
    (void)hipEventRecord(st1, s.stream);
    s.recs.push_back({4, st0, st1});
  }
  return check_launch();
}

// the ghost layers the fused block-tree sweeps read: the reconstruction's reach plus one
// (the cells of the end faces' stencils)
int fused_reach(const km_cfg_t &cfg) {
  return cfg.recon == KM_RECON_FOG ? 1 : cfg.recon == KM_RECON_PLM ? 2 : 3;
}

// The gather lists of the fused stages' ghost fill for the current tree (allocated into
// the mesh's buffers): the list entries come from the same tree walk as
// launch_amr_ghost_fill, the prolonged parents are compacted on the host.
int build_fill_list(km_sim_t &s) {
  const km_pack_t &u = s.base.pack;
  const int reach = fused_reach(s.cfg);
  AmrFillList f;
  f.depth = reach >= s.mesh.ng ? 0 : reach;
  if (Geo(u).ncell() >= (1L << 28)) {
    set_error("block too large for the ghost gather list (2^28 cells)");
    return KM_EINVAL;
  }
  f.n = amr_ghost_list_size(u, f.depth);
  int rc;
  if ((rc = s.alloc(&f.list, (size_t)f.n))) return rc;
  launch_amr_ghost_list(s.agrid, u, f.depth, f.list, s.stream);
  if (s.agrid.lmax > 0) {
    const long np = amr_prolong_parents(u, f.depth);
    DevBuf flag_d;
    KM_HIP(hipMalloc(&flag_d.p, np > 0 ? np : 1));
    launch_amr_prolong_flags(s.agrid, u, f.depth, static_cast<unsigned char *>(flag_d.p),
                             s.stream);
    if ((rc = check_launch())) return rc;
    std::vector<unsigned char> flag(np);
    KM_HIP(hipMemcpyAsync(flag.data(), flag_d.p, np, hipMemcpyDeviceToHost, s.stream));
    KM_HIP(hipStreamSynchronize(s.stream));
    std::vector<long> pl;
    for (long p = 0; p < np; ++p)
      if (flag[p]) pl.push_back(p);
    f.np = (long)pl.size();
    long *pl_d = nullptr;
    if ((rc = s.alloc(&pl_d, pl.size() + 1))) return rc;
    if (f.np)
      KM_HIP(hipMemcpy(pl_d, pl.data(), pl.size() * sizeof(long), hipMemcpyHostToDevice));
    f.plist = pl_d;
  }
  s.fill = f;
  return check_launch();
}

// the ghost zones the fused block-tree stages leave to be filled: momentum and energy, and
// density and the primitives outside the face slabs
void fill_lazy_ghosts(km_sim_t &s) {
  if (s.lazy_prims) {  // PreparePrimitive of the base interior (the same values)
    launch_prepare_primitive(s.cfg, s.base.pack, s.stream);
    s.lazy_prims = false;
  }
  if (!s.lazy_ghosts) return;
  ghost_fill(s, s.base.pack, 0, KM_TEMP(s.cfg.mhd), s.cfg.mhd == KM_MHD_CT);
  s.lazy_ghosts = false;
}

int fused_stage(km_sim_t &s, int stage, double beta) {
  const km_cfg_t &cfg = s.cfg;
  const bool last = stage == cfg.nstages;
  const FusedArgs a = fused_args(s, stage, beta);
  hipEvent_t st0 = nullptr, st1 = nullptr;
  if (s.profile) {
    st0 = s.ev();
    st1 = s.ev();
    (void)hipEventRecord(st0, s.stream);
  }
  if (int rc = stage_launch(s, a, last, s.ninst, s.est_d, sweep_recorder(s))) return rc;
  if (rk1_fused(s)) rk1_copy_back(s);
  if (s.profile) {
    (void)hipEventRecord(st1, s.stream);
    s.recs.push_back({4, st0, st1});
  }
  return check_launch();
}

// One RK stage of the fused path for constrained-transport MHD (3-D, Cartesian, single
// level). The stage input's primitives, cell-centred B and ghost zones are current (the
// previous stage's zone pass and ghost exchange, or the initial conditions), so the three
// pencil-tile sweeps read the stored recon variables (the task path's flux kernels read
// the same). They leave du and the transverse B fluxes; then one kernel forms the edge
// EMFs, FluxStokes and the face update, the zone pass (dU/dt, stage combine, cell B
// from the new faces, PreparePrimitive of u1), on the last stage the CFL estimate and
// divB, and the ghost exchange of the task path. With one stage u1 is the base state:
// every kernel after the sweeps reads and writes each zone and face in the same lane.
int ct_fused_stage(km_sim_t &s, int stage, double beta) {
  const km_cfg_t &cfg = s.cfg;
  Cont &md1 = container(s, stage);
  FusedArgs a = fused_args(s, stage, beta);
  a.u1 = md1.pack;
  hipEvent_t st0 = nullptr, st1 = nullptr;
  if (s.profile) {
    st0 = s.ev();
    st1 = s.ev();
    (void)hipEventRecord(st0, s.stream);
  }
  auto rec = sweep_recorder(s);
  for (int dir = 0; dir < 3; ++dir) {
    rec(dir, true);
    if (int rc = launch_sweep(s, a, dir, false, s.kern[dir] > 0 ? s.kern[dir] : 1)) {
      set_error("CT sweep tile shape " + std::to_string(s.kern[dir]) + " along x" +
                std::to_string(dir + 1) + " needs more LDS than a workgroup has for these blocks");
      return rc;
    }
    rec(dir, false);
  }
  // md0 with the scratch flux arrays: EMF, Stokes and face update in one kernel
  launch_ct_faces(s.base.pack, a.u0, md1.pack, beta, s.stream, s.dt_d, s.inst_d);
  int rc = cfg.precision == KM_PREC_FP32_MIXED ? fp32::launch_ct_zone(a, s.stream)
           : cfg.precision == KM_PREC_FP64_FMA  ? fma::launch_ct_zone(a, s.stream)
                                                : exact::launch_ct_zone(a, s.stream);
  if (rc) return rc;
  const bool last = stage == cfg.nstages;
  if (last) launch_fill_divb(cfg, s.base.pack, s.stream);
  hipEvent_t g0 = nullptr, g1 = nullptr;
  if (s.profile) {
    g0 = s.ev();
    g1 = s.ev();
    (void)hipEventRecord(g0, s.stream);
  }
  // the ghost zones the next sweeps read: density, cell B and the primitives within the
  // reconstruction's reach plus one of the interior, and the face field one zone deep;
  // the rest (momentum and energy, deeper layers) are filled when something reads them
  // (BMOD only in the FP64 form: the others take it from PRES after reconstruction)
  const Instances in{s.inst_d, s.grids_d};
  const int reach = fused_reach(cfg);
  const int bmod = cfg.precision == KM_PREC_FP64 ? -1 : KM_VEL0(cfg.mhd) + 4;
  // the face fills (and on the last stage the CFL estimate, which reads interior cells
  // only) run on the side stream beside the cell fills (different arrays, and each reads
  // interiors only), joined before anything else on the simulation stream
  const bool fork = s.side && s.fork_ev && s.join_ev;
  if (fork) {
    KM_HIP(hipEventRecord(s.fork_ev, s.stream));
    KM_HIP(hipStreamWaitEvent(s.side, s.fork_ev, 0));
  }
  if (last)
    launch_estimate_dt(cfg, s.base.pack, s.est_d, fork ? s.side : s.stream, s.inst_d, s.ninst,
                       s.fail_d);
  launch_ghost_fill(s.grid, md1.pack, 0, 0, s.nc, true, fork ? s.side : s.stream, in, reach, 1);
  launch_ghost_fill(s.grid, md1.pack, KM_DENS, KM_DENS + 1, s.nc, false, s.stream, in, reach);
  launch_ghost_fill(s.grid, md1.pack, KM_MAGC0, KM_TEMP(cfg.mhd), s.nc, false, s.stream, in,
                    reach, 1, bmod);
  if (fork) {
    KM_HIP(hipEventRecord(s.join_ev, s.side));
    KM_HIP(hipStreamWaitEvent(s.stream, s.join_ev, 0));
  }
  s.lazy_ghosts = true;
  s.lazy_prims = true;  // the zone pass leaves TEMP to PreparePrimitive
  if (s.profile) {
    (void)hipEventRecord(g1, s.stream);
    s.recs.push_back({3, g0, g1});
    (void)hipEventRecord(st1, s.stream);
    s.recs.push_back({4, st0, st1});
  }
  return check_launch();
}

// Replace the mesh of a block-tree simulation by new leaves: every leaf-sized buffer is
// allocated anew, base data moves by km_amr_remesh, then the old buffers are released.
int amr_remesh(km_sim_t &s, const std::vector<int32_t> &nl, const std::vector<int32_t> &origin) {
  const int n = (int)nl.size() / 4;
  AmrTreeHost t;
  const std::string err = build_amr_tree(s.mesh, nl.data(), n, t);
  if (!err.empty()) {
    set_error("remesh: " + err);
    return KM_EINVAL;
  }
  t.cyl = s.cfg.geometry == KM_GEOM_CYLINDRICAL;
  std::vector<std::unique_ptr<DevBuf>> fresh;
  s.target = &fresh;
  const km_pack_t old_pack = s.base.pack;
  s.nb = n;
  int rc;
  if ((rc = s.alloc(&s.blocks_d, n))) return rc;
  KM_HIP(hipMemcpy(s.blocks_d, t.blocks.data(), n * sizeof(km_block_t), hipMemcpyHostToDevice));
  int *lookup_d = nullptr;
  if ((rc = s.alloc(&lookup_d, t.lookup.size()))) return rc;
  KM_HIP(hipMemcpy(lookup_d, t.lookup.data(), t.lookup.size() * sizeof(int),
                   hipMemcpyHostToDevice));
  s.agrid = t.grid(lookup_d);
  if ((rc = make_cont(s, s.base, true, nullptr, true, false))) return rc;
  for (int i = 0; i < 2; ++i)
    if ((rc = make_cont(s, s.stage[i], s.own_prims, s.own_prims ? nullptr : &s.base, true,
                        false)))
      return rc;
  if ((rc = make_cont(s, s.dudt, false, nullptr, true, false))) return rc;
  if ((rc = make_cont(s, s.scratch, false, nullptr, false, true))) return rc;
  if ((rc = s.alloc(&s.tags_d, n))) return rc;
  if ((rc = s.alloc(&s.err2_d, n))) return rc;
  if (s.path == KM_PATH_FUSED) {
    const long nint = (long)s.mesh.nx[0] * (s.mesh.ndim > 1 ? s.mesh.nx[1] : 1) *
                      (s.mesh.ndim > 2 ? s.mesh.nx[2] : 1);
    if ((rc = s.alloc(&s.du, (size_t)n * 5 * nint))) return rc;
    if ((rc = build_fill_list(s))) return rc;
  }
  int32_t *origin_d = nullptr;
  if ((rc = s.alloc(&origin_d, origin.size()))) return rc;
  KM_HIP(hipMemcpy(origin_d, origin.data(), origin.size() * sizeof(int32_t),
                   hipMemcpyHostToDevice));
  const bool cyl = s.cfg.geometry == KM_GEOM_CYLINDRICAL;
  launch_amr_remesh(old_pack, s.base.pack, origin_d, 0, s.nvar, cyl, s.stream);
  if (s.cfg.mhd == KM_MHD_CT)
    launch_amr_remesh_faces(old_pack, s.base.pack, origin_d, cyl, s.mesh.xmin[0], s.stream);
  if ((rc = check_launch())) return rc;
  KM_HIP(hipStreamSynchronize(s.stream));
  s.mesh_bufs.swap(fresh);  // the old mesh's buffers go with `fresh`
  s.target = &s.bufs;
  return KM_OK;
}

// After a cycle: tag the filled base state, update the tree and, when it changed,
// remesh, refill ghost zones and re-estimate dt on the new mesh (Parthenon's
// EvolutionDriver::Execute: Step, LoadBalancingAndAdaptiveRefinement,
// InitializeBlockTimeStepsAfterAMR, then the time-step rule)
int adapt_step(km_sim_t &s) {
  const km_amr_adapt_t &a = s.adapt;
  fill_lazy_ghosts(s);  // the estimator and the remesh read ghost zones of every field
  std::vector<int32_t> tags(s.nb, KM_AMR_DEREFINE);
  if (a.ncrit > 0) {
    for (int c = 0; c < a.ncrit; ++c)
      launch_amr_tag_loehner(s.base.pack, a.crit[c], s.tags_d, c == 0, s.err2_d, s.stream);
    if (int rc = check_launch()) return rc;
    KM_HIP(hipMemcpyAsync(tags.data(), s.tags_d, s.nb * sizeof(int32_t), hipMemcpyDeviceToHost,
                          s.stream));
    KM_HIP(hipStreamSynchronize(s.stream));
  }
  std::vector<int32_t> nl, nd, org;
  bool changed = false;
  const std::string err = update_amr_tree(s.mesh, a.numlevel, a.derefine_count,
                                          s.leaves.data(), s.nb, tags.data(), s.deref.data(),
                                          nl, nd, org, changed);
  if (!err.empty()) {
    set_error(err);
    return KM_EINVAL;
  }
  if (!changed) return KM_OK;
  if (int rc = amr_remesh(s, nl, org)) return rc;
  s.leaves = std::move(nl);
  s.deref = std::move(nd);
  const bool ct = s.cfg.mhd == KM_MHD_CT;
  ghost_fill(s, s.base.pack, 0, KM_TEMP(s.cfg.mhd), ct);
  if (s.nvar > KM_NVAR(s.cfg.mhd)) ghost_fill(s, s.base.pack, KM_NVAR(s.cfg.mhd), s.nvar, false);
  if (ct) launch_fill_divb(s.cfg, s.base.pack, s.stream);  // divB of the new mesh's faces
  launch_estimate_dt(s.cfg, s.base.pack, s.est_d, s.stream, nullptr, 1, s.fail_d);
  return check_launch();
}

// isentropic_vortex parameters as km_problem_fill reads them (isentropic_vortex.cpp:55-65)
HistVortex vortex_params(const km_cfg_t &cfg, const double *params, int n) {
  auto P = [&](int i, double d) { return (params && i < n) ? params[i] : d; };
  HistVortex p;
  p.density = P(0, 1.0);
  p.pressure = P(1, 1.0);
  p.velx = P(2, 1.0);
  p.vely = P(3, 1.0);
  p.strength = P(4, 5.0);
  p.mhd_strength = P(5, 1.0);
  p.gamma = cfg.gamma;
  return p;
}

// ErrorHistory of every history variable: block sums on the device, added on the host in
// block order, over the x1-x2 extent of the domain (isentropic_vortex.hpp:158-199)
int vortex_errors(const km_cfg_t &cfg, const km_pack_t &u, const km_mesh_desc_t &mesh,
                  const HistVortex &p, hipStream_t st, double *out) {
  const int nv = cfg.mhd != KM_MHD_OFF ? 8 : 5;
  const size_t n = (size_t)(u.nblocks > 0 ? u.nblocks : 1) * nv;
  DevBuf part;
  KM_HIP(hipMalloc(&part.p, n * sizeof(double)));
  launch_history_vortex(cfg, u, p, static_cast<double *>(part.p), st);
  if (int rc = check_launch()) return rc;
  std::vector<double> h(n, 0.0);
  KM_HIP(hipMemcpyAsync(h.data(), part.p, n * sizeof(double), hipMemcpyDeviceToHost, st));
  KM_HIP(hipStreamSynchronize(st));
  const double area = (mesh.xmax[0] - mesh.xmin[0]) * (mesh.xmax[1] - mesh.xmin[1]);
  for (int v = 0; v < nv; ++v) {
    double sum = 0.0;
    for (int b = 0; b < u.nblocks; ++b) sum += h[(size_t)b * nv + v];
    out[v] = sum / area;
  }
  return KM_OK;
}

// one history row of the current base state: time, dt, cycle, block count, variables
int history_row(km_sim_t &s) {
  double t = 0., dt = 0.;
  KM_HIP(hipMemcpyAsync(&t, s.time_d, sizeof(double), hipMemcpyDeviceToHost, s.stream));
  KM_HIP(hipMemcpyAsync(&dt, s.dt_d, sizeof(double), hipMemcpyDeviceToHost, s.stream));
  KM_HIP(hipStreamSynchronize(s.stream));
  std::fprintf(s.hist, "%.14e %.14e %d %d", t, dt, s.ncycle, s.nb);
  if (s.hist_vortex) {
    // the single-level hydro fused path keeps the conserved interiors only
    if (s.path == KM_PATH_FUSED && !s.amr && s.cfg.mhd == KM_MHD_OFF && !s.prims_from_ic) {
      launch_ghost_fill(s.grid, s.base.pack, 0, s.nc, s.nc, false, s.stream,
                        Instances{s.inst_d, s.grids_d});
      launch_prepare_primitive(s.cfg, s.base.pack, s.stream);
    }
    double e[8];
    if (int rc = vortex_errors(s.cfg, s.base.pack, s.mesh, s.hist_p, s.stream, e)) return rc;
    for (int v = 0; v < (s.cfg.mhd != KM_MHD_OFF ? 8 : 5); ++v) std::fprintf(s.hist, " %.14e", e[v]);
  }
  std::fputc('\n', s.hist);
  std::fflush(s.hist);
  s.hist_last_cycle = s.ncycle;
  return KM_OK;
}

// one cycle; dt_host is the step dt when the task path needs it by value
int run_cycle(km_sim_t &s, double dt_host) {
  const int n = s.cfg.nstages;
  const double *betas = n == 1 ? kBetaRK1 : n == 2 ? kBetaRK2 : kBetaRK3;
  for (int stage = 1; stage <= n; ++stage) {
    int rc;
    if (s.path == KM_PATH_TASKS || s.prims_from_ic)
      rc = task_stage(s, stage, betas[stage - 1], dt_host);
    else if (s.amr)
      rc = amr_fused_stage(s, stage, betas[stage - 1]);
    else if (s.cfg.mhd == KM_MHD_CT)
      rc = ct_fused_stage(s, stage, betas[stage - 1]);
    else
      rc = fused_stage(s, stage, betas[stage - 1]);
    s.prims_from_ic = false;
    if (rc) return rc;
  }
  if (s.adaptive)
    if (int rc = adapt_step(s)) return rc;
  hipLaunchKernelGGL(k_dt_update, dim3(nb64(s.ninst)), dim3(64), 0, s.stream, s.dt_d, s.time_d,
                     s.est_d, s.dt_keep ? 1 : 0, s.ninst, dt_rule(s.cfg), s.fail_d + 1, s.fail_d);
  s.dt_keep = false;
  s.ncycle += 1;
  return check_launch();
}

// The device flag word, read after the cycles of an advance: a non-physical zone in a
// last-stage state (CFL pass) or dt outside [dt_min, dt_max] for too many cycles makes the
// advance fail, as the reference's run would (kamayan_driver.cpp:44-53). The flag stays set
// until the next km_sim_set_primitive.
int check_failure(km_sim_t &s) {
  int f = 0;
  KM_HIP(hipMemcpy(&f, s.fail_d, sizeof(int), hipMemcpyDeviceToHost));
  if (!f) return KM_OK;
  std::string m = "simulation failed by cycle " + std::to_string(s.ncycle) + ":";
  if (f & KM_FAIL_STATE) m += " a zone has density or pressure <= 0 or a non-finite value;";
  if (f & KM_FAIL_DT_MIN)
    m += " dt stayed below parthenon/time/dt_min for " +
         std::to_string(s.cfg.dt_min_cycle_limit) + " cycles;";
  if (f & KM_FAIL_DT_MAX)
    m += " dt stayed above parthenon/time/dt_max for " +
         std::to_string(s.cfg.dt_max_cycle_limit) + " cycles;";
  m.pop_back();
  set_error(m);
  return KM_ESTATE;
}

}  // namespace

extern "C" {

const char *km_last_error(void) { return g_err.c_str(); }
const char *km_version(void) { return "kamayan_amd 0.1 (gfx950)"; }

// ---- memory helpers ----
int km_device_count(int32_t *count) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  *count = e == hipSuccess ? n : 0;
  return KM_OK;
}
int km_malloc(void **ptr, size_t bytes) {
  KM_HIP(hipMalloc(ptr, bytes));
  return KM_OK;
}
int km_free(void *ptr) {
  KM_HIP(hipFree(ptr));
  return KM_OK;
}
int km_memcpy_h2d(void *dst, const void *src, size_t bytes) {
  KM_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return KM_OK;
}
int km_memcpy_d2h(void *dst, const void *src, size_t bytes) {
  KM_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return KM_OK;
}
int km_copy_device(void *dst, const void *src, size_t bytes, void *stream) {
  if (!dst || !src || bytes % 16 != 0 || (reinterpret_cast<uintptr_t>(dst) % 16) ||
      (reinterpret_cast<uintptr_t>(src) % 16)) {
    set_error("km_copy_device needs 16-byte aligned pointers and a multiple of 16 bytes");
    return KM_EINVAL;
  }
  const long n = (long)(bytes / 16);
  if (n == 0) return KM_OK;
  const long want = (n + 255) / 256;
  const unsigned grid = (unsigned)(want < 8192 ? want : 8192);
  hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, as_stream(stream),
                     static_cast<double2 *>(dst), static_cast<const double2 *>(src), n);
  return check_launch();
}
int km_synchronize(void) {
  KM_HIP(hipDeviceSynchronize());
  return KM_OK;
}

// ---- tasks ----
// flux tables for every active dimension, with a slot per conserved field and species
static int check_flux_tables(const km_cfg_t *cfg, const km_pack_t *u) {
  for (int d = 0; d < u->ndim; ++d)
    if (!u->flux[d]) {
      set_error("flux tables missing");
      return KM_EINVAL;
    }
  if (u->nflux < KM_NCONS(cfg->mhd) + KM_NSPEC_SLOTS(cfg->nspecies)) {
    set_error("pack has fewer flux variables than the conserved fields and species need");
    return KM_EINVAL;
  }
  return KM_OK;
}

int km_calculate_fluxes(const km_cfg_t *cfg, const km_pack_t *u, void *stream) {
  if (int rc = check_cfg(cfg, u)) return rc;
  if (int rc = check_flux_tables(cfg, u)) return rc;
  if (cfg->mhd == KM_MHD_CT)
    for (int d = 0; d < u->ndim; ++d)
      if (!u->face[d]) {
        set_error("CT needs face B tables");
        return KM_EINVAL;
      }
  launch_calculate_fluxes(*cfg, *u, flux_box(*cfg, *u), as_stream(stream));
  return check_launch();
}

int km_calculate_emf(const km_cfg_t *cfg, const km_pack_t *u, void *stream) {
  if (int rc = check_cfg(cfg, u)) return rc;
  if (cfg->mhd != KM_MHD_CT || u->ndim < 2) return KM_OK;
  if (!u->edge[2] || (u->ndim > 2 && (!u->edge[0] || !u->edge[1]))) {
    set_error("edge tables missing");
    return KM_EINVAL;
  }
  launch_calculate_emf(*cfg, *u, as_stream(stream));
  return check_launch();
}

int km_fluxes_to_dudt(const km_cfg_t *cfg, const km_pack_t *u, const km_pack_t *dudt,
                      void *stream) {
  if (int rc = check_cfg(cfg, u)) return rc;
  if (int rc = check_flux_tables(cfg, u)) return rc;
  if (!dudt) return KM_EINVAL;
  if (dudt->nvar < KM_NVAR(cfg->mhd) + KM_NSPEC_SLOTS(cfg->nspecies)) {
    set_error("dU/dt pack has fewer cell variables than the state");
    return KM_EINVAL;
  }
  launch_fluxes_to_dudt(*cfg, *u, *dudt, as_stream(stream));
  return check_launch();
}

int km_add_source_terms(const km_cfg_t *cfg, const km_pack_t *u, const km_pack_t *dudt,
                        void *stream) {
  if (int rc = check_cfg(cfg, u)) return rc;
  if (!dudt) return KM_EINVAL;
  launch_add_source_terms(*cfg, *u, *dudt, as_stream(stream));
  return check_launch();
}

int km_prepare_conserved_stage(const km_cfg_t *cfg, const km_pack_t *u, void *stream) {
  if (int rc = check_cfg(cfg, u)) return rc;
  launch_prepare_conserved_stage(*cfg, *u, as_stream(stream));
  return check_launch();
}

int km_apply_dudt(const km_cfg_t *cfg, const km_pack_t *base, const km_pack_t *u0,
                  const km_pack_t *u1, const km_pack_t *dudt, double beta, double dt,
                  void *stream) {
  if (int rc = check_cfg(cfg, u0)) return rc;
  if (!base || !u1 || !dudt) return KM_EINVAL;
  launch_apply_dudt(*cfg, *base, *u0, *u1, *dudt, beta, dt, as_stream(stream));
  return check_launch();
}

int km_prepare_primitive(const km_cfg_t *cfg, const km_pack_t *u, void *stream) {
  if (int rc = check_cfg(cfg, u)) return rc;
  launch_prepare_primitive(*cfg, *u, as_stream(stream));
  return check_launch();
}

int km_prepare_conserved(const km_cfg_t *cfg, const km_pack_t *u, void *stream) {
  if (int rc = check_cfg(cfg, u)) return rc;
  launch_prepare_conserved(*cfg, *u, as_stream(stream));
  return check_launch();
}

int km_estimate_dt(const km_cfg_t *cfg, const km_pack_t *u, double *dt_dev, void *stream) {
  if (int rc = check_cfg(cfg, u)) return rc;
  if (!dt_dev) return KM_EINVAL;
  launch_estimate_dt(*cfg, *u, dt_dev, as_stream(stream));
  return check_launch();
}

int km_fill_divb(const km_cfg_t *cfg, const km_pack_t *u, void *stream) {
  if (int rc = check_cfg(cfg, u)) return rc;
  if (cfg->mhd != KM_MHD_CT) return KM_OK;
  launch_fill_divb(*cfg, *u, as_stream(stream));
  return check_launch();
}

int km_ghost_plan_uniform(const km_pack_t *shape, const int32_t nb[3], const int32_t bc[6],
                          km_ghost_plan_t **plan) {
  if (!shape || !nb || !bc || !plan) return KM_EINVAL;
  if ((long)nb[0] * nb[1] * nb[2] != shape->nblocks) {
    set_error("block grid does not match the pack's block count");
    return KM_EINVAL;
  }
  for (int d = 0; d < shape->ndim; ++d) {
    const bool per_lo = bc[2 * d] == KM_BC_PERIODIC, per_hi = bc[2 * d + 1] == KM_BC_PERIODIC;
    if (per_lo != per_hi) {
      set_error("periodic boundaries must be paired");
      return KM_EINVAL;
    }
    if (shape->ng > shape->nx[d] && nb[d] > 1) {
      set_error("ghost width larger than a block");
      return KM_EINVAL;
    }
  }
  for (int d = 0; d < 6; ++d)
    if (bc[d] < KM_BC_PERIODIC || bc[d] > KM_BC_AXISYMMETRIC ||
        (bc[d] == KM_BC_AXISYMMETRIC && d != 0)) {
      set_error("boundary codes: periodic, outflow, reflect; axisymmetric at ix1 only");
      return KM_EINVAL;
    }
  auto *p = new km_ghost_plan_t();
  for (int d = 0; d < 3; ++d) p->grid.nb[d] = nb[d];
  for (int d = 0; d < 6; ++d) p->grid.bc[d] = bc[d];
  p->grid.first = 0;  // the pack holds this grid's blocks from block 0
  *plan = p;
  return KM_OK;
}

int km_ghost_plan_destroy(km_ghost_plan_t *plan) {
  delete plan;
  return KM_OK;
}

int km_ghost_plan_amr(const km_pack_t *shape, const km_mesh_desc_t *mesh,
                      const int32_t *leaves, int32_t nleaves, km_ghost_plan_t **plan) {
  return km_ghost_plan_amr_geom(shape, mesh, leaves, nleaves, KM_GEOM_CARTESIAN, plan);
}

int km_ghost_plan_amr_geom(const km_pack_t *shape, const km_mesh_desc_t *mesh,
                           const int32_t *leaves, int32_t nleaves, int32_t geometry,
                           km_ghost_plan_t **plan) {
  if (!shape || !mesh || !leaves || !plan) return KM_EINVAL;
  if (nleaves != shape->nblocks || shape->ndim != mesh->ndim || shape->ng != mesh->ng) {
    set_error("pack shape does not match the mesh and its leaves");
    return KM_EINVAL;
  }
  for (int d = 0; d < shape->ndim; ++d)
    if (shape->nx[d] != mesh->nx[d]) {
      set_error("pack block size does not match the mesh");
      return KM_EINVAL;
    }
  if (geometry != KM_GEOM_CARTESIAN && geometry != KM_GEOM_CYLINDRICAL) {
    set_error("dispatch parm [Geometry] not handled");
    return KM_EINVAL;
  }
  AmrTreeHost t;
  const std::string err = build_amr_tree(*mesh, leaves, nleaves, t);
  if (!err.empty()) {
    set_error(err);
    return KM_EINVAL;
  }
  t.cyl = geometry == KM_GEOM_CYLINDRICAL;
  auto p = std::make_unique<km_ghost_plan_t>();
  p->amr = true;
  KM_HIP(hipMalloc(&p->lookup_d, t.lookup.size() * sizeof(int)));
  KM_HIP(hipMemcpy(p->lookup_d, t.lookup.data(), t.lookup.size() * sizeof(int),
                   hipMemcpyHostToDevice));
  p->agrid = t.grid(p->lookup_d);
  *plan = p.release();
  return KM_OK;
}

int km_ghost_fill(const km_ghost_plan_t *plan, const km_pack_t *u, int32_t v0, int32_t v1,
                  int32_t with_faces, void *stream) {
  if (!plan || !u || v0 < 0 || v1 > u->nvar || v0 > v1) return KM_EINVAL;
  // a pack with face tables is an MHD pack: slots 5..7 hold the cell B vector
  const int nc = u->face[0] ? 8 : 5;
  if (with_faces && (!u->face[0] || !u->face[1])) {
    set_error("face tables missing");
    return KM_EINVAL;
  }
  if (plan->amr) {
    if (with_faces && (u->ng % 2 || !u->face[2])) {
      set_error("face fields across levels need an even nghost and all three face tables");
      return KM_EINVAL;
    }
    launch_amr_ghost_fill(plan->agrid, *u, v0, v1, nc, as_stream(stream));
    if (with_faces) launch_amr_ghost_faces(plan->agrid, *u, as_stream(stream));
  } else {
    launch_ghost_fill(plan->grid, *u, v0, v1, nc, with_faces != 0, as_stream(stream));
  }
  return check_launch();
}

int km_flux_correct(const km_ghost_plan_t *plan, const km_pack_t *u, void *stream) {
  if (!plan || !u) return KM_EINVAL;
  if (!plan->amr) return KM_OK;  // one level: nothing to correct
  for (int d = 0; d < u->ndim; ++d)
    if (!u->flux[d]) {
      set_error("flux tables missing");
      return KM_EINVAL;
    }
  launch_amr_flux_correct(plan->agrid, *u, u->nflux, as_stream(stream));
  // CT: the edge EMFs too (E3 always, E1 and E2 in 3-D)
  if (u->edge[2] && (u->ndim < 3 || (u->edge[0] && u->edge[1])))
    launch_amr_emf_correct(plan->agrid, *u, as_stream(stream));
  return check_launch();
}

// ---- whole-mesh driver ----
// Face-neighbour table [nb][3][2] of independent uniform block grids stored one after the
// other, each x-fastest: a block index (periodic wraps included) or KM_NBR_OUTFLOW /
// KM_NBR_REFLECT at a physical boundary (mapped inside the fused kernels).
static std::vector<int> uniform_neighbours(const km_mesh_desc_t *meshes, int ninst, int nb) {
  std::vector<int> nbr((size_t)nb * 6);
  int first = 0;
  for (int m = 0; m < ninst; ++m) {
    const km_mesh_desc_t &me = meshes[m];
    const int nbm = me.nb[0] * me.nb[1] * me.nb[2];
    for (int lb = 0; lb < nbm; ++lb) {
      const int l[3] = {lb % me.nb[0], (lb / me.nb[0]) % me.nb[1], lb / (me.nb[0] * me.nb[1])};
      for (int d = 0; d < 3; ++d)
        for (int side = 0; side < 2; ++side) {
          int mm[3] = {l[0], l[1], l[2]};
          mm[d] += side ? 1 : -1;
          int val;
          if (d >= me.ndim) {
            val = KM_NBR_OUTFLOW;
          } else if (mm[d] < 0 || mm[d] >= me.nb[d]) {
            const int bc = me.bc[2 * d + side];
            if (bc == KM_BC_PERIODIC) {
              mm[d] = (mm[d] + me.nb[d]) % me.nb[d];
              val = first + (mm[2] * me.nb[1] + mm[1]) * me.nb[0] + mm[0];
            } else {
              val = bc == KM_BC_REFLECT ? KM_NBR_REFLECT : KM_NBR_OUTFLOW;
            }
          } else {
            val = first + (mm[2] * me.nb[1] + mm[1]) * me.nb[0] + mm[0];
          }
          nbr[(size_t)(first + lb) * 6 + d * 2 + side] = val;
        }
    }
    first += nbm;
  }
  return nbr;
}

static int sim_create(const km_cfg_t *cfg, const km_mesh_desc_t *meshes, int32_t ninst,
                      int32_t path, const AmrTreeHost *tree, km_sim_t **out) {
  if (!cfg || !meshes || !out || ninst < 1) return KM_EINVAL;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    set_error("no GPU visible");
    return KM_ENODEV;
  }
  const km_mesh_desc_t *mesh = &meshes[0];
  for (int m = 1; m < ninst; ++m) {
    bool same = meshes[m].ndim == mesh->ndim && meshes[m].ng == mesh->ng;
    for (int d = 0; d < mesh->ndim; ++d) same = same && meshes[m].nx[d] == mesh->nx[d];
    if (!same) {
      set_error("ensemble instances must share ndim, nghost and cells per block");
      return KM_EINVAL;
    }
  }
  auto s = std::make_unique<km_sim_t>();
  s->cfg = *cfg;
  s->cfg.ndim = mesh->ndim;
  s->mesh = *mesh;
  s->path = path;
  s->ninst = ninst;
  s->meshes.assign(meshes, meshes + ninst);
  if (path == KM_PATH_FUSED && cfg->mhd != KM_MHD_OFF && (mesh->ndim != 3 || tree)) {
    set_error("the fused path runs CT MHD on 3-D single-level meshes; 1-D and 2-D CT and "
              "block trees with CT run on the task path");
    return KM_EINVAL;
  }
  if (path == KM_PATH_FUSED && cfg->nspecies > 1) {
    set_error("the fused stage path is built for one species; more run on the task path");
    return KM_EINVAL;
  }
  if (cfg->geometry == KM_GEOM_CYLINDRICAL && path == KM_PATH_FUSED) {
    set_error("cylindrical geometry runs on the task path");
    return KM_EINVAL;
  }
  for (int m = 0; m < ninst; ++m)
    for (int d = 1; d < 6; ++d)
      if (meshes[m].bc[d] == KM_BC_AXISYMMETRIC) {
        set_error("the axisymmetric boundary is the inner x1 boundary (ix1_bc) only");
        return KM_EINVAL;
      }
  if (meshes[0].bc[0] == KM_BC_AXISYMMETRIC && cfg->geometry != KM_GEOM_CYLINDRICAL) {
    set_error("the axisymmetric boundary needs cylindrical geometry");
    return KM_EINVAL;
  }
  if (cfg->nstages < 1 || cfg->nstages > 3) {
    set_error("integrator must be rk1, rk2 or rk3");
    return KM_EINVAL;
  }
  s->nb = 0;
  std::vector<int> first(ninst);
  for (int m = 0; m < ninst; ++m) {
    first[m] = s->nb;
    s->nb += meshes[m].nb[0] * meshes[m].nb[1] * meshes[m].nb[2];
  }
  if (tree) s->nb = (int)tree->blocks.size();
  s->nvar = KM_NVAR(cfg->mhd) + KM_NSPEC_SLOTS(cfg->nspecies);
  s->nc = KM_NCONS(cfg->mhd);
  s->nflux = s->nc + KM_NSPEC_SLOTS(cfg->nspecies);
  km_pack_t shape;
  std::memset(&shape, 0, sizeof(shape));
  shape.nblocks = s->nb;
  shape.ndim = mesh->ndim;
  shape.ng = mesh->ng;
  for (int d = 0; d < 3; ++d) shape.nx[d] = d < mesh->ndim ? mesh->nx[d] : 1;
  shape.nvar = s->nvar;
  shape.nflux = s->nflux;
  if (int rc = check_cfg(&s->cfg, &shape)) return rc;
  const Geo g(shape);
  s->ncell = g.ncell();
  s->nstag = g.nstag();
  s->grids_h.resize(ninst);
  for (int m = 0; m < ninst; ++m) {
    for (int d = 0; d < 3; ++d) s->grids_h[m].nb[d] = meshes[m].nb[d];
    for (int d = 0; d < 6; ++d) s->grids_h[m].bc[d] = meshes[m].bc[d];
    s->grids_h[m].first = first[m];
    for (int d = 0; d < meshes[m].ndim; ++d) {
      const bool plo = meshes[m].bc[2 * d] == KM_BC_PERIODIC;
      const bool phi = meshes[m].bc[2 * d + 1] == KM_BC_PERIODIC;
      if (plo != phi) {
        set_error("periodic boundaries must be paired");
        return KM_EINVAL;
      }
    }
  }
  s->grid = s->grids_h[0];
  KM_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
  if (s->path == KM_PATH_FUSED) {  // a uniform mesh may become a block tree (adaptive)
    KM_HIP(hipStreamCreateWithFlags(&s->side, hipStreamNonBlocking));
    KM_HIP(hipEventCreateWithFlags(&s->fork_ev, hipEventDisableTiming));
    KM_HIP(hipEventCreateWithFlags(&s->join_ev, hipEventDisableTiming));
  }
  std::vector<km_block_t> blocks(s->nb);
  if (tree)
    blocks = tree->blocks;
  else
    for (int m = 0; m < ninst; ++m) km_mesh_blocks(&meshes[m], blocks.data() + first[m]);
  int rc;
  if (tree) s->target = &s->mesh_bufs;
  if ((rc = s->alloc(&s->blocks_d, s->nb))) return rc;
  KM_HIP(hipMemcpy(s->blocks_d, blocks.data(), s->nb * sizeof(km_block_t), hipMemcpyHostToDevice));
  if (tree) {
    int *lookup_d = nullptr;
    if ((rc = s->alloc(&lookup_d, tree->lookup.size()))) return rc;
    KM_HIP(hipMemcpy(lookup_d, tree->lookup.data(), tree->lookup.size() * sizeof(int),
                     hipMemcpyHostToDevice));
    s->amr = true;
    s->agrid = tree->grid(lookup_d);
  }
  if (ninst > 1) {
    std::vector<int> inst(s->nb);
    for (int m = 0; m < ninst; ++m)
      for (int b = first[m]; b < first[m] + meshes[m].nb[0] * meshes[m].nb[1] * meshes[m].nb[2]; ++b)
        inst[b] = m;
    if ((rc = s->alloc(&s->inst_d, s->nb))) return rc;
    KM_HIP(hipMemcpy(s->inst_d, inst.data(), inst.size() * sizeof(int), hipMemcpyHostToDevice));
    if ((rc = s->alloc(&s->grids_d, ninst))) return rc;
    KM_HIP(hipMemcpy(s->grids_d, s->grids_h.data(), ninst * sizeof(UniformGrid),
                     hipMemcpyHostToDevice));
  }
  if ((rc = make_cont(*s, s->base, true, nullptr, true, false))) return rc;
  s->own_prims = path == KM_PATH_FUSED && tree && s->cfg.nstages > 1;
  for (int i = 0; i < 2; ++i)
    if ((rc = make_cont(*s, s->stage[i], s->own_prims, s->own_prims ? nullptr : &s->base, true,
                        false)))
      return rc;
  if ((rc = make_cont(*s, s->dudt, false, nullptr, true, false))) return rc;
  if ((rc = make_cont(*s, s->scratch, false, nullptr, false, true))) return rc;
  if (path == KM_PATH_FUSED) {
    const long nint = (long)shape.nx[0] * shape.nx[1] * shape.nx[2];
    if ((rc = s->alloc(&s->du, (size_t)s->nb * 5 * nint))) return rc;
  }
  if (path == KM_PATH_FUSED && tree)
    if ((rc = build_fill_list(*s))) return rc;
  if (path == KM_PATH_FUSED && !tree) {
    const std::vector<int> nbr = uniform_neighbours(meshes, ninst, s->nb);
    if ((rc = s->alloc(&s->nbr_d, nbr.size()))) return rc;
    KM_HIP(hipMemcpy(s->nbr_d, nbr.data(), nbr.size() * sizeof(int), hipMemcpyHostToDevice));
  }
  s->target = &s->bufs;
  if ((rc = s->alloc(&s->dt_d, 3 * (size_t)ninst))) return rc;
  s->est_d = s->dt_d + ninst;
  s->time_d = s->dt_d + 2 * ninst;
  if ((rc = s->alloc(&s->fail_d, 1 + 2 * (size_t)ninst))) return rc;
  *out = s.release();
  return KM_OK;
}

int km_sim_create_ensemble(const km_cfg_t *cfg, const km_mesh_desc_t *meshes, int32_t ninst,
                           int32_t path, km_sim_t **out) {
  return sim_create(cfg, meshes, ninst, path, nullptr, out);
}

int km_sim_create(const km_cfg_t *cfg, const km_mesh_desc_t *mesh, int32_t path,
                  km_sim_t **out) {
  return sim_create(cfg, mesh, 1, path, nullptr, out);
}

int km_sim_create_amr(const km_cfg_t *cfg, const km_mesh_desc_t *mesh, const int32_t *leaves,
                      int32_t nleaves, int32_t path, km_sim_t **out) {
  if (!cfg || !mesh || !leaves || !out) return KM_EINVAL;
  if (path != KM_PATH_TASKS && path != KM_PATH_FUSED) {
    set_error("unknown path");
    return KM_EINVAL;
  }
  AmrTreeHost t;
  const std::string err = build_amr_tree(*mesh, leaves, nleaves, t);
  if (!err.empty()) {
    set_error(err);
    return KM_EINVAL;
  }
  if (cfg->mhd == KM_MHD_CT && t.lmax > 0 && mesh->ng % 2) {
    set_error("constrained transport across refinement levels needs an even nghost");
    return KM_EINVAL;
  }
  t.cyl = cfg->geometry == KM_GEOM_CYLINDRICAL;
  const int rc = sim_create(cfg, mesh, 1, path, &t, out);
  if (rc == KM_OK) {
    (*out)->leaves.assign(leaves, leaves + 4 * (size_t)nleaves);
    (*out)->deref.assign(nleaves, 0);
  }
  return rc;
}

int km_sim_ensemble_info(const km_sim_t *s, double *times, double *dts, int32_t n) {
  if (!s || n < 1 || n > s->ninst) return KM_EINVAL;
  std::vector<double> h(3 * (size_t)s->ninst);
  KM_HIP(hipStreamSynchronize(s->stream));
  KM_HIP(hipMemcpy(h.data(), s->dt_d, h.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int m = 0; m < n; ++m) {
    if (dts) dts[m] = h[m];
    if (times) times[m] = h[2 * s->ninst + m];
  }
  return KM_OK;
}

int km_sim_destroy(km_sim_t *s) {
  if (s) (void)hipStreamSynchronize(s->stream);
  delete s;
  return KM_OK;
}

static int copy_state(km_sim_t *s, const double *cell, const double *const face[3],
                      double *cell_out, double *const face_out[3]) {
  for (int b = 0; b < s->nb; ++b) {
    for (int v = 0; v < s->nvar; ++v) {
      double *dev = s->base.cell_h[(size_t)b * s->nvar + v];
      const size_t off = ((size_t)b * s->nvar + v) * s->ncell;
      if (cell)
        KM_HIP(hipMemcpy(dev, cell + off, s->ncell * sizeof(double), hipMemcpyHostToDevice));
      if (cell_out)
        KM_HIP(hipMemcpy(cell_out + off, dev, s->ncell * sizeof(double), hipMemcpyDeviceToHost));
    }
    if (s->cfg.mhd == KM_MHD_CT)
      for (int d = 0; d < 3; ++d) {
        double *dev = s->base.face_h[d][b];
        const size_t off = (size_t)b * s->nstag;
        if (face && face[d])
          KM_HIP(hipMemcpy(dev, face[d] + off, s->nstag * sizeof(double), hipMemcpyHostToDevice));
        if (face_out && face_out[d])
          KM_HIP(hipMemcpy(face_out[d] + off, dev, s->nstag * sizeof(double),
                           hipMemcpyDeviceToHost));
      }
  }
  return KM_OK;
}

int km_sim_set_primitive(km_sim_t *s, const double *cell, const double *const face[3]) {
  if (!s || !cell) return KM_EINVAL;
  KM_HIP(hipStreamSynchronize(s->stream));
  if (int rc = copy_state(s, cell, face, nullptr, nullptr)) return rc;
  launch_prepare_conserved(s->cfg, s->base.pack, s->stream);
  ghost_fill(*s, s->base.pack, 0, KM_TEMP(s->cfg.mhd), s->cfg.mhd == KM_MHD_CT);
  if (s->nvar > KM_NVAR(s->cfg.mhd))
    ghost_fill(*s, s->base.pack, KM_NVAR(s->cfg.mhd), s->nvar, false);
  hipLaunchKernelGGL(k_set_int, dim3(nb64(1 + 2 * s->ninst)), dim3(64), 0, s->stream, s->fail_d,
                     0, 1 + 2 * s->ninst);
  launch_estimate_dt(s->cfg, s->base.pack, s->est_d, s->stream, s->inst_d, s->ninst, s->fail_d);
  hipLaunchKernelGGL(k_dt_init, dim3(nb64(s->ninst)), dim3(64), 0, s->stream, s->dt_d, s->est_d,
                     dt_rule(s->cfg), s->ninst);
  hipLaunchKernelGGL(k_set, dim3(nb64(s->ninst)), dim3(64), 0, s->stream, s->time_d, 0.0,
                     s->ninst);
  if (int rc = check_launch()) return rc;
  KM_HIP(hipStreamSynchronize(s->stream));
  s->ncycle = 0;
  s->prims_from_ic = true;
  s->dt_keep = false;
  if (s->graph_exec) {
    (void)hipGraphExecDestroy(s->graph_exec);
    (void)hipGraphDestroy(s->graph);
    s->graph_exec = nullptr;
    s->graph = nullptr;
  }
  return KM_OK;
}

int km_sim_get_state(km_sim_t *s, double *cell, double *const face[3]) {
  if (!s || !cell) return KM_EINVAL;
  KM_HIP(hipStreamSynchronize(s->stream));
  if (s->lazy_ghosts || s->lazy_prims) {
    fill_lazy_ghosts(*s);
    KM_HIP(hipStreamSynchronize(s->stream));  // the copies below run on another stream
  }
  if (s->path == KM_PATH_FUSED && !s->amr && s->cfg.mhd == KM_MHD_OFF) {
    // the single-level hydro fused path keeps only the conserved interior current; fill the
    // ghost zones and derive the primitives
    launch_ghost_fill(s->grid, s->base.pack, 0, s->nc, s->nc, false, s->stream,
                      Instances{s->inst_d, s->grids_d});
    launch_prepare_primitive(s->cfg, s->base.pack, s->stream);
    KM_HIP(hipStreamSynchronize(s->stream));
  }
  return copy_state(s, nullptr, nullptr, cell, face);
}

int km_sim_set_dt(km_sim_t *s, double dt) {
  if (!s) return KM_EINVAL;
  KM_HIP(hipStreamSynchronize(s->stream));
  KM_HIP(hipMemcpy(s->dt_d, &dt, sizeof(double), hipMemcpyHostToDevice));
  s->dt_keep = true;
  return KM_OK;
}

int km_sim_advance(km_sim_t *s, int32_t ncycles, double tlim) {
  if (!s) return KM_EINVAL;
  for (int n = 0; n < ncycles; ++n) {
    if (!s->prims_from_ic &&
        (s->kern[0] < 0 || s->kern[1] < 0 || s->kern[2] < 0 || s->stage_kern < 0))
      if (int rc = autotune(*s, fused_args(*s, 1, 1.0))) return rc;
    // a cycle needs the host when it reads dt there (tlim, an override), writes history,
    // starts from the initial primitives or may remesh; every other cycle of either path
    // is one captured graph replayed (the task path keeps dt on the device as well)
    const bool need_host = s->adaptive || s->prims_from_ic || s->dt_keep || tlim > 0. || s->hist;
    if (need_host) {
      double dt = 0., time = 0.;
      KM_HIP(hipMemcpyAsync(&dt, s->dt_d, sizeof(double), hipMemcpyDeviceToHost, s->stream));
      KM_HIP(hipMemcpyAsync(&time, s->time_d, sizeof(double), hipMemcpyDeviceToHost, s->stream));
      KM_HIP(hipStreamSynchronize(s->stream));
      if (tlim > 0. && s->ninst > 1) {
        set_error("tlim is not supported for ensembles; advance by cycles");
        return KM_EINVAL;
      }
      if (tlim > 0.) {
        if (time >= tlim) break;
        if (time + dt > tlim) {
          dt = tlim - time;
          KM_HIP(hipMemcpy(s->dt_d, &dt, sizeof(double), hipMemcpyHostToDevice));
        }
      }
      if (int rc = run_cycle(*s, dt)) return rc;
      if (s->hist) {  // history rows: at the next output time, and at tlim
        double t = 0.;
        KM_HIP(hipMemcpyAsync(&t, s->time_d, sizeof(double), hipMemcpyDeviceToHost, s->stream));
        KM_HIP(hipStreamSynchronize(s->stream));
        if (t >= s->hist_next) {
          if (int rc = history_row(*s)) return rc;
          s->hist_next += s->hist_dt;
        } else if (tlim > 0. && t >= tlim) {
          if (int rc = history_row(*s)) return rc;
        }
      }
      continue;
    }
    // no host decisions: capture one cycle in a graph and replay it (not with
    // parthenon/mesh/pack_size, whose packs are launched from the host as the reference's
    // task list does)
    if (s->use_graph && !s->profile && s->cfg.pack_size <= 0) {
      if (!s->graph_exec) {
        KM_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeRelaxed));
        const int nc0 = s->ncycle;
        int rc = run_cycle(*s, 0.0);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(s->stream, &g);
        if (rc) return rc;
        if (e != hipSuccess) {
          set_error(std::string("graph capture: ") + hipGetErrorString(e));
          return KM_EHIP;
        }
        s->graph = g;
        KM_HIP(hipGraphInstantiate(&s->graph_exec, g, nullptr, nullptr, 0));
        s->ncycle = nc0;  // captured, not yet run
      }
      KM_HIP(hipGraphLaunch(s->graph_exec, s->stream));
      s->ncycle += 1;
    } else {
      if (int rc = run_cycle(*s, 0.0)) return rc;
    }
  }
  // replayed cycles skip the host code that marks the momentum and energy ghosts stale
  if ((s->amr || s->cfg.mhd == KM_MHD_CT) && s->path == KM_PATH_FUSED && ncycles > 0)
    s->lazy_ghosts = true;
  if (((s->amr && s->own_prims) || s->cfg.mhd == KM_MHD_CT) && s->path == KM_PATH_FUSED &&
      ncycles > 0)
    s->lazy_prims = true;
  KM_HIP(hipStreamSynchronize(s->stream));
  return check_failure(*s);
}

int km_sim_set_fused_kernel(km_sim_t *s, int32_t dir, int32_t kern) {
  if (!s) return KM_EINVAL;
  if (dir == 3) {  // the whole stage
    const int nk = stage3_kernels(s->cfg, s->mesh.ndim);
    if (kern < -1 || kern > nk || (kern > 0 && !stage3_applies(*s))) {
      set_error("stage kernel: -1 (autotune), 0 (the sweeps) or a one-launch variant 1.." +
                std::to_string(nk) + " (fused path, 2-D or 3-D hydro, single-level meshes)");
      return KM_EINVAL;
    }
    KM_HIP(hipStreamSynchronize(s->stream));
    s->stage_kern = kern;
    s->drop_graph();
    return KM_OK;
  }
  if (dir < -1 || dir > 2 || kern < -1 || kern > tile_shapes(s->cfg)) {
    set_error("fused sweep kernel: direction 0..2 (or -1 for all) and kernel -1 (autotune), "
              "0 (march) or a pencil-tile shape 1.." + std::to_string(tile_shapes(s->cfg)));
    return KM_EINVAL;
  }
  if (kern == 0 && s->amr && s->path == KM_PATH_FUSED) {
    set_error("the pencil march reads face neighbours of a single-level mesh; block trees "
              "run the fused sweeps as pencil tiles");
    return KM_EINVAL;
  }
  if (kern == 0 && s->cfg.mhd == KM_MHD_CT && s->path == KM_PATH_FUSED) {
    set_error("the fused CT stage runs its sweeps as pencil tiles (1.." +
              std::to_string(tile_shapes(s->cfg)) + ")");
    return KM_EINVAL;
  }
  KM_HIP(hipStreamSynchronize(s->stream));
  for (int d = 0; d < 3; ++d)
    if (dir < 0 || d == dir) s->kern[d] = kern;
  // a fixed sweep kernel runs the stage as sweeps; autotuning picks the stage kernel too
  s->stage_kern = kern >= 0 ? 0 : -1;
  s->drop_graph();
  return KM_OK;
}

int km_sim_fused_kernel(km_sim_t *s, int32_t dir, int32_t *kern) {
  if (!s || !kern || dir < 0 || dir > 3) return KM_EINVAL;
  *kern = dir == 3 ? s->stage_kern : s->kern[dir];
  return KM_OK;
}

int km_sim_profile(km_sim_t *s, int32_t on) {
  if (!s) return KM_EINVAL;
  KM_HIP(hipStreamSynchronize(s->stream));
  for (auto &r : s->recs) {
    s->ev_free.push_back(r.a);
    s->ev_free.push_back(r.b);
  }
  s->recs.clear();
  s->profile = on != 0;
  return KM_OK;
}

int km_sim_kernel_times(km_sim_t *s, double *ms, int32_t *count, int32_t nkinds) {
  if (!s || !ms || !count) return KM_EINVAL;
  KM_HIP(hipStreamSynchronize(s->stream));
  for (int k = 0; k < nkinds; ++k) {
    ms[k] = 0.0;
    count[k] = 0;
  }
  for (auto &r : s->recs) {
    if (r.kind >= nkinds) continue;
    float t = 0.f;
    KM_HIP(hipEventElapsedTime(&t, r.a, r.b));
    ms[r.kind] += t;
    count[r.kind] += 1;
  }
  return KM_OK;
}

int km_sim_set_adaptive(km_sim_t *s, const km_amr_adapt_t *a) {
  if (!s || !a) return KM_EINVAL;
  if (!s->amr) {
    set_error("adaptive refinement needs a block-tree simulation (km_sim_create_amr)");
    return KM_EINVAL;
  }
  if (s->mesh.ng < 2) {
    set_error("the Loehner estimator reads two ghost layers: nghost must be at least 2");
    return KM_EINVAL;
  }
  if (a->ncrit < 0 || a->ncrit > KM_AMR_MAX_CRIT || a->numlevel < 1 || a->numlevel > 16 ||
      a->derefine_count < 0) {
    set_error("bad adaptive settings");
    return KM_EINVAL;
  }
  for (int c = 0; c < a->ncrit; ++c)
    if (a->crit[c].var < 0 || a->crit[c].var >= s->nvar) {
      set_error("refinement field slot out of range");
      return KM_EINVAL;
    }
  for (int b = 0; b < s->nb; ++b)
    if (s->leaves[4 * (size_t)b] > a->numlevel - 1) {
      set_error("a leaf is finer than numlevel allows");
      return KM_EINVAL;
    }
  KM_HIP(hipStreamSynchronize(s->stream));
  if (!s->tags_d) {
    s->target = &s->mesh_bufs;
    int rc = s->alloc(&s->tags_d, s->nb);
    if (!rc) rc = s->alloc(&s->err2_d, s->nb);
    s->target = &s->bufs;
    if (rc) return rc;
  }
  s->adaptive = true;
  s->adapt = *a;
  s->deref.assign(s->nb, 0);
  return KM_OK;
}

int km_sim_amr_leaves(const km_sim_t *s, int32_t *leaves, int32_t *nleaves) {
  if (!s || !nleaves || !s->amr) return KM_EINVAL;
  const int n = s->nb;
  const int cap = *nleaves;
  *nleaves = n;
  if (!leaves || cap < n) {
    set_error("leaf buffer holds " + std::to_string(cap) + " entries, " + std::to_string(n) +
              " needed");
    return KM_EINVAL;
  }
  std::memcpy(leaves, s->leaves.data(), s->leaves.size() * sizeof(int32_t));
  return KM_OK;
}

int km_history_vortex_errors(const km_cfg_t *cfg, const km_pack_t *u, const km_mesh_desc_t *mesh,
                             const double *params, int32_t nparams, double *errors,
                             void *stream) {
  if (int rc = check_cfg(cfg, u)) return rc;
  if (!mesh || !errors) return KM_EINVAL;
  return vortex_errors(*cfg, *u, *mesh, vortex_params(*cfg, params, nparams),
                       as_stream(stream), errors);
}

int km_sim_set_history(km_sim_t *s, const char *path, double dt, const char *problem,
                       const double *params, int32_t nparams) {
  if (!s || !path || !(dt > 0.)) return KM_EINVAL;
  if (s->ninst > 1) {
    set_error("history output is for single runs, not ensembles");
    return KM_EINVAL;
  }
  const std::string prob = problem ? problem : "";
  if (!prob.empty() && prob != "isentropic_vortex") {
    set_error("no history variables for problem '" + prob + "'");
    return KM_EINVAL;
  }
  if (s->hist) std::fclose(s->hist);
  s->hist = std::fopen(path, "w");
  if (!s->hist) {
    set_error(std::string("cannot open ") + path);
    return KM_EINVAL;
  }
  s->hist_vortex = prob == "isentropic_vortex";
  s->hist_p = vortex_params(s->cfg, params, nparams);
  s->hist_dt = dt;
  static const char *const names[8] = {"density error", "velx error", "vely error",
                                        "velz error", "pressure error", "magx error",
                                        "magy error", "magz error"};
  std::fprintf(s->hist, "# History data\n# [1]=time [2]=dt [3]=cycle [4]=nbtotal");
  if (s->hist_vortex)
    for (int v = 0; v < (s->cfg.mhd != KM_MHD_OFF ? 8 : 5); ++v)
      std::fprintf(s->hist, " [%d]=%s", v + 5, names[v]);
  std::fputc('\n', s->hist);
  KM_HIP(hipStreamSynchronize(s->stream));
  double t = 0.;
  KM_HIP(hipMemcpy(&t, s->time_d, sizeof(double), hipMemcpyDeviceToHost));
  if (int rc = history_row(*s)) return rc;
  s->hist_next = t + dt;
  return KM_OK;
}

// Per-block Loehner maxima of km_amr_tag_loehner_noalloc: one file-scope device array
// (each GPU holds its own copy for the life of the process). Host code uses the address
// hipGetSymbolAddress gives for the current device, never the array's name.
static __device__ __align__(256) unsigned long long scratch[KM_TAG_NOALLOC_MAX_BLOCKS];
static std::mutex g_scratch_mu;
static std::map<int, void *> g_scratch_ptr;  // device id -> address of `scratch` there

int km_amr_tag_loehner_noalloc(const km_pack_t *u, const km_amr_criterion_t *crit,
                               int32_t *tags_dev, int32_t reset, void *stream) {
  if (!u || !crit || !tags_dev || crit->var < 0 || crit->var >= u->nvar || u->ng < 2) {
    set_error("km_amr_tag_loehner_noalloc: bad pack, field slot or ghost width (needs 2)");
    return KM_EINVAL;
  }
  if (u->nblocks > KM_TAG_NOALLOC_MAX_BLOCKS) {
    set_error("km_amr_tag_loehner_noalloc: " + std::to_string(u->nblocks) +
              " blocks exceed the scratch limit of " +
              std::to_string(KM_TAG_NOALLOC_MAX_BLOCKS) + "; use km_amr_tag_loehner");
    return KM_EINVAL;
  }
  int dev = 0;
  KM_HIP(hipGetDevice(&dev));
  void *ptr = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    auto it = g_scratch_ptr.find(dev);
    if (it == g_scratch_ptr.end()) {
      KM_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
      g_scratch_ptr[dev] = ptr;
    } else {
      ptr = it->second;
    }
  }
  // the maxima are cleared by a fill kernel before the zones fold into them
  launch_amr_tag_loehner(*u, *crit, tags_dev, reset != 0, static_cast<double *>(ptr),
                         as_stream(stream));
  if (int rc = check_launch()) return rc;
  KM_HIP(hipStreamSynchronize(as_stream(stream)));
  return KM_OK;
}

int km_amr_tag_loehner(const km_pack_t *u, const km_amr_criterion_t *crit, int32_t *tags_dev,
                       int32_t reset, void *stream) {
  if (!u || !crit || !tags_dev || crit->var < 0 || crit->var >= u->nvar || u->ng < 2) {
    set_error("km_amr_tag_loehner: bad pack, field slot or ghost width (needs 2)");
    return KM_EINVAL;
  }
  DevBuf err2;  // per-block maxima, released after the stream drains
  KM_HIP(hipMalloc(&err2.p, u->nblocks * sizeof(double) + 256));
  launch_amr_tag_loehner(*u, *crit, tags_dev, reset != 0, static_cast<double *>(err2.p),
                         as_stream(stream));
  if (int rc = check_launch()) return rc;
  KM_HIP(hipStreamSynchronize(as_stream(stream)));
  return KM_OK;
}

int km_amr_remesh(const km_pack_t *old_u, const km_pack_t *new_u, const int32_t *origin_dev,
                  int32_t v0, int32_t v1, void *stream) {
  return km_amr_remesh_geom(old_u, new_u, origin_dev, v0, v1, nullptr, KM_GEOM_CARTESIAN,
                            stream);
}

int km_amr_remesh_geom(const km_pack_t *old_u, const km_pack_t *new_u, const int32_t *origin_dev,
                       int32_t v0, int32_t v1, const km_mesh_desc_t *mesh, int32_t geometry,
                       void *stream) {
  const bool cyl = geometry == KM_GEOM_CYLINDRICAL;
  if ((cyl && !mesh) || (geometry != KM_GEOM_CARTESIAN && !cyl)) {
    set_error("km_amr_remesh: cylindrical geometry needs the mesh description");
    return KM_EINVAL;
  }
  if (!old_u || !new_u || !origin_dev || old_u->ndim != new_u->ndim || old_u->ng != new_u->ng ||
      v0 < 0 || v1 > old_u->nvar || v1 > new_u->nvar) {
    set_error("km_amr_remesh: packs must share ndim, ng, nx and hold the slots");
    return KM_EINVAL;
  }
  for (int d = 0; d < 3; ++d)
    if (old_u->nx[d] != new_u->nx[d]) {
      set_error("km_amr_remesh: packs must share nx");
      return KM_EINVAL;
    }
  launch_amr_remesh(*old_u, *new_u, origin_dev, v0, v1, cyl, as_stream(stream));
  bool faces = true;
  for (int d = 0; d < 3; ++d) faces = faces && old_u->face[d] && new_u->face[d];
  if (faces)
    launch_amr_remesh_faces(*old_u, *new_u, origin_dev, cyl, cyl ? mesh->xmin[0] : 0.0,
                            as_stream(stream));
  return check_launch();
}

// ---- the fused stage on caller-owned packs ----
static int plan_checks(const km_cfg_t *cfg, const km_mesh_desc_t *mesh) {
  if (!cfg || !mesh) return KM_EINVAL;
  if (cfg->mhd != KM_MHD_OFF || cfg->nspecies > 1 || cfg->geometry != KM_GEOM_CARTESIAN) {
    set_error("the fused stage is built for Cartesian hydro with one species; MHD, species "
              "and cylindrical geometry run on the per-task entry points");
    return KM_EINVAL;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    set_error("no GPU visible");
    return KM_ENODEV;
  }
  return KM_OK;
}

static int plan_buffers(km_stage_plan_t &p) {
  const long nint = (long)p.mesh.nx[0] * (p.mesh.ndim > 1 ? p.mesh.nx[1] : 1) *
                    (p.mesh.ndim > 2 ? p.mesh.nx[2] : 1);
  KM_HIP(hipMalloc(&p.du, (size_t)p.nb * 5 * nint * sizeof(double) + 256));
  return KM_OK;
}

int km_stage_plan_uniform(const km_cfg_t *cfg, const km_mesh_desc_t *meshes, int32_t ninst,
                          km_stage_plan_t **plan) {
  if (!plan || ninst < 1) return KM_EINVAL;
  if (int rc = plan_checks(cfg, meshes)) return rc;
  auto p = std::make_unique<km_stage_plan_t>();
  p->cfg = *cfg;
  p->mesh = meshes[0];
  p->cfg.ndim = meshes[0].ndim;
  p->ninst = ninst;
  std::vector<int> inst;
  for (int m = 0; m < ninst; ++m) {
    bool same = meshes[m].ndim == p->mesh.ndim && meshes[m].ng == p->mesh.ng;
    for (int d = 0; d < p->mesh.ndim; ++d) same = same && meshes[m].nx[d] == p->mesh.nx[d];
    for (int d = 0; d < meshes[m].ndim; ++d)
      if ((meshes[m].bc[2 * d] == KM_BC_PERIODIC) != (meshes[m].bc[2 * d + 1] == KM_BC_PERIODIC))
        same = false;
    if (!same) {
      set_error("instances must share ndim, nghost and cells per block, with paired periodic "
                "boundaries");
      return KM_EINVAL;
    }
    const int nbm = meshes[m].nb[0] * meshes[m].nb[1] * meshes[m].nb[2];
    inst.insert(inst.end(), nbm, m);
    p->nb += nbm;
  }
  const std::vector<int> nbr = uniform_neighbours(meshes, ninst, p->nb);
  KM_HIP(hipMalloc(&p->nbr_d, nbr.size() * sizeof(int)));
  KM_HIP(hipMemcpy(p->nbr_d, nbr.data(), nbr.size() * sizeof(int), hipMemcpyHostToDevice));
  if (ninst > 1) {
    KM_HIP(hipMalloc(&p->inst_d, inst.size() * sizeof(int)));
    KM_HIP(hipMemcpy(p->inst_d, inst.data(), inst.size() * sizeof(int), hipMemcpyHostToDevice));
  }
  if (int rc = plan_buffers(*p)) return rc;
  *plan = p.release();
  return KM_OK;
}

int km_stage_plan_amr(const km_cfg_t *cfg, const km_mesh_desc_t *mesh, const int32_t *leaves,
                      int32_t nleaves, km_stage_plan_t **plan) {
  if (!plan || !leaves || nleaves < 1) return KM_EINVAL;
  if (int rc = plan_checks(cfg, mesh)) return rc;
  AmrTreeHost t;
  const std::string err = build_amr_tree(*mesh, leaves, nleaves, t);
  if (!err.empty()) {
    set_error(err);
    return KM_EINVAL;
  }
  auto p = std::make_unique<km_stage_plan_t>();
  p->cfg = *cfg;
  p->mesh = *mesh;
  p->cfg.ndim = mesh->ndim;
  p->nb = nleaves;
  p->amr = true;
  KM_HIP(hipMalloc(&p->lookup_d, t.lookup.size() * sizeof(int)));
  KM_HIP(hipMemcpy(p->lookup_d, t.lookup.data(), t.lookup.size() * sizeof(int),
                   hipMemcpyHostToDevice));
  p->agrid = t.grid(p->lookup_d);
  if (int rc = plan_buffers(*p)) return rc;
  *plan = p.release();
  return KM_OK;
}

int km_stage_plan_destroy(km_stage_plan_t *plan) {
  delete plan;
  return KM_OK;
}

int km_stage_plan_set_kernel(km_stage_plan_t *p, int32_t dir, int32_t kern) {
  if (!p) return KM_EINVAL;
  std::lock_guard<std::mutex> lock(p->mu);
  if (dir == 3) {
    const int nk = stage3_kernels(p->cfg, p->mesh.ndim);
    if (kern < -1 || kern > nk || (kern > 0 && !stage3_applies(*p))) {
      set_error("stage kernel: -1 (autotune), 0 (the sweeps) or a one-launch variant 1.." +
                std::to_string(nk) + " (2-D and 3-D single-level meshes)");
      return KM_EINVAL;
    }
    p->stage_kern = kern;
    return KM_OK;
  }
  if (dir < -1 || dir > 2 || kern < -1 || kern > tile_shapes(p->cfg) || (kern == 0 && p->amr)) {
    set_error("sweep kernel: direction 0..2 (or -1 for all) and kernel -1 (autotune), 0 (march, "
              "single-level meshes) or a pencil-tile shape 1.." + std::to_string(tile_shapes(p->cfg)));
    return KM_EINVAL;
  }
  for (int d = 0; d < 3; ++d)
    if (dir < 0 || d == dir) p->kern[d] = kern;
  p->stage_kern = kern >= 0 ? 0 : -1;
  return KM_OK;
}

int km_stage_plan_kernel(km_stage_plan_t *p, int32_t dir, int32_t *kern) {
  if (!p || !kern || dir < 0 || dir > 3) return KM_EINVAL;
  std::lock_guard<std::mutex> lock(p->mu);
  *kern = dir == 3 ? p->stage_kern : p->kern[dir];
  return KM_OK;
}

int km_stage_fused(const km_cfg_t *cfg, km_stage_plan_t *p, const km_pack_t *base,
                   const km_pack_t *u0, const km_pack_t *u1, double beta, const double *dt_dev,
                   double *dt_est_dev, void *stream) {
  if (!cfg || !p || !base || !u0 || !u1 || !dt_dev) {
    set_error("null argument");
    return KM_EINVAL;
  }
  if (int rc = check_cfg(cfg, u0)) return rc;
  if (cfg->mhd != KM_MHD_OFF || cfg->nspecies > 1 || cfg->geometry != KM_GEOM_CARTESIAN) {
    set_error("the fused stage is built for Cartesian hydro with one species");
    return KM_EINVAL;
  }
  for (const km_pack_t *q : {base, u0, u1}) {
    bool ok = q->nblocks == p->nb && q->ndim == p->mesh.ndim && q->ng == p->mesh.ng &&
              q->cell && q->blocks && q->nvar >= KM_NVAR(0);
    for (int d = 0; d < q->ndim; ++d) ok = ok && q->nx[d] == p->mesh.nx[d];
    if (!ok) {
      set_error("pack shape does not match the stage plan");
      return KM_EINVAL;
    }
  }
  if (u1->cell == u0->cell) {
    set_error("u1 must not be u0: the stage reads u0 across block faces while it writes u1");
    return KM_EINVAL;
  }
  if (p->amr) {
    for (int d = 0; d < u0->ndim; ++d)
      if (!u0->flux[d] || u0->nflux < 5) {
        set_error("on a block tree u0 needs flux tables (the fluxes at level jumps)");
        return KM_EINVAL;
      }
    if (dt_est_dev) {
      set_error("on a block tree the CFL estimate follows PreparePrimitive: call "
                "km_prepare_primitive and km_estimate_dt after the stage");
      return KM_EINVAL;
    }
  }
  std::lock_guard<std::mutex> lock(p->mu);
  if (cfg->precision != p->cfg.precision || cfg->recon != p->cfg.recon ||
      cfg->limiter != p->cfg.limiter || cfg->riemann != p->cfg.riemann) {
    for (int d = 0; d < 3; ++d) p->kern[d] = -1;
    p->stage_kern = -1;
  }
  p->cfg = *cfg;
  p->cfg.ndim = p->mesh.ndim;
  p->stream = as_stream(stream);
  FusedArgs a;
  a.u0 = *u0;
  a.base = *base;
  a.u1 = *u1;
  a.du = p->du;
  a.dt = dt_dev;
  a.dt_est = dt_est_dev;
  a.inst = p->inst_d;
  a.nbr = p->nbr_d;
  a.beta = beta;
  a.lim = cfg->limiter;
  a.eos = make_eos(*cfg);
  a.amr = p->amr ? 1 : 0;
  a.agrid = p->agrid;
  a.fail = nullptr;
  a.write_prims = 0;  // the caller's containers may share their primitive arrays
  if (p->kern[0] < 0 || p->kern[1] < 0 || p->kern[2] < 0 || p->stage_kern < 0)
    if (int rc = autotune(*p, a)) return rc;
  auto none = [](int, bool) {};
  const int rc = p->amr ? amr_stage_launch(*p, a, u0->nflux, none)
                        : stage_launch(*p, a, dt_est_dev != nullptr, p->ninst, dt_est_dev, none);
  if (rc) return rc;
  return check_launch();
}

int km_sim_info(const km_sim_t *s, double *time, double *dt, int32_t *ncycle) {
  if (!s) return KM_EINVAL;
  double hd = 0., ht = 0.;
  KM_HIP(hipStreamSynchronize(s->stream));
  KM_HIP(hipMemcpy(&hd, s->dt_d, sizeof(double), hipMemcpyDeviceToHost));
  KM_HIP(hipMemcpy(&ht, s->time_d, sizeof(double), hipMemcpyDeviceToHost));
  if (time) *time = ht;
  if (dt) *dt = hd;
  if (ncycle) *ncycle = s->ncycle;
  return KM_OK;
}

}  // extern "C"
