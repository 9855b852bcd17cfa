/*
 * km_hydro.h -- C ABI of libkm_hydro.so, the MI355X (gfx950) implementation of
 * kamayan's hydro/MHD flux-and-update path.
 *
 * Every entry point below replaces one task of the reference's hydro unit or grid
 * update; the reference interface it stands in for is cited beside it
 * (paths relative to the kamayan source tree).
 *
 * Conventions
 *  - All pointers held in km_pack_t (and the dt pointers) are DEVICE pointers for the
 *    km_* functions; the CPU twin used for checking takes host pointers in the same
 *    structs.
 *  - Field layout is Parthenon's MeshData block pack: pack(b, var, k, j, i), one
 *    (E3, E2, E1) array per (block, variable), i fastest, with ng ghost zones on each
 *    side of every active dimension (E = nx + 2*ng; E = 1 for inactive dimensions).
 *  - The flux through face i-1/2 of cell i is stored at index i of the flux array,
 *    which has the shape of the cell array (hydro_add_flux_tasks.cpp:99-103).
 *  - Face- and edge-centred arrays (CT face B and EMFs) all use one staggered shape
 *    (E3 + (ndim>2), E2 + (ndim>1), E1 + 1); element (k, j, i) of an x1-face array is
 *    the face at x1_{i-1/2} of cell (k, j, i).
 *  - Return value 0 == TaskStatus::complete; nonzero codes below; km_last_error()
 *    gives a thread-local message.
 *  - Kernels are enqueued on the given hipStream_t (passed as void*; NULL = default
 *    stream) and do not block the host, except where a function says so.
 */
#ifndef KM_HYDRO_H_
#define KM_HYDRO_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------------ */
#define KM_OK 0
#define KM_EINVAL 1  /* unsupported option / bad argument (dispatcher.hpp:122-130) */
#define KM_EHIP 2    /* HIP runtime error */
#define KM_ENOMEM 3  /* allocation failed */
#define KM_ENODEV 4  /* no GPU visible */
/* the state went non-physical (density or pressure <= 0, or a non-finite value), or the
 * time step stayed below parthenon/time/dt_min (above dt_max) for its cycle limit: the
 * conditions under which the reference's run fails (kamayan_driver.cpp:44-53) */
#define KM_ESTATE 5

/* ---- runtime options; value order follows the reference's POLYMORPHIC_PARM -------- */
enum { KM_RECON_FOG = 0, KM_RECON_PLM = 1, KM_RECON_PPM = 2, KM_RECON_WENOZ = 3 }; /* hydro_types.hpp:18 */
enum { KM_LIM_MINMOD = 0, KM_LIM_VAN_LEER = 1, KM_LIM_MC = 2 };                  /* hydro_types.hpp:19 */
enum { KM_RIEMANN_HLL = 0, KM_RIEMANN_HLLC = 1, KM_RIEMANN_HLLD = 2 };           /* hydro_types.hpp:20 */
enum { KM_RECONVARS_PRIMITIVE = 0 };                                             /* hydro_types.hpp:21 */
enum { KM_STRATEGY_SCRATCHPAD = 0, KM_STRATEGY_SCRATCHVAR = 1 };                 /* hydro_types.hpp:22 */
enum { KM_EMF_ARITHMETIC = 0 };                                                  /* hydro_types.hpp:24 */
enum { KM_FLUID_ONET = 0, KM_FLUID_THREET = 1 };                                 /* physics_types.hpp:7; deck values 1t, 3t */
/* eos/mode_init: the EOS call of PostMeshInitialization (eos.cpp:46-62, 165-168):
 * dens_pres fills EINT from PRES, dens_ener PRES from EINT, dens_temp EINT and PRES from
 * TEMP (plus BMOD, and TEMP where it is an output) */
enum { KM_EOS_INIT_DENS_PRES = 0, KM_EOS_INIT_DENS_ENER = 1, KM_EOS_INIT_DENS_TEMP = 2 };
enum { KM_MHD_OFF = 0, KM_MHD_CT = 1 };                                          /* physics_types.hpp:8 */
enum { KM_GEOM_CARTESIAN = 0, KM_GEOM_CYLINDRICAL = 1 };                         /* geometry_types.hpp:5 */
/* FP64: every operation in FP64, rounded in the reference's order (-ffp-contract=off);
 * FP32_MIXED: reconstruction + Riemann in FP32, everything else FP64;
 * FP64_FMA: FP64 with fused multiply-adds (agreement with FP64 to ~1e-14 relative). */
enum { KM_PREC_FP64 = 0, KM_PREC_FP32_MIXED = 1, KM_PREC_FP64_FMA = 2 };
/* AXISYMMETRIC: the inner x1 boundary of cylindrical geometry, a mirror across r = 0
 * that flips the r and phi vector components (boundary_conditions.cpp:23-60). */
enum { KM_BC_PERIODIC = 0, KM_BC_OUTFLOW = 1, KM_BC_REFLECT = 2, KM_BC_AXISYMMETRIC = 3 };

/* Parsed option set. Built from the reference's input-deck keys by km_cfg_set /
 * km_cfg_parse_deck (hydro.cpp:55-90, physics.cpp:18-24, kamayan_driver.cpp:25-27,
 * eos.cpp:37-39, grid.cpp:35-37). */
typedef struct km_cfg_t {
  int32_t recon;         /* hydro/reconstruction         default fog    */
  int32_t limiter;       /* hydro/slope_limiter          default minmod */
  int32_t riemann;       /* hydro/riemann                default hll    */
  int32_t recon_vars;    /* hydro/ReconstructionVars     default primitive */
  int32_t strategy;      /* hydro/ReconstructionStrategy default scratchpad */
  int32_t emf_averaging; /* hydro/EMF_averaging          default arithmetic */
  int32_t fluid;         /* physics/fluid                default oneT   */
  int32_t mhd;           /* physics/MHD                  default off    */
  int32_t geometry;      /* geometry/geometry            default cartesian */
  int32_t nstages;       /* parthenon/time/integrator rk1|rk2|rk3 -> 1|2|3, default rk2 */
  int32_t precision;     /* kamayan_amd/precision fp64|fp32_mixed|fp64_fma (this library's own key) */
  int32_t ndim;          /* from parthenon/mesh nx2, nx3 */
  double gamma;          /* material/single/gamma        default 5/3    */
  double abar;           /* material/single/Abar         default 1      */
  double cfl;            /* hydro/cfl                    default 0.8    */
  int32_t nspecies;      /* entries of material/species  default 1 ("single") */
  int32_t mode_init;     /* eos/mode_init                default dens_pres */
  /* material/<species[0]>/{gamma,Abar} set gamma/abar (MakeEos, eos.cpp:170-181): the
   * first name of material/species, "single" by default */
  char species0[32];
  /* parthenon/time step control of the whole-mesh driver (kamayan_driver.cpp:28-53);
   * defaults are the reference's: dt_factor 2, dt_init/dt_ceil/dt_user/dt_max DBL_MAX,
   * dt_floor DBL_MIN, dt_min/dt_force -DBL_MAX (off), dt_init_force 1, cycle limits 10
   * (dt_min) and 1 (dt_max) */
  double dt_factor, dt_init, dt_ceil, dt_floor, dt_min, dt_max, dt_user, dt_force;
  int32_t dt_init_force, dt_min_cycle_limit, dt_max_cycle_limit;
  /* parthenon/mesh/pack_size: blocks per MeshData pack on the task path, whose kernels
   * are launched pack by pack from the host every cycle (-1, the default: one pack of
   * every block, graph-replayed) */
  int32_t pack_size;
} km_cfg_t;

/* ---- cell-variable slots of a hydro state pack ------------------------------------
 * Conserved first (the reference's Conserved TypeList order, hydro_types.hpp:61-74):
 *   DENS, MOMENTUM[3], ENER, and MAGC[3] when MHD == ct   -> ncons = 5 or 8
 * then the primitives / derived fields (NonFlux + EOS outputs, hydro_types.hpp:65;
 * eos.cpp:67):
 *   VELOCITY[3], PRES, BMOD, EINT, TEMP, and DIVB when MHD == ct */
#define KM_DENS 0
#define KM_MOM0 1
#define KM_ENER 4
#define KM_MAGC0 5
#define KM_NCONS(mhd) ((mhd) == KM_MHD_CT ? 8 : 5)
#define KM_VEL0(mhd) (KM_NCONS(mhd) + 0)
#define KM_PRES(mhd) (KM_NCONS(mhd) + 3)
#define KM_BMOD(mhd) (KM_NCONS(mhd) + 4)
#define KM_EINT(mhd) (KM_NCONS(mhd) + 5)
#define KM_TEMP(mhd) (KM_NCONS(mhd) + 6)
#define KM_DIVB(mhd) (KM_NCONS(mhd) + 7)
#define KM_NVAR(mhd) ((mhd) == KM_MHD_CT ? 16 : 12)
/* With more than one species (material/species), the mass fractions MFRAC follow:
 * slot KM_MFRAC(mhd, s) = KM_NVAR(mhd) + s, and flux slot KM_NCONS(mhd) + s. They are
 * conserved fields (stage containers own them): X_s between stages, rho X_s inside a
 * stage's update (material.cpp:132-197). A pack then holds KM_NVAR(mhd) + nspecies cell
 * vars and KM_NCONS(mhd) + nspecies flux vars. */
#define KM_MFRAC(mhd, s) (KM_NVAR(mhd) + (s))
#define KM_NSPEC_SLOTS(nspecies) ((nspecies) > 1 ? (nspecies) : 0)

/* Per-block uniform-Cartesian metric: the values parthenon::UniformCartesian gives
 * to Dx<axis>(), FaceArea<axis>() and CellVolume() (geometry.hpp:35-167). */
typedef struct km_block_t {
  double xmin[3]; /* lower corner of the block's interior */
  double dx[3];   /* cell widths                          */
  double area[3]; /* dx2*dx3, dx1*dx3, dx1*dx2           */
  double vol;     /* dx1*dx2*dx3                          */
  int32_t level;  /* refinement level, 0 = root grid      */
  int32_t lx[3];  /* logical location at that level       */
} km_block_t;

/* MeshData block pack (grid.hpp:36-131). Tables are indexed [b*nvar + v] for cell,
 * [b*nflux + v] for flux, [b] for face/edge; unused tables may be NULL. */
typedef struct km_pack_t {
  int32_t nblocks;
  int32_t ndim;
  int32_t ng;      /* ghost zones per side (parthenon/mesh/nghost, grid.cpp:56) */
  int32_t nx[3];   /* interior cells per block along x1, x2, x3 (1 if inactive) */
  int32_t nvar;    /* entries per block in cell[] */
  int32_t nflux;   /* entries per block in flux[d] */
  const km_block_t *blocks;
  double *const *cell;
  double *const *flux[3];
  double *const *face[3];
  double *const *edge[3];
} km_pack_t;

/* ---- configuration ---------------------------------------------------------------- */
/* Reference defaults (hydro.cpp:55-90; eos.cpp:37-39; kamayan_driver.cpp:25-27). */
int km_cfg_default(km_cfg_t *cfg);
/* Set one "block/key" = value, e.g. ("hydro/reconstruction", "wenoz"). Unknown keys
 * are ignored (return KM_OK, as Parthenon keeps unrelated deck keys); an invalid value
 * for a known key returns KM_EINVAL like the reference's rule check
 * (runtime_parameters.hpp:157-169). */
int km_cfg_set(km_cfg_t *cfg, const char *block_key, const char *value);
/* Parse an input deck ("<block>" headers, "key = value" lines, '#' comments). */
int km_cfg_parse_deck(km_cfg_t *cfg, const char *deck_text);
/* Minimum ghost width the options need (python/kamayan/code_units/Hydro.py:14,35;
 * +1 for CT, hydro_add_flux_tasks.cpp:59-70). */
int km_cfg_min_ghost(const km_cfg_t *cfg);

/* ---- hot-path tasks (device pointers; enqueued on `stream`) --------------------------- */
/* hydro::CalculateFluxes -- Dispatcher<CalculateFluxesNested|Scratch>
 * (hydro_add_flux_tasks.cpp:29-385, called from AddFluxTasks :444-476).
 * Reads the reconstructed set DENS, VELOCITY, PRES, BMOD, EINT (+MAGC, face B for CT)
 * from u->cell / u->face and writes u->flux[d] for every conserved variable. With more
 * than one species, each mass fraction is reconstructed the same way and its flux is the
 * mass flux times the upwind side's mass fraction (:119-135); along x2 and x3 the last
 * species' flux is left unwritten, as the reference's loops end one species early
 * (:181-197, :240-256). The pack then needs KM_NCONS + nspecies flux slots. */
int km_calculate_fluxes(const km_cfg_t *cfg, const km_pack_t *u, void *stream);
/* hydro::CalculateEMF (hydro_add_flux_tasks.cpp:387-442): edge EMFs from the MAGC
 * face fluxes into u->edge[d]. No-op unless MHD == ct and ndim >= 2. */
int km_calculate_emf(const km_cfg_t *cfg, const km_pack_t *u, void *stream);
/* grid::FluxesToDuDt (grid.cpp:179-211 -> FluxDivergence grid_update.hpp:15-57 and
 * FluxStokes grid_update.hpp:80-137): dudt->cell[v] for the conserved cell vars and,
 * for CT, dudt->face[d]; with more than one species also each mass fraction's slot. */
int km_fluxes_to_dudt(const km_cfg_t *cfg, const km_pack_t *u, const km_pack_t *dudt,
                      void *stream);
/* hydro::AddSourceTerms (hydro.cpp:199-240, AddTasksOneStep, between FluxesToDuDt and
 * ApplyDuDt): in cylindrical geometry the r momentum of dudt gains
 * (rho v_phi^2 + P) / r at each cell centre (primconsflux.hpp:144-159), from the
 * primitives of u; nothing in Cartesian geometry. */
int km_add_source_terms(const km_cfg_t *cfg, const km_pack_t *u, const km_pack_t *dudt,
                        void *stream);
/* hydro::PrepareConserved, the per-stage task on md0 between AddTasksOneStep and ApplyDuDt
 * (hydro.cpp:244-286, kamayan_driver.cpp:224-231): with CT in cylindrical geometry the cell
 * B_phi becomes B_phi / r_centroid. material::PrepareConserved (material.cpp:132-155), the
 * same per-stage task: with more than one species each mass fraction X_s of u becomes
 * rho X_s. Nothing otherwise. */
int km_prepare_conserved_stage(const km_cfg_t *cfg, const km_pack_t *u, void *stream);
/* grid::ApplyDuDt (grid.cpp:213-283):
 *   u0 <- beta*base + (1-beta)*u0 ;  u1 <- u0 + (beta*dt)*dudt
 * over the interior of every conserved cell var (mass fractions included) and (CT) every
 * face element. */
int km_apply_dudt(const km_cfg_t *cfg, const km_pack_t *base, const km_pack_t *u0,
                  const km_pack_t *u1, const km_pack_t *dudt, double beta, double dt,
                  void *stream);
/* PreparePrimitive callbacks in DAG order: hydro PreparePrimitive_impl
 * (primconsflux.cpp:82-142) then eos EosWrapped(ener) (eos.cpp:71-114,164), then
 * material PreparePrimitive (material.cpp:157-197): with more than one species the mass
 * fractions are divided by their sum. */
int km_prepare_primitive(const km_cfg_t *cfg, const km_pack_t *u, void *stream);
/* PostMeshInitialization in DAG order: eos PrepareConserved with mode_init=dens_pres
 * (eos.cpp:165-168) then hydro ConvertToConserved_impl (primconsflux.cpp:19-79), then
 * material PostMeshInitialization (material.cpp:104-130): the mass fractions are divided by
 * their sum. */
int km_prepare_conserved(const km_cfg_t *cfg, const km_pack_t *u, void *stream);
/* hydro EstimateTimeStep (hydro_time_step.cpp:19-65): writes (cfl/ndim)*min(dx/(|v|+cf))
 * over the interior to *dt_dev (device). Does not block the host. */
int km_estimate_dt(const km_cfg_t *cfg, const km_pack_t *u, double *dt_dev, void *stream);
/* hydro FillDerived_impl (hydro.cpp:150-194): DIVB from face B (CT only). */
int km_fill_divb(const km_cfg_t *cfg, const km_pack_t *u, void *stream);

/* ---- ghost fill (Parthenon AddBoundaryExchangeTasks, kamayan_driver.cpp:150-151) ---- */
/* Mesh description (parthenon/mesh and parthenon/meshblock deck blocks). */
typedef struct km_mesh_desc_t {
  int32_t ndim;
  int32_t ng;
  int32_t nx[3];    /* cells per block  (parthenon/meshblock/nx*) */
  int32_t nb[3];    /* blocks per dimension on the root grid */
  double xmin[3];   /* parthenon/mesh/x*min */
  double xmax[3];   /* parthenon/mesh/x*max */
  int32_t bc[6];    /* ix1, ox1, ix2, ox2, ix3, ox3 */
} km_mesh_desc_t;

typedef struct km_ghost_plan_t km_ghost_plan_t; /* opaque, device-resident */
/* Same-level ghost fill plan for a uniform grid of nb[0] x nb[1] x nb[2] blocks stored
 * in x-fastest order; bc = {ix1, ox1, ix2, ox2, ix3, ox3} of KM_BC_*. */
int km_ghost_plan_uniform(const km_pack_t *shape, const int32_t nb[3], const int32_t bc[6],
                          km_ghost_plan_t **plan);
int km_ghost_plan_destroy(km_ghost_plan_t *plan);
/* Ghost fill plan for a block-tree (AMR) mesh: the leaves of Parthenon's MeshBlockTree,
 * leaves[4*b..4*b+3] = {level, lx1, lx2, lx3} of pack block b, over the root grid
 * mesh->nb with the domain and boundaries of `mesh`. Leaves must tile the domain; leaves
 * that touch (faces, edges, corners, across periodic boundaries) differ by at most one
 * level; with more than one level every active block dimension holds an even number of
 * at least max(2 ng, 2 ((ng+1)/2 + 1)) cells (8 for nghost 4), which keeps ghost zones
 * and prolongation stencils inside blocks adjacent to the block being filled. */
int km_ghost_plan_amr(const km_pack_t *shape, const km_mesh_desc_t *mesh,
                      const int32_t *leaves, int32_t nleaves, km_ghost_plan_t **plan);
/* The same plan for a geometry (KM_GEOM_*; km_ghost_plan_amr is Cartesian). Cylindrical
 * (r, z) meshes take kamayan's geometry-templated operators (refinement_operations.hpp:
 * 52-398): restriction weighted by the r-dependent cell volumes and face areas (and phi
 * edge lengths |r| 2 pi), r and z faces prolongated as r B (values times their radius,
 * the child divided by its radius + DBL_EPSILON) and Toth-Roe on r B with the phi
 * spacing 2 pi; an axisymmetric inner x1 boundary mirrors zones and x1 faces
 * (boundary_conditions.cpp:23-71). Flux correction on such a plan weights the fine
 * fluxes by their cylindrical face areas. */
int km_ghost_plan_amr_geom(const km_pack_t *shape, const km_mesh_desc_t *mesh,
                           const int32_t *leaves, int32_t nleaves, int32_t geometry,
                           km_ghost_plan_t **plan);
/* Fill every ghost zone (faces, edges, corners) of cell vars [v0, v1) and, when
 * with_faces != 0, of the CT face arrays. With a block-tree plan a ghost zone is copied
 * from a same-level neighbour, restricted from a finer one (RestrictAverage,
 * refinement_operations.hpp:52-121) or prolongated from a coarser one
 * (ProlongateSharedMinMod, :123-268). CT face arrays across levels (even nghost) take
 * RestrictAverage / ProlongateSharedMinMod on faces and ProlongateInternalTothAndRoe
 * (:270-398) for fine faces inside coarse cells; a block's own faces are never written.
 * A pack with face tables (u->face[0] set) is an MHD pack: reflecting and axisymmetric
 * boundaries negate its cell B vector (slots 5..7) as well as the momentum and velocity. */
int km_ghost_fill(const km_ghost_plan_t *plan, const km_pack_t *u, int32_t v0, int32_t v1,
                  int32_t with_faces, void *stream);
/* Flux correction (Parthenon AddFluxCorrectionTasks, kamayan_driver.cpp:198-207): every
 * block face shared with finer leaves takes the area-weighted restriction of the fine
 * fluxes, for the u->nflux conserved variables, and, when u->edge tables are present
 * (CT), each edge EMF lying in such a face takes the length-weighted restriction of the
 * fine edges on it. No-op for a uniform plan. */
int km_flux_correct(const km_ghost_plan_t *plan, const km_pack_t *u, void *stream);

/* ---- one RK stage as fused kernels on caller-owned packs ---------------------------- */
/* The body of one BuildTaskListRKStage (kamayan_driver.cpp:186-248) from the fluxes to the
 * stage combine -- hydro::CalculateFluxes (hydro_add_flux_tasks.cpp:444-476), on block
 * trees the flux correction (kamayan_driver.cpp:198-207), grid::FluxesToDuDt and
 * grid::ApplyDuDt (grid.cpp:179-283) -- as the fused kernels: on a single-level mesh one
 * stage kernel or three directional sweeps reading u0's conserved interior directly (no
 * ghost exchange, no stored primitives or fluxes), on a block tree pencil tiles reading the
 * stored primitives. The worklist (plan) holds what the kernels need beyond the packs: the
 * face-neighbour table and instance map of single-level meshes or ensembles, the leaf
 * lookup of a block tree, the FP64 flux-divergence scratch of the sweeps and the kernel
 * choices (autotuned at the first call, as km_sim_set_fused_kernel describes). Cartesian
 * hydro with one species. */
typedef struct km_stage_plan_t km_stage_plan_t; /* opaque, device-resident */
/* Single-level meshes: ninst independent uniform block grids (one for a plain mesh), their
 * blocks one after the other in the packs, each x-fastest (as km_sim_create_ensemble). */
int km_stage_plan_uniform(const km_cfg_t *cfg, const km_mesh_desc_t *meshes, int32_t ninst,
                          km_stage_plan_t **plan);
/* A block-tree mesh with the leaves described at km_ghost_plan_amr, in pack order. */
int km_stage_plan_amr(const km_cfg_t *cfg, const km_mesh_desc_t *mesh, const int32_t *leaves,
                      int32_t nleaves, km_stage_plan_t **plan);
int km_stage_plan_destroy(km_stage_plan_t *plan);
/* Kernel choice of the plan's stages: dir 0..2 (-1: all) a sweep kernel as for
 * km_sim_set_fused_kernel; dir 3 the whole stage: -1 autotune, 0 three sweeps, k >= 1 the
 * one-launch stage kernel variant k (2-D and 3-D single-level meshes; 2-D meshes have
 * variants 1-4, for 16-, 32- and 64-cell block rows). No reference counterpart
 * (the reference fixes its loop shapes in source). */
int km_stage_plan_set_kernel(km_stage_plan_t *plan, int32_t dir, int32_t kernel);
int km_stage_plan_kernel(km_stage_plan_t *plan, int32_t dir, int32_t *kernel);
/* One stage: u1 = (beta base + (1 - beta) u0) + beta dt_dev[inst] dudt(u0) over the
 * conserved interior (grid.cpp:232-235; u0's in-place combine is not written back, since
 * nothing reads md0 afterwards). dt_dev: device step dt per instance. dt_est_dev (device,
 * per instance, or NULL): on a single-level mesh the CFL estimate of u1 times cfl/ndim, the
 * last stage's EstimateTimestep (hydro_time_step.cpp:19-65); block trees take NULL and run
 * km_prepare_primitive then km_estimate_dt after the stage, as the task order does. u0
 * must not be u1 (the kernels read u0 across block faces while writing u1); base may be
 * u0. Single-level meshes read u0's conserved interior only; block trees read u0's stored
 * primitives and ghost zones (PreparePrimitive and the ghost exchange of the previous
 * stage) and need u0's flux tables, which receive the fluxes at level jumps. u1's ghost
 * zones are stale after the call, as after ApplyDuDt; some kernels write the x1 ghost
 * zones of the rows they update with the edge zone's value (rows in whole cache lines),
 * which the next ghost exchange overwrites. Calls on one plan are serialised; the first
 * call (and the first after a change of method or form) runs the autotuner on the call's
 * data and stream. */
int km_stage_fused(const km_cfg_t *cfg, km_stage_plan_t *plan, const km_pack_t *base,
                   const km_pack_t *u0, const km_pack_t *u1, double beta, const double *dt_dev,
                   double *dt_est_dev, void *stream);

/* ---- whole-mesh driver (KamayanDriver RK stages, kamayan_driver.cpp:121-248) ----- */
enum {
  KM_PATH_TASKS = 0, /* one kernel per reference task, same sequence as BuildTaskListRKStage */
  KM_PATH_FUSED = 1  /* fused sweeps per RK stage (hydro; CT MHD on 3-D single-level meshes) */
};

typedef struct km_sim_t km_sim_t; /* opaque: device mesh, stage containers, dt state */
int km_sim_create(const km_cfg_t *cfg, const km_mesh_desc_t *mesh, int32_t path,
                  km_sim_t **sim);
/* An ensemble of ninst independent runs advanced together: instance m is the uniform mesh
 * meshes[m] (all instances share ndim, nghost and cells per block; block counts,
 * extents and boundaries may differ) with its own dt and time. Its blocks follow those of
 * instances 0..m-1 in the host buffers. Every kernel of a stage covers all instances in
 * one launch. Ensembles advance by cycle count (tlim must be 0). */
int km_sim_create_ensemble(const km_cfg_t *cfg, const km_mesh_desc_t *meshes, int32_t ninst,
                           int32_t path, km_sim_t **sim);
/* A block-tree (AMR) mesh with the leaves described at km_ghost_plan_amr. On the task
 * path (hydro or CT MHD; CT needs an even nghost) each stage adds flux (and EMF)
 * correction after the fluxes and fills ghost zones across levels. On the fused path
 * (hydro, one species, Cartesian) the sweeps run as pencil tiles over every leaf in one
 * launch per direction, reading the exchanged primitives; faces bordering a finer leaf
 * take the restricted fine flux, computed first; the results are the task path's bits.
 * Host buffers hold the blocks in leaf order. */
int km_sim_create_amr(const km_cfg_t *cfg, const km_mesh_desc_t *mesh, const int32_t *leaves,
                      int32_t nleaves, int32_t path, km_sim_t **sim);
/* time and dt of the first n instances (blocks the host) */
int km_sim_ensemble_info(const km_sim_t *sim, double *times, double *dts, int32_t n);
int km_sim_destroy(km_sim_t *sim);
/* Host buffers in pack layout: cell = [nblocks][nvar][E3][E2][E1] with nvar =
 * KM_NVAR(mhd) + KM_NSPEC_SLOTS(nspecies), face[d] = [nblocks][S3][S2][S1] (NULL unless
 * CT). km_sim_set_primitive takes DENS, VELOCITY, PRES (+MAGC, face B, mass fractions) on
 * the interior, as a problem generator leaves them, and runs PostMeshInitialization, the
 * initial ghost fill and the initial dt estimate. */
int km_sim_set_primitive(km_sim_t *sim, const double *cell, const double *const face[3]);
int km_sim_get_state(km_sim_t *sim, double *cell, double *const face[3]);
/* Advance ncycles cycles (or until t >= tlim if tlim > 0). Blocks the host. */
int km_sim_advance(km_sim_t *sim, int32_t ncycles, double tlim);
int km_sim_info(const km_sim_t *sim, double *time, double *dt, int32_t *ncycle);
/* Override the next step's dt (used to replay a fixed dt sequence). */
int km_sim_set_dt(km_sim_t *sim, double dt);
/* Record HIP events around every fused-stage kernel from now on (on = 1) and discard
 * earlier records. Profiling runs cycles eagerly (no graph replay). */
int km_sim_profile(km_sim_t *sim, int32_t on);
/* Sum of recorded kernel times (ms) and launch counts per kind: 0..2 = sweep along
 * x1..x3 (the last one also updates u1), 3 = ghost fill, 4 = whole stage. Blocks. */
int km_sim_kernel_times(km_sim_t *sim, double *ms, int32_t *count, int32_t nkinds);
/* Kernel of the fused path's sweep along x(dir+1) (dir = -1: all three): -1 = chosen by
 * the library's autotuner (the default; at the first cycle that starts from conserved
 * values it times every candidate on the simulation's own data and keeps the choice per
 * arithmetic form, method, block shape and direction for the process), 0 = the pencil
 * march (single-level meshes only), k >= 1 = pencil-tile shape k of 5. The reference has no counterpart:
 * CalculateFluxesNested/Scratch (hydro_add_flux_tasks.cpp:29-385) fix their loop shape
 * in source. */
int km_sim_set_fused_kernel(km_sim_t *sim, int32_t dir, int32_t kernel);
/* The kernel in use for the sweep along x(dir+1) (-1 until the autotuner has run). */
int km_sim_fused_kernel(km_sim_t *sim, int32_t dir, int32_t *kernel);
/* ---- problem generators (host-side; fill primitives on the interior) -------------- */
/* names: "isentropic_vortex" (isentropic_vortex.cpp:105-175), "smooth3d", "mhd_random",
 * "kelvin_helmholtz", "sedov" (sedov.cpp:75-143, 3-D form), "mhd_blast"
 * (mhd_blast.cpp:47-160; Cartesian, and r-z cylindrical). Buffers as for
 * km_sim_set_primitive; `seed` feeds the random modes where a problem has them. With
 * cfg->nspecies > 1 the mass fractions are set too: sedov puts species 0 inside the blast
 * radius and species 1 outside (sedov.cpp:138-141), the others species 0 everywhere. */
int km_problem_fill(const char *name, const km_cfg_t *cfg, const km_mesh_desc_t *mesh,
                    uint64_t seed, const double *params, int32_t nparams, double *cell,
                    double *const face[3]);
/* The same generators over an explicit block table (e.g. the leaves of a block tree). */
int km_problem_fill_blocks(const char *name, const km_cfg_t *cfg, const km_mesh_desc_t *mesh,
                           const km_block_t *blocks, int32_t nblocks, uint64_t seed,
                           const double *params, int32_t nparams, double *cell,
                           double *const face[3]);
/* Block metric table for a uniform mesh (blocks x-fastest). */
int km_mesh_blocks(const km_mesh_desc_t *mesh, km_block_t *blocks_host);
/* Block metric table of block-tree leaves (checked as in km_ghost_plan_amr). */
int km_amr_blocks(const km_mesh_desc_t *mesh, const int32_t *leaves, int32_t nleaves,
                  km_block_t *blocks_host);
/* Leaves of the root grid refined to level lmax wherever a block meets the ball
 * |x - center| <= radius, then refined further until touching leaves differ by at most
 * one level. *nleaves: capacity of `leaves` (entries of 4) in, leaf count out; a buffer
 * too small returns KM_EINVAL with the count needed. */
int km_amr_refine_ball(const km_mesh_desc_t *mesh, int32_t lmax, const double center[3],
                       double radius, int32_t *leaves, int32_t *nleaves);

/* ---- adaptive refinement: tagging task (kamayan_driver.cpp:178-181) and remesh ------ */
enum { KM_AMR_DEREFINE = -1, KM_AMR_SAME = 0, KM_AMR_REFINE = 1 };
/* One <kamayan/refinementN> block (grid.cpp:96-124) with method = loehner. */
typedef struct km_amr_criterion_t {
  int32_t var;         /* cell slot of field (+ comp4), e.g. KM_PRES(mhd) for field = pres */
  int32_t max_level;   /* max_level, default numlevel (grid.cpp:112-113) */
  double refine_tol;   /* refine_tol, default 0.8 */
  double derefine_tol; /* derefine_tol, default 0.2 */
  double filter;       /* filter, default 0.01 */
} km_amr_criterion_t;
#define KM_AMR_MAX_CRIT 4
typedef struct km_amr_adapt_t {
  int32_t numlevel;       /* parthenon/mesh/numlevel, default 1 */
  int32_t derefine_count; /* parthenon/mesh/derefine_count, default 10 */
  int32_t ncrit;          /* criteria in use */
  km_amr_criterion_t crit[KM_AMR_MAX_CRIT];
} km_amr_adapt_t;
/* numlevel 1, derefine_count 10, no criteria. */
int km_amr_adapt_default(km_amr_adapt_t *adapt);
/* Set one deck key for physics/MHD = mhd: "parthenon/mesh/numlevel",
 * "parthenon/mesh/derefine_count", "kamayan/refinementN/<key>" with key field (dens,
 * momentum, ener, magc, velocity, pres, "bulk modulus", eint, temp), comp4, method
 * (loehner; Parthenon's derivative_order_1/2 are not provided: KM_EINVAL), refine_tol,
 * derefine_tol, filter, max_level. Criterion N exists once its field is set. */
int km_amr_adapt_set(km_amr_adapt_t *adapt, int32_t mhd, const char *block_key,
                     const char *value);
/* grid::AMRLoehner::operator() (grid_refinement.cpp:30-152): per block, the largest
 * Loehner error over the interior of cell slot crit->var gives flag = refine (error >
 * refine_tol and level < max_level), derefine (error < derefine_tol) or same, and
 * tags_dev[b] = max(tags_dev[b], flag) (Kokkos ScatterMax). reset != 0 first sets every
 * tag to KM_AMR_DEREFINE, Parthenon's starting value (CheckAllRefinement). Reads two ghost
 * layers, so ghost zones must be filled. Blocks the host until the tags are written. */
int km_amr_tag_loehner(const km_pack_t *u, const km_amr_criterion_t *crit, int32_t *tags_dev,
                       int32_t reset, void *stream);
/* km_amr_tag_loehner without allocation: no counterpart in the reference. For callers that
 * cannot allocate device memory (code under hipGraph capture, serving loops, harnesses that
 * forbid hipMalloc). The per-block maxima live in one library-held device array of
 * KM_TAG_NOALLOC_MAX_BLOCKS entries (512 KiB, held by each GPU the library runs on until
 * the process ends; cleared by a fill kernel at each call); packs with more blocks are
 * refused with KM_EINVAL. Calls that overlap on one GPU share that array and race:
 * serialise them. Same arguments and results as km_amr_tag_loehner; blocks the host. */
#define KM_TAG_NOALLOC_MAX_BLOCKS 65536
int km_amr_tag_loehner_noalloc(const km_pack_t *u, const km_amr_criterion_t *crit,
                               int32_t *tags_dev, int32_t reset, void *stream);
/* Parthenon's MeshRefinement::SetRefinement for every leaf (refine below level
 * numlevel-1; derefine after derefine_count consecutive derefine tags and only without a
 * finer neighbour), then Mesh::UpdateMeshBlockTree: leaves whose 2^d siblings are all
 * flagged merge into their parent, flagged leaves split (refining coarser neighbours first
 * to keep touching leaves within one level); refinements go in increasing level order,
 * then merges in decreasing level order, each merge refused when a leaf two levels finer
 * would touch the parent. Host arrays. In: leaves[4*nleaves], tags[nleaves];
 * deref_count[nleaves] is updated in place. Out: new_leaves (capacity *nnew in, count
 * out) in Morton order over the finest level, new_deref per new leaf (kept leaves keep
 * theirs, new ones start at 0), origin[9*i] = {0, old leaf} kept, {1, old parent} split,
 * {2, old children x1-fastest} merged; *changed = 0 when nothing changes. */
int km_amr_update_tree(const km_mesh_desc_t *mesh, int32_t numlevel, int32_t derefine_count,
                       const int32_t *leaves, int32_t nleaves, const int32_t *tags,
                       int32_t *deref_count, int32_t *new_leaves, int32_t *new_deref,
                       int32_t *origin, int32_t *nnew, int32_t *changed);
/* Block data of the new mesh from the old one (Parthenon RedistributeAndRefineMeshBlocks)
 * for cell slots [v0, v1): a kept leaf copies every zone, a split leaf's children take
 * ProlongateSharedMinMod (refinement_operations.hpp:123-268) of the parent's zones (its
 * ghost zones give the slopes at its edges) over their interior, a merged parent the
 * RestrictAverage (:52-121) of its children over its interior. When both packs carry all
 * three face tables (CT), the face field moves too: kept faces are copied, a child's faces
 * on parent faces take ProlongateSharedMinMod with slopes across the face, those inside
 * parent cells ProlongateInternalTothAndRoe (:270-398), and a merged parent's faces the
 * area-weighted RestrictAverage of the child faces on them (the upper child's copy of a
 * face two children share). Ghost zones and ghost faces of new leaves are left for the
 * ghost fill. origin_dev: the origin table of km_amr_update_tree in device memory. Both
 * packs share ndim, ng and nx. */
int km_amr_remesh(const km_pack_t *old_u, const km_pack_t *new_u, const int32_t *origin_dev,
                  int32_t v0, int32_t v1, void *stream);
/* The same for a geometry (km_amr_remesh is Cartesian); cylindrical geometry takes the
 * operators of km_ghost_plan_amr_geom and needs the mesh description (its x1min places
 * the radii). */
int km_amr_remesh_geom(const km_pack_t *old_u, const km_pack_t *new_u, const int32_t *origin_dev,
                       int32_t v0, int32_t v1, const km_mesh_desc_t *mesh, int32_t geometry,
                       void *stream);
/* Adaptive refinement of a block-tree simulation (hydro or CT): after every cycle each leaf
 * is tagged by the criteria (km_amr_tag_loehner on the filled base state), the tree is
 * updated (km_amr_update_tree) and, when it changed, every cell field and the CT face
 * field move to the new leaves (km_amr_remesh), ghost zones (and faces) are refilled, divB
 * is recomputed (CT) and dt is re-estimated on the new mesh before the time-step rule
 * (EvolutionDriver::Execute and InitializeBlockTimeStepsAfterAMR in Parthenon). Resets
 * every derefine count. */
int km_sim_set_adaptive(km_sim_t *sim, const km_amr_adapt_t *adapt);
/* Current leaves of a block-tree simulation in block order: *nleaves capacity in (leaf
 * entries of 4), count out; host buffers of get/set_state follow this count. */
int km_sim_amr_leaves(const km_sim_t *sim, int32_t *leaves, int32_t *nleaves);

/* ---- history outputs (Parthenon's "hst" output type) ------------------------------- */
/* The isentropic vortex's history variables (isentropic_vortex.cpp:74-101): errors[v] =
 * ErrorHistory (isentropic_vortex.hpp:150-200) of DENS, VEL0, VEL1, VEL2, PRES and, with
 * MHD, MAGC0..2: sum over interior cells of |exact - q| * V / ((x1max - x1min) *
 * (x2max - x2min)), exact = the vortex state at the cell centroid (the radial centroid in
 * cylindrical geometry) with the parameters of km_problem_fill("isentropic_vortex"). u holds
 * primitives on the device; mesh gives the domain. 5 or 8 values; blocks the host. */
int km_history_vortex_errors(const km_cfg_t *cfg, const km_pack_t *u, const km_mesh_desc_t *mesh,
                             const double *params, int32_t nparams, double *errors,
                             void *stream);
/* A history file for a simulation (parthenon/outputN/file_type = hst, dt): '#' header
 * lines, then one row per output of time, dt (of the next step), cycle, block count and
 * the problem's history variables ("isentropic_vortex": the errors above; NULL or "": none).
 * A row is written now, after every cycle that reaches the next output time (which then
 * advances by dt), and when km_sim_advance reaches tlim. Single runs only (no ensembles). */
int km_sim_set_history(km_sim_t *sim, const char *path, double dt, const char *problem,
                       const double *params, int32_t nparams);

/* ---- device memory helpers (so callers need no other runtime) --------------------- */
int km_device_count(int32_t *count);
int km_malloc(void **ptr, size_t bytes);
int km_free(void *ptr);
int km_memcpy_h2d(void *dst, const void *src, size_t bytes);
int km_memcpy_d2h(void *dst, const void *src, size_t bytes);
/* Device-to-device copy by a plain streaming kernel (16-byte aligned, multiple of 16 B);
 * also the reference kernel for calibrating HBM byte counters. */
int km_copy_device(void *dst, const void *src, size_t bytes, void *stream);
int km_synchronize(void);

const char *km_last_error(void);
const char *km_version(void);

#ifdef __cplusplus
}
#endif
#endif /* KM_HYDRO_H_ */
