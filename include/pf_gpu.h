/*
 * pf_gpu.h -- C ABI of the MI355X (gfx950) sparse two-phase Phase-Field-FLIP step.
 *
 * The reference (lsfs, JangHyun-bin/adaptive-pf-flip) has no FFI: its seam is four
 * free C++ functions called by SparseSim3DTP::step().  Each entry point below names
 * the reference interface it replaces (paths relative to the reference root).  The
 * C++ host class in adaptive-pf-flip_amd/host/sparse_sim3d_tp_gpu.h keeps the
 * reference's SparseSim3DTP field names and step() on top of this ABI.
 *
 * Conventions: plain pointers and sizes, int status (PF_OK == 0), no exceptions
 * cross the boundary.  One context = one grid + one particle set + one HIP stream on
 * the current device; calls on a context are stream-ordered; downloads synchronise.
 * A context is not re-entrant (as a SparseMacGrid3D is not).
 */
#ifndef PF_GPU_H
#define PF_GPU_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PF_ABI_VERSION 5

enum pf_status {
  PF_OK = 0,
  PF_ERR_ARG = 1,      /* invalid argument (null pointer, bad size, particle outside the grid) */
  PF_ERR_HIP = 2,      /* a HIP runtime call failed */
  PF_ERR_STATE = 3,    /* call out of order (e.g. no particles uploaded) */
  PF_ERR_NOMEM = 4     /* device allocation failed */
};

/* Field ids: the eight SparseBlockGrid3D<4> members of SparseMacGrid3D<4>
 * (src/grid/sparse_mac_grid3d.h:9-11), plus the step's pre-gravity copy of u,v,w
 * (the `saved` grid of sparse_sim3d_tp.cpp:866). */
enum pf_field {
  PF_FIELD_U = 0, PF_FIELD_V = 1, PF_FIELD_W = 2,
  PF_FIELD_MU = 3, PF_FIELD_MV = 4, PF_FIELD_MW = 5,
  PF_FIELD_P = 6, PF_FIELD_MARKER = 7,
  PF_FIELD_SAVED_U = 8, PF_FIELD_SAVED_V = 9, PF_FIELD_SAVED_W = 10
};

enum pf_stop_rule {
  PF_STOP_REF_MAXABS = 0, /* |b - A x|_inf < tol, the reference rule (sparse_ops3d_tp.cpp:340-357) */
  PF_STOP_REL2 = 1        /* ||b - A x||_2 / ||b||_2 <= tol, both measured in FP64 */
};

/* SparseMacGrid3D<4>(nx, ny, nz, dx) with origin (src/grid/sparse_mac_grid3d.h:5-17) */
typedef struct pf_grid_desc {
  int32_t nx, ny, nz, reserved;
  double dx, ox, oy, oz;
} pf_grid_desc;

/* PhaseParams (src/physics/phasefield.h:4-9) */
typedef struct pf_phase {
  double rho_l, rho_g, alpha_phi, rho_tilde_0;
} pf_phase;

/* Pressure-solve configuration.  The reference passes cg_iters and cg_tol
 * (sparse_ops3d_tp.h:9-10) and solves in FP64 with Jacobi PCG.  Here the inner
 * Jacobi PCG runs FP32 stencils with FP64 dot products and scalars, inside an FP64
 * iterative-refinement loop whose true residual decides the stop. */
typedef struct pf_solve_cfg {
  int32_t max_iters;        /* cap on total inner PCG iterations (reference cg_iters) */
  int32_t stop_rule;        /* pf_stop_rule */
  double tol;               /* reference cg_tol (absolute) or relative tolerance */
  double inner_rel_tol;     /* inner FP32 PCG reduction per refinement pass (e.g. 1e-3) */
  int32_t max_outer;        /* cap on FP64 refinement passes */
  int32_t warm_start;       /* 1: start from the stored pressure instead of 0 */
  int32_t check_every;      /* inner iterations between host convergence checks */
  int32_t precond;          /* pf_precond */
} pf_solve_cfg;

enum pf_precond {
  PF_PRECOND_JACOBI = 0,    /* the reference's diagonal preconditioner (sparse_ops3d_tp.cpp:322-327) */
  PF_PRECOND_LEARNED = 1,   /* the multigrid V-cycle with per-aggregate correction weights from
                               the attached network (pf_lpc_attach), evaluated once per solve;
                               falls back to Jacobi when it stagnates (stats.fallback_used) */
  PF_PRECOND_MULTIGRID = 2  /* a V-cycle of Galerkin aggregation multigrid (2x2x2 aggregates,
                               damped-Jacobi smoothing; pf_set_multigrid), sparse path; no
                               reference counterpart.  The host checks for convergence at
                               least every 4 iterations (check_every is capped there) */
};

/* Cycle of the multigrid preconditioner (no reference counterpart).  Levels halve
 * the grid down to at most 4 cells a side; each level smooths `sweeps` times before
 * and after its coarse correction, the coarsest `coarsest_sweeps` times. */
typedef struct pf_mg_cfg {
  int32_t sweeps;           /* Jacobi sweeps before and after the coarse correction, 1..16 */
  int32_t coarsest_sweeps;  /* damped-Jacobi sweeps (omega) on the coarsest level, 1..1024 */
  float omega;              /* Jacobi damping, in (0, 1) */
  float correction;         /* weight of the coarse-grid correction, in (0, 2) */
  float chebyshev_ratio;    /* 0: every sweep damped by omega; r in (1, 1000]: the sweeps of a
                               smoothing pass take the reciprocals of the Chebyshev roots of
                               degree `sweeps` on [2/r, 2] (2 bounds the spectrum of D^-1 A) */
} pf_mg_cfg;

/* Mirrors MRPressureSolveStats3D's core fields (src/pressure/multires_pressure3d.h:60-133) */
typedef struct pf_solve_stats {
  int32_t iterations, outer_iterations, converged, breakdown;
  int32_t pinned, fallback_used, reserved0, reserved1;
  int64_t fluid_cells, pinned_cell;
  double initial_max_residual, final_max_residual;
  double true_rel_residual, true_max_residual;
  double rhs_norm2;
  double ms;
} pf_solve_stats;

/* ParticleEscapeStats3D (src/driver/particle_escape3d.h:37-80) */
typedef struct pf_escape_stats {
  int32_t clamped_liquid, clamped_gas, x_lo, x_hi, y_lo, y_hi, z_lo, z_hi;
} pf_escape_stats;

enum pf_model {
  PF_MODEL_TWO_PHASE = 0,    /* SparseSim3DTP::step, sparse_sim3d_tp.cpp:854-884 */
  PF_MODEL_SINGLE_PHASE = 1  /* SparseSim3D::step, sparse_sim3d.cpp:106-115 */
};

/* The particle adaptivity controls of SparseSim3DTP / MRSim3DTP
 * (sparse_sim3d_tp.h:29-60, multires_sim3d_tp.h:80-111), under the same names.
 * pf_adapt_defaults gives the reference's defaults (every stage off). */
typedef struct pf_adapt_cfg {
  int32_t narrow_band_air, narrow_band_air_radius;
  int32_t gas_particle_coarsening, gas_particles_per_cell_target;
  uint32_t gas_particle_coarsening_seed;
  int32_t liquid_particle_coarsening, liquid_particles_per_cell_target;
  uint32_t liquid_particle_coarsening_seed;
  int32_t liquid_particle_refill, liquid_particle_refill_interface_only;
  int32_t liquid_particle_refill_interface_radius, liquid_particle_refill_max_added_per_step;
  int32_t liquid_refill_particles_per_cell_target;
  uint32_t liquid_particle_refill_seed;
} pf_adapt_cfg;

/* NarrowBandAirResult3D, GasParticleCoarseningResult3D (gas, then liquid) and
 * LiquidParticleRefillResult3D (particle_adaptivity3d.h:50-73): the *_last fields */
typedef struct pf_adapt_stats {
  int32_t narrow_band_air_removed, narrow_band_air_liquid_cells, narrow_band_air_gas_before,
      narrow_band_air_gas_after;
  int32_t gas_removed, gas_cells, gas_overfull_cells, gas_before, gas_after;
  int32_t liquid_removed, liquid_cells, liquid_overfull_cells, liquid_before, liquid_after;
  int32_t refill_added, refill_cells, refill_interface_cells, refill_underfull_cells, refill_budget_limited,
      refill_before, refill_after;
  int32_t reserved;
} pf_adapt_stats;

/* The SparseSim3DTP fields that step() reads (src/driver/sparse_sim3d_tp.h:12-28).
 * For PF_MODEL_SINGLE_PHASE (SparseSim3D, sparse_sim3d.h:5-15): alpha_liquid is the
 * one FLIP ratio, phase, Vp, alpha_gas and c_div are unused. */
typedef struct pf_step_cfg {
  double dt, gravity, Vp, alpha_liquid, alpha_gas;
  double adaptive_cfl, adaptive_min_dt;
  double c_div;              /* divergenceCorrection passed to the projection */
  int32_t adaptive_timestep, advection_order;
  int32_t timing;            /* 1: record per-stage HIP-event times into pf_step_stats */
  int32_t model;             /* pf_model */
  pf_phase phase;
  pf_solve_cfg solve;
  /* CSF surface tension (SparseSim3DTP fields, sparse_sim3d_tp.h:156-159; the
   * reference's defaults are off, 0.02, 0.25 and 0) */
  int32_t surface_tension, surface_tension_smoothing_radius;
  double surface_tension_strength, surface_tension_max_delta_speed;
  /* 1: leave pf_step_stats.iface zero instead of computing the interface
   * diagnostics the reference computes every step (sparse_sim3d_tp.cpp:865) */
  int32_t skip_interface_diagnostics, reserved;
  /* updateVolumeCorrectionStats (sparse_sim3d_tp.cpp:368-379): when set, the step
   * computes c_div from the liquid volume after particle adaptivity and the step's
   * dt, and the c_div field above is ignored */
  int32_t c_div_volume_correction, reserved2;
  double c_div_strength, liquid_volume_target;
  pf_adapt_cfg adapt;        /* applyParticleAdaptivity, first in the step (:855) */
} pf_step_cfg;

enum pf_stage {
  PF_STAGE_BIN = 0, PF_STAGE_MARK, PF_STAGE_P2G, PF_STAGE_GRID, PF_STAGE_ASSEMBLE,
  PF_STAGE_SOLVE, PF_STAGE_VELOCITY, PF_STAGE_G2P_ADVECT, PF_NSTAGES
};

/* InterfaceDiagnostics3D, src/driver/interface_diagnostics3d.h:7-19 */
typedef struct pf_iface_stats {
  int32_t sample_cells, interface_cells, finite, surface_tension_candidate;
  double phi_min, phi_max, phi_mean, grad_mean, grad_max, curvature_abs_mean, curvature_abs_max;
} pf_iface_stats;

/* SurfaceTensionStats3D, src/driver/interface_diagnostics3d.h:21-36 */
typedef struct pf_surface_stats {
  int32_t enabled, applied_cells, finite, curvature_smoothing_radius, capillary_stable, reserved;
  double strength, max_delta_speed_limit, mean_delta_speed, max_delta_speed;
  double raw_curvature_abs_mean, raw_curvature_abs_max;
  double smoothed_curvature_abs_mean, smoothed_curvature_abs_max;
  double capillary_dt_limit;
} pf_surface_stats;

typedef struct pf_step_stats {
  double effective_dt, max_particle_speed, cfl_limit_dt;
  int32_t limited, reserved;
  pf_solve_stats solve;
  pf_escape_stats escape;
  double stage_ms[PF_NSTAGES];   /* PF_STAGE_GRID includes the interface work */
  double total_ms;
  pf_iface_stats iface;          /* SparseSim3DTP::interface_diagnostics_last */
  pf_surface_stats surface;      /* SparseSim3DTP::surface_tension_stats_last */
  double liquid_volume, c_div;   /* liquid_volume_current_last, c_div_last */
  pf_adapt_stats adapt;          /* the adaptivity *_last fields */
} pf_step_stats;

typedef struct pf_ctx pf_ctx;

/* ---- lifecycle (replaces constructing SparseSim3DTP / SparseMacGrid3D<4>,
 *      src/driver/sparse_sim3d_tp.h:162) */
int pf_create(pf_ctx** out, const pf_grid_desc* desc);
int pf_destroy(pf_ctx* ctx);
const char* pf_last_error(const pf_ctx* ctx);
int pf_abi_version(void);
int pf_synchronize(pf_ctx* ctx);

/* ---- particles (replaces Particles3DTP, src/particles/particles3d_tp.h:4-35).
 * Host arrays of n entries; volume may be NULL (all 1.0).  Positions must lie in the
 * cell range [o, o + n*dx) on each axis; the caller's order is kept for download. */
int pf_particles_upload(pf_ctx* ctx, int64_t n, const float* x, const float* y, const float* z,
                        const float* u, const float* v, const float* w, const uint8_t* type,
                        const float* volume);
int pf_particles_download(pf_ctx* ctx, float* x, float* y, float* z, float* u, float* v, float* w);
/* particle types and volumes in the caller's order (either may be NULL; volumes are
 * 1.0 when none were given and no adaptivity changed them) */
int pf_particles_download_ext(pf_ctx* ctx, uint8_t* type, float* volume);
/* liquidVolume(particles, 1.0) (sparse_sim3d_tp.cpp:353-359): the summed volume of the
 * liquid particles */
double pf_liquid_volume(const pf_ctx* ctx);
int64_t pf_particle_count(const pf_ctx* ctx);
/* Seeded synthetic scene generated on the device (same construction as the oracle's
 * scene generator): 0 Rayleigh-Taylor, 1 bubble tank, 2 dam-break. */
int pf_scene_generate(pf_ctx* ctx, int32_t kind, uint64_t seed);

/* ---- step stages, in SparseSim3DTP::step order (sparse_sim3d_tp.cpp:854-884) */
/* markCells, sparse_sim3d_tp.cpp:22-50 (also bins particles by cell) */
int pf_mark_cells(pf_ctx* ctx);
/* spP2G3D_tp, sparse_ops3d_tp.cpp:95-153 / sparse_ops3d_tp.h:8 */
int pf_p2g(pf_ctx* ctx, const pf_phase* phase, double Vp);
/* spP2G3D, sparse_ops3d.cpp:72-125 / sparse_ops3d.h:6: the single-phase trilinear
 * splat with unit particle mass */
int pf_p2g_single(pf_ctx* ctx);
/* `SparseMacGrid3D<4> saved = grid`, sparse_sim3d_tp.cpp:866 (velocities only) */
int pf_save_velocity(pf_ctx* ctx);
/* applyGravity, sparse_sim3d_tp.cpp:52-70 */
int pf_apply_gravity(pf_ctx* ctx, double dt, double gravity);
/* applyWallBoundary, sparse_sim3d_tp.cpp:72-97 */
int pf_apply_walls(pf_ctx* ctx);
/* spProjectStepVC3D, sparse_ops3d_tp.cpp:234-423 / sparse_ops3d_tp.h:9-10 */
int pf_project(pf_ctx* ctx, const pf_phase* phase, double dt, const pf_solve_cfg* cfg,
               double div_correction, pf_solve_stats* stats);
/* spProjectStep3D, sparse_ops3d.cpp:170-319 / sparse_ops3d.h:9: the single-phase
 * (rho = 1) projection: unit face coefficients, no pinned cell */
int pf_project_single(pf_ctx* ctx, double dt, const pf_solve_cfg* cfg, pf_solve_stats* stats);
/* spG2P3D_tp, sparse_ops3d_tp.cpp:155-172 (saved = the pf_save_velocity copy); the
 * single-phase spG2P3D (sparse_ops3d.cpp:127-142) is alpha_liquid == alpha_gas */
int pf_g2p(pf_ctx* ctx, double alpha_liquid, double alpha_gas);
/* spAdvect3D_tp, sparse_ops3d_tp.cpp:201-232 / sparse_ops3d_tp.h:12-15; order 2 is
 * also the single-phase spAdvect3D (sparse_ops3d.cpp:144-168) */
int pf_advect(pf_ctx* ctx, double dt, int32_t order, pf_escape_stats* stats);
/* diagnoseSparseInterface3D, src/driver/interface_diagnostics3d.cpp:362-379 /
 * interface_diagnostics3d.h:38: statistics of the face-averaged phase field, its
 * gradient and curvature over the FLUID cells (masses from the last P2G) */
int pf_interface_diagnostics(pf_ctx* ctx, const pf_phase* phase, pf_iface_stats* out);
/* applySparseSurfaceTension3D, interface_diagnostics3d.cpp:404-440 / .h:44: CSF
 * velocity increments added to the faces with positive mass; out may be NULL */
int pf_apply_surface_tension(pf_ctx* ctx, const pf_phase* phase, double dt, double strength,
                             double max_delta_speed, int32_t smoothing_radius, pf_surface_stats* out);
/* SparseSim3DTP::applyParticleAdaptivity, sparse_sim3d_tp.cpp:847-852 (the same in
 * MRSim3DTP, multires_sim3d_tp.cpp:570-575): applyNarrowBandAir,
 * applyGasParticleCoarsening, applyTypedParticleCoarsening for liquid and
 * applyLiquidParticleRefill (particle_adaptivity3d.h:112-374) on the context's
 * particles, each stage only when enabled.  Survivors keep their order and refilled
 * particles follow them, so downloads return the reference's particle order. */
int pf_particle_adaptivity(pf_ctx* ctx, const pf_adapt_cfg* cfg, pf_adapt_stats* stats);
void pf_adapt_defaults(pf_adapt_cfg* cfg);
/* computeAdaptiveParticleTimestep3D, src/driver/timestep3d.h:16-50 */
int pf_timestep(pf_ctx* ctx, double requested_dt, int32_t enabled, double cfl, double min_dt,
                double* effective_dt, double* max_speed, double* cfl_limit_dt, int32_t* limited);
/* pf_timestep without a context and without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  u, v, w are caller-owned device
 * arrays of n velocities; the reduction runs on `stream` (a hipStream_t, NULL = the
 * default stream) through a fixed device scratch array of 512 KiB per device, so n is
 * at most 268,435,456 (PF_ERR_ARG beyond).  The outputs are pf_timestep's.  Calls
 * that overlap on one device share that array and race. */
int pf_timestep_noalloc(const float* u, const float* v, const float* w, int64_t n, double dx,
                        double requested_dt, int32_t enabled, double cfl, double min_dt, void* stream,
                        double* eff, double* vmax, double* cfl_dt, int32_t* limited);
/* SparseSim3DTP::step, sparse_sim3d_tp.cpp:854-884 (no particle adaptivity or
 * secondary particles; interface diagnostics and optional surface tension as the
 * reference), with G2P and advection fused;
 * cfg->model = PF_MODEL_SINGLE_PHASE runs SparseSim3D::step, sparse_sim3d.cpp:106-115 */
int pf_step(pf_ctx* ctx, const pf_step_cfg* cfg, pf_step_stats* stats);

/* ---- field access (SparseBlockGrid3D::get / ref, sparse_block_grid3d.h:46-55).
 * Dense arrays in i + fnx*(j + fny*k) order over the field's own extents
 * (faces: (nx+1,ny,nz), (nx,ny+1,nz), (nx,ny,nz+1); cells: (nx,ny,nz)). */
int pf_field_dims(const pf_ctx* ctx, int32_t field, int32_t dims[3]);
int pf_field_download(pf_ctx* ctx, int32_t field, float* dense);
/* Writes a field and activates blocks as ref() would: every block listed in
 * block_active (one byte per block id), or if NULL every block holding a non-zero. */
int pf_field_upload(pf_ctx* ctx, int32_t field, const float* dense, const uint8_t* block_active);
/* Active block ids of a field in ascending order (the set of
 * SparseBlockGrid3D::active_block_ids, sparse_block_grid3d.h:16); count returned in *n. */
int pf_active_blocks(pf_ctx* ctx, int32_t field, int32_t* bids, int64_t capacity, int64_t* n);
/* Particle count of every cell from the binning, dense i + nx*(j + ny*k) order: the
 * cell of a particle is markCells' (int)((x - o)/dx) per axis
 * (src/driver/sparse_sim3d_tp.cpp:43-46).  Bins the current particles if needed. */
int pf_bin_counts(pf_ctx* ctx, uint32_t* dense);
/* The P2G colour partition: for colour c = color8(bx, by, bz) (src/grid/block_color.h:7,
 * runColor8 in src/driver/sparse_ops3d_common.h:17-28), the cell blocks holding
 * particles, ascending, concatenated colour 0..7 into lists (capacity entries);
 * counts[c] = length of colour c's list (lists may be NULL to query the counts). */
int pf_color_lists(pf_ctx* ctx, int32_t* lists, int64_t capacity, int64_t counts[8]);
/* FP64 phase field phi of the current raw face masses of one axis (0 u, 1 v, 2 w),
 * computed on the device by the function the assembly's beta uses
 * (phiFromRawDensity, src/physics/phasefield.h:10-14); dense over the face field. */
int pf_face_phi(pf_ctx* ctx, const pf_phase* phase, int32_t axis, double* dense);

/* ---- host-side physics helpers (no device work) */
/* calibrateRhoTilde0, src/transfer/transfer3d_tp.cpp:29-36 */
double pf_calibrate_rho_tilde0(const pf_phase* phase, double Vp);
/* phiFromRawDensity and betaFromPhi, src/physics/phasefield.h:10-18 */
double pf_phi_from_raw(double raw, const pf_phase* phase);
double pf_beta_from_phi(double phi, const pf_phase* phase);

/* ---- render-cache export (src/driver/render_cache3d.h) */
/* RenderCacheCamera3D, render_cache3d.h:19-27 */
typedef struct pf_render_camera {
  double position[3], target[3], up[3];
  double fov_degrees, near_clip, far_clip, focal_length_mm;
} pf_render_camera;
/* the arguments of writeSparseRenderCache3D / writeMRRenderCache3D besides the sim
 * (render_cache3d.h:592-711): frame and time, the sim's effective_dt_last, Vp and
 * phase, and the camera */
typedef struct pf_render_frame {
  int32_t frame, reserved;
  double time, dt, Vp;
  pf_phase phase;
  pf_render_camera camera;
} pf_render_frame;
/* RenderCacheManifestFrame3D, render_cache3d.h:39-45 */
typedef struct pf_render_manifest_frame {
  int32_t frame, step;
  double time;
  const char* path;
  int64_t bytes;
} pf_render_manifest_frame;
/* defaultRenderCacheCamera3D, render_cache3d.h:48-62 */
void pf_render_default_camera(int32_t nx, int32_t ny, int32_t nz, double dx, pf_render_camera* out);
/* writeSparseRenderCache3D (a context without a multires layout) or
 * writeMRRenderCache3D (one with): the JSONL frame file of the context's grid and
 * particles, sections in the reference's order and number format.  Sparse phase
 * cells follow the marker blocks in the order markCells allocated them; that order
 * is known while the particles are still in the binning the last marking used (after
 * a step, a marking or a re-marking), else the blocks follow in id order.  *bytes
 * (may be NULL) receives the file size. */
int pf_write_render_cache(pf_ctx* ctx, const char* path, const pf_render_frame* frame, int64_t* bytes);
/* writeRenderCacheManifest3D, render_cache3d.h:713-755 */
int pf_write_render_manifest(const char* path, const char* sim_kind, int32_t nx, int32_t ny, int32_t nz, double dx,
                             const pf_render_manifest_frame* frames, int64_t nframes);

/* ---- measurement hooks (no reference counterpart: the reference times whole
 * steps with steady_clock, apps/bench_large_scale3d_tp.cpp:559-563) */
enum pf_kernel {
  PF_K_BIN_COUNT = 0, PF_K_BIN_SCATTER, PF_K_P2G, PF_K_FACE_OPS, PF_K_ASSEMBLE,
  PF_K_PCG_SPMV, PF_K_PCG_UPDATE, PF_K_RESIDUAL64, PF_K_VELOCITY, PF_K_G2P_ADVECT,
  PF_K_LPC_APPLY,      /* one application of the learned preconditioner (5 kernels) */
  PF_K_INTERFACE,      /* interface fields and diagnostics (5 kernels) */
  PF_K_SURFACE,        /* CSF increments and face updates (5 kernels) */
  PF_K_MG_APPLY,       /* one multigrid V-cycle (PF_PRECOND_MULTIGRID) */
  PF_NKERNELS
};
/* The library's HIP stream (hipStream_t). */
void* pf_stream(pf_ctx* ctx);
/* enable != 0: bracket every launch of the kernels above with HIP events on the
 * library's stream (counters reset); 0: stop. */
int pf_kernel_timing(pf_ctx* ctx, int32_t enable);
/* the same for the kernels whose bit (1 << PF_K_*) is set in mask (0: stop) */
int pf_kernel_timing_mask(pf_ctx* ctx, uint32_t mask);
/* launches and summed event time of one kernel since timing was enabled (synchronises) */
int pf_kernel_time(pf_ctx* ctx, int32_t kernel, int64_t* launches, double* total_ms);
/* plain device copy of n floats (n a multiple of 4) between two buffers allocated
 * for the call, best of four event-timed passes; a known-size kernel for
 * calibrating HBM traffic counters */
int pf_copy_probe(pf_ctx* ctx, int64_t n, double* ms);

/* ---- multires two-phase path (MRSim3DTP, src/driver/multires_sim3d_tp.h:8-191).
 * The context's grid is the fine grid; a layout of 2-level leaves (4^3 blocks at
 * level 0 = fine or 1 = coarse, MRLayout3D, src/grid/multires_layout3d.h:27-162)
 * turns it into the reference's multires grid: fields stay dense fine arrays (faces
 * the reference does not enumerate hold zero), markers and pressure are per leaf
 * cell (broadcast over its fine cells in downloads).  Positions are relative to the
 * origin 0, as the reference's multires code assumes. */
enum pf_mr_solve_mode {
  PF_MR_SOLVE_FAST = 0,       /* FP32 Jacobi PCG with FP64 refinement; the option fields below are ignored */
  PF_MR_SOLVE_REFERENCE = 1,  /* the reference's FP64 iteration (multires_pressure3d.cpp:846-1449) with the
                                 options below, validated as validateSolveConfig (:307-333) */
  PF_MR_SOLVE_MULTIGRID = 2   /* no reference counterpart: FP32 flexible PCG around an aggregation V-cycle on
                                 V A (first coarse level: the 4^3 blocks) with FP64 refinement; leaves of
                                 level 0..2, more than 8 cells per axis; option fields ignored */
};

typedef struct pf_mr_solve_cfg {    /* MRPressureSolveConfig3D (multires_pressure3d.h:137-166) */
  int32_t max_iterations;           /* cap on inner iterations (reference max_iterations) */
  int32_t check_every;              /* inner iterations between host convergence checks */
  double absolute_tolerance;        /* stop when |rhs - A x|_inf < max(abs, rel * |rhs|_inf) */
  double relative_tolerance;
  double inner_rel_tol;             /* inner FP32 PCG reduction per refinement pass */
  int32_t max_outer;
  int32_t mode;                     /* pf_mr_solve_mode */
  /* PF_MR_SOLVE_REFERENCE: the reference's options under their names */
  int32_t use_jacobi_preconditioner, use_flexible_cg_beta, adaptive_restart, relaxation_sweeps;
  double restart_growth_threshold, relaxation_omega, relaxation_min_omega;
  int32_t use_coarse_correction, coarse_correction_iterations, coarse_correction_sweeps;
  int32_t use_coarse_preconditioner;
  double coarse_correction_absolute_tolerance, coarse_correction_relative_tolerance;
  double coarse_correction_min_scale;
  int32_t coarse_preconditioner_iterations, coarse_preconditioner_auto_disable;
  int32_t coarse_preconditioner_auto_disable_after, reserved;
  double coarse_preconditioner_absolute_tolerance, coarse_preconditioner_relative_tolerance;
  double coarse_preconditioner_scale, coarse_preconditioner_min_rz_gain;
  double coarse_preconditioner_max_work_ratio;
} pf_mr_solve_cfg;

/* MRPressureSolveStats3D fields beyond pf_solve_stats (multires_pressure3d.h:60-136),
 * filled by the last multires projection in PF_MR_SOLVE_REFERENCE mode (zero, and
 * reference_order 0, after a PF_MR_SOLVE_FAST solve) */
typedef struct pf_mr_solve_detail {
  int32_t reference_order, restarts, beta_resets, relaxation_accepted, relaxation_rejected;
  int32_t coarse_correction_cells, coarse_correction_iterations, coarse_correction_sweeps;
  int32_t coarse_correction_accepted_sweeps, coarse_correction_rejected_sweeps;
  int32_t coarse_correction_accepted, coarse_correction_converged, coarse_correction_breakdown;
  int32_t coarse_preconditioner_cells, coarse_preconditioner_applications;
  int32_t coarse_preconditioner_accepted_applications, coarse_preconditioner_rejected_applications;
  int32_t coarse_preconditioner_skipped_applications, coarse_preconditioner_budget_limited_applications;
  int32_t coarse_preconditioner_iterations, coarse_preconditioner_auto_disabled;
  int32_t coarse_preconditioner_auto_disable_iteration, coarse_preconditioner_breakdown;
  int32_t coarse_preconditioner_budget_exhausted;
  double initial_residual, final_residual, effective_tolerance, relaxation_final_omega;
  double coarse_correction_effective_tolerance, coarse_correction_initial_residual;
  double coarse_correction_final_residual, coarse_correction_last_scale;
  double coarse_preconditioner_effective_tolerance, coarse_preconditioner_last_rz_gain;
} pf_mr_solve_detail;

typedef struct pf_mr_refine_cfg {   /* MRSim3DTP dynamic refinement fields (multires_sim3d_tp.h:60-68) */
  int32_t dynamic_refinement, particle_padding, gas_padding, hysteresis_cells, max_fine_leaves, reserved;
} pf_mr_refine_cfg;

typedef struct pf_mr_step_cfg {     /* the MRSim3DTP fields step() reads */
  double dt, gravity, Vp, alpha_liquid, alpha_gas;
  double adaptive_cfl, adaptive_min_dt;
  double c_div;
  int32_t adaptive_timestep, advection_order;
  int32_t timing, reserved;
  pf_phase phase;
  pf_mr_solve_cfg solve;
  pf_mr_refine_cfg refine;
  /* CSF surface tension and the diagnostics switch, as in pf_step_cfg (MRSim3DTP
   * fields, multires_sim3d_tp.h:170-175) */
  int32_t surface_tension, surface_tension_smoothing_radius;
  double surface_tension_strength, surface_tension_max_delta_speed;
  int32_t skip_interface_diagnostics, reserved2;
  /* volume correction and particle adaptivity, as in pf_step_cfg */
  int32_t c_div_volume_correction, reserved3;
  double c_div_strength, liquid_volume_target;
  pf_adapt_cfg adapt;
} pf_mr_step_cfg;

/* MRSim3DTP's solver defaults (multires_sim3d_tp.h:34-60) in PF_MR_SOLVE_FAST mode:
 * max 160 iterations, |r|_inf < max(1e-7, 1e-6 |r0|_inf), Jacobi, adaptive restart
 * (growth 10), no relaxation, coarse options off (32 iterations, 1 sweep, 1e-3, 1/64;
 * 8 iterations, 1e-2, scale 1, auto-disable after 4) */
void pf_mr_solve_defaults(pf_mr_solve_cfg* cfg);
/* the MRPressureSolveStats3D counters of the last multires projection */
int pf_mr_last_solve_detail(pf_ctx* ctx, pf_mr_solve_detail* detail);

/* MRLayout3D::leaf_blocks (multires_layout3d.h:30): n (level, bx, by, bz) leaves that
 * tile the domain; they are sorted, the grid fields cleared (MRMacGrid3D(layout)) */
int pf_mr_set_layout(pf_ctx* ctx, int64_t n, const int32_t* leaves);
int pf_mr_get_layout(pf_ctx* ctx, int32_t* leaves, int64_t capacity, int64_t* n);
/* layoutWithFineBox (multires_sim3d_tp.cpp:172-180): setCoarseEverywhere(level) then
 * refineFineCellBox(box = x0, y0, z0, x1, y1, z1) when box_valid; returns the leaf
 * count, writes at most capacity leaves (host only) */
int64_t pf_mr_layout_fine_box(int32_t nx, int32_t ny, int32_t nz, int32_t level, int32_t box_valid,
                              const int32_t* box, int32_t* leaves, int64_t capacity);
/* refineFineCellBox (multires_layout3d.h:62-97) applied to n leaves: leaves of level
 * > 0 overlapping the box (clamped to the domain) split into their in-domain children;
 * returns the new count, writes at most capacity leaves (host only) */
int64_t pf_mr_layout_refine(int32_t nx, int32_t ny, int32_t nz, const int32_t* in, int64_t n, const int32_t* box,
                            int32_t* leaves, int64_t capacity);
/* The multires multi-bubble tank, a three-level generalisation of
 * MRSim3DTP::initBubbleTankInterfaceBand (multires_sim3d_tp.cpp:457-500): water in the
 * fine cells j < ny/2, nbubbles (at most 64) gas bubbles drawn from the seed with radius
 * U[0.02, 0.06] nx dx, and a static layout of three levels by distance to the nearest
 * interface (level 0 within band fine cells, level 1 within band + 16, level 2 beyond;
 * band >= 1).  Sets the layout and seeds 8 particles per leaf cell over the interior
 * water, gas inside the bubbles.  bubbles (may be NULL) receives 4 floats per bubble:
 * centre and radius in domain units.  The origin must be 0. */
int pf_mr_scene_bubbles(pf_ctx* ctx, uint64_t seed, int32_t nbubbles, int32_t band, float* bubbles);
/* the enumerated faces of an axis (MRMacGrid3D::uFaces etc., multires_mac_grid3d.h:43-45)
 * as one byte per fine face, dense in the face field's extents */
int pf_mr_face_valid(pf_ctx* ctx, int32_t axis, uint8_t* dense);
/* markCells, multires_sim3d_tp.cpp:39-70 (also bins particles) */
int pf_mr_mark_cells(pf_ctx* ctx);
/* mrP2G3D_tp, multires_ops3d_tp.cpp:213-278 / multires_ops3d_tp.h:8 */
int pf_mr_p2g(pf_ctx* ctx, const pf_phase* phase, double Vp);
/* projectMR3D(grid, phase, dt, config, stats), multires_pressure3d.cpp:747-1524 /
 * multires_pressure3d.h:194-196; gravity and walls are pf_apply_gravity /
 * pf_apply_walls, G2P and advection pf_g2p / pf_advect (mrG2P3D_tp, mrAdvect3D_tp) */
int pf_mr_project(pf_ctx* ctx, const pf_phase* phase, double dt, const pf_mr_solve_cfg* cfg, double c_div,
                  pf_solve_stats* stats);
/* MRSim3DTP::updateDynamicRefinement, multires_sim3d_tp.cpp:577-669 (state kept in
 * the context); reports the fine leaf count and whether the budget cut the box.  With
 * cfg->dynamic_refinement 0 nothing changes and the last update's figures are reported
 * (pf_mr_step's own update included). */
int pf_mr_update_refinement(pf_ctx* ctx, const pf_mr_refine_cfg* cfg, int32_t* fine_leaves, int32_t* budget_limited);
/* the refinement box the last update kept (MRSim3DTP::dynamic_retained_box_valid and
 * dynamic_retained_x0 .. z1, multires_sim3d_tp.h:68-74) as x0, y0, z0, x1, y1, z1 */
int pf_mr_retained_box(pf_ctx* ctx, int32_t* box, int32_t* valid);
/* diagnoseMRInterface3D and applyMRSurfaceTension3D, interface_diagnostics3d.cpp:381-401,
 * :447-496 / interface_diagnostics3d.h:41, 51: the pf_interface_diagnostics /
 * pf_apply_surface_tension stencils on the fine face patches, one sample per FLUID
 * leaf cell at its fine centre, in the reference's leaf order */
int pf_mr_interface_diagnostics(pf_ctx* ctx, const pf_phase* phase, pf_iface_stats* out);
int pf_mr_apply_surface_tension(pf_ctx* ctx, const pf_phase* phase, double dt, double strength,
                                double max_delta_speed, int32_t smoothing_radius, pf_surface_stats* out);
/* MRSim3DTP::step, multires_sim3d_tp.cpp:671-734 (no particle adaptivity or secondary
 * particles; interface diagnostics and optional surface tension as the reference) */
int pf_mr_step(pf_ctx* ctx, const pf_mr_step_cfg* cfg, pf_step_stats* stats);

/* ---- learned preconditioner (no reference counterpart: the reference
 * preconditions with Jacobi, sparse_ops3d_tp.cpp:322-327).  A 3-layer 3x3x3
 * convolutional network (widths 16-32-16 and a 1x1x1 head), BF16 on MFMA with FP32
 * accumulation, over 16^3 tiles of the multigrid cycle's 2x2x2 aggregates with a
 * 1-cell halo.  Its inputs per aggregate come from beta and phi of the assembled
 * system (mean beta relative to the tile, mean phi, the beta jump inside the
 * aggregate, the Dirichlet share of the diagonal); its output o sets the weight
 * w = exp(clamp(o, -2, 2)) of that aggregate's coarse correction, applied
 * symmetrically (P S B S P^T, S = diag(sqrt(w))), so the preconditioner is a fixed
 * linear operator for the solve and the network runs once per solve.  Trained on the
 * GPU with hand-written backward kernels on residuals recorded from the solver's own
 * runs, loss ||r - A M(r)||^2 / ||r||^2 (exact gradient through the cycle), Adam. */
typedef struct pf_lpc pf_lpc;
int pf_lpc_create(pf_lpc** out, uint64_t seed);       /* He-initialised layers, zero head */
int pf_lpc_destroy(pf_lpc* lpc);
const char* pf_lpc_last_error(const pf_lpc* lpc);
int64_t pf_lpc_num_weights(void);                     /* 29,392 */
/* flat FP32 weights: W1[27][4][16], W2[27][16][32], W3[27][32][16], Wh[16];
 * tap = dx + 3 dy + 9 dz */
int pf_lpc_get_weights(pf_lpc* lpc, float* w, int64_t n);
int pf_lpc_set_weights(pf_lpc* lpc, const float* w, int64_t n);
/* network on host tiles: X[ntiles][18][18][18][4] (z, y, x, channel) -> O[ntiles][16^3] */
int pf_lpc_forward_tiles(pf_lpc* lpc, int64_t ntiles, const float* X, float* O);
/* gradient of sum(O * gO) w.r.t. the flat weights, gW[pf_lpc_num_weights()] */
int pf_lpc_grad_tiles(pf_lpc* lpc, int64_t ntiles, const float* X, const float* gO, float* gW);
/* attach a network to a context (NULL detaches); solves with precond =
 * PF_PRECOND_LEARNED use it; the network must outlive its use */
int pf_lpc_attach(pf_ctx* ctx, pf_lpc* lpc);
/* record up to max_snapshots FP32 residuals, one per `every` inner iterations, during
 * each following solve (the set is replaced by every solve); 0 stops recording */
int pf_lpc_record(pf_ctx* ctx, int32_t max_snapshots, int32_t every);
int32_t pf_lpc_snapshot_count(const pf_ctx* ctx);
/* `steps` Adam steps on the attached context's last system and its snapshots
 * (uniformly drawn from the seed); *mean_loss = mean of ||r - A M r||^2/||r||^2 */
int pf_lpc_train(pf_lpc* lpc, pf_ctx* ctx, int32_t steps, double lr, uint64_t seed, double* mean_loss);
/* the loss with the network's weights and with the plain cycle (w = 1) on one
 * recorded snapshot */
int pf_lpc_loss(pf_lpc* lpc, pf_ctx* ctx, int32_t snapshot, double* loss_learned, double* loss_plain);
/* the gradient of that loss w.r.t. the flat weights (grad[pf_lpc_num_weights()]) and
 * the loss, on one recorded snapshot */
int pf_lpc_grad(pf_lpc* lpc, pf_ctx* ctx, int32_t snapshot, float* grad, double* loss);
/* the training objective of pf_lpc_train / pf_lpc_loss / pf_lpc_grad.
 * PF_LPC_OBJECTIVE_RESIDUAL (default): ||r - A M r||^2 / ||r||^2 with r the snapshot.
 * PF_LPC_OBJECTIVE_ENERGY: the energy-norm error reduction of one preconditioned
 * Richardson step, e1^T A e1 / e0^T A e0 with e1 = e0 - M A e0, where e0 is the
 * snapshot after power_steps such steps with the same M (0..64): the slowest modes of
 * the cycle, whose reduction bounds the PCG iteration count. */
enum { PF_LPC_OBJECTIVE_RESIDUAL = 0, PF_LPC_OBJECTIVE_ENERGY = 1 };
int pf_lpc_set_objective(pf_lpc* lpc, int32_t objective, int32_t power_steps);

/* ---- multigrid preconditioner (no reference counterpart; PF_PRECOND_MULTIGRID) */
/* the built-in cycle (2 Chebyshev sweeps on [1/3, 2], 32 coarsest sweeps at omega
 * 0.9, correction 1.8); every context starts with it */
void pf_mg_defaults(pf_mg_cfg* cfg);
/* the cycle of the following multigrid solves on ctx (PF_ERR_ARG out of range) */
int pf_set_multigrid(pf_ctx* ctx, const pf_mg_cfg* cfg);
int pf_get_multigrid(const pf_ctx* ctx, pf_mg_cfg* cfg);
/* Per-aggregate weights w > 0 of the level-1 coarse correction of the following
 * multigrid solves (no reference counterpart): the cycle uses P S B S P^T with
 * S = diag(sqrt(w)) in place of P B P^T, so it stays symmetric; w = 1 everywhere is
 * the plain cycle.  w is dense over the aggregates, ceil(nx/2) x ceil(ny/2) x
 * ceil(nz/2), index i + ax*(j + ay*k); NULL clears it. */
int pf_mg_set_correction_weights(pf_ctx* ctx, const float* w, int64_t n);

#ifdef __cplusplus
}
#endif
#endif
