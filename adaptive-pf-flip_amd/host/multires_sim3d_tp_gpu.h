// multires_sim3d_tp_gpu.h -- C++17 drop-in for the reference's MRSim3DTP
// (src/driver/multires_sim3d_tp.h:10-191) whose step runs on one MI355X through the C
// ABI in include/pf_gpu.h (pf_mr_*).
//
// Keeps the fields step() reads and the per-step stats it writes under the
// reference's names, the layout (as MRLayout3D leaf blocks), initBubbleTankInterfaceBand,
// updateDynamicRefinement and step().  Particles live on the device between steps
// (see sparse_sim3d_tp_gpu.h); particle adaptivity and surface tension use the
// reference's fields (off by default, as in the reference); secondary particles are
// not built.
#pragma once

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "pf_gpu.h"
#include "sparse_sim3d_tp_gpu.h"

namespace pfgpu {

struct MRBlockKey3D {
  int level = 0, bx = 0, by = 0, bz = 0;
};

// the MRPressureSolveStats3D fields this library reports (multires_pressure3d.h:60-136);
// the option counters are filled when the solve ran in the reference's order
struct MRPressureSolveStats3D {
  int active_cells = 0, pinned_cell = -1, iterations = 0;
  int restarts = 0, beta_resets = 0, relaxation_accepted = 0, relaxation_rejected = 0;
  int coarse_correction_cells = 0, coarse_correction_iterations = 0, coarse_correction_sweeps = 0;
  int coarse_correction_accepted_sweeps = 0, coarse_correction_rejected_sweeps = 0;
  int coarse_preconditioner_cells = 0, coarse_preconditioner_applications = 0;
  int coarse_preconditioner_accepted_applications = 0, coarse_preconditioner_rejected_applications = 0;
  int coarse_preconditioner_skipped_applications = 0, coarse_preconditioner_iterations = 0;
  int coarse_preconditioner_auto_disable_iteration = -1;
  double initial_residual = 0.0, final_residual = 0.0, effective_tolerance = 0.0;
  double coarse_correction_last_scale = 0.0, coarse_preconditioner_last_rz_gain = 0.0;
  bool converged = false, breakdown = false, coarse_correction_accepted = false;
  bool coarse_preconditioner_auto_disabled = false;
};

class MRSim3DTP : public ParticleAdaptivityFields {
 public:
  Particles3DTP particles;
  PhaseParams phase;
  double dt = 0.02, gravity = -9.81, Vp = 1.0, alpha_liquid = 0.95, alpha_gas = 0.95;
  bool adaptive_timestep = false;
  double adaptive_cfl = 0.5, adaptive_min_dt = 1e-5;
  double effective_dt_last = 0.02, max_particle_speed_last = 0.0, cfl_limit_dt_last = 0.02;
  int adaptive_timestep_limited_last = 0;
  int advection_order = 2;
  bool c_div_volume_correction = false;
  double c_div_strength = 1.0, liquid_volume_target = 0.0, liquid_volume_current_last = 0.0;
  double liquid_volume_error_last = 0.0, c_div_last = 0.0;
  int cg_iters = 160;
  double cg_tol = 1e-7, cg_rel_tol = 1e-6;
  // the fast solve with the multires multigrid cycle (PF_MR_SOLVE_MULTIGRID, no
  // reference counterpart; leaves of level 0 to 2)
  bool cg_multigrid = false;
  // the reference's solver options (multires_sim3d_tp.h:37-60); any non-default one
  // runs the solve in the reference's FP64 order (PF_MR_SOLVE_REFERENCE)
  bool cg_jacobi_preconditioner = true, cg_flexible_beta = false, cg_adaptive_restart = true;
  double cg_restart_growth = 10.0;
  int cg_relaxation_sweeps = 0;
  double cg_relaxation_omega = 0.67, cg_relaxation_min_omega = 0.05;
  bool cg_coarse_correction = false;
  int cg_coarse_correction_iters = 32, cg_coarse_correction_sweeps = 1;
  double cg_coarse_correction_abs_tol = 0.0, cg_coarse_correction_rel_tol = 1e-3;
  double cg_coarse_correction_min_scale = 1.0 / 64.0;
  bool cg_coarse_preconditioner = false;
  int cg_coarse_preconditioner_iters = 8;
  double cg_coarse_preconditioner_abs_tol = 0.0, cg_coarse_preconditioner_rel_tol = 1e-2;
  double cg_coarse_preconditioner_scale = 1.0, cg_coarse_preconditioner_min_rz_gain = 0.0;
  double cg_coarse_preconditioner_max_work_ratio = 0.0;
  bool cg_coarse_preconditioner_auto_disable = false;
  int cg_coarse_preconditioner_auto_disable_after = 4;
  // no reference counterpart: run the reference's order even with default options
  bool cg_reference_order = false;
  bool dynamic_refinement = true;
  int dynamic_particle_padding = 1, dynamic_gas_padding = 2, dynamic_hysteresis_cells = 1;
  int dynamic_max_fine_leaves = 0;
  bool dynamic_budget_limited = false;
  int dynamic_last_fine_leaves = 0;
  bool dynamic_retained_box_valid = false;
  int dynamic_retained_x0 = 0, dynamic_retained_y0 = 0, dynamic_retained_z0 = 0;
  int dynamic_retained_x1 = 0, dynamic_retained_y1 = 0, dynamic_retained_z1 = 0;
  int particle_boundary_clamped_liquid_last = 0, particle_boundary_clamped_gas_last = 0;
  InterfaceDiagnostics3D interface_diagnostics_last = defaultInterfaceDiagnostics();
  bool surface_tension = false;
  double surface_tension_strength = 0.02;
  double surface_tension_max_delta_speed = 0.25;
  int surface_tension_curvature_smoothing_radius = 0;
  SurfaceTensionStats3D surface_tension_stats_last = defaultSurfaceTensionStats();
  // false: skip the per-step interface statistics (the reference always computes them)
  bool interface_diagnostics = true;
  MRPressureSolveStats3D last_pressure_stats;
  pf_step_stats stats_last{};

  MRSim3DTP(int nx, int ny, int nz, double dx);
  ~MRSim3DTP();
  MRSim3DTP(const MRSim3DTP&) = delete;
  MRSim3DTP& operator=(const MRSim3DTP&) = delete;

  void initBubbleTankInterfaceBand();
  // the three-level multi-bubble tank (pf_mr_scene_bubbles; no reference counterpart):
  // static layout by distance to the interfaces, particles seeded on the device; returns
  // the bubbles as centre x, y, z and radius per bubble
  std::vector<float> initMultiBubbleTank(uint64_t seed, int nbubbles = 16, int band = 6);
  void updateDynamicRefinement();
  void step();

  // the layout (MRLayout3D::leaves / countLevel / leafAtFineCell)
  const std::vector<MRBlockKey3D>& leaves() const { return leaves_; }
  size_t countLevel(int level) const;
  int levelAtFineCell(int x, int y, int z) const;
  void setLayout(const std::vector<MRBlockKey3D>& leaves);
  int activePressureCellCount() const;
  int uFaceCount();
  int vFaceCount();
  int wFaceCount();

  void markParticlesDirty() { dirty_ = true; }
  void syncParticlesToHost();
  pf_ctx* context() const { return ctx_; }

 private:
  void check(int status, const char* what);
  void upload();
  void fetchLayout();
  void fetchRetained();
  int faceCount(int axis);
  pf_ctx* ctx_ = nullptr;
  int nx_, ny_, nz_;
  double dx_;
  bool dirty_ = true;
  std::vector<MRBlockKey3D> leaves_;
};

}  // namespace pfgpu
