// multires_sim3d_tp_gpu.cpp -- host side of the GPU MRSim3DTP (see the header).
#include "multires_sim3d_tp_gpu.h"

#include <algorithm>
#include <cmath>

namespace pfgpu {

namespace {

constexpr int kBlock = 4;  // MRLayout3D<4>: leaves are 4^3 cells of their level

std::vector<MRBlockKey3D> toKeys(const std::vector<int32_t>& flat) {
  std::vector<MRBlockKey3D> out(flat.size() / 4);
  for (size_t a = 0; a < out.size(); ++a)
    out[a] = MRBlockKey3D{flat[4 * a], flat[4 * a + 1], flat[4 * a + 2], flat[4 * a + 3]};
  return out;
}

}  // namespace

MRSim3DTP::MRSim3DTP(int nx, int ny, int nz, double dx) : nx_(nx), ny_(ny), nz_(nz), dx_(dx) {
  pf_grid_desc d{};
  d.nx = nx;
  d.ny = ny;
  d.nz = nz;
  d.dx = dx;
  int st = pf_create(&ctx_, &d);
  if (st != PF_OK) throw std::runtime_error("pf_create failed with status " + std::to_string(st));
  // like MRLayout3D's constructor (multires_layout3d.h:33-34) the layout starts empty:
  // initBubbleTankInterfaceBand, setLayout or a refinement update gives it leaves
}

MRSim3DTP::~MRSim3DTP() { pf_destroy(ctx_); }

void MRSim3DTP::check(int status, const char* what) {
  if (status != PF_OK)
    throw std::runtime_error(std::string(what) + ": status " + std::to_string(status) + ": " + pf_last_error(ctx_));
}

void MRSim3DTP::setLayout(const std::vector<MRBlockKey3D>& leaves) {
  std::vector<int32_t> flat;
  flat.reserve(4 * leaves.size());
  for (const MRBlockKey3D& b : leaves) flat.insert(flat.end(), {b.level, b.bx, b.by, b.bz});
  check(pf_mr_set_layout(ctx_, (int64_t)leaves.size(), flat.data()), "pf_mr_set_layout");
  fetchLayout();
}

void MRSim3DTP::fetchLayout() {
  int64_t n = 0;
  check(pf_mr_get_layout(ctx_, nullptr, 0, &n), "pf_mr_get_layout");
  std::vector<int32_t> flat(4 * (size_t)n);
  check(pf_mr_get_layout(ctx_, flat.data(), n, &n), "pf_mr_get_layout");
  leaves_ = toKeys(flat);
}

size_t MRSim3DTP::countLevel(int level) const {
  return (size_t)std::count_if(leaves_.begin(), leaves_.end(), [&](const MRBlockKey3D& b) { return b.level == level; });
}

int MRSim3DTP::levelAtFineCell(int x, int y, int z) const {
  for (const MRBlockKey3D& b : leaves_) {
    const int s = kBlock << b.level;
    if (x / s == b.bx && y / s == b.by && z / s == b.bz) return b.level;
  }
  return -1;
}

// grid.p.leafCells().size(): the leaf cells whose fine footprint meets the domain
// (multires_scalar_grid3d.h:75-88)
int MRSim3DTP::activePressureCellCount() const {
  int64_t cells = 0;
  for (const MRBlockKey3D& b : leaves_) {
    const int step = 1 << b.level, s = kBlock * step;
    auto inside = [&](int origin, int n) {
      int c = 0;
      for (int l = 0; l < kBlock; ++l) c += origin + l * step < n ? 1 : 0;
      return c;
    };
    cells += (int64_t)inside(b.bx * s, nx_) * inside(b.by * s, ny_) * inside(b.bz * s, nz_);
  }
  return (int)cells;
}

// grid.uFaceRefs().size() etc.: the enumerated fine faces (multires_mac_grid3d.h:110-151)
int MRSim3DTP::faceCount(int axis) {
  if (leaves_.empty()) return 0;
  const size_t n = (size_t)(nx_ + (axis == 0)) * (ny_ + (axis == 1)) * (nz_ + (axis == 2));
  std::vector<uint8_t> valid(n);
  check(pf_mr_face_valid(ctx_, axis, valid.data()), "pf_mr_face_valid");
  return (int)std::count(valid.begin(), valid.end(), (uint8_t)1);
}
int MRSim3DTP::uFaceCount() { return faceCount(0); }
int MRSim3DTP::vFaceCount() { return faceCount(1); }
int MRSim3DTP::wFaceCount() { return faceCount(2); }

// initBubbleTankInterfaceBand, multires_sim3d_tp.cpp:457-503: coarse layout with the
// refineBubbleStaticBand boxes, the bubble-tank lattice, then one refinement update
void MRSim3DTP::initBubbleTankInterfaceBand() {
  particles = Particles3DTP();
  pf_phase ph{phase.rho_l, phase.rho_g, phase.alpha_phi, phase.rho_tilde_0};
  phase.rho_tilde_0 = pf_calibrate_rho_tilde0(&ph, Vp);
  dynamic_budget_limited = false;
  dynamic_last_fine_leaves = 0;
  dynamic_retained_box_valid = false;
  particle_boundary_clamped_liquid_last = particle_boundary_clamped_gas_last = 0;
  effective_dt_last = dt;
  max_particle_speed_last = 0.0;
  cfl_limit_dt_last = dt;
  adaptive_timestep_limited_last = 0;

  const int waterLevel = ny_ / 2;
  // refineBubbleStaticBand (multires_sim3d_tp.cpp:138-149): the water band, then the
  // bubble region
  const int32_t band[6] = {0, 0, 0, nx_, std::min(waterLevel + 2, ny_), nz_};
  int64_t n = pf_mr_layout_fine_box(nx_, ny_, nz_, 1, 1, band, nullptr, 0);
  std::vector<int32_t> first(4 * (size_t)n);
  pf_mr_layout_fine_box(nx_, ny_, nz_, 1, 1, band, first.data(), n);
  const int32_t bubble[6] = {nx_ / 2 - nx_ / 4, waterLevel / 4, nz_ / 2 - nz_ / 4,
                             nx_ / 2 + nx_ / 4, waterLevel,     nz_ / 2 + nz_ / 4};
  int64_t m = pf_mr_layout_refine(nx_, ny_, nz_, first.data(), n, bubble, nullptr, 0);
  if (m < 0) throw std::runtime_error("pf_mr_layout_refine failed");
  std::vector<int32_t> flat(4 * (size_t)m);
  pf_mr_layout_refine(nx_, ny_, nz_, first.data(), n, bubble, flat.data(), m);
  check(pf_mr_set_layout(ctx_, m, flat.data()), "pf_mr_set_layout");
  leaves_ = toKeys(flat);

  const double cx = nx_ * 0.5, cy = waterLevel * 0.375, cz = nz_ * 0.5;
  const double r = std::max(1.25, std::min(nx_, nz_) * 0.1875);
  for (int k = 1; k < nz_ - 1; ++k)
    for (int j = 1; j < waterLevel; ++j)
      for (int i = 1; i < nx_ - 1; ++i) {
        const double ddx = (i + 0.5) - cx, ddy = (j + 0.5) - cy, ddz = (k + 0.5) - cz;
        const unsigned char t = ddx * ddx + ddy * ddy + ddz * ddz < r * r ? 1 : 0;
        for (int s = 0; s < 8; ++s)
          particles.add(Vec3{(i + 0.25 + 0.5 * (s & 1)) * dx_, (j + 0.25 + 0.5 * ((s >> 1) & 1)) * dx_,
                             (k + 0.25 + 0.5 * ((s >> 2) & 1)) * dx_},
                        Vec3{}, t);
      }
  // resetVolumeCorrectionStats, multires_sim3d_tp.cpp:432-437
  liquid_volume_current_last = 0.0;
  for (size_t p = 0; p < particles.size(); ++p)
    if (particles.type[p] == 0) liquid_volume_current_last += particles.volume[p] * Vp;
  liquid_volume_target = liquid_volume_current_last;
  liquid_volume_error_last = 0.0;
  c_div_last = 0.0;
  dirty_ = true;
  if (dynamic_refinement) updateDynamicRefinement();
}

std::vector<float> MRSim3DTP::initMultiBubbleTank(uint64_t seed, int nbubbles, int band) {
  pf_phase ph{phase.rho_l, phase.rho_g, phase.alpha_phi, phase.rho_tilde_0};
  phase.rho_tilde_0 = pf_calibrate_rho_tilde0(&ph, Vp);
  dynamic_budget_limited = false;
  dynamic_last_fine_leaves = 0;
  dynamic_retained_box_valid = false;
  dynamic_refinement = false;  // the layout is static
  particle_boundary_clamped_liquid_last = particle_boundary_clamped_gas_last = 0;
  effective_dt_last = dt;
  max_particle_speed_last = 0.0;
  cfl_limit_dt_last = dt;
  adaptive_timestep_limited_last = 0;
  std::vector<float> bubbles(4 * (size_t)std::max(nbubbles, 0));
  check(pf_mr_scene_bubbles(ctx_, seed, nbubbles, band, bubbles.data()), "pf_mr_scene_bubbles");
  fetchLayout();
  syncParticlesToHost();
  liquid_volume_current_last = 0.0;
  for (size_t p = 0; p < particles.size(); ++p)
    if (particles.type[p] == 0) liquid_volume_current_last += particles.volume[p] * Vp;
  liquid_volume_target = liquid_volume_current_last;
  liquid_volume_error_last = 0.0;
  c_div_last = 0.0;
  dirty_ = false;
  return bubbles;
}

void MRSim3DTP::upload() {
  const size_t n = particles.size();
  std::vector<float> x(n), y(n), z(n), u(n), v(n), w(n), vol(n);
  bool unit = true;
  for (size_t p = 0; p < n; ++p) {
    x[p] = (float)particles.pos[p].x; y[p] = (float)particles.pos[p].y; z[p] = (float)particles.pos[p].z;
    u[p] = (float)particles.vel[p].x; v[p] = (float)particles.vel[p].y; w[p] = (float)particles.vel[p].z;
    vol[p] = (float)particles.volume[p];
    unit = unit && particles.volume[p] == 1.0;
  }
  check(pf_particles_upload(ctx_, (int64_t)n, x.data(), y.data(), z.data(), u.data(), v.data(), w.data(),
                            particles.type.data(), unit ? nullptr : vol.data()),
        "pf_particles_upload");
  dirty_ = false;
}

void MRSim3DTP::syncParticlesToHost() { downloadParticles(ctx_, particles); }

// updateDynamicRefinement, multires_sim3d_tp.cpp:577-669
void MRSim3DTP::updateDynamicRefinement() {
  if (dirty_) upload();
  pf_mr_refine_cfg rc{};
  rc.dynamic_refinement = dynamic_refinement ? 1 : 0;
  rc.particle_padding = dynamic_particle_padding;
  rc.gas_padding = dynamic_gas_padding;
  rc.hysteresis_cells = dynamic_hysteresis_cells;
  rc.max_fine_leaves = dynamic_max_fine_leaves;
  int32_t fine = 0, limited = 0;
  check(pf_mr_update_refinement(ctx_, &rc, &fine, &limited), "pf_mr_update_refinement");
  if (dynamic_refinement) {
    dynamic_last_fine_leaves = fine;
    dynamic_budget_limited = limited != 0;
  }
  fetchLayout();
  fetchRetained();
}

void MRSim3DTP::fetchRetained() {
  int32_t box[6], valid = 0;
  check(pf_mr_retained_box(ctx_, box, &valid), "pf_mr_retained_box");
  dynamic_retained_box_valid = valid != 0;
  dynamic_retained_x0 = box[0]; dynamic_retained_y0 = box[1]; dynamic_retained_z0 = box[2];
  dynamic_retained_x1 = box[3]; dynamic_retained_y1 = box[4]; dynamic_retained_z1 = box[5];
}

// MRSim3DTP::step, multires_sim3d_tp.cpp:671-734
void MRSim3DTP::step() {
  if (dirty_) upload();
  pf_mr_step_cfg cfg{};
  cfg.dt = dt;
  cfg.gravity = gravity;
  cfg.Vp = Vp;
  cfg.alpha_liquid = alpha_liquid;
  cfg.alpha_gas = alpha_gas;
  cfg.adaptive_cfl = adaptive_cfl;
  cfg.adaptive_min_dt = adaptive_min_dt;
  cfg.adaptive_timestep = adaptive_timestep ? 1 : 0;
  cfg.advection_order = advection_order;
  cfg.phase = pf_phase{phase.rho_l, phase.rho_g, phase.alpha_phi, phase.rho_tilde_0};
  pf_mr_solve_cfg& sc = cfg.solve;
  pf_mr_solve_defaults(&sc);
  sc.max_iterations = cg_iters;
  sc.absolute_tolerance = cg_tol;
  sc.relative_tolerance = cg_rel_tol;
  sc.use_jacobi_preconditioner = cg_jacobi_preconditioner;
  sc.use_flexible_cg_beta = cg_flexible_beta;
  sc.adaptive_restart = cg_adaptive_restart;
  sc.restart_growth_threshold = cg_restart_growth;
  sc.relaxation_sweeps = cg_relaxation_sweeps;
  sc.relaxation_omega = cg_relaxation_omega;
  sc.relaxation_min_omega = cg_relaxation_min_omega;
  sc.use_coarse_correction = cg_coarse_correction;
  sc.coarse_correction_iterations = cg_coarse_correction_iters;
  sc.coarse_correction_sweeps = cg_coarse_correction_sweeps;
  sc.coarse_correction_absolute_tolerance = cg_coarse_correction_abs_tol;
  sc.coarse_correction_relative_tolerance = cg_coarse_correction_rel_tol;
  sc.coarse_correction_min_scale = cg_coarse_correction_min_scale;
  sc.use_coarse_preconditioner = cg_coarse_preconditioner;
  sc.coarse_preconditioner_iterations = cg_coarse_preconditioner_iters;
  sc.coarse_preconditioner_absolute_tolerance = cg_coarse_preconditioner_abs_tol;
  sc.coarse_preconditioner_relative_tolerance = cg_coarse_preconditioner_rel_tol;
  sc.coarse_preconditioner_scale = cg_coarse_preconditioner_scale;
  sc.coarse_preconditioner_min_rz_gain = cg_coarse_preconditioner_min_rz_gain;
  sc.coarse_preconditioner_max_work_ratio = cg_coarse_preconditioner_max_work_ratio;
  sc.coarse_preconditioner_auto_disable = cg_coarse_preconditioner_auto_disable;
  sc.coarse_preconditioner_auto_disable_after = cg_coarse_preconditioner_auto_disable_after;
  const bool options = !cg_jacobi_preconditioner || cg_flexible_beta || cg_relaxation_sweeps > 0 ||
                       cg_coarse_correction || cg_coarse_preconditioner;
  sc.mode = (options || cg_reference_order) ? PF_MR_SOLVE_REFERENCE
                                            : (cg_multigrid ? PF_MR_SOLVE_MULTIGRID : PF_MR_SOLVE_FAST);
  cfg.refine.dynamic_refinement = dynamic_refinement ? 1 : 0;
  cfg.refine.particle_padding = dynamic_particle_padding;
  cfg.refine.gas_padding = dynamic_gas_padding;
  cfg.refine.hysteresis_cells = dynamic_hysteresis_cells;
  cfg.refine.max_fine_leaves = dynamic_max_fine_leaves;
  cfg.surface_tension = surface_tension ? 1 : 0;
  cfg.surface_tension_strength = surface_tension_strength;
  cfg.surface_tension_max_delta_speed = surface_tension_max_delta_speed;
  cfg.surface_tension_smoothing_radius = surface_tension_curvature_smoothing_radius;
  cfg.skip_interface_diagnostics = interface_diagnostics ? 0 : 1;
  cfg.adapt = adaptCfg();
  // applyParticleAdaptivity and updateVolumeCorrectionStats (multires_sim3d_tp.cpp:439-450)
  // run inside the step
  cfg.c_div_volume_correction = 1;
  cfg.c_div_strength = c_div_volume_correction ? c_div_strength : 0.0;
  cfg.liquid_volume_target = liquid_volume_target;
  check(pf_mr_step(ctx_, &cfg, &stats_last), "pf_mr_step");
  storeAdaptStats(stats_last.adapt);
  liquid_volume_current_last = stats_last.liquid_volume;
  liquid_volume_error_last = liquid_volume_target - liquid_volume_current_last;
  c_div_last = stats_last.c_div;
  fetchLayout();
  if (dynamic_refinement) {
    int32_t fine = 0, limited = 0;
    pf_mr_refine_cfg off{};
    check(pf_mr_update_refinement(ctx_, &off, &fine, &limited), "pf_mr_update_refinement");
    dynamic_last_fine_leaves = fine;
    dynamic_budget_limited = limited != 0;
    fetchRetained();
  }
  effective_dt_last = stats_last.effective_dt;
  max_particle_speed_last = stats_last.max_particle_speed;
  cfl_limit_dt_last = stats_last.cfl_limit_dt;
  adaptive_timestep_limited_last = stats_last.limited;
  interface_diagnostics_last = stats_last.iface;
  surface_tension_stats_last = stats_last.surface;
  const pf_solve_stats& s = stats_last.solve;
  last_pressure_stats.active_cells = (int)s.fluid_cells;
  last_pressure_stats.pinned_cell = (int)s.pinned_cell;
  last_pressure_stats.iterations = s.iterations;
  last_pressure_stats.initial_residual = s.initial_max_residual;
  last_pressure_stats.final_residual = s.true_max_residual;
  last_pressure_stats.effective_tolerance = std::max(cg_tol, cg_rel_tol * s.initial_max_residual);
  last_pressure_stats.converged = s.converged != 0;
  last_pressure_stats.breakdown = s.breakdown != 0;
  pf_mr_solve_detail d{};
  check(pf_mr_last_solve_detail(ctx_, &d), "pf_mr_last_solve_detail");
  if (d.reference_order) {
    MRPressureSolveStats3D& L = last_pressure_stats;
    L.final_residual = d.final_residual;
    L.effective_tolerance = d.effective_tolerance;
    L.restarts = d.restarts;
    L.beta_resets = d.beta_resets;
    L.relaxation_accepted = d.relaxation_accepted;
    L.relaxation_rejected = d.relaxation_rejected;
    L.coarse_correction_cells = d.coarse_correction_cells;
    L.coarse_correction_iterations = d.coarse_correction_iterations;
    L.coarse_correction_sweeps = d.coarse_correction_sweeps;
    L.coarse_correction_accepted_sweeps = d.coarse_correction_accepted_sweeps;
    L.coarse_correction_rejected_sweeps = d.coarse_correction_rejected_sweeps;
    L.coarse_correction_accepted = d.coarse_correction_accepted != 0;
    L.coarse_correction_last_scale = d.coarse_correction_last_scale;
    L.coarse_preconditioner_cells = d.coarse_preconditioner_cells;
    L.coarse_preconditioner_applications = d.coarse_preconditioner_applications;
    L.coarse_preconditioner_accepted_applications = d.coarse_preconditioner_accepted_applications;
    L.coarse_preconditioner_rejected_applications = d.coarse_preconditioner_rejected_applications;
    L.coarse_preconditioner_skipped_applications = d.coarse_preconditioner_skipped_applications;
    L.coarse_preconditioner_iterations = d.coarse_preconditioner_iterations;
    L.coarse_preconditioner_auto_disabled = d.coarse_preconditioner_auto_disabled != 0;
    L.coarse_preconditioner_auto_disable_iteration = d.coarse_preconditioner_auto_disable_iteration;
    L.coarse_preconditioner_last_rz_gain = d.coarse_preconditioner_last_rz_gain;
  }
  particle_boundary_clamped_liquid_last = stats_last.escape.clamped_liquid;
  particle_boundary_clamped_gas_last = stats_last.escape.clamped_gas;
}

}  // namespace pfgpu
