// sparse_sim3d_tp_gpu.cpp -- host side of the GPU SparseSim3DTP (see the header).
#include "sparse_sim3d_tp_gpu.h"

#include <cmath>

namespace pfgpu {

SparseSim3DTP::SparseSim3DTP(int nx, int ny, int nz, double dx) : nx_(nx), ny_(ny), nz_(nz), dx_(dx) {
  pf_grid_desc d{};
  d.nx = nx;
  d.ny = ny;
  d.nz = nz;
  d.dx = dx;
  int st = pf_create(&ctx_, &d);
  if (st != PF_OK) throw std::runtime_error("pf_create failed with status " + std::to_string(st));
}

SparseSim3DTP::~SparseSim3DTP() { pf_destroy(ctx_); }

void SparseSim3DTP::check(int status, const char* what) {
  if (status != PF_OK)
    throw std::runtime_error(std::string(what) + ": status " + std::to_string(status) + ": " + pf_last_error(ctx_));
}

// seedCell, src/driver/sparse_sim3d_tp.cpp:12-20: eight particles on the 2x2x2
// sub-lattice of a cell
void SparseSim3DTP::seedCell(int i, int j, int k, unsigned char type) {
  for (int s = 0; s < 8; ++s) {
    Vec3 p{(i + 0.25 + 0.5 * (s & 1)) * dx_, (j + 0.25 + 0.5 * ((s >> 1) & 1)) * dx_,
           (k + 0.25 + 0.5 * ((s >> 2) & 1)) * dx_};
    particles.add(p, Vec3{}, type);
  }
}

double SparseSim3DTP::liquidVolume() const {
  double v = 0.0;
  for (size_t p = 0; p < particles.size(); ++p)
    if (particles.type[p] == 0) v += particles.volume[p] * Vp;
  return v;
}

// resetSceneState + resetVolumeCorrectionStats, sparse_sim3d_tp.cpp:353-397
void SparseSim3DTP::resetSceneState() {
  particles = Particles3DTP();
  particle_boundary_clamped_liquid_total = 0;
  particle_boundary_clamped_gas_total = 0;
  effective_dt_last = dt;
  max_particle_speed_last = 0.0;
  cfl_limit_dt_last = dt;
  adaptive_timestep_limited_last = 0;
  pf_phase ph{phase.rho_l, phase.rho_g, phase.alpha_phi, phase.rho_tilde_0};
  phase.rho_tilde_0 = pf_calibrate_rho_tilde0(&ph, Vp);
}

// initTwoPhaseDamBreak, sparse_sim3d_tp.cpp:401-414
void SparseSim3DTP::initTwoPhaseDamBreak() {
  resetSceneState();
  const int wx = nx_ * 4 / 10, hy = ny_ * 7 / 10;
  for (int k = 1; k < nz_ - 1; ++k)
    for (int j = 1; j < ny_ - 1; ++j)
      for (int i = 1; i < nx_ - 1; ++i) seedCell(i, j, k, (i < wx && j < hy) ? 0 : 1);
  liquid_volume_current_last = liquidVolume();
  liquid_volume_target = liquid_volume_current_last;
  dirty_ = true;
}

// initRayleighTaylor, sparse_sim3d_tp.cpp:739-753
void SparseSim3DTP::initRayleighTaylor() {
  resetSceneState();
  const int mid = ny_ / 2;
  const double pi = 3.14159265358979323846;
  for (int k = 1; k < nz_ - 1; ++k)
    for (int j = 1; j < ny_ - 1; ++j)
      for (int i = 1; i < nx_ - 1; ++i) {
        double pert = std::cos(2.0 * pi * i / nx_) * std::cos(2.0 * pi * k / nz_);
        seedCell(i, j, k, (double)j > (mid + pert) ? 0 : 1);
      }
  liquid_volume_current_last = liquidVolume();
  liquid_volume_target = liquid_volume_current_last;
  dirty_ = true;
}

// initBubbleTank, sparse_sim3d_tp.cpp:755-776
void SparseSim3DTP::initBubbleTank() {
  resetSceneState();
  const int waterLevel = ny_ / 2;
  const double cx = nx_ * 0.5, cy = waterLevel * 0.375, cz = nz_ * 0.5;
  const double r = std::max(1.25, std::min(nx_, nz_) * 0.1875);
  for (int k = 1; k < nz_ - 1; ++k)
    for (int j = 1; j < waterLevel; ++j)
      for (int i = 1; i < nx_ - 1; ++i) {
        double dxc = (i + 0.5) - cx, dyc = (j + 0.5) - cy, dzc = (k + 0.5) - cz;
        seedCell(i, j, k, dxc * dxc + dyc * dyc + dzc * dzc < r * r ? 1 : 0);
      }
  liquid_volume_current_last = liquidVolume();
  liquid_volume_target = liquid_volume_current_last;
  dirty_ = true;
}

void SparseSim3DTP::upload() {
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

pf_adapt_cfg ParticleAdaptivityFields::adaptCfg() const {
  pf_adapt_cfg c{};
  c.narrow_band_air = narrow_band_air ? 1 : 0;
  c.narrow_band_air_radius = narrow_band_air_radius;
  c.gas_particle_coarsening = gas_particle_coarsening ? 1 : 0;
  c.gas_particles_per_cell_target = gas_particles_per_cell_target;
  c.gas_particle_coarsening_seed = gas_particle_coarsening_seed;
  c.liquid_particle_coarsening = liquid_particle_coarsening ? 1 : 0;
  c.liquid_particles_per_cell_target = liquid_particles_per_cell_target;
  c.liquid_particle_coarsening_seed = liquid_particle_coarsening_seed;
  c.liquid_particle_refill = liquid_particle_refill ? 1 : 0;
  c.liquid_particle_refill_interface_only = liquid_particle_refill_interface_only ? 1 : 0;
  c.liquid_particle_refill_interface_radius = liquid_particle_refill_interface_radius;
  c.liquid_particle_refill_max_added_per_step = liquid_particle_refill_max_added_per_step;
  c.liquid_refill_particles_per_cell_target = liquid_refill_particles_per_cell_target;
  c.liquid_particle_refill_seed = liquid_particle_refill_seed;
  return c;
}

void ParticleAdaptivityFields::storeAdaptStats(const pf_adapt_stats& s) {
  narrow_band_air_removed_last = s.narrow_band_air_removed;
  narrow_band_air_liquid_cells_last = s.narrow_band_air_liquid_cells;
  narrow_band_air_gas_particles_before_last = s.narrow_band_air_gas_before;
  narrow_band_air_gas_particles_after_last = s.narrow_band_air_gas_after;
  narrow_band_air_removed_total += s.narrow_band_air_removed;
  gas_particle_coarsening_removed_last = s.gas_removed;
  gas_particle_coarsening_cells_last = s.gas_cells;
  gas_particle_coarsening_overfull_cells_last = s.gas_overfull_cells;
  gas_particle_coarsening_before_last = s.gas_before;
  gas_particle_coarsening_after_last = s.gas_after;
  gas_particle_coarsening_removed_total += s.gas_removed;
  liquid_particle_coarsening_removed_last = s.liquid_removed;
  liquid_particle_coarsening_cells_last = s.liquid_cells;
  liquid_particle_coarsening_overfull_cells_last = s.liquid_overfull_cells;
  liquid_particle_coarsening_before_last = s.liquid_before;
  liquid_particle_coarsening_after_last = s.liquid_after;
  liquid_particle_coarsening_removed_total += s.liquid_removed;
  liquid_particle_refill_added_last = s.refill_added;
  liquid_particle_refill_cells_last = s.refill_cells;
  liquid_particle_refill_interface_cells_last = s.refill_interface_cells;
  liquid_particle_refill_underfull_cells_last = s.refill_underfull_cells;
  liquid_particle_refill_budget_limited_last = s.refill_budget_limited;
  liquid_particle_refill_before_last = s.refill_before;
  liquid_particle_refill_after_last = s.refill_after;
  liquid_particle_refill_added_total += s.refill_added;
}

void downloadParticles(pf_ctx* ctx, Particles3DTP& particles) {
  const int64_t n = pf_particle_count(ctx);
  if (n < 0) throw std::runtime_error("pf_particle_count: no context");
  std::vector<float> x(n), y(n), z(n), u(n), v(n), w(n), vol(n);
  std::vector<unsigned char> t(n);
  if (pf_particles_download(ctx, x.data(), y.data(), z.data(), u.data(), v.data(), w.data()) != PF_OK ||
      pf_particles_download_ext(ctx, t.data(), vol.data()) != PF_OK)
    throw std::runtime_error(std::string("pf_particles_download: ") + pf_last_error(ctx));
  particles.pos.resize(n);
  particles.vel.resize(n);
  particles.type.assign(t.begin(), t.end());
  particles.volume.resize(n);
  for (int64_t p = 0; p < n; ++p) {
    particles.pos[p] = Vec3{x[p], y[p], z[p]};
    particles.vel[p] = Vec3{u[p], v[p], w[p]};
    particles.volume[p] = vol[p];
  }
}

void SparseSim3DTP::syncParticlesToHost() { downloadParticles(ctx_, particles); }

size_t SparseSim3DTP::activeCellBlocks() {
  int64_t n = 0;
  check(pf_active_blocks(ctx_, PF_FIELD_P, nullptr, 0, &n), "pf_active_blocks");
  return (size_t)n;
}

size_t SparseSim3DTP::totalCellBlocks() const {
  return (size_t)((nx_ + 3) / 4) * ((ny_ + 3) / 4) * ((nz_ + 3) / 4);
}

// SparseSim3DTP::step, sparse_sim3d_tp.cpp:854-884
void SparseSim3DTP::step() {
  if (dirty_) upload();
  pf_step_cfg cfg{};
  cfg.dt = dt;
  cfg.gravity = gravity;
  cfg.Vp = Vp;
  cfg.alpha_liquid = alpha_liquid;
  cfg.alpha_gas = alpha_gas;
  cfg.adaptive_cfl = adaptive_cfl;
  cfg.adaptive_min_dt = adaptive_min_dt;
  cfg.adaptive_timestep = adaptive_timestep ? 1 : 0;
  cfg.advection_order = advection_order;
  cfg.timing = stage_timing ? 1 : 0;
  cfg.phase = pf_phase{phase.rho_l, phase.rho_g, phase.alpha_phi, phase.rho_tilde_0};
  cfg.solve.max_iters = cg_iters;
  cfg.solve.stop_rule = stop_rule;
  cfg.solve.tol = cg_tol;
  cfg.solve.inner_rel_tol = 0.1;
  cfg.solve.max_outer = 10000;
  cfg.solve.warm_start = warm_start;
  cfg.solve.check_every = 8;
  cfg.solve.precond = precond;
  cfg.surface_tension = surface_tension ? 1 : 0;
  cfg.surface_tension_strength = surface_tension_strength;
  cfg.surface_tension_max_delta_speed = surface_tension_max_delta_speed;
  cfg.surface_tension_smoothing_radius = surface_tension_curvature_smoothing_radius;
  cfg.skip_interface_diagnostics = interface_diagnostics ? 0 : 1;
  // applyParticleAdaptivity and updateVolumeCorrectionStats run inside the step
  // (sparse_sim3d_tp.cpp:855-862)
  cfg.adapt = adaptCfg();
  cfg.c_div_volume_correction = 1;
  cfg.c_div_strength = c_div_volume_correction ? c_div_strength : 0.0;
  cfg.liquid_volume_target = liquid_volume_target;
  check(pf_step(ctx_, &cfg, &stats_last), "pf_step");
  storeAdaptStats(stats_last.adapt);
  liquid_volume_current_last = stats_last.liquid_volume;
  liquid_volume_error_last = liquid_volume_target - liquid_volume_current_last;
  c_div_last = stats_last.c_div;
  effective_dt_last = stats_last.effective_dt;
  max_particle_speed_last = stats_last.max_particle_speed;
  cfl_limit_dt_last = stats_last.cfl_limit_dt;
  adaptive_timestep_limited_last = stats_last.limited;
  interface_diagnostics_last = stats_last.iface;
  surface_tension_stats_last = stats_last.surface;
  // storeParticleBoundaryStats, sparse_sim3d_tp.cpp:116-131
  const pf_escape_stats& e = stats_last.escape;
  particle_boundary_clamped_liquid_last = e.clamped_liquid;
  particle_boundary_clamped_gas_last = e.clamped_gas;
  particle_boundary_clamped_liquid_total += e.clamped_liquid;
  particle_boundary_clamped_gas_total += e.clamped_gas;
  particle_boundary_clamped_x_lo_last = e.x_lo;
  particle_boundary_clamped_x_hi_last = e.x_hi;
  particle_boundary_clamped_y_lo_last = e.y_lo;
  particle_boundary_clamped_y_hi_last = e.y_hi;
  particle_boundary_clamped_z_lo_last = e.z_lo;
  particle_boundary_clamped_z_hi_last = e.z_hi;
}

}  // namespace pfgpu
