// sparse_sim3d_tp_gpu.h -- C++17 drop-in for the reference's SparseSim3DTP whose step
// runs on one MI355X through the C ABI in include/pf_gpu.h.
//
// Mirrors src/driver/sparse_sim3d_tp.h:8-177 (the fields step() reads and the
// per-step stats it writes, under the same names), Particles3DTP
// (src/particles/particles3d_tp.h:4-35), PhaseParams (src/physics/phasefield.h:4-9)
// and the scene initialisers initRayleighTaylor / initBubbleTank /
// initTwoPhaseDamBreak (src/driver/sparse_sim3d_tp.cpp:401-414, 739-776).
//
// Particles live on the device between steps.  The host `particles` member is the
// caller's view: edits to it are uploaded by the next step() after
// markParticlesDirty(); syncParticlesToHost() brings the device state back (caller
// order preserved).  Errors from the ABI are thrown as std::runtime_error, as the
// reference throws from its multires solver (multires_pressure3d.cpp:307-333).
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "pf_gpu.h"

namespace pfgpu {

struct Vec3 {
  double x = 0.0, y = 0.0, z = 0.0;
};

struct Particles3DTP {
  std::vector<Vec3> pos, vel;
  std::vector<unsigned char> type;  // 0 = liquid, 1 = gas
  std::vector<double> volume;       // volume multiplier (1.0 = legacy)
  size_t size() const { return pos.size(); }
  void add(const Vec3& p, const Vec3& v, unsigned char t, double vol = 1.0) {
    pos.push_back(p);
    vel.push_back(v);
    type.push_back(t);
    volume.push_back(vol);
  }
};

// InterfaceDiagnostics3D and SurfaceTensionStats3D
// (src/driver/interface_diagnostics3d.h:7-36): the ABI structs carry the same field
// names; these give them the reference's default values
using InterfaceDiagnostics3D = pf_iface_stats;
using SurfaceTensionStats3D = pf_surface_stats;
inline InterfaceDiagnostics3D defaultInterfaceDiagnostics() {
  InterfaceDiagnostics3D d{};
  d.finite = 1;
  return d;
}
inline SurfaceTensionStats3D defaultSurfaceTensionStats() {
  SurfaceTensionStats3D s{};
  s.finite = 1;
  s.capillary_stable = 1;
  return s;
}

// The particle adaptivity fields of SparseSim3DTP and MRSim3DTP (sparse_sim3d_tp.h:29-66,
// multires_sim3d_tp.h:80-117): the controls and the *_last / *_total results under
// the reference's names and defaults.  Both classes inherit them.
struct ParticleAdaptivityFields {
  bool narrow_band_air = false;
  int narrow_band_air_radius = 2;
  int narrow_band_air_removed_last = 0, narrow_band_air_removed_total = 0;
  int narrow_band_air_liquid_cells_last = 0;
  int narrow_band_air_gas_particles_before_last = 0, narrow_band_air_gas_particles_after_last = 0;
  bool gas_particle_coarsening = false;
  int gas_particles_per_cell_target = 4;
  unsigned int gas_particle_coarsening_seed = 0x9e3779b9u;
  int gas_particle_coarsening_removed_last = 0, gas_particle_coarsening_removed_total = 0;
  int gas_particle_coarsening_cells_last = 0, gas_particle_coarsening_overfull_cells_last = 0;
  int gas_particle_coarsening_before_last = 0, gas_particle_coarsening_after_last = 0;
  bool liquid_particle_coarsening = false;
  int liquid_particles_per_cell_target = 4;
  unsigned int liquid_particle_coarsening_seed = 0x7f4a7c15u;
  int liquid_particle_coarsening_removed_last = 0, liquid_particle_coarsening_removed_total = 0;
  int liquid_particle_coarsening_cells_last = 0, liquid_particle_coarsening_overfull_cells_last = 0;
  int liquid_particle_coarsening_before_last = 0, liquid_particle_coarsening_after_last = 0;
  bool liquid_particle_refill = false;
  bool liquid_particle_refill_interface_only = false;
  int liquid_particle_refill_interface_radius = 1;
  int liquid_particle_refill_max_added_per_step = 0;
  int liquid_refill_particles_per_cell_target = 4;
  unsigned int liquid_particle_refill_seed = 0x6a09e667u;
  int liquid_particle_refill_added_last = 0, liquid_particle_refill_added_total = 0;
  int liquid_particle_refill_cells_last = 0, liquid_particle_refill_interface_cells_last = 0;
  int liquid_particle_refill_underfull_cells_last = 0, liquid_particle_refill_budget_limited_last = 0;
  int liquid_particle_refill_before_last = 0, liquid_particle_refill_after_last = 0;

  pf_adapt_cfg adaptCfg() const;
  // the store*Stats helpers of sparse_sim3d_tp.cpp:778-845
  void storeAdaptStats(const pf_adapt_stats& s);
};

struct PhaseParams {
  double rho_l = 1.0;
  double rho_g = 0.01;
  double alpha_phi = 1.0;
  double rho_tilde_0 = 1.0;
};

// Resizes `particles` to the context's count and fills positions, velocities, types
// and volumes in the caller's order (after particle adaptivity the count changes).
void downloadParticles(pf_ctx* ctx, Particles3DTP& particles);

class SparseSim3DTP : public ParticleAdaptivityFields {
 public:
  Particles3DTP particles;
  PhaseParams phase;
  double dt = 0.02, gravity = -9.81, Vp = 1.0, alpha_liquid = 0.95, alpha_gas = 0.95;
  bool adaptive_timestep = false;
  double adaptive_cfl = 0.5;
  double adaptive_min_dt = 1e-5;
  double effective_dt_last = 0.02;
  double max_particle_speed_last = 0.0;
  double cfl_limit_dt_last = 0.02;
  int adaptive_timestep_limited_last = 0;
  int advection_order = 2;
  bool c_div_volume_correction = false;
  double c_div_strength = 1.0;
  double liquid_volume_target = 0.0;
  double liquid_volume_current_last = 0.0;
  double liquid_volume_error_last = 0.0;
  double c_div_last = 0.0;
  int cg_iters = 600;
  double cg_tol = 1e-7;
  int particle_boundary_clamped_liquid_last = 0;
  int particle_boundary_clamped_gas_last = 0;
  int particle_boundary_clamped_liquid_total = 0;
  int particle_boundary_clamped_gas_total = 0;
  int particle_boundary_clamped_x_lo_last = 0;
  int particle_boundary_clamped_x_hi_last = 0;
  int particle_boundary_clamped_y_lo_last = 0;
  int particle_boundary_clamped_y_hi_last = 0;
  int particle_boundary_clamped_z_lo_last = 0;
  int particle_boundary_clamped_z_hi_last = 0;
  InterfaceDiagnostics3D interface_diagnostics_last = defaultInterfaceDiagnostics();
  bool surface_tension = false;
  double surface_tension_strength = 0.02;
  double surface_tension_max_delta_speed = 0.25;
  int surface_tension_curvature_smoothing_radius = 0;
  SurfaceTensionStats3D surface_tension_stats_last = defaultSurfaceTensionStats();
  // false: skip the per-step interface statistics (the reference always computes them)
  bool interface_diagnostics = true;

  // GPU solver controls (no reference counterpart): the reference's rule is
  // |r|_inf < cg_tol; PF_STOP_REL2 stops on ||b - Ax||_2/||b||_2 <= cg_tol in FP64
  int32_t stop_rule = PF_STOP_REF_MAXABS;
  int32_t warm_start = 0;
  // PF_PRECOND_JACOBI (the reference's), PF_PRECOND_LEARNED or PF_PRECOND_MULTIGRID
  int32_t precond = PF_PRECOND_JACOBI;
  bool stage_timing = false;
  pf_step_stats stats_last{};

  SparseSim3DTP(int nx, int ny, int nz, double dx);
  ~SparseSim3DTP();
  SparseSim3DTP(const SparseSim3DTP&) = delete;
  SparseSim3DTP& operator=(const SparseSim3DTP&) = delete;

  void initTwoPhaseDamBreak();
  void initRayleighTaylor();
  void initBubbleTank();
  void step();

  // device <-> host particle state
  void markParticlesDirty() { dirty_ = true; }
  void syncParticlesToHost();
  // SparseMacGrid3D::activeCellBlocks / totalCellBlocks (sparse_mac_grid3d.h:45-46)
  size_t activeCellBlocks();
  size_t totalCellBlocks() const;
  int nx() const { return nx_; }
  int ny() const { return ny_; }
  int nz() const { return nz_; }
  double dx() const { return dx_; }
  pf_ctx* context() const { return ctx_; }

 private:
  void check(int status, const char* what);
  void upload();
  void resetSceneState();
  void seedCell(int i, int j, int k, unsigned char type);
  double liquidVolume() const;

  pf_ctx* ctx_ = nullptr;
  int nx_, ny_, nz_;
  double dx_;
  bool dirty_ = true;
};

}  // namespace pfgpu
