// ref_harness.cpp -- C entry points around the reference's own sparse two-phase
// step, so the tests can run the real reference (built from its sources by
// oracle/Makefile into oracle/_ref/) on the same inputs as the oracle and the HIP
// path.  TEST INFRASTRUCTURE ONLY; nothing in the product links it.
//
// Wraps: SparseSim3DTP (src/driver/sparse_sim3d_tp.h:8-177), spP2G3D_tp,
// spProjectStepVC3D, spG2P3D_tp, spAdvect3D_tp (src/driver/sparse_ops3d_tp.h:8-15),
// calibrateRhoTilde0 (src/transfer/transfer3d_tp.h:5), phiFromRawDensity /
// betaFromPhi (src/physics/phasefield.h:10-18), computeAdaptiveParticleTimestep3D
// (src/driver/timestep3d.h:16); and the single-phase SparseSim3D
// (src/driver/sparse_sim3d.h:5-15) with spP2G3D / spProjectStep3D
// (src/driver/sparse_ops3d.h:5-9); and the multires MRSim3DTP
// (src/driver/multires_sim3d_tp.h) with mrP2G3D_tp / mrG2P3D_tp / mrAdvect3D_tp
// (src/driver/multires_ops3d_tp.h) and projectMR3D (src/pressure/multires_pressure3d.h);
// the particle adaptivity of src/driver/particle_adaptivity3d.h; the render-cache
// writers of src/driver/render_cache3d.h.
#include "driver/multires_ops3d_tp.h"
#include "driver/particle_adaptivity3d.h"
#include "driver/render_cache3d.h"
#include "driver/multires_sim3d_tp.h"
#include "driver/sparse_ops3d.h"
#include "driver/sparse_ops3d_tp.h"
#include "driver/sparse_sim3d.h"
#include "driver/sparse_sim3d_tp.h"
#include "driver/timestep3d.h"
#include "grid/sparse_mac_grid3d.h"
#include "physics/phasefield.h"
#include "pressure/multires_pressure3d.h"
#include "transfer/transfer3d_tp.h"

#include <chrono>
#include <cstdint>
#include <cstring>

namespace {
SparseBlockGrid3D<4>& fieldOf(SparseMacGrid3D<4>& g, int f) {
  switch (f) {
    case 0: return g.uf;
    case 1: return g.vf;
    case 2: return g.wf;
    case 3: return g.muf;
    case 4: return g.mvf;
    case 5: return g.mwf;
    case 6: return g.pf;
    default: return g.mkf;
  }
}
PhaseParams phaseOf(const double* ph) {
  PhaseParams p;
  p.rho_l = ph[0]; p.rho_g = ph[1]; p.alpha_phi = ph[2]; p.rho_tilde_0 = ph[3];
  return p;
}
// the adaptivity fields of SparseSim3DTP / MRSim3DTP (cfg: the 14 words of orc_adapt_cfg) and their
// *_last results (22 counters)
template <class S>
void set_adapt(S* s, const uint32_t* c) {
  s->narrow_band_air = c[0] != 0; s->narrow_band_air_radius = (int)c[1];
  s->gas_particle_coarsening = c[2] != 0; s->gas_particles_per_cell_target = (int)c[3];
  s->gas_particle_coarsening_seed = c[4];
  s->liquid_particle_coarsening = c[5] != 0; s->liquid_particles_per_cell_target = (int)c[6];
  s->liquid_particle_coarsening_seed = c[7];
  s->liquid_particle_refill = c[8] != 0; s->liquid_particle_refill_interface_only = c[9] != 0;
  s->liquid_particle_refill_interface_radius = (int)c[10]; s->liquid_particle_refill_max_added_per_step = (int)c[11];
  s->liquid_refill_particles_per_cell_target = (int)c[12]; s->liquid_particle_refill_seed = c[13];
}
template <class S>
void get_adapt(const S* s, int32_t* o) {
  const int32_t v[22] = {s->narrow_band_air_removed_last, s->narrow_band_air_liquid_cells_last,
                         s->narrow_band_air_gas_particles_before_last, s->narrow_band_air_gas_particles_after_last,
                         s->gas_particle_coarsening_removed_last, s->gas_particle_coarsening_cells_last,
                         s->gas_particle_coarsening_overfull_cells_last, s->gas_particle_coarsening_before_last,
                         s->gas_particle_coarsening_after_last, s->liquid_particle_coarsening_removed_last,
                         s->liquid_particle_coarsening_cells_last, s->liquid_particle_coarsening_overfull_cells_last,
                         s->liquid_particle_coarsening_before_last, s->liquid_particle_coarsening_after_last,
                         s->liquid_particle_refill_added_last, s->liquid_particle_refill_cells_last,
                         s->liquid_particle_refill_interface_cells_last, s->liquid_particle_refill_underfull_cells_last,
                         s->liquid_particle_refill_budget_limited_last, s->liquid_particle_refill_before_last,
                         s->liquid_particle_refill_after_last, 0};
  std::memcpy(o, v, sizeof(v));
}
}  // namespace

extern "C" {

void* ref_sim_create(int nx, int ny, int nz, double dx) { return new SparseSim3DTP(nx, ny, nz, dx); }
void ref_sim_destroy(void* s) { delete static_cast<SparseSim3DTP*>(s); }
void* ref_sim_clone(void* s) { return new SparseSim3DTP(*static_cast<SparseSim3DTP*>(s)); }

// cfg: dt, gravity, Vp, alpha_liquid, alpha_gas, adaptive_cfl, adaptive_min_dt, cg_tol,
//      c_div_strength, liquid_volume_target ; icfg: adaptive_timestep, advection_order,
//      cg_iters, c_div_volume_correction ; phase: rho_l, rho_g, alpha_phi, rho_tilde_0
void ref_sim_config(void* sp, const double* cfg, const int32_t* icfg, const double* phase) {
  auto* s = static_cast<SparseSim3DTP*>(sp);
  s->dt = cfg[0]; s->gravity = cfg[1]; s->Vp = cfg[2]; s->alpha_liquid = cfg[3];
  s->alpha_gas = cfg[4]; s->adaptive_cfl = cfg[5]; s->adaptive_min_dt = cfg[6];
  s->cg_tol = cfg[7]; s->c_div_strength = cfg[8]; s->liquid_volume_target = cfg[9];
  s->adaptive_timestep = icfg[0] != 0; s->advection_order = icfg[1]; s->cg_iters = icfg[2];
  s->c_div_volume_correction = icfg[3] != 0;
  s->phase = phaseOf(phase);
}

void ref_sim_set_particles(void* sp, int64_t n, const double* x, const double* y, const double* z,
                           const double* u, const double* v, const double* w, const uint8_t* t,
                           const double* vol) {
  auto* s = static_cast<SparseSim3DTP*>(sp);
  s->particles = Particles3DTP();
  for (int64_t p = 0; p < n; ++p)
    s->particles.add({x[p], y[p], z[p]}, {u[p], v[p], w[p]}, t[p], vol ? vol[p] : 1.0);
}

void ref_sim_get_particles(void* sp, double* x, double* y, double* z, double* u, double* v,
                           double* w) {
  auto* s = static_cast<SparseSim3DTP*>(sp);
  for (size_t p = 0; p < s->particles.size(); ++p) {
    x[p] = s->particles.pos[p].x; y[p] = s->particles.pos[p].y; z[p] = s->particles.pos[p].z;
    u[p] = s->particles.vel[p].x; v[p] = s->particles.vel[p].y; w[p] = s->particles.vel[p].z;
  }
}

// out: effective_dt, max_particle_speed, c_div_last, liquid_volume_current_last,
//      clamped_liquid, clamped_gas, x_lo, x_hi, y_lo, y_hi, z_lo, z_hi, ms
void ref_sim_step(void* sp, double* out) {
  auto* s = static_cast<SparseSim3DTP*>(sp);
  auto t0 = std::chrono::steady_clock::now();
  s->step();
  auto t1 = std::chrono::steady_clock::now();
  if (out) {
    out[0] = s->effective_dt_last; out[1] = s->max_particle_speed_last; out[2] = s->c_div_last;
    out[3] = s->liquid_volume_current_last;
    out[4] = s->particle_boundary_clamped_liquid_last; out[5] = s->particle_boundary_clamped_gas_last;
    out[6] = s->particle_boundary_clamped_x_lo_last; out[7] = s->particle_boundary_clamped_x_hi_last;
    out[8] = s->particle_boundary_clamped_y_lo_last; out[9] = s->particle_boundary_clamped_y_hi_last;
    out[10] = s->particle_boundary_clamped_z_lo_last; out[11] = s->particle_boundary_clamped_z_hi_last;
    out[12] = std::chrono::duration<double, std::milli>(t1 - t0).count();
  }
}

static void iface_out(const InterfaceDiagnostics3D& d, double* out) {
  out[0] = d.sample_cells; out[1] = d.interface_cells; out[2] = d.finite; out[3] = d.surface_tension_candidate;
  out[4] = d.phi_min; out[5] = d.phi_max; out[6] = d.phi_mean; out[7] = d.grad_mean; out[8] = d.grad_max;
  out[9] = d.curvature_abs_mean; out[10] = d.curvature_abs_max;
}
static void st_out(const SurfaceTensionStats3D& t, double* out) {
  out[0] = t.enabled; out[1] = t.applied_cells; out[2] = t.finite; out[3] = t.curvature_smoothing_radius;
  out[4] = t.capillary_stable; out[5] = t.strength; out[6] = t.max_delta_speed_limit; out[7] = t.mean_delta_speed;
  out[8] = t.max_delta_speed; out[9] = t.raw_curvature_abs_mean; out[10] = t.raw_curvature_abs_max;
  out[11] = t.smoothed_curvature_abs_mean; out[12] = t.smoothed_curvature_abs_max; out[13] = t.capillary_dt_limit;
}
// diagnoseSparseInterface3D on the sim's grid (11 values, InterfaceDiagnostics3D order)
void ref_interface_diagnostics(void* sp, double* out) {
  auto* s = static_cast<SparseSim3DTP*>(sp);
  iface_out(diagnoseSparseInterface3D(s->grid, s->phase), out);
}
// applySparseSurfaceTension3D on the sim's grid (14 values, SurfaceTensionStats3D order)
void ref_surface_tension(void* sp, double dt, double strength, double maxd, int radius, double* out) {
  auto* s = static_cast<SparseSim3DTP*>(sp);
  st_out(applySparseSurfaceTension3D(s->grid, s->phase, dt, strength, maxd, radius), out);
}
// the sim's surface tension flags (sparse_sim3d_tp.h:156-159)
void ref_sim_surface_tension(void* sp, int on, double strength, double maxd, int radius) {
  auto* s = static_cast<SparseSim3DTP*>(sp);
  s->surface_tension = on != 0;
  s->surface_tension_strength = strength;
  s->surface_tension_max_delta_speed = maxd;
  s->surface_tension_curvature_smoothing_radius = radius;
}
// interface_diagnostics_last (11) then surface_tension_stats_last (14)
void ref_sim_last_interface(void* sp, double* out) {
  auto* s = static_cast<SparseSim3DTP*>(sp);
  iface_out(s->interface_diagnostics_last, out);
  st_out(s->surface_tension_stats_last, out + 11);
}

void ref_field_dims(void* sp, int f, int32_t* dims) {
  auto& F = fieldOf(static_cast<SparseSim3DTP*>(sp)->grid, f);
  dims[0] = F.nx; dims[1] = F.ny; dims[2] = F.nz;
}

void ref_field_dense(void* sp, int f, double* out) {
  auto& F = fieldOf(static_cast<SparseSim3DTP*>(sp)->grid, f);
  for (int k = 0; k < F.nz; ++k)
    for (int j = 0; j < F.ny; ++j)
      for (int i = 0; i < F.nx; ++i)
        out[i + (int64_t)F.nx * (j + (int64_t)F.ny * k)] = F.get(i, j, k);
}

int64_t ref_active_blocks(void* sp, int f, int32_t* out) {
  auto& F = fieldOf(static_cast<SparseSim3DTP*>(sp)->grid, f);
  if (out) for (size_t a = 0; a < F.active_block_ids.size(); ++a) out[a] = F.active_block_ids[a];
  return (int64_t)F.active_block_ids.size();
}

void ref_set(void* sp, int f, int i, int j, int k, double v) {
  auto* s = static_cast<SparseSim3DTP*>(sp);
  if (f == 7) s->grid.setCell(i, j, k, (int)v);
  else fieldOf(s->grid, f).ref(i, j, k) = (float)v;
}

void ref_p2g(void* sp) {
  auto* s = static_cast<SparseSim3DTP*>(sp);
  spP2G3D_tp(s->grid, s->particles, s->phase, s->Vp);
}

void ref_project(void* sp, double dt, int iters, double tol, double cdiv, double* ms) {
  auto* s = static_cast<SparseSim3DTP*>(sp);
  auto t0 = std::chrono::steady_clock::now();
  spProjectStepVC3D(s->grid, s->phase, dt, iters, tol, cdiv);
  auto t1 = std::chrono::steady_clock::now();
  if (ms) *ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
}

void ref_g2p(void* sp, void* savedp, double aL, double aG) {
  auto* s = static_cast<SparseSim3DTP*>(sp);
  auto* sv = static_cast<SparseSim3DTP*>(savedp);
  spG2P3D_tp(s->grid, s->particles, sv->grid, aL, aG);
}

// stats: clamped_liquid, clamped_gas, x_lo, x_hi, y_lo, y_hi, z_lo, z_hi, escaped records
void ref_advect(void* sp, double dt, int order, int32_t* stats) {
  auto* s = static_cast<SparseSim3DTP*>(sp);
  ParticleEscapeStats3D es;
  ParticleEscapeBuffer3D eb;
  spAdvect3D_tp(s->particles, s->grid, dt, &es, order, &eb);
  if (stats) {
    stats[0] = es.clamped_liquid; stats[1] = es.clamped_gas; stats[2] = es.clamped_x_lo;
    stats[3] = es.clamped_x_hi; stats[4] = es.clamped_y_lo; stats[5] = es.clamped_y_hi;
    stats[6] = es.clamped_z_lo; stats[7] = es.clamped_z_hi; stats[8] = (int32_t)eb.records.size();
  }
}

// The reference's own scene initialisers (sparse_sim3d_tp.cpp:401-414, 739-776):
// kind 0 initRayleighTaylor, 1 initBubbleTank, 2 initTwoPhaseDamBreak.
int64_t ref_sim_init(void* sp, int kind) {
  auto* s = static_cast<SparseSim3DTP*>(sp);
  if (kind == 0) s->initRayleighTaylor();
  else if (kind == 1) s->initBubbleTank();
  else s->initTwoPhaseDamBreak();
  return (int64_t)s->particles.size();
}

void ref_sim_get_types(void* sp, uint8_t* t, double* phase) {
  auto* s = static_cast<SparseSim3DTP*>(sp);
  for (size_t p = 0; p < s->particles.size(); ++p) t[p] = s->particles.type[p];
  phase[0] = s->phase.rho_l; phase[1] = s->phase.rho_g;
  phase[2] = s->phase.alpha_phi; phase[3] = s->phase.rho_tilde_0;
}

double ref_calibrate_rho_tilde0(const double* phase, double Vp) {
  return calibrateRhoTilde0(phaseOf(phase), Vp);
}
double ref_phi(double raw, const double* phase) { return phiFromRawDensity(raw, phaseOf(phase)); }
double ref_beta(double phi, const double* phase) { return betaFromPhi(phi, phaseOf(phase)); }

void ref_timestep(int64_t n, const double* u, const double* v, const double* w, double dx,
                  double dt, int enabled, double cfl, double mindt, double* out) {
  Particles3DTP ps;
  for (int64_t p = 0; p < n; ++p) ps.add({0, 0, 0}, {u[p], v[p], w[p]}, 0);
  TimestepStats3D st = computeAdaptiveParticleTimestep3D(ps, dx, dt, enabled != 0, cfl, mindt);
  out[0] = st.effective_dt; out[1] = st.max_particle_speed; out[2] = st.cfl_limit_dt;
  out[3] = st.limited;
}

// ---- single-phase SparseSim3D
void* ref1_create(int nx, int ny, int nz, double dx) { return new SparseSim3D(nx, ny, nz, dx); }
void ref1_destroy(void* s) { delete static_cast<SparseSim3D*>(s); }
// cfg: dt, alpha, gravity, cg_tol ; cg_iters
void ref1_config(void* sp, const double* cfg, int cg_iters) {
  auto* s = static_cast<SparseSim3D*>(sp);
  s->dt = cfg[0]; s->alpha = cfg[1]; s->gravity = cfg[2]; s->cg_tol = cfg[3]; s->cg_iters = cg_iters;
}
int64_t ref1_init_dambreak(void* sp) {
  auto* s = static_cast<SparseSim3D*>(sp);
  s->initDamBreak();
  return (int64_t)s->particles.size();
}
void ref1_set_particles(void* sp, int64_t n, const double* x, const double* y, const double* z,
                        const double* u, const double* v, const double* w) {
  auto* s = static_cast<SparseSim3D*>(sp);
  s->particles = Particles3D();
  for (int64_t p = 0; p < n; ++p) s->particles.add({x[p], y[p], z[p]}, {u[p], v[p], w[p]}, 0);
}
void ref1_get_particles(void* sp, double* x, double* y, double* z, double* u, double* v, double* w) {
  auto* s = static_cast<SparseSim3D*>(sp);
  for (size_t p = 0; p < s->particles.size(); ++p) {
    x[p] = s->particles.pos[p].x; y[p] = s->particles.pos[p].y; z[p] = s->particles.pos[p].z;
    u[p] = s->particles.vel[p].x; v[p] = s->particles.vel[p].y; w[p] = s->particles.vel[p].z;
  }
}
double ref1_step(void* sp) {
  auto* s = static_cast<SparseSim3D*>(sp);
  auto t0 = std::chrono::steady_clock::now();
  s->step();
  auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double, std::milli>(t1 - t0).count();
}
void ref1_field_dense(void* sp, int f, double* out) {
  auto& F = fieldOf(static_cast<SparseSim3D*>(sp)->grid, f);
  for (int k = 0; k < F.nz; ++k)
    for (int j = 0; j < F.ny; ++j)
      for (int i = 0; i < F.nx; ++i)
        out[i + (int64_t)F.nx * (j + (int64_t)F.ny * k)] = F.get(i, j, k);
}
int64_t ref1_active_blocks(void* sp, int f, int32_t* out) {
  auto& F = fieldOf(static_cast<SparseSim3D*>(sp)->grid, f);
  if (out) for (size_t a = 0; a < F.active_block_ids.size(); ++a) out[a] = F.active_block_ids[a];
  return (int64_t)F.active_block_ids.size();
}
void ref1_set(void* sp, int f, int i, int j, int k, double v) {
  auto* s = static_cast<SparseSim3D*>(sp);
  if (f == 7) s->grid.setCell(i, j, k, (int)v);
  else fieldOf(s->grid, f).ref(i, j, k) = (float)v;
}
void ref1_p2g(void* sp) {
  auto* s = static_cast<SparseSim3D*>(sp);
  spP2G3D(s->grid, s->particles);
}
void ref1_project(void* sp, double dt, int iters, double tol) {
  auto* s = static_cast<SparseSim3D*>(sp);
  spProjectStep3D(s->grid, dt, iters, tol);
}

// ---- multires MRSim3DTP
namespace {
MRSim3DTP* mrs(void* p) { return static_cast<MRSim3DTP*>(p); }
std::map<MRFaceKey3D, float>& mrFace(MRMacGrid3D<4>& g, int f) {
  switch (f) {
    case 0: return g.ufield;
    case 1: return g.vfield;
    case 2: return g.wfield;
    case 3: return g.mu;
    case 4: return g.mv;
    default: return g.mw;
  }
}
}  // namespace

void* ref_mr_create(int nx, int ny, int nz, double dx) { return new MRSim3DTP(nx, ny, nz, dx); }
void ref_mr_destroy(void* s) { delete mrs(s); }
void* ref_mr_clone(void* s) { return new MRSim3DTP(*mrs(s)); }

void ref_mr_set_leaves(void* sp, int n, const int32_t* lv) {
  auto* s = mrs(sp);
  s->layout.leaf_blocks.clear();
  for (int a = 0; a < n; ++a) s->layout.leaf_blocks.push_back({lv[4 * a], lv[4 * a + 1], lv[4 * a + 2], lv[4 * a + 3]});
  s->layout.enforceTwoToOneBalance();
  s->grid = MRMacGrid3D<4>(s->layout);
}
int64_t ref_mr_get_leaves(void* sp, int32_t* out) {
  const auto& L = mrs(sp)->grid.layout.leaves();
  if (out)
    for (size_t a = 0; a < L.size(); ++a) {
      out[4 * a] = L[a].level; out[4 * a + 1] = L[a].bx; out[4 * a + 2] = L[a].by; out[4 * a + 3] = L[a].bz;
    }
  return (int64_t)L.size();
}
// layoutWithFineBox through the public layout API (setCoarseEverywhere + refineFineCellBox)
int64_t ref_mr_layout_fine_box(int nx, int ny, int nz, double dx, int valid, const int32_t* b, int32_t* out) {
  MRLayout3D<4> L(nx, ny, nz, dx);
  L.setCoarseEverywhere(1);
  if (valid) {
    int x0 = std::max(0, std::min(nx, b[0])), y0 = std::max(0, std::min(ny, b[1])), z0 = std::max(0, std::min(nz, b[2]));
    int x1 = std::max(0, std::min(nx, b[3])), y1 = std::max(0, std::min(ny, b[4])), z1 = std::max(0, std::min(nz, b[5]));
    if (x0 < x1 && y0 < y1 && z0 < z1) L.refineFineCellBox(x0, y0, z0, x1, y1, z1);
  }
  L.enforceTwoToOneBalance();
  const auto& lv = L.leaves();
  if (out)
    for (size_t a = 0; a < lv.size(); ++a) {
      out[4 * a] = lv[a].level; out[4 * a + 1] = lv[a].bx; out[4 * a + 2] = lv[a].by; out[4 * a + 3] = lv[a].bz;
    }
  return (int64_t)lv.size();
}
// cfg: dt, gravity, Vp, aL, aG, cfl, mindt, cg_tol, cg_rel_tol, restart_growth, relax_omega,
//      relax_min_omega, c_div_strength, liquid_volume_target
// icfg: adaptive_timestep, advection_order, cg_iters, jacobi, flexible, adaptive_restart, relax_sweeps,
//       dynamic_refinement, particle_padding, gas_padding, hysteresis, max_fine_leaves, c_div_volume_correction
void ref_mr_config(void* sp, const double* c, const int32_t* ic, const double* phase) {
  auto* s = mrs(sp);
  s->dt = c[0]; s->gravity = c[1]; s->Vp = c[2]; s->alpha_liquid = c[3]; s->alpha_gas = c[4];
  s->adaptive_cfl = c[5]; s->adaptive_min_dt = c[6]; s->cg_tol = c[7]; s->cg_rel_tol = c[8];
  s->cg_restart_growth = c[9]; s->cg_relaxation_omega = c[10]; s->cg_relaxation_min_omega = c[11];
  s->c_div_strength = c[12]; s->liquid_volume_target = c[13];
  s->adaptive_timestep = ic[0] != 0; s->advection_order = ic[1]; s->cg_iters = ic[2];
  s->cg_jacobi_preconditioner = ic[3] != 0; s->cg_flexible_beta = ic[4] != 0; s->cg_adaptive_restart = ic[5] != 0;
  s->cg_relaxation_sweeps = ic[6]; s->dynamic_refinement = ic[7] != 0; s->dynamic_particle_padding = ic[8];
  s->dynamic_gas_padding = ic[9]; s->dynamic_hysteresis_cells = ic[10]; s->dynamic_max_fine_leaves = ic[11];
  s->c_div_volume_correction = ic[12] != 0;
  s->phase = phaseOf(phase);
}
void ref_mr_set_particles(void* sp, int64_t n, const double* x, const double* y, const double* z, const double* u,
                          const double* v, const double* w, const uint8_t* t, const double* vol) {
  auto* s = mrs(sp);
  s->particles = Particles3DTP();
  for (int64_t p = 0; p < n; ++p) s->particles.add({x[p], y[p], z[p]}, {u[p], v[p], w[p]}, t[p], vol ? vol[p] : 1.0);
}
int64_t ref_mr_get_particles(void* sp, double* x, double* y, double* z, double* u, double* v, double* w, uint8_t* t) {
  auto* s = mrs(sp);
  if (x)
    for (size_t p = 0; p < s->particles.size(); ++p) {
      x[p] = s->particles.pos[p].x; y[p] = s->particles.pos[p].y; z[p] = s->particles.pos[p].z;
      u[p] = s->particles.vel[p].x; v[p] = s->particles.vel[p].y; w[p] = s->particles.vel[p].z;
      if (t) t[p] = s->particles.type[p];
    }
  return (int64_t)s->particles.size();
}
int64_t ref_mr_init_bubble(void* sp) {
  auto* s = mrs(sp);
  s->initBubbleTankInterfaceBand();
  return (int64_t)s->particles.size();
}
// out: effective_dt, c_div_last, iterations, converged, final_residual, restarts, active cells,
//      pinned cell, clamped liquid, clamped gas, ms, dynamic_last_fine_leaves, budget_limited
void ref_mr_step(void* sp, double* out) {
  auto* s = mrs(sp);
  auto t0 = std::chrono::steady_clock::now();
  s->step();
  auto t1 = std::chrono::steady_clock::now();
  if (out) {
    const auto& st = s->last_pressure_stats;
    out[0] = s->effective_dt_last; out[1] = s->c_div_last; out[2] = st.iterations; out[3] = st.converged;
    out[4] = st.final_residual; out[5] = st.restarts; out[6] = st.active_cells; out[7] = st.pinned_cell;
    out[8] = s->particle_boundary_clamped_liquid_last; out[9] = s->particle_boundary_clamped_gas_last;
    out[10] = std::chrono::duration<double, std::milli>(t1 - t0).count();
    out[11] = s->dynamic_last_fine_leaves; out[12] = s->dynamic_budget_limited;
  }
}
void ref_mr_dynamic(void* sp) { mrs(sp)->updateDynamicRefinement(); }
void ref_mr_p2g(void* sp) {
  auto* s = mrs(sp);
  mrP2G3D_tp(s->grid, s->particles, s->phase, s->Vp);
}
void ref_mr_g2p(void* sp, void* savedp, double aL, double aG) {
  mrG2P3D_tp(mrs(sp)->grid, mrs(sp)->particles, mrs(savedp)->grid, aL, aG);
}
void ref_mr_advect(void* sp, double dt, int order, int32_t* stats) {
  ParticleEscapeStats3D es;
  mrAdvect3D_tp(mrs(sp)->particles, mrs(sp)->grid, dt, &es, order, nullptr);
  if (stats) {
    stats[0] = es.clamped_liquid; stats[1] = es.clamped_gas; stats[2] = es.clamped_x_lo; stats[3] = es.clamped_x_hi;
    stats[4] = es.clamped_y_lo; stats[5] = es.clamped_y_hi; stats[6] = es.clamped_z_lo; stats[7] = es.clamped_z_hi;
  }
}
// cfg as orc_mr_solve_cfg: ic = max_iterations, jacobi, flexible, restart, relaxation sweeps,
// coarse correction (on, iterations, sweeps), coarse preconditioner (on, iterations, auto disable,
// auto disable after); c = abs tol, rel tol, growth, omega, min omega, divergence correction,
// coarse correction (abs tol, rel tol, min scale), coarse preconditioner (abs tol, rel tol, scale,
// min rz gain, max work ratio).
// out: iterations, restarts, beta_resets, relax accepted, relax rejected, converged, breakdown,
// initial residual, final residual, effective tolerance, active cells, pinned cell, then the coarse
// statistics (see the assignments below)
void ref_mr_project(void* sp, double dt, const int32_t* ic, const double* c, double* out) {
  auto* s = mrs(sp);
  MRPressureSolveConfig3D cf;
  cf.max_iterations = ic[0]; cf.use_jacobi_preconditioner = ic[1] != 0; cf.use_flexible_cg_beta = ic[2] != 0;
  cf.adaptive_restart = ic[3] != 0; cf.relaxation_sweeps = ic[4];
  cf.use_coarse_correction = ic[5] != 0; cf.coarse_correction_iterations = ic[6]; cf.coarse_correction_sweeps = ic[7];
  cf.use_coarse_preconditioner = ic[8] != 0; cf.coarse_preconditioner_iterations = ic[9];
  cf.coarse_preconditioner_auto_disable = ic[10] != 0; cf.coarse_preconditioner_auto_disable_after = ic[11];
  cf.absolute_tolerance = c[0]; cf.relative_tolerance = c[1]; cf.restart_growth_threshold = c[2];
  cf.relaxation_omega = c[3]; cf.relaxation_min_omega = c[4]; cf.divergence_correction = c[5];
  cf.coarse_correction_absolute_tolerance = c[6]; cf.coarse_correction_relative_tolerance = c[7];
  cf.coarse_correction_min_scale = c[8]; cf.coarse_preconditioner_absolute_tolerance = c[9];
  cf.coarse_preconditioner_relative_tolerance = c[10]; cf.coarse_preconditioner_scale = c[11];
  cf.coarse_preconditioner_min_rz_gain = c[12]; cf.coarse_preconditioner_max_work_ratio = c[13];
  MRPressureSolveStats3D st;
  projectMR3D(s->grid, s->phase, dt, cf, &st);
  if (out) {
    out[0] = st.iterations; out[1] = st.restarts; out[2] = st.beta_resets; out[3] = st.relaxation_accepted;
    out[4] = st.relaxation_rejected; out[5] = st.converged; out[6] = st.breakdown; out[7] = st.initial_residual;
    out[8] = st.final_residual; out[9] = st.effective_tolerance; out[10] = st.active_cells; out[11] = st.pinned_cell;
    out[12] = st.coarse_correction_cells; out[13] = st.coarse_correction_iterations;
    out[14] = st.coarse_correction_sweeps; out[15] = st.coarse_correction_accepted_sweeps;
    out[16] = st.coarse_correction_rejected_sweeps; out[17] = st.coarse_correction_accepted;
    out[18] = st.coarse_correction_last_scale; out[19] = st.coarse_correction_final_residual;
    out[20] = st.coarse_preconditioner_cells; out[21] = st.coarse_preconditioner_applications;
    out[22] = st.coarse_preconditioner_accepted_applications; out[23] = st.coarse_preconditioner_rejected_applications;
    out[24] = st.coarse_preconditioner_skipped_applications;
    out[25] = st.coarse_preconditioner_budget_limited_applications;
    out[26] = st.coarse_preconditioner_iterations; out[27] = st.coarse_preconditioner_auto_disabled;
    out[28] = st.coarse_preconditioner_auto_disable_iteration; out[29] = st.coarse_preconditioner_last_rz_gain;
  }
}
// dense views: faces over the fine face grids; P / MK per fine cell (its leaf cell's value)
void ref_mr_field_dense(void* sp, int f, double* out) {
  auto& g = mrs(sp)->grid;
  const int nx = g.layout.nx, ny = g.layout.ny, nz = g.layout.nz;
  if (f <= 5) {
    const int ax = f % 3;
    const int fx = nx + (ax == 0), fy = ny + (ax == 1), fz = nz + (ax == 2);
    std::fill(out, out + (int64_t)fx * fy * fz, 0.0);
    for (const auto& kv : mrFace(g, f))
      out[kv.first.fineX + (int64_t)fx * (kv.first.fineY + (int64_t)fy * kv.first.fineZ)] = kv.second;
    return;
  }
  const MRScalarGrid3D<4>& sg = f == 6 ? g.p : g.marker;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) out[i + (int64_t)nx * (j + (int64_t)ny * k)] = sg.get(sg.cellAtFineCell(i, j, k));
}
void ref_mr_face_valid(void* sp, int axis, uint8_t* out) {
  auto& g = mrs(sp)->grid;
  const int nx = g.layout.nx, ny = g.layout.ny, nz = g.layout.nz;
  const int fx = nx + (axis == 0), fy = ny + (axis == 1), fz = nz + (axis == 2);
  std::fill(out, out + (int64_t)fx * fy * fz, (uint8_t)0);
  const auto& faces = axis == 0 ? g.uFaceRefs() : (axis == 1 ? g.vFaceRefs() : g.wFaceRefs());
  for (const auto& k : faces) out[k.fineX + (int64_t)fx * (k.fineY + (int64_t)fy * k.fineZ)] = 1;
}
// diagnoseMRInterface3D / applyMRSurfaceTension3D on the sim's grid, the surface
// tension flags, and interface_diagnostics_last (11) then surface_tension_stats_last (14)
void ref_mr_interface_diagnostics(void* sp, double* out) {
  auto* s = mrs(sp);
  iface_out(diagnoseMRInterface3D(s->grid, s->phase), out);
}
void ref_mr_surface_tension(void* sp, double dt, double strength, double maxd, int radius, double* out) {
  auto* s = mrs(sp);
  st_out(applyMRSurfaceTension3D(s->grid, s->phase, dt, strength, maxd, radius), out);
}
void ref_mr_surface_flags(void* sp, int on, double strength, double maxd, int radius) {
  auto* s = mrs(sp);
  s->surface_tension = on != 0;
  s->surface_tension_strength = strength;
  s->surface_tension_max_delta_speed = maxd;
  s->surface_tension_curvature_smoothing_radius = radius;
}
void ref_mr_last_interface(void* sp, double* out) {
  auto* s = mrs(sp);
  iface_out(s->interface_diagnostics_last, out);
  st_out(s->surface_tension_stats_last, out + 11);
}
// particle adaptivity (particle_adaptivity3d.h) on arrays, in applyParticleAdaptivity
// order.  dom: nx, ny, nz and dx, ox, oy, oz; cfg: the 14 words of orc_adapt_cfg;
// st: 22 counters in orc_adapt_stats order.  Returns the new count, -1 past cap.
int64_t ref_pa_apply(const int32_t* dn, const double* dd, const uint32_t* cfg, int64_t n, int64_t cap, double* x,
                     double* y, double* z, double* u, double* v, double* w, uint8_t* t, double* vol, int32_t* st) {
  const pa3d::ParticleCellDomain3D dom{dn[0], dn[1], dn[2], dd[0], dd[1], dd[2], dd[3]};
  Particles3DTP P;
  for (int64_t p = 0; p < n; ++p) P.add({x[p], y[p], z[p]}, {u[p], v[p], w[p]}, t[p], vol[p]);
  const auto nb = pa3d::applyNarrowBandAir(P, dom, cfg[0] != 0, (int)cfg[1]);
  const auto gc = pa3d::applyGasParticleCoarsening(P, dom, cfg[2] != 0, (int)cfg[3], cfg[4]);
  const auto lc = pa3d::applyTypedParticleCoarsening(P, dom, cfg[5] != 0, 0, (int)cfg[6], cfg[7]);
  const auto rf = pa3d::applyLiquidParticleRefill(P, dom, cfg[8] != 0, (int)cfg[12], cfg[13], cfg[9] != 0,
                                                  (int)cfg[10], (int)cfg[11]);
  const int32_t o[22] = {nb.removed, nb.liquidCells, nb.gasBefore, nb.gasAfter,
                         gc.removed, gc.cells, gc.overfullCells, gc.particlesBefore, gc.particlesAfter,
                         lc.removed, lc.cells, lc.overfullCells, lc.particlesBefore, lc.particlesAfter,
                         rf.added, rf.cells, rf.interfaceCells, rf.underfullCells, rf.budgetLimited,
                         rf.particlesBefore, rf.particlesAfter, 0};
  std::memcpy(st, o, sizeof(o));
  if ((int64_t)P.size() > cap) return -1;
  for (size_t p = 0; p < P.size(); ++p) {
    x[p] = P.pos[p].x; y[p] = P.pos[p].y; z[p] = P.pos[p].z;
    u[p] = P.vel[p].x; v[p] = P.vel[p].y; w[p] = P.vel[p].z;
    t[p] = P.type[p]; vol[p] = P.volume[p];
  }
  return (int64_t)P.size();
}

// the sim's whole particle state (count first: adaptivity changes it)
int64_t ref_sim_particles_full(void* sp, double* x, double* y, double* z, double* u, double* v, double* w,
                               uint8_t* t, double* vol) {
  const Particles3DTP& P = static_cast<SparseSim3DTP*>(sp)->particles;
  if (x)
    for (size_t p = 0; p < P.size(); ++p) {
      x[p] = P.pos[p].x; y[p] = P.pos[p].y; z[p] = P.pos[p].z;
      u[p] = P.vel[p].x; v[p] = P.vel[p].y; w[p] = P.vel[p].z;
      t[p] = P.type[p]; vol[p] = P.volume[p];
    }
  return (int64_t)P.size();
}
void ref_sim_set_adapt(void* sp, const uint32_t* c) { set_adapt(static_cast<SparseSim3DTP*>(sp), c); }
void ref_sim_get_adapt(void* sp, int32_t* o) { get_adapt(static_cast<SparseSim3DTP*>(sp), o); }
int64_t ref_mr_particles_full(void* sp, double* x, double* y, double* z, double* u, double* v, double* w,
                              uint8_t* t, double* vol) {
  const Particles3DTP& P = mrs(sp)->particles;
  if (x)
    for (size_t p = 0; p < P.size(); ++p) {
      x[p] = P.pos[p].x; y[p] = P.pos[p].y; z[p] = P.pos[p].z;
      u[p] = P.vel[p].x; v[p] = P.vel[p].y; w[p] = P.vel[p].z;
      t[p] = P.type[p]; vol[p] = P.volume[p];
    }
  return (int64_t)P.size();
}
void ref_mr_set_adapt(void* sp, const uint32_t* c) { set_adapt(mrs(sp), c); }
void ref_mr_get_adapt(void* sp, int32_t* o) { get_adapt(mrs(sp), o); }

// writeSparseRenderCache3D / writeMRRenderCache3D with the default camera; 0, or -1
// when the writer throws
int ref_render_sparse(void* sp, const char* path, int frame, double time) {
  auto* s = static_cast<SparseSim3DTP*>(sp);
  try {
    writeSparseRenderCache3D(*s, path, frame, time,
                             defaultRenderCacheCamera3D(s->grid.nx, s->grid.ny, s->grid.nz, s->grid.dx));
  } catch (const std::exception&) {
    return -1;
  }
  return 0;
}
int ref_render_mr(void* sp, const char* path, int frame, double time) {
  auto* s = mrs(sp);
  try {
    writeMRRenderCache3D(*s, path, frame, time,
                         defaultRenderCacheCamera3D(s->layout.nx, s->layout.ny, s->layout.nz, s->layout.dx));
  } catch (const std::exception&) {
    return -1;
  }
  return 0;
}
// writeRenderCacheManifest3D; 0, or -1 when it throws
int ref_render_manifest(const char* path, const char* kind, int nx, int ny, int nz, double dx, int64_t n,
                        const int32_t* frame, const int32_t* step, const double* time, const char* const* paths,
                        const int64_t* bytes) {
  std::vector<RenderCacheManifestFrame3D> fr((size_t)n);
  for (int64_t a = 0; a < n; ++a) fr[(size_t)a] = RenderCacheManifestFrame3D{frame[a], step[a], time[a], paths[a], bytes[a]};
  try {
    writeRenderCacheManifest3D(path, kind, nx, ny, nz, dx, fr);
  } catch (const std::exception&) {
    return -1;
  }
  return 0;
}
void ref_render_default_camera(int nx, int ny, int nz, double dx, double* out) {
  const RenderCacheCamera3D c = defaultRenderCacheCamera3D(nx, ny, nz, dx);
  const double v[13] = {c.position.x, c.position.y, c.position.z, c.target.x, c.target.y, c.target.z,
                        c.up.x, c.up.y, c.up.z, c.fov_degrees, c.near_clip, c.far_clip, c.focal_length_mm};
  std::memcpy(out, v, sizeof(v));
}

void ref_mr_set(void* sp, int f, int i, int j, int k, double v) {
  auto& g = mrs(sp)->grid;
  if (f <= 5) {
    mrFace(g, f)[MRFaceKey3D{f % 3, i, j, k, 1, 1}] = (float)v;
    return;
  }
  MRScalarGrid3D<4>& sg = f == 6 ? g.p : g.marker;
  MRCellKey3D c = sg.cellAtFineCell(i, j, k);
  if (c.block.level >= 0) sg.ref(c) = (float)v;
}

}  // extern "C"
