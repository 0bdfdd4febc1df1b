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
// (src/driver/sparse_ops3d.h:5-9).
#include "driver/sparse_ops3d.h"
#include "driver/sparse_ops3d_tp.h"
#include "driver/sparse_sim3d.h"
#include "driver/sparse_sim3d_tp.h"
#include "driver/timestep3d.h"
#include "grid/sparse_mac_grid3d.h"
#include "physics/phasefield.h"
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

}  // extern "C"
