// sim_demo.cpp -- the reference's own SparseSim3DTP integration tests, run against
// the GPU drop-in class with the same code a reference caller would write.
//   tests/test_sparse_sim3d_tp.cpp:110-148  RT step stays stable (12 steps)
//   tests/test_sparse_sim3d_tp.cpp:749-786  bubble rises, headspace stays sparse (24 steps)
//   tests/test_sparse_sim3d_tp.cpp:150-169  adaptive timestep reports the effective dt
//   tests/test_sparse_sim3d.cpp:6-36         single-phase dam-break falls, stays sparse (20 steps)
// and the reference's MRSim3DTP tests against the multires drop-in class:
//   tests/test_multires_sim3d_tp.cpp:136-148  bubble tank: refined lower band, coarse headspace
//   tests/test_multires_sim3d_tp.cpp:345-401  dynamic refinement: occupancy, hysteresis, budget
//   tests/test_multires_sim3d_tp.cpp:403-420  bubble tank step conserves particles (4 steps)
//   tests/test_multires_sim3d_tp.cpp:422-439  adaptive timestep reports the effective dt
//   tests/test_multires_sim3d_tp.cpp:558-573  c_div uses the liquid volume error
//   tests/test_multires_sim3d_tp.cpp:575-594  high density ratio pressure converges
// plus the RT steps with the multigrid preconditioner against the Jacobi ones, and a
// step with the reference's coarse correction and coarse preconditioner on
// Host particle edits are followed by markParticlesDirty() (particles live on the
// device) and device state is read back with syncParticlesToHost().
// Exit status 0 when every check passes.
#include <algorithm>
#include <cmath>
#include <cstdio>

#include "multires_sim3d_tp_gpu.h"
#include "render_cache3d_gpu.h"
#include "sparse_sim3d_gpu.h"
#include "sparse_sim3d_tp_gpu.h"

using pfgpu::MRSim3DTP;
using pfgpu::SparseSim3D;
using pfgpu::SparseSim3DTP;
using pfgpu::RenderCacheCamera3D;
using pfgpu::RenderCacheManifestFrame3D;
using pfgpu::defaultRenderCacheCamera3D;
using pfgpu::writeRenderCacheManifest3D;
using pfgpu::writeSparseRenderCache3D;

static int failures = 0;
#define CHECK(cond)                                                   \
  do {                                                                \
    if (!(cond)) {                                                    \
      std::fprintf(stderr, "CHECK failed at line %d: %s\n", __LINE__, #cond); \
      ++failures;                                                     \
    }                                                                 \
  } while (0)

static bool close(double a, double b) { return std::fabs(a - b) <= 1e-5 * std::max(1.0, std::fabs(b)); }

template <class Sim>
static double meanY(const Sim& sim, unsigned char type) {
  double s = 0.0;
  int n = 0;
  for (size_t i = 0; i < sim.particles.size(); ++i)
    if (sim.particles.type[i] == type) {
      s += sim.particles.pos[i].y;
      ++n;
    }
  return n ? s / n : 0.0;
}

template <class Sim>
static bool finite(const Sim& sim) {
  for (const auto& p : sim.particles.pos)
    if (!std::isfinite(p.x) || !std::isfinite(p.y) || !std::isfinite(p.z)) return false;
  return true;
}

int main() {
  {
    SparseSim3DTP sim(8, 12, 8, 1.0);
    sim.dt = 0.03;
    sim.initRayleighTaylor();
    const size_t n0 = sim.particles.size();
    CHECK(n0 > 0);
    const double heavy0 = meanY(sim, 0), gas0 = meanY(sim, 1);
    CHECK(heavy0 > gas0);
    size_t maxActive = 0;
    for (int s = 0; s < 12; ++s) {
      sim.step();
      maxActive = std::max(maxActive, sim.activeCellBlocks());
    }
    sim.syncParticlesToHost();
    CHECK(sim.particles.size() == n0);
    CHECK(finite(sim));
    CHECK(meanY(sim, 0) < heavy0);
    CHECK(maxActive > 0);
    std::printf("rt: heavy mean y %.4f -> %.4f, max active cell blocks %zu\n", heavy0, meanY(sim, 0), maxActive);
  }
  {
    // the same RT steps with the reference's Jacobi preconditioner and with the
    // multigrid one (PF_PRECOND_MULTIGRID): both stop on the same FP64 residual
    SparseSim3DTP a(32, 32, 32, 1.0 / 32), b(32, 32, 32, 1.0 / 32);
    int it[2] = {0, 0};
    SparseSim3DTP* sims[2] = {&a, &b};
    for (int q = 0; q < 2; ++q) {
      SparseSim3DTP& sim = *sims[q];
      sim.dt = 2e-3;
      sim.stop_rule = PF_STOP_REL2;
      sim.cg_tol = 1e-9;
      sim.cg_iters = 20000;
      sim.precond = q == 0 ? PF_PRECOND_JACOBI : PF_PRECOND_MULTIGRID;
      sim.initRayleighTaylor();
      for (int s = 0; s < 3; ++s) {
        sim.step();
        CHECK(sim.stats_last.solve.converged == 1);
        it[q] += sim.stats_last.solve.iterations;
      }
      sim.syncParticlesToHost();
    }
    double dmax = 0.0;
    for (size_t i = 0; i < a.particles.size(); ++i)
      dmax = std::max(dmax, std::fabs(a.particles.pos[i].y - b.particles.pos[i].y));
    CHECK(a.particles.size() == b.particles.size());
    CHECK(dmax < 1e-5);
    CHECK(it[1] * 3 <= it[0]);
    std::printf("rt 32^3 x3: jacobi %d iterations, multigrid %d, max |dy| %.2e\n", it[0], it[1], dmax);
  }
  {
    SparseSim3DTP sim(8, 12, 8, 1.0);
    sim.dt = 0.03;
    sim.initBubbleTank();
    const size_t n0 = sim.particles.size();
    const double gas0 = meanY(sim, 1);
    CHECK(gas0 > 0.0);
    size_t maxActive = 0;
    for (int s = 0; s < 24; ++s) {
      sim.step();
      maxActive = std::max(maxActive, sim.activeCellBlocks());
    }
    sim.syncParticlesToHost();
    CHECK(sim.particles.size() == n0);
    CHECK(finite(sim));
    CHECK(meanY(sim, 1) > gas0 + 0.2);
    CHECK(maxActive > 0 && maxActive < sim.totalCellBlocks());
    std::printf("bubble: gas mean y %.4f -> %.4f, max active %zu of %zu\n", gas0, meanY(sim, 1), maxActive,
                sim.totalCellBlocks());
  }
  {
    SparseSim3DTP sim(8, 12, 8, 1.0);
    sim.dt = 0.2;
    sim.adaptive_timestep = true;
    sim.adaptive_cfl = 0.5;
    sim.adaptive_min_dt = 1e-5;
    sim.initBubbleTank();
    sim.particles.vel[0] = {100.0, 0.0, 0.0};
    sim.markParticlesDirty();
    sim.step();
    CHECK(std::fabs(sim.max_particle_speed_last - 100.0) < 1e-3);
    CHECK(std::fabs(sim.effective_dt_last - 0.005) < 1e-9);
    CHECK(sim.adaptive_timestep_limited_last == 1);
    std::printf("adaptive dt: %.6f\n", sim.effective_dt_last);
  }
  {
    SparseSim3D sim(16, 16, 16, 1.0);
    sim.initDamBreak();
    const size_t n0 = sim.particles.size();
    CHECK(n0 > 0);
    double my0 = 0.0;
    for (const auto& p : sim.particles.pos) my0 += p.y;
    my0 /= (double)n0;
    size_t maxActive = 0;
    for (int s = 0; s < 20; ++s) {
      sim.step();
      maxActive = std::max(maxActive, sim.activeCellBlocks());
    }
    sim.syncParticlesToHost();
    CHECK(sim.particles.size() == n0);
    bool fin = true;
    double my1 = 0.0;
    for (const auto& p : sim.particles.pos) {
      fin = fin && std::isfinite(p.x) && std::isfinite(p.y) && std::isfinite(p.z);
      my1 += p.y;
    }
    my1 /= (double)n0;
    CHECK(fin);
    CHECK(my1 < my0);
    CHECK(maxActive > 0 && maxActive < sim.totalCellBlocks());
    std::printf("single-phase dam-break: mean y %.4f -> %.4f, max active %zu of %zu\n", my0, my1, maxActive,
                sim.totalCellBlocks());
  }
  {
    MRSim3DTP sim(8, 12, 8, 1.0);
    sim.initBubbleTankInterfaceBand();
    CHECK(sim.countLevel(0) > 0);
    CHECK(sim.countLevel(1) > 0);
    CHECK(sim.particles.size() > 0);
    CHECK(meanY(sim, 1) > 0.0);
    CHECK(sim.activePressureCellCount() < 8 * 12 * 8);
    CHECK(sim.uFaceCount() > 0 && sim.vFaceCount() > 0 && sim.wFaceCount() > 0);
    std::printf("mr bubble band: %zu fine / %zu coarse leaves, %d cells, faces %d %d %d\n", sim.countLevel(0),
                sim.countLevel(1), sim.activePressureCellCount(), sim.uFaceCount(), sim.vFaceCount(),
                sim.wFaceCount());
  }
  {
    MRSim3DTP sim(16, 16, 16, 1.0);
    sim.dynamic_particle_padding = 0;
    sim.dynamic_gas_padding = 1;
    sim.particles.add({2.5, 2.5, 2.5}, {0.0, 0.0, 0.0}, 0);
    sim.particles.add({3.5, 2.5, 2.5}, {0.0, 0.0, 0.0}, 1);
    sim.updateDynamicRefinement();
    CHECK(sim.levelAtFineCell(2, 2, 2) == 0);
    CHECK(sim.levelAtFineCell(14, 14, 14) == 1);
    sim.particles.pos[0] = {13.5, 13.5, 13.5};
    sim.particles.pos[1] = {12.5, 13.5, 13.5};
    sim.markParticlesDirty();
    sim.updateDynamicRefinement();
    CHECK(sim.levelAtFineCell(13, 13, 13) == 0);
    CHECK(sim.levelAtFineCell(2, 2, 2) == 1);
  }
  {
    MRSim3DTP sim(16, 16, 16, 1.0);
    sim.dynamic_particle_padding = 0;
    sim.dynamic_gas_padding = 0;
    sim.dynamic_hysteresis_cells = 2;
    sim.particles.add({2.5, 2.5, 2.5}, {0.0, 0.0, 0.0}, 0);
    sim.updateDynamicRefinement();
    const size_t initialLeaves = sim.countLevel(0);
    sim.particles.pos[0] = {4.5, 2.5, 2.5};
    sim.markParticlesDirty();
    sim.updateDynamicRefinement();
    CHECK(sim.dynamic_retained_box_valid);
    CHECK(sim.countLevel(0) == initialLeaves);
    CHECK(sim.levelAtFineCell(2, 2, 2) == 0);
    CHECK(sim.levelAtFineCell(4, 2, 2) == 0);
  }
  {
    MRSim3DTP sim(32, 32, 32, 1.0);
    sim.dynamic_particle_padding = 8;
    sim.dynamic_gas_padding = 0;
    sim.dynamic_hysteresis_cells = 0;
    sim.dynamic_max_fine_leaves = 8;
    sim.particles.add({2.5, 2.5, 2.5}, {0.0, 0.0, 0.0}, 0);
    sim.particles.add({29.5, 29.5, 29.5}, {0.0, 0.0, 0.0}, 0);
    sim.updateDynamicRefinement();
    CHECK(sim.dynamic_budget_limited);
    CHECK(sim.dynamic_last_fine_leaves == (int)sim.countLevel(0));
    CHECK(sim.countLevel(0) <= (size_t)sim.dynamic_max_fine_leaves);
    std::printf("mr budget: %zu fine leaves (limit 8)\n", sim.countLevel(0));
  }
  {
    MRSim3DTP sim(8, 12, 8, 1.0);
    sim.dt = 0.02;
    sim.cg_iters = 120;
    sim.initBubbleTankInterfaceBand();
    const size_t n0 = sim.particles.size();
    const double gas0 = meanY(sim, 1);
    for (int s = 0; s < 4; ++s) sim.step();
    sim.syncParticlesToHost();
    CHECK(sim.particles.size() == n0);
    CHECK(finite(sim));
    CHECK(std::isfinite(meanY(sim, 1)));
    CHECK(meanY(sim, 1) >= gas0 - 0.05);
    CHECK(sim.activePressureCellCount() < 8 * 12 * 8);
    std::printf("mr bubble step: gas mean y %.4f -> %.4f, %d pressure cells\n", gas0, meanY(sim, 1),
                sim.activePressureCellCount());
  }
  {
    MRSim3DTP sim(8, 12, 8, 1.0);
    sim.dt = 0.2;
    sim.cg_iters = 20;
    sim.adaptive_timestep = true;
    sim.adaptive_cfl = 0.5;
    sim.adaptive_min_dt = 1e-5;
    sim.initBubbleTankInterfaceBand();
    CHECK(!sim.particles.vel.empty());
    sim.particles.vel[0] = {100.0, 0.0, 0.0};
    sim.markParticlesDirty();
    sim.step();
    CHECK(close(sim.dt, 0.2));
    CHECK(close(sim.max_particle_speed_last, 100.0));
    CHECK(close(sim.effective_dt_last, 0.005));
    CHECK(sim.adaptive_timestep_limited_last == 1);
    std::printf("mr adaptive dt: %.6f\n", sim.effective_dt_last);
  }
  {
    MRSim3DTP sim(8, 12, 8, 1.0);
    sim.dt = 0.02;
    sim.cg_iters = 120;
    sim.c_div_volume_correction = true;
    sim.c_div_strength = 1.0;
    sim.initBubbleTankInterfaceBand();
    double liquidVolume = 0.0;
    for (size_t i = 0; i < sim.particles.size(); ++i)
      if (sim.particles.type[i] == 0) liquidVolume += sim.particles.volume[i];
    sim.liquid_volume_target = liquidVolume + 1.0;
    sim.step();
    CHECK(close(sim.liquid_volume_current_last, liquidVolume));
    CHECK(close(sim.liquid_volume_error_last, 1.0));
    CHECK(close(sim.c_div_last, 1.0 / (sim.dt * sim.liquid_volume_target)));
  }
  {
    MRSim3DTP sim(8, 12, 8, 1.0);
    sim.phase.rho_l = 1000.0;
    sim.phase.rho_g = 1.0;
    sim.dt = 0.02;
    sim.cg_iters = 160;
    sim.cg_rel_tol = 1e-5;
    sim.initBubbleTankInterfaceBand();
    const size_t n0 = sim.particles.size();
    sim.step();
    sim.syncParticlesToHost();
    CHECK(sim.particles.size() == n0);
    CHECK(finite(sim));
    CHECK(sim.last_pressure_stats.converged);
    CHECK(sim.last_pressure_stats.final_residual <= sim.last_pressure_stats.effective_tolerance);
    CHECK(!sim.last_pressure_stats.breakdown);
    std::printf("mr rho ratio 1000: %d iterations, residual %.3e <= %.3e\n", sim.last_pressure_stats.iterations,
                sim.last_pressure_stats.final_residual, sim.last_pressure_stats.effective_tolerance);
  }
  {
    // the reference's coarse options through the class fields: the solve runs in the
    // reference's FP64 order and reports the option counters
    MRSim3DTP sim(16, 16, 16, 1.0);
    sim.dt = 0.02;
    sim.cg_iters = 400;
    sim.cg_coarse_correction = true;
    sim.cg_coarse_preconditioner = true;
    sim.initBubbleTankInterfaceBand();
    const size_t n0 = sim.particles.size();
    sim.step();
    sim.syncParticlesToHost();
    const auto& st = sim.last_pressure_stats;
    CHECK(sim.particles.size() == n0);
    CHECK(finite(sim));
    CHECK(st.converged);
    CHECK(st.final_residual < st.effective_tolerance);
    CHECK(st.coarse_correction_cells > 0 && st.coarse_correction_sweeps > 0);
    CHECK(st.coarse_preconditioner_applications > 0);
    std::printf("mr coarse options: %d iterations, %d coarse sweeps (%d accepted), %d coarse preconditioner "
                "applications (%d accepted)\n", st.iterations, st.coarse_correction_sweeps,
                st.coarse_correction_accepted_sweeps, st.coarse_preconditioner_applications,
                st.coarse_preconditioner_accepted_applications);
  }
  {
    // surface tension through the class fields (sparse and multires): the step
    // reports interface diagnostics every step and CSF statistics when enabled
    SparseSim3DTP sim(16, 16, 16, 1.0);
    sim.dt = 0.01;
    sim.initBubbleTank();
    sim.step();
    CHECK(sim.interface_diagnostics_last.interface_cells > 0 && sim.interface_diagnostics_last.finite == 1);
    CHECK(sim.surface_tension_stats_last.enabled == 0 && sim.surface_tension_stats_last.finite == 1);
    sim.surface_tension = true;
    sim.surface_tension_strength = 0.5;
    sim.surface_tension_curvature_smoothing_radius = 1;
    sim.step();
    const auto& st = sim.surface_tension_stats_last;
    CHECK(st.enabled == 1 && st.applied_cells > 0 && st.finite == 1 && st.curvature_smoothing_radius == 1);
    CHECK(st.max_delta_speed <= sim.surface_tension_max_delta_speed);
    MRSim3DTP mr(16, 16, 16, 1.0);
    mr.dt = 0.01;
    mr.surface_tension = true;
    mr.initBubbleTankInterfaceBand();
    mr.step();
    CHECK(mr.interface_diagnostics_last.interface_cells > 0);
    CHECK(mr.surface_tension_stats_last.enabled == 1 && mr.surface_tension_stats_last.applied_cells > 0);
    std::printf("surface tension: sparse %d interface cells, %d CSF cells (max |dv| %.3e); multires %d / %d\n",
                sim.interface_diagnostics_last.interface_cells, st.applied_cells, st.max_delta_speed,
                mr.interface_diagnostics_last.interface_cells, mr.surface_tension_stats_last.applied_cells);
  }
  {
    // particle adaptivity through the class fields: coarsening thins the 2x2x2 seeding,
    // refill tops underfull liquid cells back up, and the host view follows the new count
    SparseSim3DTP sim(12, 12, 12, 1.0);
    sim.dt = 0.01;
    sim.initBubbleTank();
    const size_t n0 = sim.particles.size();
    sim.gas_particle_coarsening = true;
    sim.liquid_particle_coarsening = true;
    sim.liquid_particles_per_cell_target = 5;
    sim.liquid_particle_refill = true;
    sim.liquid_refill_particles_per_cell_target = 6;
    sim.step();
    sim.syncParticlesToHost();
    const int removed = sim.gas_particle_coarsening_removed_last + sim.liquid_particle_coarsening_removed_last;
    CHECK(removed > 0 && sim.liquid_particle_refill_added_last > 0);
    CHECK(sim.particles.size() == n0 - (size_t)removed + (size_t)sim.liquid_particle_refill_added_last);
    CHECK(sim.particles.volume.size() == sim.particles.size() && finite(sim));
    double lv = 0.0;
    for (size_t p = 0; p < sim.particles.size(); ++p)
      if (sim.particles.type[p] == 0) lv += sim.particles.volume[p];
    CHECK(std::fabs(lv - sim.liquid_volume_current_last) <= 1e-4 * lv);
    std::printf("adaptivity: %zu -> %zu particles (%d removed, %d refilled), liquid volume %.3f\n", n0,
                sim.particles.size(), removed, sim.liquid_particle_refill_added_last, lv);
  }
  {
    // render-cache export (export_render_cache3d.cpp's loop): frames and a manifest
    SparseSim3DTP sim(12, 12, 12, 1.0);
    sim.dt = 0.01;
    sim.initBubbleTank();
    std::vector<RenderCacheManifestFrame3D> frames;
    const RenderCacheCamera3D cam = defaultRenderCacheCamera3D(12, 12, 12, 1.0);
    for (int f = 0; f < 2; ++f) {
      sim.step();
      const std::string path = "/tmp/pf_demo_frame_" + std::to_string(f) + ".jsonl";
      const long long bytes = writeSparseRenderCache3D(sim, path, f, 0.01 * (f + 1), cam);
      frames.push_back(RenderCacheManifestFrame3D{f, f + 1, 0.01 * (f + 1), path, bytes});
      CHECK(bytes > 0);
    }
    writeRenderCacheManifest3D("/tmp/pf_demo_manifest.json", "sparse3d_tp", 12, 12, 12, 1.0, frames);
    bool threw = false;
    RenderCacheCamera3D bad = cam;
    bad.far_clip = bad.near_clip;
    try {
      writeSparseRenderCache3D(sim, "/tmp/pf_demo_bad.jsonl", 0, 0.0, bad);
    } catch (const std::invalid_argument&) {
      threw = true;
    }
    CHECK(threw);
    std::printf("render cache: 2 frames, %lld and %lld bytes\n", frames[0].bytes, frames[1].bytes);
  }
  {
    // the three-level multi-bubble tank with the multires multigrid cycle (no reference
    // counterpart): levels 0 to 2 present, the solve converges to the reference's stop
    // rule, particles stay finite
    MRSim3DTP sim(96, 96, 96, 1.0);
    sim.phase.rho_l = 1.0;
    sim.phase.rho_g = 0.01;
    sim.dt = 0.02;
    sim.cg_iters = 2000;
    sim.cg_tol = 0.0;
    sim.cg_rel_tol = 1e-6;
    sim.cg_multigrid = true;
    const std::vector<float> bubbles = sim.initMultiBubbleTank(7, 4, 6);
    CHECK(bubbles.size() == 16);
    CHECK(sim.countLevel(0) > 0 && sim.countLevel(1) > 0 && sim.countLevel(2) > 0);
    const size_t n0 = sim.particles.size();
    CHECK(n0 > 0);
    sim.step();
    CHECK(sim.stats_last.solve.converged == 1);
    CHECK(sim.particles.size() == n0);
    CHECK(finite(sim));
    std::printf("multires 3 levels: %zu particles, multigrid solve %d iterations\n", n0,
                sim.stats_last.solve.iterations);
  }
  std::printf(failures ? "FAILED %d checks\n" : "all checks passed\n", failures);
  return failures ? 1 : 0;
}
