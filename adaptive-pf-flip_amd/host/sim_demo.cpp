// sim_demo.cpp -- the reference's own SparseSim3DTP integration tests, run against
// the GPU drop-in class with the same code a reference caller would write.
//   tests/test_sparse_sim3d_tp.cpp:110-148  RT step stays stable (12 steps)
//   tests/test_sparse_sim3d_tp.cpp:749-786  bubble rises, headspace stays sparse (24 steps)
//   tests/test_sparse_sim3d_tp.cpp:150-169  adaptive timestep reports the effective dt
//   tests/test_sparse_sim3d.cpp:6-36         single-phase dam-break falls, stays sparse (20 steps)
// Exit status 0 when every check passes.
#include <cmath>
#include <cstdio>

#include "sparse_sim3d_gpu.h"
#include "sparse_sim3d_tp_gpu.h"

using pfgpu::SparseSim3D;
using pfgpu::SparseSim3DTP;

static int failures = 0;
#define CHECK(cond)                                                   \
  do {                                                                \
    if (!(cond)) {                                                    \
      std::fprintf(stderr, "CHECK failed at line %d: %s\n", __LINE__, #cond); \
      ++failures;                                                     \
    }                                                                 \
  } while (0)

static double meanY(const SparseSim3DTP& sim, unsigned char type) {
  double s = 0.0;
  int n = 0;
  for (size_t i = 0; i < sim.particles.size(); ++i)
    if (sim.particles.type[i] == type) {
      s += sim.particles.pos[i].y;
      ++n;
    }
  return n ? s / n : 0.0;
}

static bool finite(const SparseSim3DTP& sim) {
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
  std::printf(failures ? "FAILED %d checks\n" : "all checks passed\n", failures);
  return failures ? 1 : 0;
}
