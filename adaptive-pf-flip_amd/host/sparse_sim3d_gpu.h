// sparse_sim3d_gpu.h -- C++17 drop-in for the reference's single-phase SparseSim3D
// (src/driver/sparse_sim3d.h:5-15) whose step runs on one MI355X through the C ABI
// in include/pf_gpu.h (pf_step with PF_MODEL_SINGLE_PHASE).
//
// Same public fields as the reference (particles, dt, alpha, gravity, cg_iters,
// cg_tol), initDamBreak() and step().  Particles (Particles3D,
// src/particles/particles3d.h:4-10) live on the device between steps; edits to the
// host copy are uploaded after markParticlesDirty(), syncParticlesToHost() brings
// the device state back.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "pf_gpu.h"
#include "sparse_sim3d_tp_gpu.h"

namespace pfgpu {

struct Particles3D {
  std::vector<Vec3> pos, vel;
  std::vector<unsigned char> type;  // all 0 (liquid)
  size_t size() const { return pos.size(); }
  void add(const Vec3& p, const Vec3& v, unsigned char t = 0) {
    pos.push_back(p);
    vel.push_back(v);
    type.push_back(t);
  }
};

class SparseSim3D {
 public:
  Particles3D particles;
  double dt = 0.05, alpha = 0.95, gravity = -9.81;
  int cg_iters = 300;
  double cg_tol = 1e-6;

  // GPU solver controls (no reference counterpart): stop rule and CFL-limited dt
  // (computeAdaptiveParticleTimestep3D, src/driver/timestep3d.h:16-50)
  int32_t stop_rule = PF_STOP_REF_MAXABS;
  bool adaptive_timestep = false;
  double adaptive_cfl = 1.0, adaptive_min_dt = 1e-5;
  pf_step_stats stats_last{};

  SparseSim3D(int nx, int ny, int nz, double dx);
  ~SparseSim3D();
  SparseSim3D(const SparseSim3D&) = delete;
  SparseSim3D& operator=(const SparseSim3D&) = delete;

  void initDamBreak();
  void step();

  void markParticlesDirty() { dirty_ = true; }
  void syncParticlesToHost();
  size_t activeCellBlocks();
  size_t totalCellBlocks() const;
  pf_ctx* context() { return ctx_; }

 private:
  void check(int status, const char* what);
  pf_ctx* ctx_ = nullptr;
  int nx_, ny_, nz_;
  double dx_;
  bool dirty_ = true;
};

}  // namespace pfgpu
