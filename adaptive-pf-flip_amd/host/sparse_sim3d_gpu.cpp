// sparse_sim3d_gpu.cpp -- host side of the GPU single-phase SparseSim3D (see the header).
#include "sparse_sim3d_gpu.h"

namespace pfgpu {

SparseSim3D::SparseSim3D(int nx, int ny, int nz, double dx) : nx_(nx), ny_(ny), nz_(nz), dx_(dx) {
  pf_grid_desc d{};
  d.nx = nx;
  d.ny = ny;
  d.nz = nz;
  d.dx = dx;
  int st = pf_create(&ctx_, &d);
  if (st != PF_OK) throw std::runtime_error("pf_create failed with status " + std::to_string(st));
}

SparseSim3D::~SparseSim3D() { pf_destroy(ctx_); }

void SparseSim3D::check(int status, const char* what) {
  if (status != PF_OK)
    throw std::runtime_error(std::string(what) + ": status " + std::to_string(status) + ": " + pf_last_error(ctx_));
}

// initDamBreak, src/driver/sparse_sim3d.cpp:4-24: a 2x2x2 sub-lattice in every cell
// of the column i < 0.4 nx, j < 0.7 ny (interior cells only)
void SparseSim3D::initDamBreak() {
  const int wx = nx_ * 4 / 10, hy = ny_ * 7 / 10;
  for (int k = 1; k < nz_ - 1; ++k)
    for (int j = 1; j < hy; ++j)
      for (int i = 1; i < wx; ++i)
        for (int s = 0; s < 8; ++s)
          particles.add(Vec3{(i + 0.25 + 0.5 * (s & 1)) * dx_, (j + 0.25 + 0.5 * ((s >> 1) & 1)) * dx_,
                             (k + 0.25 + 0.5 * ((s >> 2) & 1)) * dx_},
                        Vec3{});
  dirty_ = true;
}

void SparseSim3D::syncParticlesToHost() {
  const size_t n = particles.size();
  std::vector<float> x(n), y(n), z(n), u(n), v(n), w(n);
  check(pf_particles_download(ctx_, x.data(), y.data(), z.data(), u.data(), v.data(), w.data()),
        "pf_particles_download");
  for (size_t p = 0; p < n; ++p) {
    particles.pos[p] = Vec3{x[p], y[p], z[p]};
    particles.vel[p] = Vec3{u[p], v[p], w[p]};
  }
}

size_t SparseSim3D::activeCellBlocks() {
  int64_t n = 0;
  check(pf_active_blocks(ctx_, PF_FIELD_P, nullptr, 0, &n), "pf_active_blocks");
  return (size_t)n;
}

size_t SparseSim3D::totalCellBlocks() const {
  return (size_t)((nx_ + 3) / 4) * ((ny_ + 3) / 4) * ((nz_ + 3) / 4);
}

// SparseSim3D::step, src/driver/sparse_sim3d.cpp:106-115
void SparseSim3D::step() {
  if (dirty_) {
    const size_t n = particles.size();
    std::vector<float> x(n), y(n), z(n), u(n), v(n), w(n);
    for (size_t p = 0; p < n; ++p) {
      x[p] = (float)particles.pos[p].x; y[p] = (float)particles.pos[p].y; z[p] = (float)particles.pos[p].z;
      u[p] = (float)particles.vel[p].x; v[p] = (float)particles.vel[p].y; w[p] = (float)particles.vel[p].z;
    }
    check(pf_particles_upload(ctx_, (int64_t)n, x.data(), y.data(), z.data(), u.data(), v.data(), w.data(),
                              particles.type.data(), nullptr),
          "pf_particles_upload");
    dirty_ = false;
  }
  pf_step_cfg cfg{};
  cfg.dt = dt;
  cfg.gravity = gravity;
  cfg.Vp = 1.0;
  cfg.alpha_liquid = alpha;
  cfg.alpha_gas = alpha;
  cfg.adaptive_timestep = adaptive_timestep ? 1 : 0;
  cfg.adaptive_cfl = adaptive_cfl;
  cfg.adaptive_min_dt = adaptive_min_dt;
  cfg.advection_order = 2;
  cfg.model = PF_MODEL_SINGLE_PHASE;
  cfg.phase = pf_phase{1.0, 1.0, 1.0, 1.0};
  cfg.solve.max_iters = cg_iters;
  cfg.solve.stop_rule = stop_rule;
  cfg.solve.tol = cg_tol;
  cfg.solve.inner_rel_tol = 0.1;
  cfg.solve.max_outer = 10000;
  cfg.solve.check_every = 8;
  check(pf_step(ctx_, &cfg, &stats_last), "pf_step");
}

}  // namespace pfgpu
