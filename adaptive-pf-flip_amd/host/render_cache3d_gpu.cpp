// render_cache3d_gpu.cpp -- see render_cache3d_gpu.h.
#include "render_cache3d_gpu.h"

#include <stdexcept>

namespace pfgpu {

namespace {

pf_render_camera toAbi(const RenderCacheCamera3D& c) {
  return pf_render_camera{{c.position.x, c.position.y, c.position.z}, {c.target.x, c.target.y, c.target.z},
                          {c.up.x, c.up.y, c.up.z}, c.fov_degrees, c.near_clip, c.far_clip, c.focal_length_mm};
}

long long writeFrame(pf_ctx* ctx, const std::string& path, int frame, double time, double dt, double Vp,
                     const PhaseParams& phase, const RenderCacheCamera3D& camera) {
  if (path.empty()) throw std::runtime_error("render cache path is empty");
  pf_render_frame f{};
  f.frame = frame;
  f.time = time;
  f.dt = dt;
  f.Vp = Vp;
  f.phase = pf_phase{phase.rho_l, phase.rho_g, phase.alpha_phi, phase.rho_tilde_0};
  f.camera = toAbi(camera);
  int64_t bytes = 0;
  const int st = pf_write_render_cache(ctx, path.c_str(), &f, &bytes);
  if (st != PF_OK) {
    const std::string msg = pf_last_error(ctx);
    if (msg == "invalid render cache metadata") throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
  }
  return (long long)bytes;
}

}  // namespace

RenderCacheCamera3D defaultRenderCacheCamera3D(int nx, int ny, int nz, double dx) {
  pf_render_camera c{};
  pf_render_default_camera(nx, ny, nz, dx, &c);
  RenderCacheCamera3D r;
  r.position = Vec3{c.position[0], c.position[1], c.position[2]};
  r.target = Vec3{c.target[0], c.target[1], c.target[2]};
  r.up = Vec3{c.up[0], c.up[1], c.up[2]};
  r.fov_degrees = c.fov_degrees;
  r.near_clip = c.near_clip;
  r.far_clip = c.far_clip;
  r.focal_length_mm = c.focal_length_mm;
  return r;
}

long long writeSparseRenderCache3D(const SparseSim3DTP& sim, const std::string& path, int frame, double time,
                                   const RenderCacheCamera3D& camera) {
  return writeFrame(sim.context(), path, frame, time, sim.effective_dt_last, sim.Vp, sim.phase, camera);
}

long long writeMRRenderCache3D(const MRSim3DTP& sim, const std::string& path, int frame, double time,
                               const RenderCacheCamera3D& camera) {
  return writeFrame(sim.context(), path, frame, time, sim.effective_dt_last, sim.Vp, sim.phase, camera);
}

void writeRenderCacheManifest3D(const std::string& path, const char* simKind, int nx, int ny, int nz, double dx,
                                const std::vector<RenderCacheManifestFrame3D>& frames) {
  if (path.empty()) throw std::runtime_error("render cache path is empty");
  std::vector<pf_render_manifest_frame> f(frames.size());
  for (size_t a = 0; a < frames.size(); ++a)
    f[a] = pf_render_manifest_frame{frames[a].frame, frames[a].step, frames[a].time, frames[a].path.c_str(),
                                    (int64_t)frames[a].bytes};
  if (pf_write_render_manifest(path.c_str(), simKind, nx, ny, nz, dx, f.data(), (int64_t)f.size()) != PF_OK)
    throw std::invalid_argument("invalid render cache manifest metadata or write failure");
}

}  // namespace pfgpu
