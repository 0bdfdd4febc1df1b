// render_cache3d_gpu.h -- the render-cache writers of src/driver/render_cache3d.h for the
// GPU classes: the same names, arguments and file format, through the C ABI
// (pf_write_render_cache, pf_write_render_manifest).  Errors throw, as the reference's
// writers do: std::invalid_argument for invalid metadata, std::runtime_error for I/O.
#pragma once

#include <string>
#include <vector>

#include "multires_sim3d_tp_gpu.h"
#include "sparse_sim3d_tp_gpu.h"

namespace pfgpu {

// RenderCacheCamera3D, render_cache3d.h:19-27
struct RenderCacheCamera3D {
  Vec3 position;
  Vec3 target;
  Vec3 up;
  double fov_degrees = 45.0;
  double near_clip = 0.1;
  double far_clip = 1000.0;
  double focal_length_mm = 50.0;
};

// RenderCacheManifestFrame3D, render_cache3d.h:39-45
struct RenderCacheManifestFrame3D {
  int frame = 0;
  int step = 0;
  double time = 0.0;
  std::string path;
  long long bytes = 0;
};

// defaultRenderCacheCamera3D, render_cache3d.h:48-62
RenderCacheCamera3D defaultRenderCacheCamera3D(int nx, int ny, int nz, double dx);

// writeSparseRenderCache3D / writeMRRenderCache3D, render_cache3d.h:615-711; return the
// bytes written
long long writeSparseRenderCache3D(const SparseSim3DTP& sim, const std::string& path, int frame, double time,
                                   const RenderCacheCamera3D& camera);
long long writeMRRenderCache3D(const MRSim3DTP& sim, const std::string& path, int frame, double time,
                               const RenderCacheCamera3D& camera);

// writeRenderCacheManifest3D, render_cache3d.h:713-755
void writeRenderCacheManifest3D(const std::string& path, const char* simKind, int nx, int ny, int nz, double dx,
                                const std::vector<RenderCacheManifestFrame3D>& frames);

}  // namespace pfgpu
