// pf_speed.h -- particle speed as the CFL timestep reads it, shared by every kernel
// that takes the maximum (so each gives the same bits for the same velocities).
#pragma once

#include <hip/hip_runtime.h>

namespace pf {

// |v| in FP32 for a particle with finite velocity, 0 otherwise
// (computeAdaptiveParticleTimestep3D skips non-finite velocities, timestep3d.h:27-32)
__device__ inline float particle_speed(float u, float v, float w) {
#pragma clang fp contract(off)
  if (!(isfinite(u) && isfinite(v) && isfinite(w))) return 0.f;
  return sqrtf(u * u + v * v + w * w);
}

}  // namespace pf
