// pf_noalloc.hip -- entry points that allocate nothing, for callers that cannot call
// hipMalloc (code under HIP graph capture, tight serving loops, harnesses that forbid
// allocation).
//
// pf_timestep_noalloc is pf_timestep (computeAdaptiveParticleTimestep3D,
// src/driver/timestep3d.h:16-50) on caller-owned device arrays.  Its reduction
// scratch -- one partial maximum per chunk of kChunk particles plus the final
// maximum -- lives in the file-scope device array `scratch` below, of a fixed
// maximum size; each device holds its own copy until the process ends.  Host code
// never passes the array's name to a kernel (in host code it is only a stand-in
// symbol): the device address of each device's copy is looked up once, with that
// device current, by hipGetSymbolAddress, and only that pointer is used.
//
// Every call on one device shares that device's array and leaves the last call's
// partials in it; nothing has to start at zero (every slot a call reads was written
// by the same call).  Calls that overlap on one device race on the array.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <mutex>

#include "pf_gpu.h"
#include "pf_speed.h"

namespace {

constexpr int kThreads = 256;
constexpr int kItems = 8;                          // particles per thread
constexpr long long kChunk = (long long)kThreads * kItems;  // particles per workgroup
constexpr int kMaxPartials = 1 << 17;              // 131,072 partial maxima
constexpr int kMaxDevices = 64;

// [0, kMaxPartials): per-chunk maxima; [kMaxPartials]: the maximum over all chunks
static __device__ __align__(256) float scratch[kMaxPartials + 64];

std::mutex g_mu;
float* g_scratch[kMaxDevices] = {};

// the device address of `scratch` on the current device (looked up once per device)
hipError_t scratch_ptr(float** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[dev] = static_cast<float*>(p);
  }
  *out = g_scratch[dev];
  return hipSuccess;
}

__device__ inline float wave_max(float s) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s = fmaxf(s, __shfl_xor(s, o, 64));
  return s;
}

__device__ inline float block_max(float s) {
  __shared__ float w[kThreads / 64];
  s = wave_max(s);
  if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = s;
  __syncthreads();
  float m = w[0];
#pragma unroll
  for (int q = 1; q < kThreads / 64; ++q) m = fmaxf(m, w[q]);
  return m;
}

// part[b] = max speed over particles [b*kChunk, (b+1)*kChunk) (coalesced: item q of
// thread t reads particle b*kChunk + q*kThreads + t)
__global__ void __launch_bounds__(kThreads) k_speed_chunks(long long n, const float* __restrict__ u,
                                                           const float* __restrict__ v, const float* __restrict__ w,
                                                           float* __restrict__ part) {
  const long long base = (long long)blockIdx.x * kChunk + threadIdx.x;
  float s = 0.f;
#pragma unroll
  for (int q = 0; q < kItems; ++q) {
    const long long p = base + (long long)q * kThreads;
    if (p < n) s = fmaxf(s, pf::particle_speed(u[p], v[p], w[p]));
  }
  s = block_max(s);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// out[0] = max(part[0..nparts)), one workgroup
__global__ void __launch_bounds__(kThreads) k_speed_final(int nparts, const float* __restrict__ part,
                                                          float* __restrict__ out) {
  float s = 0.f;
  for (int b = threadIdx.x; b < nparts; b += kThreads) s = fmaxf(s, part[b]);
  s = block_max(s);
  if (threadIdx.x == 0) out[0] = s;
}

}  // namespace

extern "C" int pf_timestep_noalloc(const float* u, const float* v, const float* w, int64_t n, double dx,
                                   double requested_dt, int32_t enabled, double cfl, double min_dt, void* stream,
                                   double* eff, double* vmax_out, double* cfl_dt, int32_t* limited) {
  if (n < 0 || n > (int64_t)kMaxPartials * kChunk) return PF_ERR_ARG;
  if (n > 0 && (!u || !v || !w)) return PF_ERR_ARG;
  hipStream_t s = static_cast<hipStream_t>(stream);
  float vm = 0.f;
  if (n > 0) {
    float* sc = nullptr;
    if (scratch_ptr(&sc) != hipSuccess) return PF_ERR_HIP;
    const int nparts = (int)((n + kChunk - 1) / kChunk);  // <= kMaxPartials by the size check
    hipLaunchKernelGGL(k_speed_chunks, dim3(nparts), dim3(kThreads), 0, s, (long long)n, u, v, w, sc);
    hipLaunchKernelGGL(k_speed_final, dim3(1), dim3(kThreads), 0, s, nparts, sc, sc + kMaxPartials);
    if (hipGetLastError() != hipSuccess) return PF_ERR_HIP;
    if (hipMemcpyAsync(&vm, sc + kMaxPartials, sizeof(float), hipMemcpyDeviceToHost, s) != hipSuccess)
      return PF_ERR_HIP;
    if (hipStreamSynchronize(s) != hipSuccess) return PF_ERR_HIP;
  }
  // computeAdaptiveParticleTimestep3D, timestep3d.h:22-49 (as pf_timestep)
  const double vmax = vm;
  double e = requested_dt, cd = requested_dt;
  int lim = 0;
  if (enabled && requested_dt > 0.0 && dx > 0.0 && cfl > 0.0 && vmax > 0.0) {
    cd = cfl * dx / vmax;
    double l = std::min(requested_dt, cd);
    if (min_dt > 0.0) l = std::max(min_dt, l);
    e = std::min(requested_dt, l);
    lim = e < requested_dt ? 1 : 0;
  }
  if (eff) *eff = e;
  if (vmax_out) *vmax_out = vmax;
  if (cfl_dt) *cfl_dt = cd;
  if (limited) *limited = lim;
  return PF_OK;
}
