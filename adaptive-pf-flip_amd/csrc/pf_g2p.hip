// pf_g2p.hip -- G2P FLIP/PIC blend and RK2/RK3 advection with boundary clamps.
//
// Reference: spG2P3D_tp (src/driver/sparse_ops3d_tp.cpp:155-172), sampleField /
// sampleU/V/W (src/driver/sparse_ops3d_common.h:30-63), spAdvect3D_tp with
// advectMidpoint / advectRK3 (sparse_ops3d_tp.cpp:185-232), ParticleEscapeStats3D
// (src/driver/particle_escape3d.h:37-80).
//
// One thread per particle in cell-sorted order, so a wavefront's trilinear taps hit
// a handful of neighbouring 4^3 blocks (256-byte lines) and stay in L2.  When G2P
// and advection run fused (the step), the first RK stage reuses the PIC sample at
// the particle (same post-projection grid, same position), saving 24 taps.
#include "pf_internal.h"

namespace pf {

struct G2PArgs {
  Geom g;
  Dims df[3];
  const float* vel[3];    // post-projection u, v, w
  const float* saved[3];  // pre-gravity u, v, w
  PBuf P;
  long long n;
  float inv_dx, ox, oy, oz;
  float lo[3], hi[3];
  float aL, aG, dt;
  int order;
  int do_g2p, do_advect;
  int32_t* esc;           // 8 clamp counters
  uint32_t* vmax_bits;    // max |v| (float bits) after the update
};

// trilinear sample with per-tap index clamp, sparse_ops3d_common.h:30-51
__device__ inline float sample(const float* __restrict__ f, const Dims& d, float gx, float gy, float gz) {
  float fi = floorf(gx), fj = floorf(gy), fk = floorf(gz);
  int i0 = (int)fi, j0 = (int)fj, k0 = (int)fk;
  float fx = gx - fi, fy = gy - fj, fz = gz - fk;
  float wx[2] = {1.f - fx, fx}, wy[2] = {1.f - fy, fy}, wz[2] = {1.f - fz, fz};
  float s = 0.f;
#pragma unroll
  for (int dk = 0; dk < 2; ++dk)
#pragma unroll
    for (int dj = 0; dj < 2; ++dj)
#pragma unroll
      for (int di = 0; di < 2; ++di) {
        int i = min(max(i0 + di, 0), d.nx - 1);
        int j = min(max(j0 + dj, 0), d.ny - 1);
        int k = min(max(k0 + dk, 0), d.nz - 1);
        s = fmaf(wx[di] * wy[dj] * wz[dk], f[bm_index(d, i, j, k)], s);
      }
  return s;
}

__device__ inline void sample_vel(const G2PArgs& a, const float* const* f, float x, float y, float z, float out[3]) {
  float px = (x - a.ox) * a.inv_dx, py = (y - a.oy) * a.inv_dx, pz = (z - a.oz) * a.inv_dx;
  out[0] = sample(f[0], a.df[0], px, py - 0.5f, pz - 0.5f);
  out[1] = sample(f[1], a.df[1], px - 0.5f, py, pz - 0.5f);
  out[2] = sample(f[2], a.df[2], px - 0.5f, py - 0.5f, pz);
}

__global__ void __launch_bounds__(256) k_g2p_advect(G2PArgs a) {
  long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = p < a.n;
  bool cl[6] = {false, false, false, false, false, false};
  bool liquid = false, gas = false;
  float speed = 0.f;
  if (live) {
    float x = a.P.x[p], y = a.P.y[p], z = a.P.z[p];
    uint8_t ty = a.P.type[p];
    float k1[3];
    sample_vel(a, a.vel, x, y, z, k1);
    if (a.do_g2p) {
      float sv[3];
      sample_vel(a, a.saved, x, y, z, sv);
      float al = ty == 0 ? a.aL : a.aG;
      float u = a.P.u[p], v = a.P.v[p], w = a.P.w[p];
      u = (u + (k1[0] - sv[0])) * al + k1[0] * (1.f - al);
      v = (v + (k1[1] - sv[1])) * al + k1[1] * (1.f - al);
      w = (w + (k1[2] - sv[2])) * al + k1[2] * (1.f - al);
      a.P.u[p] = u; a.P.v[p] = v; a.P.w[p] = w;
      if (isfinite(u) && isfinite(v) && isfinite(w)) speed = sqrtf(u * u + v * v + w * w);
    }
    if (a.do_advect) {
      float pos[3] = {x, y, z}, nxt[3], k2[3];
      const float dt = a.dt;
      if (a.order == 3) {
        float q[3], k3[3];
        for (int c = 0; c < 3; ++c) q[c] = pos[c] + k1[c] * (0.5f * dt);
        sample_vel(a, a.vel, q[0], q[1], q[2], k2);
        for (int c = 0; c < 3; ++c) q[c] = (pos[c] - k1[c] * dt) + k2[c] * (2.f * dt);
        sample_vel(a, a.vel, q[0], q[1], q[2], k3);
        for (int c = 0; c < 3; ++c) nxt[c] = pos[c] + ((k1[c] + k2[c] * 4.f) + k3[c]) * (dt / 6.f);
      } else {
        float m[3];
        for (int c = 0; c < 3; ++c) m[c] = pos[c] + k1[c] * (0.5f * dt);
        sample_vel(a, a.vel, m[0], m[1], m[2], k2);
        for (int c = 0; c < 3; ++c) nxt[c] = pos[c] + k2[c] * dt;
      }
      for (int c = 0; c < 3; ++c) {
        cl[2 * c] = nxt[c] < a.lo[c];
        cl[2 * c + 1] = nxt[c] > a.hi[c];
        nxt[c] = fmaxf(a.lo[c], fminf(a.hi[c], nxt[c]));
      }
      a.P.x[p] = nxt[0]; a.P.y[p] = nxt[1]; a.P.z[p] = nxt[2];
      bool any = cl[0] || cl[1] || cl[2] || cl[3] || cl[4] || cl[5];
      liquid = any && ty == 0;
      gas = any && ty == 1;
    }
  }
  // wave-aggregated counters
  const int lane = threadIdx.x & 63;
  int cnt[8];
  cnt[0] = __popcll(__ballot(liquid));
  cnt[1] = __popcll(__ballot(gas));
#pragma unroll
  for (int q = 0; q < 6; ++q) cnt[2 + q] = __popcll(__ballot(cl[q]));
  float m = speed;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
  if (lane == 0) {
    if (a.do_advect)
#pragma unroll
      for (int q = 0; q < 8; ++q)
        if (cnt[q]) atomicAdd(&a.esc[q], cnt[q]);
    if (a.do_g2p) atomicMax(a.vmax_bits, __float_as_uint(m));
  }
}

// max |v| over particles with finite velocity (timestep3d.h:27-32)
__global__ void __launch_bounds__(256) k_max_speed(long long n, PBuf P, uint32_t* vmax_bits) {
  long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  float s = 0.f;
  if (p < n) {
    float u = P.u[p], v = P.v[p], w = P.w[p];
    if (isfinite(u) && isfinite(v) && isfinite(w)) s = sqrtf(u * u + v * v + w * w);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s = fmaxf(s, __shfl_xor(s, o, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(vmax_bits, __float_as_uint(s));
}

}  // namespace pf
