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
//
// The kernel is bound by its gathers (72 dword taps per particle: the new and saved
// velocity at the particle, the new one at the midpoint), not by VALU or occupancy.
// So before a G2P each face axis is packed into quads (vel, vel+x, saved, saved+x) of a
// face and its +x neighbour: the two x-taps of a (j, k) pair are one 16-byte gather for
// both fields and one 8-byte gather for vel alone, 24 gathers per particle instead of
// 72, with the same values, weights and order of operations (the same bits).
#include "pf_internal.h"
#include "pf_speed.h"

namespace pf {

struct G2PArgs {
  Geom g;
  Dims df[3];
  const float* vel[3];    // post-projection u, v, w
  const float* saved[3];  // pre-gravity u, v, w
  const float4* quad[3];  // per face (vel, vel+x, saved, saved+x): it and its +x neighbour
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

// trilinear sample with per-tap index clamp, sparse_ops3d_common.h:30-51.  The
// block-major offset is separable, bm_index(i, j, k) = X(i) + Y(j) + Z(k), so the 8
// taps take 6 clamped terms instead of 8 full index computations.
__device__ inline float sample(const float* __restrict__ f, const Dims& d, float gx, float gy, float gz) {
  float fi = floorf(gx), fj = floorf(gy), fk = floorf(gz);
  int i0 = (int)fi, j0 = (int)fj, k0 = (int)fk;
  float fx = gx - fi, fy = gy - fj, fz = gz - fk;
  float wx[2] = {1.f - fx, fx}, wy[2] = {1.f - fy, fy}, wz[2] = {1.f - fz, fz};
  const long long sy = (long long)d.nbx * kBlockVol, sz = sy * d.nby;
  long long X[2], Y[2], Z[2];
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    const int i = min(max(i0 + e, 0), d.nx - 1);
    const int j = min(max(j0 + e, 0), d.ny - 1);
    const int k = min(max(k0 + e, 0), d.nz - 1);
    X[e] = (long long)(i >> 2) * kBlockVol + (i & 3);
    Y[e] = (long long)(j >> 2) * sy + 4 * (j & 3);
    Z[e] = (long long)(k >> 2) * sz + 16 * (k & 3);
  }
  float s = 0.f;
#pragma unroll
  for (int dk = 0; dk < 2; ++dk)
#pragma unroll
    for (int dj = 0; dj < 2; ++dj)
#pragma unroll
      for (int di = 0; di < 2; ++di) s = fmaf(wx[di] * wy[dj] * wz[dk], f[X[di] + Y[dj] + Z[dk]], s);
  return s;
}

// The same trilinear sample from a quad field Q[q] = (vel[q], vel[q+], saved[q], saved[q+]),
// q+ the face one step along x (q itself at the last x): one 16-byte gather serves the
// two x-taps of a (j, k) pair for both fields, one 8-byte gather for vel alone.  The per-tap clamps are the sample's:
// for i0 >= nx - 1 the partner is the face itself (as packed); for i0 < 0 both taps are
// face 0, taken from the quad's first half.  Same weights and order of operations as
// two calls of sample: the same bits.  ws: also the saved field's sample.
template <bool WS>
__device__ inline float2 sample_q(const float4* __restrict__ f, const Dims& d, float gx, float gy, float gz) {
  float fi = floorf(gx), fj = floorf(gy), fk = floorf(gz);
  int i0 = (int)fi, j0 = (int)fj, k0 = (int)fk;
  float fx = gx - fi, fy = gy - fj, fz = gz - fk;
  float wx[2] = {1.f - fx, fx}, wy[2] = {1.f - fy, fy}, wz[2] = {1.f - fz, fz};
  const long long sy = (long long)d.nbx * kBlockVol, sz = sy * d.nby;
  const int ia = min(max(i0, 0), d.nx - 1);
  const bool lo = i0 < 0;
  const long long X = (long long)(ia >> 2) * kBlockVol + (ia & 3);
  long long Y[2], Z[2];
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    const int j = min(max(j0 + e, 0), d.ny - 1);
    const int k = min(max(k0 + e, 0), d.nz - 1);
    Y[e] = (long long)(j >> 2) * sy + 4 * (j & 3);
    Z[e] = (long long)(k >> 2) * sz + 16 * (k & 3);
  }
  float sv = 0.f, ss = 0.f;
#pragma unroll
  for (int dk = 0; dk < 2; ++dk)
#pragma unroll
    for (int dj = 0; dj < 2; ++dj) {
      const float4 q = f[X + Y[dj] + Z[dk]];
      const float v1 = lo ? q.x : q.y, s1 = lo ? q.z : q.w;
      const float w0 = wx[0] * wy[dj] * wz[dk], w1 = wx[1] * wy[dj] * wz[dk];
      sv = fmaf(w0, q.x, sv);
      if (WS) ss = fmaf(w0, q.z, ss);
      sv = fmaf(w1, v1, sv);
      if (WS) ss = fmaf(w1, s1, ss);
    }
  return make_float2(sv, ss);
}

__device__ inline void sample_vel(const G2PArgs& a, const float* const* f, float x, float y, float z, float out[3]) {
  float px = (x - a.ox) * a.inv_dx, py = (y - a.oy) * a.inv_dx, pz = (z - a.oz) * a.inv_dx;
  out[0] = sample(f[0], a.df[0], px, py - 0.5f, pz - 0.5f);
  out[1] = sample(f[1], a.df[1], px - 0.5f, py, pz - 0.5f);
  out[2] = sample(f[2], a.df[2], px - 0.5f, py - 0.5f, pz);
}

// u, v, w (and, WS, the saved u, v, w) at a point from the quad fields
template <bool WS>
__device__ inline void sample_vel_q(const G2PArgs& a, float x, float y, float z, float out[3], float sav[3]) {
  float px = (x - a.ox) * a.inv_dx, py = (y - a.oy) * a.inv_dx, pz = (z - a.oz) * a.inv_dx;
  const float2 s0 = sample_q<WS>(a.quad[0], a.df[0], px, py - 0.5f, pz - 0.5f);
  const float2 s1 = sample_q<WS>(a.quad[1], a.df[1], px - 0.5f, py, pz - 0.5f);
  const float2 s2 = sample_q<WS>(a.quad[2], a.df[2], px - 0.5f, py - 0.5f, pz);
  out[0] = s0.x; out[1] = s1.x; out[2] = s2.x;
  if (WS) { sav[0] = s0.y; sav[1] = s1.y; sav[2] = s2.y; }
}

// per-thread counters -> one atomic per counter per workgroup
__device__ inline void block_flush_counters(int (&cnt)[8], float speed, int32_t* esc, uint32_t* vmax_bits,
                                            bool do_esc, bool do_vmax) {
  __shared__ int s_cnt[4][8];
  __shared__ float s_v[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < 8; ++q)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) cnt[q] += __shfl_xor(cnt[q], o, 64);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) speed = fmaxf(speed, __shfl_xor(speed, o, 64));
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < 8; ++q) s_cnt[wave][q] = cnt[q];
    s_v[wave] = speed;
  }
  __syncthreads();
  if (threadIdx.x < 8 && do_esc) {
    int t = s_cnt[0][threadIdx.x] + s_cnt[1][threadIdx.x] + s_cnt[2][threadIdx.x] + s_cnt[3][threadIdx.x];
    if (t) atomicAdd(&esc[threadIdx.x], t);
  }
  if (threadIdx.x == 8 && do_vmax) {
    float m = fmaxf(fmaxf(s_v[0], s_v[1]), fmaxf(s_v[2], s_v[3]));
    atomicMax(vmax_bits, __float_as_uint(m));
  }
}

// grid-stride over particles (fixed grid), counters flushed once per workgroup.  QUAD
// (the G2P with the quad fields packed): every sample from the quads.
template <bool QUAD>
__global__ void __launch_bounds__(256) k_g2p_advect(G2PArgs a) {
  int cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  float speed = 0.f;
  // Workgroups are dealt round-robin to the 8 XCDs; workgroup w (XCD w % 8) walks
  // a contiguous eighth of the cell-sorted particles, so each XCD's L2 holds the
  // grid around one region instead of every XCD reading the whole grid.
  const long long chunk = (a.n + 7) / 8;
  const long long p_end = min(a.n, (long long)((blockIdx.x & 7) + 1) * chunk);
  const long long stride = (long long)(gridDim.x >> 3) * blockDim.x;
  for (long long p = (long long)(blockIdx.x & 7) * chunk + (long long)(blockIdx.x >> 3) * blockDim.x + threadIdx.x;
       p < p_end; p += stride) {
    float x = a.P.x[p], y = a.P.y[p], z = a.P.z[p];
    uint8_t ty = a.P.type[p];
    float k1[3], sv[3];
    if (QUAD)
      sample_vel_q<true>(a, x, y, z, k1, sv);
    else
      sample_vel(a, a.vel, x, y, z, k1);
    if (QUAD) {
      float al = ty == 0 ? a.aL : a.aG;
      float u = a.P.u[p], v = a.P.v[p], w = a.P.w[p];
      u = (u + (k1[0] - sv[0])) * al + k1[0] * (1.f - al);
      v = (v + (k1[1] - sv[1])) * al + k1[1] * (1.f - al);
      w = (w + (k1[2] - sv[2])) * al + k1[2] * (1.f - al);
      a.P.u[p] = u; a.P.v[p] = v; a.P.w[p] = w;
      speed = fmaxf(speed, particle_speed(u, v, w));
    }
    if (a.do_advect) {
      float pos[3] = {x, y, z}, nxt[3], k2[3];
      const float dt = a.dt;
      if (a.order == 3) {
        float q[3], k3[3];
        for (int c = 0; c < 3; ++c) q[c] = pos[c] + k1[c] * (0.5f * dt);
        if (QUAD) sample_vel_q<false>(a, q[0], q[1], q[2], k2, nullptr);
        else sample_vel(a, a.vel, q[0], q[1], q[2], k2);
        for (int c = 0; c < 3; ++c) q[c] = (pos[c] - k1[c] * dt) + k2[c] * (2.f * dt);
        if (QUAD) sample_vel_q<false>(a, q[0], q[1], q[2], k3, nullptr);
        else sample_vel(a, a.vel, q[0], q[1], q[2], k3);
        for (int c = 0; c < 3; ++c) nxt[c] = pos[c] + ((k1[c] + k2[c] * 4.f) + k3[c]) * (dt / 6.f);
      } else {
        float m[3];
        for (int c = 0; c < 3; ++c) m[c] = pos[c] + k1[c] * (0.5f * dt);
        if (QUAD) sample_vel_q<false>(a, m[0], m[1], m[2], k2, nullptr);
        else sample_vel(a, a.vel, m[0], m[1], m[2], k2);
        for (int c = 0; c < 3; ++c) nxt[c] = pos[c] + k2[c] * dt;
      }
      bool any = false;
      for (int c = 0; c < 3; ++c) {
        bool lo = nxt[c] < a.lo[c], hi = nxt[c] > a.hi[c];
        cnt[2 + 2 * c] += lo;
        cnt[3 + 2 * c] += hi;
        any = any || lo || hi;
        nxt[c] = fmaxf(a.lo[c], fminf(a.hi[c], nxt[c]));
      }
      a.P.x[p] = nxt[0]; a.P.y[p] = nxt[1]; a.P.z[p] = nxt[2];
      cnt[0] += any && ty == 0;
      cnt[1] += any && ty == 1;
    }
  }
  block_flush_counters(cnt, speed, a.esc, a.vmax_bits, a.do_advect, a.do_g2p);
}

// the quad field of one face axis: Q[q] = (vel, saved) of face q and of its +x
// neighbour (of q itself at the last x and in the padding), as (vel, vel+x, saved, saved+x)
__global__ void __launch_bounds__(256) k_pack_quad(Dims d, const float* __restrict__ vel,
                                                   const float* __restrict__ saved, float4* __restrict__ out) {
  for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < d.nblocks * kBlockVol;
       q += (long long)gridDim.x * 256) {
    int bx, by, bz;
    block_coords(d, q >> 6, bx, by, bz);
    const int l = (int)(q & 63), i = bx * 4 + (l & 3);
    const long long n = i + 1 < d.nx ? ((l & 3) < 3 ? q + 1 : q + kBlockVol - 3) : q;
    out[q] = make_float4(vel[q], vel[n], saved[q], saved[n]);
  }
}

// max |v| over particles with finite velocity (timestep3d.h:27-32)
__global__ void __launch_bounds__(256) k_max_speed(long long n, PBuf P, uint32_t* vmax_bits) {
  float s = 0.f;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride) {
    s = fmaxf(s, particle_speed(P.u[p], P.v[p], P.w[p]));
  }
  int cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  block_flush_counters(cnt, s, nullptr, vmax_bits, false, true);
}

}  // namespace pf
