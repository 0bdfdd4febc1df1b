// pf_lpc_cw.hip -- the learned part of the preconditioner: per-aggregate weights of the
// multigrid cycle's level-1 correction, predicted by the conv network of pf_lpc.hip.
// (No reference counterpart: the reference preconditions with Jacobi only,
// src/driver/sparse_ops3d_tp.cpp:322-327.)
//
// The network runs on 16^3 tiles of level-1 cells (the 2x2x2 aggregates of the
// multigrid cycle) holding FLUID aggregates, with a 1-cell halo, once per solve.  Its
// four input channels per aggregate come from the assembled system and the phase
// field:
//   0  beta_hat = 1 + log10(mean child beta / max beta of the tile) / 3, clamped to [0, 1]
//      (0 marks an aggregate without FLUID children)
//   1  mean child phi
//   2  log10(max child beta / min child beta) / 3, clamped to [0, 1] (the jump inside)
//   3  the Dirichlet share of the children's diagonal (cd over cd + couplings)
// Its output o per aggregate gives the weight w = exp(clamp(o, -kCwMax, kCwMax)); the
// cycle applies s = sqrt(w) to the restricted residual and to the coarse correction,
// P S B S P^T in place of P B P^T, so the preconditioner stays symmetric and linear in
// the residual.  A zero network is the plain cycle.
//
// Training minimises ||r - A M(r)||^2 / ||r||^2 over residuals the solver recorded, or
// the energy-norm factor (e0 - z)^T A (e0 - z) / e0^T A e0 with z = M(A e0) over the
// cycle's slowest modes e0 (recorded residuals after a few preconditioned Richardson
// steps).  With q = r - A z (r = A e0 for the energy form), z = T u + (terms without s)
// for the post-smoothing T, u = x + w_c P g and g = S B S y (y the restricted residual):
//   dL/dg = -2 w_c P^T T^T A q / ||r||^2     (residual form)
//   dL/dg = -2 w_c P^T T^T q / e0^T A e0      (energy form: A e1 = q)
//   dL/ds = dL/dg (.) (B S y) + y (.) B (S dL/dg)     (B symmetric)
//   dL/do = dL/ds (.) s / 2  where |o| < kCwMax.
#include "pf_internal.h"

namespace pf {

constexpr float kCwMax = 2.0f;

struct CwArgs {
  Dims dc;                 // fine cells
  Geom g;
  Dims d1;                 // level-1 slab
  int nx1, ny1, nz1;       // level-1 extents
  int ntx, nty;            // level-1 tiles per axis (x, y)
  const uint32_t* tiles;   // tile ids tx + ntx (ty + nty tz)
  long long ntiles;
  const uint8_t* marker;
  const float* fbeta;      // cell beta feature on FLUID cells
  const float* fphi;       // cell phi feature on FLUID cells
  const float* cd;         // FP32 operator: Dirichlet part of the diagonal
  const float* invd;       // FP32 1 / diagonal (0 outside the system)
  __bf16* X;               // [tile][18^3][4]
  const int* done;
};

__device__ inline void cw_tile_origin(const CwArgs& a, long long t, int& I0, int& J0, int& K0) {
  const uint32_t id = a.tiles[t];
  I0 = (int)(id % (uint32_t)a.ntx) * 16;
  J0 = (int)((id / (uint32_t)a.ntx) % (uint32_t)a.nty) * 16;
  K0 = (int)(id / ((uint32_t)a.ntx * (uint32_t)a.nty)) * 16;
}

// level-1 tile flags from the fine FLUID-block list: fine block (bx, by, bz) holds the
// aggregates 2bx..2bx+1 (per axis), in level-1 tile bx/8
__global__ void __launch_bounds__(256) k_lpc_cw_tile_flags(const uint32_t* __restrict__ list, long long n, Dims dc,
                                                           int ntx, int nty, uint8_t* __restrict__ flags) {
  const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
  if (q >= n) return;
  int bx, by, bz;
  block_coords(dc, list[q], bx, by, bz);
  flags[(bx >> 3) + (long long)ntx * ((by >> 3) + (long long)nty * (bz >> 3))] = 1;
}

// the features of aggregate (I, J, K) from its FLUID children
struct CwFeat {
  int n;
  float bsum, bmin, bmax, psum, cdsum, dsum;
};

__device__ inline CwFeat cw_children(const CwArgs& a, int I, int J, int K) {
  CwFeat f{0, 0.f, 3.4e38f, 0.f, 0.f, 0.f, 0.f};
  for (int q = 0; q < 8; ++q) {
    const int i = 2 * I + (q & 1), j = 2 * J + ((q >> 1) & 1), k = 2 * K + (q >> 2);
    if (i >= a.g.nx || j >= a.g.ny || k >= a.g.nz) continue;
    const long long c = bm_index(a.dc, i, j, k);
    if (a.marker[c] != CELL_FLUID) continue;
    const float b = a.fbeta[c];
    f.n++;
    f.bsum += b;
    f.bmin = fminf(f.bmin, b);
    f.bmax = fmaxf(f.bmax, b);
    f.psum += a.fphi[c];
    f.cdsum += a.cd[c];
    const float iv = a.invd[c];
    f.dsum += iv > 0.f ? 1.f / iv : 0.f;
  }
  return f;
}

// one workgroup per tile: the tile's largest child beta, then the 18^3 x 4 input
__global__ void __launch_bounds__(256) k_lpc_cw_gather(CwArgs a) {
  __shared__ float s_max[256];
  if (a.done && *a.done) return;
  const long long t = blockIdx.x;
  if (t >= a.ntiles) return;
  int I0, J0, K0;
  cw_tile_origin(a, t, I0, J0, K0);
  float bmax = 0.f;
  for (int q = threadIdx.x; q < 32 * 32 * 32; q += 256) {
    const int i = 2 * I0 + (q & 31), j = 2 * J0 + ((q >> 5) & 31), k = 2 * K0 + (q >> 10);
    if (i >= a.g.nx || j >= a.g.ny || k >= a.g.nz) continue;
    const long long c = bm_index(a.dc, i, j, k);
    if (a.marker[c] == CELL_FLUID) bmax = fmaxf(bmax, a.fbeta[c]);
  }
  s_max[threadIdx.x] = bmax;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) s_max[threadIdx.x] = fmaxf(s_max[threadIdx.x], s_max[threadIdx.x + w]);
    __syncthreads();
  }
  const float tmax = s_max[0];
  __bf16* X = a.X + t * 5832LL * 4;
  for (int q = threadIdx.x; q < 5832; q += 256) {
    const int I = I0 + q % 18 - 1, J = J0 + (q / 18) % 18 - 1, K = K0 + q / 324 - 1;
    float v0 = 0.f, v1 = 0.f, v2 = 0.f, v3 = 0.f;
    if (I >= 0 && J >= 0 && K >= 0 && I < a.nx1 && J < a.ny1 && K < a.nz1) {
      const CwFeat f = cw_children(a, I, J, K);
      if (f.n > 0 && tmax > 0.f) {
        const float bm = f.bsum / (float)f.n;
        v0 = fminf(fmaxf(1.f + log10f(fmaxf(bm / tmax, 1e-30f)) * (1.f / 3.f), 0.f), 1.f);
        v1 = f.psum / (float)f.n;
        v2 = fminf(fmaxf(log10f(f.bmax / fmaxf(f.bmin, 1e-30f)) * (1.f / 3.f), 0.f), 1.f);
        v3 = f.dsum > 0.f ? fminf(fmaxf(f.cdsum / f.dsum, 0.f), 1.f) : 0.f;
      }
    }
    X[q * 4 + 0] = (__bf16)v0;
    X[q * 4 + 1] = (__bf16)v1;
    X[q * 4 + 2] = (__bf16)v2;
    X[q * 4 + 3] = (__bf16)v3;
  }
}

// s = exp(clamp(o, -kCwMax, kCwMax) / 2) for the tiles' aggregates (the slab holds 1
// elsewhere)
__global__ void __launch_bounds__(256) k_lpc_cw_scatter(CwArgs a, const float* __restrict__ o,
                                                        float* __restrict__ s) {
  if (a.done && *a.done) return;
  for (long long t = blockIdx.x; t < a.ntiles; t += gridDim.x) {
    int I0, J0, K0;
    cw_tile_origin(a, t, I0, J0, K0);
    for (int q = threadIdx.x; q < 4096; q += 256) {
      const int I = I0 + (q & 15), J = J0 + ((q >> 4) & 15), K = K0 + (q >> 8);
      if (I >= a.nx1 || J >= a.ny1 || K >= a.nz1) continue;
      const float ov = fminf(fmaxf(o[t * 4096 + q], -kCwMax), kCwMax);
      s[bm_index(a.d1, I, J, K)] = expf(0.5f * ov);
    }
  }
}

// go[tile][q] = dL/ds * s / 2 where the clamp is inactive (the chain rule through the scatter)
__global__ void __launch_bounds__(256) k_lpc_cw_gout(CwArgs a, const float* __restrict__ o,
                                                     const float* __restrict__ gs, const float* __restrict__ s,
                                                     float* __restrict__ go) {
  for (long long t = blockIdx.x; t < a.ntiles; t += gridDim.x) {
    int I0, J0, K0;
    cw_tile_origin(a, t, I0, J0, K0);
    for (int q = threadIdx.x; q < 4096; q += 256) {
      const int I = I0 + (q & 15), J = J0 + ((q >> 4) & 15), K = K0 + (q >> 8);
      float v = 0.f;
      if (I < a.nx1 && J < a.ny1 && K < a.nz1) {
        const long long c = bm_index(a.d1, I, J, K);
        const float ov = o[t * 4096 + q];
        if (ov > -kCwMax && ov < kCwMax) v = gs[c] * s[c] * 0.5f;
      }
      go[t * 4096 + q] = v;
    }
  }
}

// t = invd (.) v over a dense slab (zero outside the system, where invd = 0)
__global__ void __launch_bounds__(256) k_lpc_dscale(long long n, const float* __restrict__ v,
                                                    const float* __restrict__ invd, float* __restrict__ t) {
  for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < n; q += (long long)gridDim.x * 256)
    t[q] = invd[q] * v[q];
}

// v += a y over a dense slab
__global__ void __launch_bounds__(256) k_lpc_axpy(long long n, float* __restrict__ v, const float* __restrict__ y,
                                                  float a) {
  for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < n; q += (long long)gridDim.x * 256)
    v[q] = fmaf(a, y[q], v[q]);
}

// out = scale * P^T v: each aggregate sums its children in the system (invd > 0)
__global__ void __launch_bounds__(256) k_lpc_restrict_sum(Dims dc, Geom g, const float* __restrict__ v,
                                                          const float* __restrict__ invd, Dims d1, int nx1, int ny1,
                                                          int nz1, float scale, float* __restrict__ out) {
  const long long n1 = (long long)nx1 * ny1 * nz1;
  for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < n1; q += (long long)gridDim.x * 256) {
    const int I = (int)(q % nx1), J = (int)((q / nx1) % ny1), K = (int)(q / ((long long)nx1 * ny1));
    float sum = 0.f;
    for (int c8 = 0; c8 < 8; ++c8) {
      const int i = 2 * I + (c8 & 1), j = 2 * J + ((c8 >> 1) & 1), k = 2 * K + (c8 >> 2);
      if (i >= g.nx || j >= g.ny || k >= g.nz) continue;
      const long long c = bm_index(dc, i, j, k);
      if (invd[c] > 0.f) sum += v[c];
    }
    out[bm_index(d1, I, J, K)] = scale * sum;
  }
}

// level 1: gs = dg (.) h + y (.) bw; and the input of the extra coarse solve, s (.) dg
__global__ void __launch_bounds__(256) k_lpc_cw_grad(long long n, const float* __restrict__ dg,
                                                     const float* __restrict__ h, const float* __restrict__ y,
                                                     const float* __restrict__ bw, float* __restrict__ gs) {
  for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < n; q += (long long)gridDim.x * 256)
    gs[q] = fmaf(dg[q], h[q], y[q] * bw[q]);
}

// partial sums over the list for the energy objective: e.q, z.q and e.r (slots 0..2)
__global__ void __launch_bounds__(256) k_lpc_dots3(VecArgs a, const float* __restrict__ e, const float* __restrict__ q,
                                                   const float* __restrict__ z, const float* __restrict__ r) {
  const int lane = threadIdx.x & 63;
  double eq = 0.0, zq = 0.0, er = 0.0;
  FOR_LIST_BLOCKS(a, b, lane) {
    const long long c = b * kBlockVol + lane;
    const double qc = q[c], ec = e[c];
    eq += ec * qc;
    zq += (double)z[c] * qc;
    er += ec * (double)r[c];
  }
  double v[3] = {eq, zq, er};
  const bool mx[3] = {false, false, false};
  block_reduce_store<3>(v, mx, a.part, kMaxPartials);
}

__global__ void __launch_bounds__(256) k_lpc_mul(long long n, const float* __restrict__ a, const float* __restrict__ b,
                                                 float* __restrict__ out) {
  for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < n; q += (long long)gridDim.x * 256)
    out[q] = a[q] * b[q];
}

}  // namespace pf
