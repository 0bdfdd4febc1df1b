// pf_mr_mg.hip -- the aggregation multigrid on the multires operator S = V A (no
// reference counterpart; the reference's multires solve is Jacobi PCG).
//
// The first coarse space is the uniform grid of 4^3 fine blocks.  With leaves of level
// 0..2 every leaf cell lies inside one block (a level-2 cell is a whole block), so the
// piecewise-constant prolongation P from blocks to leaf cells and R = P^T give a
// Galerkin operator R S P that couples face-adjacent blocks only: a 7-point
// conductance Laplacian on the (n/4)^3 grid whose face weight is the sum of the patch
// conductances on the plane between two blocks, and whose extra diagonal is the sum of
// the blocks' Dirichlet parts.  That is the shape of the uniform hierarchy's level 2,
// so the cycle continues with the uniform coarse levels from there.  The smoother is
// weighted Jacobi (or its Chebyshev weights) on S, with the row sums spread over a
// block's lanes as in the multires SpMV.

namespace pf {

// R S P on the block grid: one wave per fine block (all blocks), lane = fine cell; the
// face weights are the conductances of the 16 fine faces on the block's + planes
__global__ void __launch_bounds__(256) k_mr_mg_coarse(MrOp op, Dims dcc, float* __restrict__ cd,
                                                      float* __restrict__ cx, float* __restrict__ cy,
                                                      float* __restrict__ cz) {
  const long long b = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (b >= op.m.dc.nblocks) return;  // whole waves
  int bx, by, bz;
  block_coords(op.m.dc, b, bx, by, bz);
  const int li = lane & 3, lj = (lane >> 2) & 3, lk = lane >> 4;
  const int i = bx * 4 + li, j = by * 4 + lj, k = bz * 4 + lk;
  const bool in = i < op.m.g.nx && j < op.m.g.ny && k < op.m.g.nz;
  float d = in ? op.cd[b * kBlockVol + lane] : 0.f;
  float wx = (in && li == 3) ? op.cf[0][bm_index(op.m.du, i + 1, j, k)] : 0.f;
  float wy = (in && lj == 3) ? op.cf[1][bm_index(op.m.dv, i, j + 1, k)] : 0.f;
  float wz = (in && lk == 3) ? op.cf[2][bm_index(op.m.dw, i, j, k + 1)] : 0.f;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    d += __shfl_xor(d, o, 64);
    wx += __shfl_xor(wx, o, 64);
    wy += __shfl_xor(wy, o, 64);
    wz += __shfl_xor(wz, o, 64);
  }
  if (lane == 0) {
    const long long q = bm_index(dcc, bx, by, bz);
    cd[q] = d;
    cx[q] = wx;
    cy[q] = wy;
    cz[q] = wz;
  }
}

// the residual r - S x at the anchor of this lane's cell (0 on other lanes), S x by
// the lane-spread row sum; every lane of the wave must call it (shuffles)
__device__ inline float mr_mg_residual(const MrOp& op, const float* __restrict__ r, const float* __restrict__ x,
                                       long long c, int i, int j, int k, int L, bool in, bool anchor) {
  const int s = 1 << L;
  const long long anc = bm_index(op.m.dc, i & ~(s - 1), j & ~(s - 1), k & ~(s - 1));
  const float xa = in ? x[anc] : 0.f;
  const float sum = mr_apply_row_coop<float>(op.m, op.cf, i, j, k, L, in, xa, [&](long long n) { return x[n]; });
  return anchor ? r[c] - (op.cd[c] * xa + sum) : 0.f;
}

#define MR_MG_LANE                                                                     \
  const long long c = b * kBlockVol + lane;                                            \
  int bx, by, bz;                                                                      \
  block_coords(op.m.dc, b, bx, by, bz);                                                \
  const int i = bx * 4 + (lane & 3), j = by * 4 + ((lane >> 2) & 3), k = bz * 4 + (lane >> 4); \
  const bool in = i < op.m.g.nx && j < op.m.g.ny && k < op.m.g.nz;                     \
  const int L = mr_level(op.m, bx * 4, by * 4, bz * 4);                                \
  if (L < 0) continue; /* wave-uniform */                                              \
  const bool anchor = in && ((i | j | k) & ((1 << L) - 1)) == 0;

// one smoothing sweep on the listed blocks: xo = xi + omega D^-1 (r - S xi) on anchors,
// xo = omega D^-1 r when xi is null, 0 on the other lanes; with part, the partial
// r . xo (one per workgroup); skipped while *done
__global__ void __launch_bounds__(256) k_mr_mg_sweep(VecArgs a, MrOp op, const int* done, const float* __restrict__ inv,
                                                     const float* __restrict__ r, const float* __restrict__ xi,
                                                     float* __restrict__ xo, float omega, double* __restrict__ part) {
  if (done && *done) return;
  const int lane = threadIdx.x & 63;
  double rz = 0.0;
  FOR_LIST_BLOCKS(a, b, lane) {
    MR_MG_LANE
    float v = 0.f;
    if (xi) {
      const float res = mr_mg_residual(op, r, xi, c, i, j, k, L, in, anchor);
      if (anchor) v = fmaf(omega * inv[c], res, xi[c]);
    } else if (anchor) {
      v = omega * inv[c] * r[c];
    }
    xo[c] = v;
    if (anchor) rz += (double)r[c] * (double)v;
  }
  if (part) {
    double v[1] = {rz};
    const bool mx[1] = {false};
    block_reduce_store<1>(v, mx, part, kMaxPartials);
  }
}

// R (r - S x): the block sums of the anchor residuals into the block grid's slab
__global__ void __launch_bounds__(256) k_mr_mg_restrict(VecArgs a, MrOp op, const int* done,
                                                        const float* __restrict__ r, const float* __restrict__ x,
                                                        Dims dcc, float* __restrict__ rc) {
  if (done && *done) return;
  const int lane = threadIdx.x & 63;
  FOR_LIST_BLOCKS(a, b, lane) {
    MR_MG_LANE
    float res = mr_mg_residual(op, r, x, c, i, j, k, L, in, anchor);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) res += __shfl_xor(res, o, 64);
    if (lane == 0) rc[bm_index(dcc, bx, by, bz)] = res;
  }
}

// x += omega_c P xc on the anchors of the listed blocks (in place)
__global__ void __launch_bounds__(256) k_mr_mg_prolong(VecArgs a, MrOp op, const int* done, Dims dcc,
                                                       const float* __restrict__ xc, float* __restrict__ x,
                                                       float omega_c) {
  if (done && *done) return;
  const int lane = threadIdx.x & 63;
  FOR_LIST_BLOCKS(a, b, lane) {
    MR_MG_LANE
    if (anchor) x[c] = fmaf(omega_c, xc[bm_index(dcc, bx, by, bz)], x[c]);
  }
}

#undef MR_MG_LANE

// splits the listed blocks by level: f01 (levels 0 and 1) and f2 (level 2)
__global__ void __launch_bounds__(256) k_mr_mg_split(const uint32_t* __restrict__ list, long long nlist, MrGrid m,
                                                     uint8_t* __restrict__ f01, uint8_t* __restrict__ f2) {
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= nlist) return;
  const uint32_t b = list[t];
  int bx, by, bz;
  block_coords(m.dc, b, bx, by, bz);
  const int L = mr_level(m, bx * 4, by * 4, bz * 4);
  f01[b] = (L == 0 || L == 1) ? 1 : 0;
  f2[b] = L == 2 ? 1 : 0;
}

// Level-2 blocks hold one leaf cell each (anchored at the block origin), so a wave
// takes four of them: 16 lanes per block, lane r of a block the patches r, r + 16, ...,
// r + 80 of its 96, summed by xor shuffles inside the 16 lanes.  The cell's values sit
// in lane 0 of the block's slab entries; the other 63 entries of a level-2 block are
// left alone (the cycle's arrays hold zeros there).
#define MR_MG_L2_WALK                                                                      \
  const int lane = threadIdx.x & 63, rl = lane & 15;                                       \
  for (long long t = (long long)blockIdx.x * 4 + (threadIdx.x >> 6); t * 4 < n;            \
       t += (long long)gridDim.x * 4)

#define MR_MG_L2_BLOCK                                                                     \
  const long long e = t * 4 + (lane >> 4);                                                 \
  const bool live = e < n;                                                                 \
  const uint32_t b = live ? list[e] : 0u;                                                  \
  int bx, by, bz;                                                                          \
  block_coords(op.m.dc, b, bx, by, bz);                                                    \
  const long long c = (long long)b * kBlockVol;

__device__ inline float mr_l2_sum(float v) {
  v += __shfl_xor(v, 1, 64);
  v += __shfl_xor(v, 2, 64);
  v += __shfl_xor(v, 4, 64);
  v += __shfl_xor(v, 8, 64);
  return v;
}

// the level-2 sweep (see k_mr_mg_sweep); with part, the workgroup's partial r . xo is
// added to part[blockIdx.x], which the sweep over the other blocks wrote just before
__global__ void __launch_bounds__(256) k_mr_mg_sweep2(const uint32_t* __restrict__ list, long long n, MrOp op,
                                                      const int* done, const float* __restrict__ inv,
                                                      const float* __restrict__ r, const float* __restrict__ xi,
                                                      float* __restrict__ xo, float omega, double* __restrict__ part) {
  if (done && *done) return;
  double rz = 0.0;
  MR_MG_L2_WALK {
    MR_MG_L2_BLOCK
    const float xa = (live && xi) ? xi[c] : 0.f;
    float sum = 0.f;
    if (xi)
      sum = mr_l2_sum(mr_row_slots<float>(op.m, op.cf, bx * 4, by * 4, bz * 4, 4, rl, 16, live, xa,
                                          [&](long long q) { return xi[q]; }));
    if (live && rl == 0) {
      const float rc = r[c];
      const float v = xi ? fmaf(omega * inv[c], rc - (op.cd[c] * xa + sum), xa) : omega * inv[c] * rc;
      xo[c] = v;
      rz += (double)rc * (double)v;
    }
  }
  if (part) {
    __shared__ double s_rz[4];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) rz += __shfl_xor(rz, o, 64);
    if ((threadIdx.x & 63) == 0) s_rz[threadIdx.x >> 6] = rz;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(&part[blockIdx.x], s_rz[0] + s_rz[1] + s_rz[2] + s_rz[3]);
  }
}

// the level-2 restriction: the one cell's residual to its block-grid cell
__global__ void __launch_bounds__(256) k_mr_mg_restrict2(const uint32_t* __restrict__ list, long long n, MrOp op,
                                                         const int* done, const float* __restrict__ r,
                                                         const float* __restrict__ x, Dims dcc,
                                                         float* __restrict__ rc) {
  if (done && *done) return;
  MR_MG_L2_WALK {
    MR_MG_L2_BLOCK
    const float xa = live ? x[c] : 0.f;
    const float sum = mr_l2_sum(mr_row_slots<float>(op.m, op.cf, bx * 4, by * 4, bz * 4, 4, rl, 16, live, xa,
                                                    [&](long long q) { return x[q]; }));
    if (live && rl == 0) rc[bm_index(dcc, bx, by, bz)] = r[c] - (op.cd[c] * xa + sum);
  }
}

// the level-2 prolongation: x += omega_c xc at the cell
__global__ void __launch_bounds__(256) k_mr_mg_prolong2(const uint32_t* __restrict__ list, long long n, MrOp op,
                                                        const int* done, Dims dcc, const float* __restrict__ xc,
                                                        float* __restrict__ x, float omega_c) {
  if (done && *done) return;
  MR_MG_L2_WALK {
    MR_MG_L2_BLOCK
    if (live && rl == 0) x[c] = fmaf(omega_c, xc[bm_index(dcc, bx, by, bz)], x[c]);
  }
}

// the level-2 part of the multires SpMV (k_mr_spmv on the other blocks): p_new = z +
// beta p_old at the cell, A p there, the workgroup's partial p.Ap added to
// a.part[blockIdx.x], which the launch over the other blocks wrote just before
__global__ void __launch_bounds__(256) k_mr_spmv2(const uint32_t* __restrict__ list, long long n, VecArgs a, MrOp op,
                                                  const float* __restrict__ z, float* __restrict__ pA,
                                                  float* __restrict__ pB, float* __restrict__ Ap) {
  if (a.sc->done) return;
  const bool odd = a.sc->iter & 1;
  const float* __restrict__ pold = odd ? pB : pA;
  float* __restrict__ pnew = odd ? pA : pB;
  const float beta = (float)a.sc->beta;
  auto pval = [&](long long q) { return fmaf(beta, pold[q], z[q]); };
  double pap = 0.0;
  MR_MG_L2_WALK {
    MR_MG_L2_BLOCK
    const float pc = live ? pval(c) : 0.f;
    const float sum = mr_l2_sum(mr_row_slots<float>(op.m, op.cf, bx * 4, by * 4, bz * 4, 4, rl, 16, live, pc, pval));
    if (live && rl == 0) {
      const float ap = op.cd[c] * pc + sum;
      pnew[c] = pc;
      Ap[c] = ap;
      pap += (double)pc * (double)ap;
    }
  }
  __shared__ double s_pap[4];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) pap += __shfl_xor(pap, o, 64);
  if ((threadIdx.x & 63) == 0) s_pap[threadIdx.x >> 6] = pap;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(&a.part[blockIdx.x], s_pap[0] + s_pap[1] + s_pap[2] + s_pap[3]);
}

#undef MR_MG_L2_WALK
#undef MR_MG_L2_BLOCK

// the PCG update on the level-2 blocks' cells (k_pcg_update / k_pcg_update_flex on the
// other blocks), one lane per block: x += alpha p, r -= alpha A p, and z = D^-1 r
// unless FLEX; the workgroup's partials r.z (FLEX: r.z of the cycle's z) and r.r are
// added to a.part slots 0 and 1 at blockIdx.x, which the other launch wrote just before
template <bool FLEX>
__global__ void __launch_bounds__(256) k_mr_update2(const uint32_t* __restrict__ list, long long n, VecArgs a,
                                                    float* __restrict__ x, float* __restrict__ r,
                                                    float* __restrict__ z, const float* __restrict__ pA,
                                                    const float* __restrict__ pB, const float* __restrict__ Ap,
                                                    const float* __restrict__ invd) {
  if (a.sc->done) return;
  const float* __restrict__ p = (a.sc->iter & 1) ? pA : pB;
  const float alpha = (float)a.sc->alpha;
  double v0 = 0.0, v1 = 0.0;
  for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long long)gridDim.x * 256) {
    const long long c = (long long)list[e] * kBlockVol;
    const float xn = fmaf(alpha, p[c], x[c]);
    const float rn = fmaf(-alpha, Ap[c], r[c]);
    x[c] = xn;
    r[c] = rn;
    if (FLEX) {
      v0 += (double)rn * (double)z[c];
    } else {
      const float zn = invd[c] * rn;
      z[c] = zn;
      v0 += (double)rn * (double)zn;
    }
    v1 += (double)rn * (double)rn;
  }
  __shared__ double s_v[4][2];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    v0 += __shfl_xor(v0, o, 64);
    v1 += __shfl_xor(v1, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    s_v[threadIdx.x >> 6][0] = v0;
    s_v[threadIdx.x >> 6][1] = v1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicAdd(&a.part[blockIdx.x], s_v[0][0] + s_v[1][0] + s_v[2][0] + s_v[3][0]);
    atomicAdd(&a.part[kMaxPartials + blockIdx.x], s_v[0][1] + s_v[1][1] + s_v[2][1] + s_v[3][1]);
  }
}

}  // namespace pf
