// pf_lpc.hip -- the learned preconditioner's network: a 3-layer 3x3x3 convolutional
// network over 16^3 tiles with a 1-cell input halo, run on MFMA in BF16 with FP32
// accumulation, and its hand-written backward pass for training on the GPU.
// (No reference counterpart: the reference preconditions with Jacobi only,
// src/driver/sparse_ops3d_tp.cpp:322-327.)  pf_lpc_cw.hip feeds it the features of
// the multigrid cycle's aggregates and turns its output into correction weights.
//
// Network (weights W*[tap][ci][co], tap = dx + 3 dy + 9 dz, no biases):
//   X[18^3][4]                     four input channels (pf_lpc_cw.hip)
//   h1 = relu(conv(X, W1))        4 -> 16  (X carries the halo)
//   h2 = relu(conv(pad0 h1, W2)) 16 -> 32  (zero padding at the tile boundary)
//   h3 = relu(conv(pad0 h2, W3)) 32 -> 16
//   o  = h3 . Wh                  16 -> 1
// Activations are stored in BF16, channels last: [tile][cell][channel] with
// cell = x + 16 (y + 16 z); the input X is [tile][18^3][4] with the halo.
//
// Convolution as an implicit GEMM on v_mfma_f32_16x16x32_bf16: one MFMA row block is
// 16 cells along x (M), the K dimension is (tap, ci) in tap-major order, N is 16
// output channels.  Lane l of a wavefront supplies A[m = l % 16][k = 8 (l / 16) + i]
// and B[k = 8 (l / 16) + i][n = l % 16], i = 0..7, and receives
// D[m = 4 (l / 16) + i][n = l % 16], i = 0..3.  A workgroup (4 waves) computes 4
// z-planes of one tile, a wave one plane (16 rows of 16 cells).  The layer's
// weights sit in LDS as WT[n][k] (k contiguous); the A fragments come straight from
// global memory: for CIN >= 8 each lane loads 8 consecutive channels of one tap
// (16 bytes), and the 16 lanes of a group read 16 neighbouring cells.
//
// The backward pass through a convolution is the same kernel on the zero-padded
// output gradient with the weights tap-flipped and transposed
// (W'[26 - tap][co][ci] = W[tap][ci][co]), followed by the ReLU mask of the
// layer below.  Weight gradients (a reduction over all cells) run on the VALU in
// FP32 from LDS-staged planes, with per-workgroup partial sums reduced in a fixed
// order, so training is deterministic.
#include "pf_internal.h"

namespace pf {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kTile = 16, kHalo = 18, kTileVol = 4096, kHaloVol = 5832;
constexpr int kC0 = 4;
constexpr int kNW1 = 27 * 4 * 16, kNW2 = 27 * 16 * 32, kNW3 = 27 * 32 * 16, kNWH = 16;
constexpr int kNW = kNW1 + kNW2 + kNW3 + kNWH;   // 29,392 weights
constexpr int kOffW1 = 0, kOffW2 = kNW1, kOffW3 = kNW1 + kNW2, kOffWH = kNW1 + kNW2 + kNW3;

__host__ __device__ constexpr int kpad(int cin) { return ((27 * cin + 31) / 32) * 32; }

enum { EPI_RELU = 0, EPI_HEAD = 1, EPI_MASK = 2 };

struct ConvArgs {
  const __bf16* in;      // [tile][vol][CIN], vol = 18^3 (padded input) or 16^3
  const __bf16* wt;      // [COUT][KP] (k = tap * CIN + ci)
  __bf16* out;           // EPI_RELU / EPI_MASK: [tile][4096][COUT]; EPI_HEAD: h3 (may be null)
  const __bf16* mask;    // EPI_MASK: [tile][4096][COUT], the layer-below activation (> 0 passes)
  const __bf16* wh;      // EPI_HEAD: head weights (16)
  float* o;              // EPI_HEAD: [tile][4096]
  long long ntiles;
  const int* done;       // optional: skip when *done (solver paused / finished)
};

// one wave: output rows y = 0..15 of plane z of tile t; K chunks outer
template <int CIN, int COUT, bool PADDED, int EPI>
__global__ void __launch_bounds__(256) k_lpc_conv(ConvArgs a) {
  constexpr int KP = kpad(CIN);
  constexpr int NCH = KP / 32;
  constexpr int NB = COUT / 16;
  __shared__ __attribute__((aligned(16))) __bf16 s_w[COUT * KP];
  if (a.done && *a.done) return;
  const long long t = blockIdx.x >> 2;
  if (t >= a.ntiles) return;
  for (int q = threadIdx.x; q < COUT * KP / 8; q += 256)
    reinterpret_cast<bf16x8*>(s_w)[q] = reinterpret_cast<const bf16x8*>(a.wt)[q];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int z = (blockIdx.x & 3) * 4 + wave;
  const int m = lane & 15, g = lane >> 4;
  const long long in_base = t * (long long)(PADDED ? kHaloVol : kTileVol) * CIN;

  f32x4 acc[16][NB];
#pragma unroll
  for (int y = 0; y < 16; ++y)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) acc[y][nb] = f32x4{0.f, 0.f, 0.f, 0.f};

  for (int c = 0; c < NCH; ++c) {
    bf16x8 bw[NB];
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
      bw[nb] = *reinterpret_cast<const bf16x8*>(&s_w[(nb * 16 + m) * KP + c * 32 + g * 8]);
    const int k0 = c * 32 + g * 8;
#pragma unroll
    for (int y = 0; y < 16; ++y) {
      bf16x8 av;
      if constexpr (CIN >= 8) {
        const int tap = k0 / CIN, ci0 = k0 % CIN;
        const int dx = tap % 3, dy = (tap / 3) % 3, dz = tap / 9;
        const int xi = m + dx - 1, yi = y + dy - 1, zi = z + dz - 1;
        bool ok = tap < 27;
        long long pos;
        if constexpr (PADDED) {
          pos = ((long long)(zi + 1) * kHalo + (yi + 1)) * kHalo + (xi + 1);
        } else {
          ok = ok && xi >= 0 && xi < kTile && yi >= 0 && yi < kTile && zi >= 0 && zi < kTile;
          pos = ((long long)zi * kTile + yi) * kTile + xi;
        }
        if (ok) {
          av = *reinterpret_cast<const bf16x8*>(a.in + in_base + pos * CIN + ci0);
        } else {
#pragma unroll
          for (int i = 0; i < 8; ++i) av[i] = (__bf16)0.f;
        }
      } else {
        // CIN == 4: the 8 k values are two taps of 4 channels
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int tap = k0 / 4 + h;
          const int dx = tap % 3, dy = (tap / 3) % 3, dz = tap / 9;
          const int xi = m + dx - 1, yi = y + dy - 1, zi = z + dz - 1;
          bf16x4 v4;
          if (tap < 27) {
            long long pos;
            if constexpr (PADDED) {
              pos = ((long long)(zi + 1) * kHalo + (yi + 1)) * kHalo + (xi + 1);
              v4 = *reinterpret_cast<const bf16x4*>(a.in + in_base + pos * 4);
            } else {
              if (xi >= 0 && xi < kTile && yi >= 0 && yi < kTile && zi >= 0 && zi < kTile) {
                pos = ((long long)zi * kTile + yi) * kTile + xi;
                v4 = *reinterpret_cast<const bf16x4*>(a.in + in_base + pos * 4);
              } else {
                for (int i = 0; i < 4; ++i) v4[i] = (__bf16)0.f;
              }
            }
          } else {
            for (int i = 0; i < 4; ++i) v4[i] = (__bf16)0.f;
          }
#pragma unroll
          for (int i = 0; i < 4; ++i) av[h * 4 + i] = v4[i];
        }
      }
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) acc[y][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bw[nb], acc[y][nb], 0, 0, 0);
    }
  }

  // epilogue: lane holds channel nb*16 + m of cells x = 4g + i, i = 0..3
  const long long out_base = t * (long long)kTileVol;
#pragma unroll
  for (int y = 0; y < 16; ++y) {
    if constexpr (EPI == EPI_HEAD) {
      const float wh = (float)a.wh[m];
      float part[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const __bf16 h = (__bf16)fmaxf(acc[y][0][i], 0.f);
        const long long cell = ((long long)z * kTile + y) * kTile + 4 * g + i;
        if (a.out) a.out[(out_base + cell) * 16 + m] = h;
        part[i] = (float)h * wh;
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) {
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) part[i] += __shfl_xor(part[i], o, 64);
      }
      if (m < 4) {
        const long long cell = ((long long)z * kTile + y) * kTile + 4 * g + m;
        float v = m == 0 ? part[0] : (m == 1 ? part[1] : (m == 2 ? part[2] : part[3]));
        a.o[out_base + cell] = v;
      }
    } else {
#pragma unroll
      for (int nb = 0; nb < NB; ++nb)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const long long cell = ((long long)z * kTile + y) * kTile + 4 * g + i;
          const long long idx = (out_base + cell) * COUT + nb * 16 + m;
          float v = acc[y][nb][i];
          if constexpr (EPI == EPI_RELU) v = fmaxf(v, 0.f);
          if constexpr (EPI == EPI_MASK) v = (float)a.mask[idx] > 0.f ? v : 0.f;
          a.out[idx] = (__bf16)v;
        }
    }
  }
}

// Weight gradient of one layer: gw[tap][ci][co] = sum_cells in[cell + off(tap)][ci] * g[cell][co].
// Work unit = (tile, z-plane); workgroup w takes units w, w + G, ...; thread t owns
// the (tap, ci) rows t, t + 256, ... with all COUT outputs in registers.  Partial
// sums per workgroup go to part[w][27*CIN*COUT].
template <int CIN, int COUT, bool PADDED>
__global__ void __launch_bounds__(256) k_lpc_wgrad(const __bf16* __restrict__ in, const __bf16* __restrict__ gr,
                                                   long long ntiles, float* __restrict__ part) {
  constexpr int ROWS = 27 * CIN;
  constexpr int RPT = (ROWS + 255) / 256;
  __shared__ float s_in[3 * kHalo * kHalo * CIN];
  __shared__ float s_g[256 * COUT];
  float acc[RPT][COUT];
#pragma unroll
  for (int r = 0; r < RPT; ++r)
#pragma unroll
    for (int o = 0; o < COUT; ++o) acc[r][o] = 0.f;
  const long long units = ntiles * 16;
  for (long long u = blockIdx.x; u < units; u += gridDim.x) {
    const long long t = u >> 4;
    const int z = (int)(u & 15);
    __syncthreads();
    // planes z-1..z+1 with the x/y halo (zero outside an unpadded tile)
    for (int q = threadIdx.x; q < 3 * kHalo * kHalo * CIN; q += 256) {
      const int ci = q % CIN, p = q / CIN;
      const int xh = p % kHalo, yh = (p / kHalo) % kHalo, dz = p / (kHalo * kHalo);
      float v = 0.f;
      if (PADDED) {
        v = (float)in[(t * kHaloVol + ((long long)(z + dz) * kHalo + yh) * kHalo + xh) * CIN + ci];
      } else {
        const int xi = xh - 1, yi = yh - 1, zi = z + dz - 1;
        if (xi >= 0 && xi < kTile && yi >= 0 && yi < kTile && zi >= 0 && zi < kTile)
          v = (float)in[(t * kTileVol + ((long long)zi * kTile + yi) * kTile + xi) * CIN + ci];
      }
      s_in[q] = v;
    }
    for (int q = threadIdx.x; q < 256 * COUT; q += 256)
      s_g[q] = (float)gr[(t * kTileVol + (long long)z * 256) * COUT + q];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      const int row = threadIdx.x + 256 * r;
      if (row >= ROWS) continue;
      const int tap = row / CIN, ci = row % CIN;
      const int dx = tap % 3, dy = (tap / 3) % 3, dz = tap / 9;
      for (int cell = 0; cell < 256; ++cell) {
        const int x = cell & 15, y = cell >> 4;
        const float xv = s_in[((dz * kHalo + (y + dy)) * kHalo + (x + dx)) * CIN + ci];
        const float* gp = &s_g[cell * COUT];
#pragma unroll
        for (int o = 0; o < COUT; ++o) acc[r][o] = fmaf(xv, gp[o], acc[r][o]);
      }
    }
  }
  float* dst = part + (long long)blockIdx.x * (ROWS * COUT);
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    const int row = threadIdx.x + 256 * r;
    if (row >= ROWS) continue;
#pragma unroll
    for (int o = 0; o < COUT; ++o) dst[row * COUT + o] = acc[r][o];
  }
}

// head gradient: dh3 = bf16(go * Wh ⊙ (h3 > 0)); partial gWh[co] = sum_cells h3 * go
__global__ void __launch_bounds__(256) k_lpc_head_bwd(const __bf16* __restrict__ h3, const float* __restrict__ go,
                                                      const __bf16* __restrict__ wh, long long ncells,
                                                      __bf16* __restrict__ dh3, float* __restrict__ part) {
  __shared__ float s[16][17];
  float acc = 0.f;
  const int co = threadIdx.x & 15;
  const float w = (float)wh[co];
  for (long long q = (long long)blockIdx.x * 16 + (threadIdx.x >> 4); q < ncells; q += (long long)gridDim.x * 16) {
    const float h = (float)h3[q * 16 + co];
    const float gv = go[q];
    acc = fmaf(h, gv, acc);
    dh3[q * 16 + co] = (__bf16)(h > 0.f ? gv * w : 0.f);
  }
  s[threadIdx.x >> 4][co] = acc;
  __syncthreads();
  if (threadIdx.x < 16) {
    float r = 0.f;
    for (int k = 0; k < 16; ++k) r += s[k][threadIdx.x];
    part[(long long)blockIdx.x * 16 + threadIdx.x] = r;
  }
}

// out[i] = sum_w part[w * n + i] in workgroup order (deterministic)
__global__ void __launch_bounds__(256) k_lpc_sum_parts(const float* __restrict__ part, int nparts, int n,
                                                       float* __restrict__ out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int w = 0; w < nparts; ++w) s += (double)part[(long long)w * n + i];
  out[i] = (float)s;
}

// FP32 master weights -> BF16 layer matrices: forward WT[co][tap*CIN + ci] and the
// backward-data (tap-flipped, transposed) WB[ci][tap*COUT + co]; K padding is zero
template <int CIN, int COUT>
__device__ inline void prep_layer(const float* w, __bf16* wt, __bf16* wb, int idx) {
  constexpr int KPF = kpad(CIN), KPB = kpad(COUT);
  if (idx < COUT * KPF) {
    const int co = idx / KPF, k = idx % KPF;
    float v = 0.f;
    if (k < 27 * CIN) {
      const int tap = k / CIN, ci = k % CIN;
      v = w[(tap * CIN + ci) * COUT + co];
    }
    wt[idx] = (__bf16)v;
  }
  if (idx < CIN * KPB) {
    const int ci = idx / KPB, k = idx % KPB;
    float v = 0.f;
    if (k < 27 * COUT) {
      const int tapf = k / COUT, co = k % COUT;
      v = w[((26 - tapf) * CIN + ci) * COUT + co];
    }
    wb[idx] = (__bf16)v;
  }
}

struct LpcW {
  __bf16 *w1f, *w1b, *w2f, *w2b, *w3f, *w3b, *wh;
};

__global__ void __launch_bounds__(256) k_lpc_prep(const float* __restrict__ w, LpcW o) {
  const int idx = blockIdx.x * 256 + threadIdx.x;
  prep_layer<4, 16>(w + kOffW1, o.w1f, o.w1b, idx);
  prep_layer<16, 32>(w + kOffW2, o.w2f, o.w2b, idx);
  prep_layer<32, 16>(w + kOffW3, o.w3f, o.w3b, idx);
  if (idx < kNWH) o.wh[idx] = (__bf16)w[kOffWH + idx];
}

// Adam (Kingma and Ba, 2015) on the FP32 master weights
__global__ void __launch_bounds__(256) k_lpc_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                                                  const float* __restrict__ g, int n, float lr, float b1, float b2,
                                                  float eps, float bc1, float bc2) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

__global__ void __launch_bounds__(256) k_f32_to_bf16(const float* __restrict__ in, __bf16* __restrict__ out,
                                                     long long n) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
    out[i] = (__bf16)in[i];
}

// ---------------------------------------------------------------- solver coupling
// (the network's inputs and outputs in the solver are built in pf_lpc_cw.hip)

// y = A x over the list (conductance form, FP32); optional q = r - y with partial
// sums q.q and r.r (slots 0, 1)
__global__ void __launch_bounds__(256) k_lpc_apply_a(VecArgs a, const float* __restrict__ x, const float* __restrict__ r,
                                                     float* __restrict__ y, const float* __restrict__ cd,
                                                     const float* __restrict__ cx, const float* __restrict__ cy,
                                                     const float* __restrict__ cz, float scale_out) {
  const int lane = threadIdx.x & 63;
  double qq = 0.0, rr = 0.0;
  FOR_LIST_BLOCKS(a, b, lane) {
    long long c = b * kBlockVol + lane;
    CellNbr n = neighbours(a.dc, b, lane);
    const float xc = x[c];
    float ax = cd[c] * xc;
    ax = fmaf(cx[c], xc - x[n.xp], ax);
    ax = fmaf(cx[n.xm], xc - x[n.xm], ax);
    ax = fmaf(cy[c], xc - x[n.yp], ax);
    ax = fmaf(cy[n.ym], xc - x[n.ym], ax);
    ax = fmaf(cz[c], xc - x[n.zp], ax);
    ax = fmaf(cz[n.zm], xc - x[n.zm], ax);
    if (r) {
      const float q = r[c] - ax;
      y[c] = q;
      qq += (double)q * (double)q;
      rr += (double)r[c] * (double)r[c];
    } else {
      y[c] = scale_out * ax;
    }
  }
  double v[2] = {qq, rr};
  const bool mx[2] = {false, false};
  block_reduce_store<2>(v, mx, a.part, kMaxPartials);
}

}  // namespace pf
