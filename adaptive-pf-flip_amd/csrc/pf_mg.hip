// pf_mg.hip -- aggregation multigrid preconditioner for the sparse pressure solve
// (no reference counterpart: the reference preconditions with Jacobi,
// sparse_ops3d_tp.cpp:322-327; the solve still stops on the FP64 residual).
//
// Operator form.  A cell's row is cd_c x_c + sum over its faces f of w_f (x_c - x_nbr),
// with w stored at the lower cell of each face (cx, cy, cz) and the Dirichlet part
// in cd; cells outside the FLUID set have zero coefficients.  Aggregating 2x2x2 cells
// with piecewise-constant prolongation P keeps that form exactly: the Galerkin
// operator P^T A P has cd_I = sum of the children's cd and, between adjacent
// aggregates, the sum of the fine couplings that cross from one to the other.  So
// every coarse level is built from the one below by sums, in one pass per level, and
// stays symmetric positive (semi)definite like the fine system.
//
// Cycle.  V-cycle from a zero guess: nu Jacobi sweeps down (weights: one damping
// factor, or the reciprocal Chebyshev roots of degree nu on [2/r, 2]), residual
// restricted by summing the children (P^T), the coarse solve, the correction added
// back (P, over-weighted), nu sweeps up with the same weights in reverse order; many
// damped sweeps on the coarsest level.  The sweeps of a pass commute, so the cycle is
// a fixed symmetric preconditioner inside the flexible-beta PCG.  Level 0 walks the
// FLUID block list and reads an FP16 copy of the operator; coarse levels are dense
// block-major slabs (each 1/8 of the one below) in FP32, which the same row kernels
// walk whole through their own scaled FP16 copies.
#include "pf_internal.h"

namespace pf {

struct MgLevel {
  Dims d;                       // block-major slab of the level's cells
  int nx, ny, nz;               // cell extents
  const float *cd, *cx, *cy, *cz, *inv;
};

// (A x)_c for the cell at (i, j, k) of slab index c
__device__ inline float mg_ax(const MgLevel& L, const float* __restrict__ x, long long c, int i, int j, int k) {
  const float xc = x[c];
  float ax = L.cd[c] * xc;
  if (i + 1 < L.nx) ax = fmaf(L.cx[c], xc - x[bm_index(L.d, i + 1, j, k)], ax);
  if (i > 0) {
    const long long m = bm_index(L.d, i - 1, j, k);
    ax = fmaf(L.cx[m], xc - x[m], ax);
  }
  if (j + 1 < L.ny) ax = fmaf(L.cy[c], xc - x[bm_index(L.d, i, j + 1, k)], ax);
  if (j > 0) {
    const long long m = bm_index(L.d, i, j - 1, k);
    ax = fmaf(L.cy[m], xc - x[m], ax);
  }
  if (k + 1 < L.nz) ax = fmaf(L.cz[c], xc - x[bm_index(L.d, i, j, k + 1)], ax);
  if (k > 0) {
    const long long m = bm_index(L.d, i, j, k - 1);
    ax = fmaf(L.cz[m], xc - x[m], ax);
  }
  return ax;
}

__device__ inline void mg_coords(const Dims& d, long long c, int& i, int& j, int& k) {
  int bx, by, bz;
  block_coords(d, c >> 6, bx, by, bz);
  const int l = (int)(c & 63);
  i = bx * 4 + (l & 3);
  j = by * 4 + ((l >> 2) & 3);
  k = bz * 4 + (l >> 4);
}

// the cells a level-kernel visits: the FLUID list on level 0, the whole slab above
struct MgWalk {
  const uint32_t* list;   // null: dense
  long long nlist;
  const int* done;        // skip everything while *done (may be null)
};

#define MG_FOR_CELLS(W, L, c, i, j, k)                                                                      \
  if ((W).list) {                                                                                         \
    const int _lane = threadIdx.x & 63;                                                                   \
    for (long long _t = (long long)blockIdx.x * 4 + (threadIdx.x >> 6); _t < (W).nlist;                   \
         _t += (long long)gridDim.x * 4) {                                                                \
      const long long c = (long long)list_at((W).list, wave_uniform_u32((uint32_t)_t)) * kBlockVol + _lane; \
      int i, j, k;                                                                                        \
      mg_coords((L).d, c, i, j, k);                                                                       \
      MG_BODY                                                                                             \
    }                                                                                                     \
  } else {                                                                                                \
    for (long long c = (long long)blockIdx.x * 256 + threadIdx.x; c < (L).d.nblocks * kBlockVol;           \
         c += (long long)gridDim.x * 256) {                                                               \
      int i, j, k;                                                                                        \
      mg_coords((L).d, c, i, j, k);                                                                       \
      MG_BODY                                                                                             \
    }                                                                                                     \
  }

// x = omega D^-1 r (the first sweep from a zero guess)
__global__ void __launch_bounds__(256) k_mg_first(MgWalk w, MgLevel L, const float* __restrict__ r,
                                                  float* __restrict__ x, float omega) {
  if (w.done && *w.done) return;
#define MG_BODY x[c] = omega * L.inv[c] * r[c];
  MG_FOR_CELLS(w, L, c, i, j, k)
#undef MG_BODY
}

// xo = xi + omega D^-1 (r - A xi); with part: partial r . xo into part (slot 2 of the
// solver's partial buffer, one per workgroup)
__global__ void __launch_bounds__(256) k_mg_sweep(MgWalk w, MgLevel L, const float* __restrict__ r,
                                                  const float* __restrict__ xi, float* __restrict__ xo, float omega,
                                                  double* __restrict__ part) {
  if (w.done && *w.done) return;
  double rz = 0.0;
#define MG_BODY                                                       \
  {                                                                   \
    const float rc = r[c];                                            \
    const float v = fmaf(omega * L.inv[c], rc - mg_ax(L, xi, c, i, j, k), xi[c]); \
    xo[c] = v;                                                        \
    rz += (double)rc * (double)v;                                     \
  }
  MG_FOR_CELLS(w, L, c, i, j, k)
#undef MG_BODY
  if (part) {
    double v[1] = {rz};
    const bool mx[1] = {false};
    block_reduce_store<1>(v, mx, part, kMaxPartials);
  }
}

// r_c = P^T (r_f - A_f x_f), walking the fine level: a wave holds one 4x4x4 fine
// block, which is 2x2x2 whole aggregates; the eight children of an aggregate are the
// lanes that differ in bits 0, 2 and 4 of the lane index (the x, y and z low bits),
// summed by three lane exchanges.  Aggregates with no child in the walk keep their
// old value: their diagonal is zero, and every reader masks them out.
__global__ void __launch_bounds__(256) k_mg_restrict(MgWalk w, MgLevel F, const float* __restrict__ rf,
                                                     const float* __restrict__ xf, MgLevel Cl,
                                                     float* __restrict__ rc) {
  if (w.done && *w.done) return;
#define MG_BODY                                                                          \
  {                                                                                      \
    float s = F.inv[c] != 0.0f ? rf[c] - mg_ax(F, xf, c, i, j, k) : 0.0f;                \
    s += __shfl_xor(s, 1);                                                               \
    s += __shfl_xor(s, 4);                                                               \
    s += __shfl_xor(s, 16);                                                              \
    if ((threadIdx.x & 21) == 0) rc[bm_index(Cl.d, i >> 1, j >> 1, k >> 1)] = s;         \
  }
  MG_FOR_CELLS(w, F, c, i, j, k)
#undef MG_BODY
}

// the coarsest level (at most 4 cells a side: one block, one wave): `sweeps` damped
// Jacobi sweeps from zero in LDS, x written once at the end
__global__ void __launch_bounds__(64) k_mg_coarsest(const int* done, MgLevel L, const float* __restrict__ r,
                                                    float* __restrict__ x, float omega, int sweeps) {
  if (done && *done) return;
  __shared__ float xs[2][64];
  const int c = threadIdx.x;
  int i, j, k;
  mg_coords(L.d, c, i, j, k);
  const float rc = r[c], inv = L.inv[c], cd = L.cd[c];
  // couplings and neighbour lanes (a missing neighbour couples with weight 0 to itself)
  float wv[6];
  int nb[6];
  const int di[6] = {1, -1, 4, -4, 16, -16};
  const bool in[6] = {i + 1 < L.nx, i > 0, j + 1 < L.ny, j > 0, k + 1 < L.nz, k > 0};
  const float* cw[3] = {L.cx, L.cy, L.cz};
  for (int f = 0; f < 6; ++f) {
    nb[f] = in[f] ? c + di[f] : c;
    wv[f] = in[f] ? ((f & 1) ? cw[f >> 1][c + di[f]] : cw[f >> 1][c]) : 0.0f;
  }
  float xc = omega * inv * rc;
  xs[0][c] = xc;
  __syncthreads();
  for (int s2 = 1; s2 < sweeps; ++s2) {
    const float* xo = xs[(s2 - 1) & 1];
    float ax = cd * xc;
    for (int f = 0; f < 6; ++f) ax = fmaf(wv[f], xc - xo[nb[f]], ax);
    xc = fmaf(omega * inv, rc - ax, xc);
    xs[s2 & 1][c] = xc;
    __syncthreads();
  }
  x[c] = xc;
}

// x_f += omega_c P x_c on the cells of the system
__global__ void __launch_bounds__(256) k_mg_prolong(MgWalk w, MgLevel F, MgLevel Cl, const float* __restrict__ xc,
                                                    float* __restrict__ xf, float omega_c) {
  if (w.done && *w.done) return;
#define MG_BODY                                                              \
  if (F.inv[c] != 0.0f) xf[c] = fmaf(omega_c, xc[bm_index(Cl.d, i >> 1, j >> 1, k >> 1)], xf[c]);
  MG_FOR_CELLS(w, F, c, i, j, k)
#undef MG_BODY
}

// the Galerkin coarse operator of level F: sums of the children's cd and of the fine
// couplings that cross each coarse face
__global__ void __launch_bounds__(256) k_mg_coarsen(MgLevel F, MgLevel Cl, float* __restrict__ cd,
                                                    float* __restrict__ cx, float* __restrict__ cy,
                                                    float* __restrict__ cz) {
  for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < Cl.d.nblocks * kBlockVol;
       q += (long long)gridDim.x * 256) {
    int I, J, K;
    mg_coords(Cl.d, q, I, J, K);
    float d = 0.0f, wx = 0.0f, wy = 0.0f, wz = 0.0f;
    if (I < Cl.nx && J < Cl.ny && K < Cl.nz)
      for (int e = 0; e < 8; ++e) {
        const int a = e & 1, b = (e >> 1) & 1, g = e >> 2;
        const int i = 2 * I + a, j = 2 * J + b, k = 2 * K + g;
        if (i >= F.nx || j >= F.ny || k >= F.nz) continue;
        const long long c = bm_index(F.d, i, j, k);
        d += F.cd[c];
        // the high-side children's +faces lead into the next aggregate
        if (a == 1 && i + 1 < F.nx) wx += F.cx[c];
        if (b == 1 && j + 1 < F.ny) wy += F.cy[c];
        if (g == 1 && k + 1 < F.nz) wz += F.cz[c];
      }
    cd[q] = d;
    cx[q] = wx;
    cy[q] = wy;
    cz[q] = wz;
  }
}

// inv = 1 / A_II (0 outside the system)
__global__ void __launch_bounds__(256) k_mg_inverse(MgLevel L, float* __restrict__ inv) {
  for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < L.d.nblocks * kBlockVol;
       q += (long long)gridDim.x * 256) {
    int i, j, k;
    mg_coords(L.d, q, i, j, k);
    float a = 0.0f;
    if (i < L.nx && j < L.ny && k < L.nz) {
      a = L.cd[q] + L.cx[q] + L.cy[q] + L.cz[q];
      if (i > 0) a += L.cx[bm_index(L.d, i - 1, j, k)];
      if (j > 0) a += L.cy[bm_index(L.d, i, j - 1, k)];
      if (k > 0) a += L.cz[bm_index(L.d, i, j, k - 1)];
    }
    inv[q] = a > 0.0f ? 1.0f / a : 0.0f;
  }
}

// ---- FP16 copies of the level-0 operator for the cycle.  The cycle only has to be
// an approximate inverse (the solve stops on the FP64 residual), so its level-0
// passes read the coefficients and the diagonal inverse as FP16, each group scaled
// by a power of two that brings its maximum to [0.5, 1]: a third fewer bytes per
// pass.  Zeros stay zero, so the cells outside the system stay out.

// mx[0] = max of cd, cx, cy, cz; mx[1] = max of inv (float bits; all values >= 0)
__global__ void __launch_bounds__(256) k_mg_absmax(long long n, const float* __restrict__ cd,
                                                   const float* __restrict__ cx, const float* __restrict__ cy,
                                                   const float* __restrict__ cz, const float* __restrict__ inv,
                                                   uint32_t* mx) {
  float ma = 0.0f, mi = 0.0f;
  for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < n; q += (long long)gridDim.x * 256) {
    ma = fmaxf(ma, fmaxf(fmaxf(cd[q], cx[q]), fmaxf(cy[q], cz[q])));
    mi = fmaxf(mi, inv[q]);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    ma = fmaxf(ma, __shfl_xor(ma, o, 64));
    mi = fmaxf(mi, __shfl_xor(mi, o, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMax(&mx[0], __float_as_uint(ma));
    atomicMax(&mx[1], __float_as_uint(mi));
  }
}

// 2^e with e = the exponent that brings m to [0.5, 1) (1 for m = 0 or not finite)
__device__ inline float mg_pow2_above(float m) {
  if (!(m > 0.0f) || !isfinite(m)) return 1.0f;
  int e;
  frexpf(m, &e);
  return ldexpf(1.0f, e);
}

__global__ void __launch_bounds__(256) k_mg_tohalf(long long n, const float* __restrict__ cd,
                                                   const float* __restrict__ cx, const float* __restrict__ cy,
                                                   const float* __restrict__ cz, const float* __restrict__ inv,
                                                   _Float16* __restrict__ hcd, _Float16* __restrict__ hcx,
                                                   _Float16* __restrict__ hcy, _Float16* __restrict__ hcz,
                                                   _Float16* __restrict__ hinv, const uint32_t* mx,
                                                   float* __restrict__ unscale) {
  const float ua = mg_pow2_above(__uint_as_float(mx[0])), ui = mg_pow2_above(__uint_as_float(mx[1]));
  const float sa = 1.0f / ua, si = 1.0f / ui;  // powers of two: exact
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    unscale[0] = ua;
    unscale[1] = ui;
  }
  for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < n; q += (long long)gridDim.x * 256) {
    hcd[q] = (_Float16)(cd[q] * sa);
    hcx[q] = (_Float16)(cx[q] * sa);
    hcy[q] = (_Float16)(cy[q] * sa);
    hcz[q] = (_Float16)(cz[q] * sa);
    // a cell is in the system where its inverse is non-zero (the level-0 kernels test
    // the FP16 value): a tiny inverse that would round to 0 keeps the least subnormal
    _Float16 hi = (_Float16)(inv[q] * si);
    if (inv[q] > 0.f && (float)hi == 0.f) hi = (_Float16)5.9604644775390625e-8f;  // 2^-24
    hinv[q] = hi;
  }
}

// x *= s elementwise over a coarse slab (the per-aggregate correction scaling)
__global__ void __launch_bounds__(256) k_mg_scale(long long n, float* __restrict__ x, const float* __restrict__ s,
                                                  const int* done) {
  if (done && *done) return;
  for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < n; q += (long long)gridDim.x * 256)
    x[q] *= s[q];
}

typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
__device__ inline float mg_h2f(uint32_t bits) { return (float)__builtin_bit_cast(_Float16, (uint16_t)bits); }
// four FP16 values of a row at the FP32 byte offset off (masked offsets read 0)
__device__ inline F4 mg_ldh4(__amdgpu_buffer_rsrc_t r, uint32_t off, float scale) {
  const u32x2 q = __builtin_amdgcn_raw_buffer_load_b64(r, off >> 1, 0, 0);
  F4 f;
  f.v[0] = mg_h2f(q.x & 0xffffu) * scale;
  f.v[1] = mg_h2f(q.x >> 16) * scale;
  f.v[2] = mg_h2f(q.y & 0xffffu) * scale;
  f.v[3] = mg_h2f(q.y >> 16) * scale;
  return f;
}
__device__ inline float mg_ldh1(__amdgpu_buffer_rsrc_t r, uint32_t off, float scale) {
  return mg_h2f(__builtin_amdgcn_raw_buffer_load_b16(r, off >> 1, 0, 0)) * scale;
}

// ---- level 0 by cell rows (arrays below 2 GiB)
//
// The layout of k_pcg_spmv: a lane holds one x-row of 4 cells (row = ly + 4 lz, 16
// lanes per 4^3 block) and a wave the four list blocks of one slot; y- and z-neighbour
// rows inside the block come from lanes +-1 and +-4, rows on a block face are read
// from the neighbouring block, whose list position does not matter.  Modes:
//   MG_FIRST2    two sweeps from zero: u = omega D^-1 r, out = u + omega D^-1 (r - A u)
//   MG_SWEEP     out = u + omega D^-1 (r - A u) with u = x
//   MG_PSWEEP    the same with u = x + omega_c P x_c, the coarse correction added on
//                the cells of the system (a neighbour outside the system couples with
//                weight 0, so its value is taken as it is)
//   MG_RESTRICT  level 1's residual P^T (r - A x): the two aggregates of a row are
//                pairs of its registers, their y and z partners lanes ^1 and ^4
// SWEEP and PSWEEP add the partial r . out when part is set.
enum { MG_FIRST2 = 0, MG_SWEEP = 1, MG_PSWEEP = 2, MG_RESTRICT = 3 };

struct MgRowArgs {
  const uint32_t* list;                      // the blocks to walk (null: the whole slab)
  long long nlist;
  const int* done;
  Dims d, dc;                                // level 0 and level 1 slabs
  const _Float16 *cd, *cx, *cy, *cz, *inv;   // FP16 level-0 operator (scaled)
  const float* unscale;                      // [0] coefficients, [1] diagonal inverse
  const float *r, *x;
  float* out;                                // the new iterate, or level 1's residual
  const float* xc;                           // level 1's correction (PSWEEP)
  float omega, omega_c;
  float omega_u;                             // FIRST2: the first sweep's weight (omega: the second's)
  double* part;
};

struct MgRowBufs {
  __amdgpu_buffer_rsrc_t cd, cx, cy, cz, inv, r, x, out;
};

// the stencil input u of a row / of one cell at byte offset off (ui: the diagonal
// inverse's scale)
template <int MODE>
__device__ inline F4 mg_u4(const MgRowBufs& B, float omega, float ui, uint32_t off) {
  if (MODE == MG_FIRST2) {
    const F4 i4 = mg_ldh4(B.inv, off, ui), r4 = buf_ld4(B.r, off);
    F4 u;
#pragma unroll
    for (int i = 0; i < 4; ++i) u.v[i] = omega * i4.v[i] * r4.v[i];
    return u;
  }
  return buf_ld4(B.x, off);
}
template <int MODE>
__device__ inline float mg_u1(const MgRowBufs& B, float omega, float ui, uint32_t off) {
  if (MODE == MG_FIRST2) return omega * mg_ldh1(B.inv, off, ui) * buf_ld(B.r, off);
  return buf_ld(B.x, off);
}

// level 1's correction at aggregates (I, J, K) and (I + 1, J, K), I even: adjacent
// in the slab
__device__ inline float2 mg_cpair(const MgRowArgs& a, int I, int J, int K) {
  return *reinterpret_cast<const float2*>(a.xc + bm_index(a.dc, I, J, K));
}

template <int MODE>
__global__ void __launch_bounds__(256) k_mg_row(MgRowArgs a) {
  if (a.done && *a.done) return;
  const int lane = threadIdx.x & 63;
  const int qb = lane >> 4, row = lane & 15, ly = row & 3, lz = row >> 2;
  const uint32_t bytes = (uint32_t)(a.d.nblocks * kBlockVol * 4), bytes16 = bytes / 2;
  MgRowBufs B;
  B.cd = buf_rsrc(a.cd, bytes16);
  B.cx = buf_rsrc(a.cx, bytes16);
  B.cy = buf_rsrc(a.cy, bytes16);
  B.cz = buf_rsrc(a.cz, bytes16);
  B.inv = buf_rsrc(a.inv, bytes16);
  const float ua = a.unscale[0], ui = a.unscale[1];
  B.r = buf_rsrc(a.r, bytes);
  B.x = buf_rsrc(a.x, bytes);
  B.out = buf_rsrc(a.out, bytes);
  const uint32_t sy = (uint32_t)a.d.nbx * kBlockVol * 4;
  const uint32_t sz = (uint32_t)a.d.nbx * a.d.nby * kBlockVol * 4;
  const float om = a.omega, omc = a.omega_c, omu = a.omega_u;
  double rz = 0.0;
  // the XCD segments of FOR_LIST_BLOCKS, walked in slots of four list entries
  const long long seg = (a.nlist + 7) / 8, xcd = blockIdx.x & 7;
  const long long s0 = xcd * seg, e = min((long long)a.nlist, (xcd + 1) * seg);
  const long long nslot = e > s0 ? (e - s0 + 3) / 4 : 0, step = (long long)(gridDim.x >> 3) * 4;
  const uint32_t last = e > 0 ? (uint32_t)(e - 1) : 0;
  for (long long sl = (long long)(blockIdx.x >> 3) * 4 + (threadIdx.x >> 6); sl < nslot; sl += step) {
    const uint32_t t0 = wave_uniform_u32((uint32_t)(s0 + sl * 4));
    // (no list: every block of the slab, in block order)
    const uint32_t b0 = a.list ? list_at(a.list, t0) : t0;
    const uint32_t b1 = a.list ? list_at(a.list, min(t0 + 1, last)) : min(t0 + 1, last);
    const uint32_t b2 = a.list ? list_at(a.list, min(t0 + 2, last)) : min(t0 + 2, last);
    const uint32_t b3 = a.list ? list_at(a.list, min(t0 + 3, last)) : min(t0 + 3, last);
    const bool valid = (long long)t0 + qb < e;
    const uint32_t b = qb == 0 ? b0 : (qb == 1 ? b1 : (qb == 2 ? b2 : b3));
    int bx, by, bz;
    block_coords(a.d, b, bx, by, bz);
    const uint32_t o = valid ? (b * kBlockVol + 4 * row) * 4 : kNoLoad;
    const bool fxp = valid && bx + 1 < a.d.nbx, fxm = valid && bx > 0;
    const bool fyp = valid && ly == 3 && by + 1 < a.d.nby, fym = valid && ly == 0 && by > 0;
    const bool fzp = valid && lz == 3 && bz + 1 < a.d.nbz, fzm = valid && lz == 0 && bz > 0;
    const uint32_t oxp = fxp ? o + kBlockVol * 4 : kNoLoad, oxm = fxm ? o - kBlockVol * 4 + 12 : kNoLoad;
    const uint32_t oyp = fyp ? o + sy - 48 : kNoLoad, oym = fym ? o - sy + 48 : kNoLoad;
    const uint32_t ozp = fzp ? o + sz - 192 : kNoLoad, ozm = fzm ? o - sz + 192 : kNoLoad;
    const F4 cd = mg_ldh4(B.cd, o, ua), cx = mg_ldh4(B.cx, o, ua), cy = mg_ldh4(B.cy, o, ua);
    const F4 cz = mg_ldh4(B.cz, o, ua);
    const F4 iv = mg_ldh4(B.inv, o, ui), rr = buf_ld4(B.r, o);
    F4 u = mg_u4<MODE>(B, omu, ui, o);
    float uxp = mg_u1<MODE>(B, omu, ui, oxp), uxm = mg_u1<MODE>(B, omu, ui, oxm);
    const float cxxm = mg_ldh1(B.cx, oxm, ua);
    F4 uyp = mg_u4<MODE>(B, omu, ui, oyp), uym = mg_u4<MODE>(B, omu, ui, oym);
    F4 uzp = mg_u4<MODE>(B, omu, ui, ozp), uzm = mg_u4<MODE>(B, omu, ui, ozm);
    const F4 cyym = mg_ldh4(B.cy, oym, ua), czzm = mg_ldh4(B.cz, ozm, ua);
    const int J = (4 * by + ly) >> 1, K = (4 * bz + lz) >> 1, I0 = 2 * bx;
    if (MODE == MG_PSWEEP) {
      if (valid) {
        const float2 c0 = mg_cpair(a, I0, J, K);
#pragma unroll
        for (int i = 0; i < 4; ++i)
          if (iv.v[i] != 0.0f) u.v[i] = fmaf(omc, i < 2 ? c0.x : c0.y, u.v[i]);
      }
      if (fxp) uxp = fmaf(omc, a.xc[bm_index(a.dc, I0 + 2, J, K)], uxp);
      if (fxm) uxm = fmaf(omc, a.xc[bm_index(a.dc, I0 - 1, J, K)], uxm);
      if (fyp) {
        const float2 c = mg_cpair(a, I0, 2 * by + 2, K);
        uyp.v[0] = fmaf(omc, c.x, uyp.v[0]); uyp.v[1] = fmaf(omc, c.x, uyp.v[1]);
        uyp.v[2] = fmaf(omc, c.y, uyp.v[2]); uyp.v[3] = fmaf(omc, c.y, uyp.v[3]);
      }
      if (fym) {
        const float2 c = mg_cpair(a, I0, 2 * by - 1, K);
        uym.v[0] = fmaf(omc, c.x, uym.v[0]); uym.v[1] = fmaf(omc, c.x, uym.v[1]);
        uym.v[2] = fmaf(omc, c.y, uym.v[2]); uym.v[3] = fmaf(omc, c.y, uym.v[3]);
      }
      if (fzp) {
        const float2 c = mg_cpair(a, I0, J, 2 * bz + 2);
        uzp.v[0] = fmaf(omc, c.x, uzp.v[0]); uzp.v[1] = fmaf(omc, c.x, uzp.v[1]);
        uzp.v[2] = fmaf(omc, c.y, uzp.v[2]); uzp.v[3] = fmaf(omc, c.y, uzp.v[3]);
      }
      if (fzm) {
        const float2 c = mg_cpair(a, I0, J, 2 * bz - 1);
        uzm.v[0] = fmaf(omc, c.x, uzm.v[0]); uzm.v[1] = fmaf(omc, c.x, uzm.v[1]);
        uzm.v[2] = fmaf(omc, c.y, uzm.v[2]); uzm.v[3] = fmaf(omc, c.y, uzm.v[3]);
      }
    }
    // (A u) per cell, in mg_ax's order of terms
    float res[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float syp = __shfl(u.v[i], lane + 1, 64), sym = __shfl(u.v[i], lane - 1, 64);
      const float szp = __shfl(u.v[i], lane + 4, 64), szm = __shfl(u.v[i], lane - 4, 64);
      const float scy = __shfl(cy.v[i], lane - 1, 64), scz = __shfl(cz.v[i], lane - 4, 64);
      const float uc = u.v[i];
      const float pxp = i < 3 ? u.v[i + 1] : (fxp ? uxp : uc);
      const float pxm = i > 0 ? u.v[i - 1] : (fxm ? uxm : uc);
      const float cxm = i > 0 ? cx.v[i - 1] : (fxm ? cxxm : 0.f);
      const float pyp = ly < 3 ? syp : (fyp ? uyp.v[i] : uc);
      const float pym = ly > 0 ? sym : (fym ? uym.v[i] : uc);
      const float cym = ly > 0 ? scy : (fym ? cyym.v[i] : 0.f);
      const float pzp = lz < 3 ? szp : (fzp ? uzp.v[i] : uc);
      const float pzm = lz > 0 ? szm : (fzm ? uzm.v[i] : uc);
      const float czm = lz > 0 ? scz : (fzm ? czzm.v[i] : 0.f);
      float ax = cd.v[i] * uc;
      ax = fmaf(cx.v[i], uc - pxp, ax);
      ax = fmaf(cxm, uc - pxm, ax);
      ax = fmaf(cy.v[i], uc - pyp, ax);
      ax = fmaf(cym, uc - pym, ax);
      ax = fmaf(cz.v[i], uc - pzp, ax);
      ax = fmaf(czm, uc - pzm, ax);
      res[i] = rr.v[i] - ax;
    }
    if (MODE == MG_RESTRICT) {
      float s0v = (iv.v[0] != 0.0f ? res[0] : 0.0f) + (iv.v[1] != 0.0f ? res[1] : 0.0f);
      float s1v = (iv.v[2] != 0.0f ? res[2] : 0.0f) + (iv.v[3] != 0.0f ? res[3] : 0.0f);
      s0v += __shfl_xor(s0v, 1);
      s1v += __shfl_xor(s1v, 1);
      s0v += __shfl_xor(s0v, 4);
      s1v += __shfl_xor(s1v, 4);
      if (valid && (ly & 1) == 0 && (lz & 1) == 0)
        *reinterpret_cast<float2*>(a.out + bm_index(a.dc, I0, J, K)) = make_float2(s0v, s1v);
    } else {
      float ov[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        ov[i] = fmaf(om * iv.v[i], res[i], u.v[i]);
        rz += (double)rr.v[i] * (double)ov[i];
      }
      buf_st4(B.out, o, ov);
    }
  }
  if (MODE == MG_SWEEP || MODE == MG_PSWEEP) {
    if (a.part) {
      double v[1] = {rz};
      const bool mx[1] = {false};
      block_reduce_store<1>(v, mx, a.part, kMaxPartials);
    }
  }
}

}  // namespace pf
