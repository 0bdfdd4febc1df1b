// pf_iface.hip -- interface diagnostics and CSF surface tension on the sparse grid
// (src/driver/interface_diagnostics3d.cpp).
//
// Per cell, in FP64 and in the reference's operation order:
//   phi   = (sum of the six face phi) / 6                              (:137-148)
//   grad  = central differences of phi with clamped indices            (:158-175)
//   n     = grad * (1 / |grad|), zero below 1e-12                      (:177-186)
//   kappa = sum over axes of (n_+ - n_-) / (span dx)                   (:188-208)
// The fields are dense over the cell slab (cells outside the domain hold zero), so
// each stencil is a few loads.  Diagnostics reduce over the FLUID cells; the CSF
// increment grad * (-strength * kappa_s * dt), limited to max_delta_speed, goes half
// to each face of the cell (:404-440) -- a face adds its lower cell's share first,
// as the reference's loop over sorted cells does, and no cell at or after the first
// non-finite increment adds anything.
//
// Multires grids (diagnoseMRInterface3D / applyMRSurfaceTension3D, :381-401, :447-496)
// use the same fine-grid stencils over the fine face patches (faces the layout does
// not enumerate hold zero mass, as the reference's map lookups return), but sample
// each FLUID leaf cell once, at its fine centre clamped into the grid, and visit
// leaf cells in leafCells order: level, then block (bx, by, bz), then lz, ly, lx.
#include "pf_internal.h"

namespace pf {

constexpr double kIfacePhiLo = 0.05, kIfacePhiHi = 0.95;
constexpr double kIfaceGradEps = 1e-5, kIfaceNormalEps = 1e-12;
constexpr int kIfaceSlots = 10;

struct IfaceArgs {
  Dims dc, du, dv, dw;
  Geom g;
  const float *mu, *mv, *mw;
  const uint8_t* marker;
  PhaseK ph;
  double* phi;      // per cell
  double* gmag;     // |grad phi| per cell
  double* nrm[3];   // unit normal per cell
  double* kap;      // curvature (FLUID or interface cells)
  const uint8_t* lvl;  // multires: leaf level per fine cell (255 = none); null on the sparse grid
};

__device__ inline void cell_of(const Dims& dc, long long q, int& i, int& j, int& k) {
  int bx, by, bz;
  block_coords(dc, q >> 6, bx, by, bz);
  const int lane = (int)(q & 63);
  i = bx * 4 + (lane & 3);
  j = by * 4 + ((lane >> 2) & 3);
  k = bz * 4 + (lane >> 4);
}

__device__ inline int clampd(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__device__ inline bool iface_phi(double p) { return p > kIfacePhiLo && p < kIfacePhiHi; }

// whether fine cell (i, j, k) is where its leaf cell is sampled: every cell on the
// sparse grid and in a level-0 leaf; in a level-1 leaf (2^3 fine cells from an even
// corner) the centre corner + 1, clamped into the grid (mrCellFineCenter, :350-358)
__device__ inline bool iface_sample_point(const IfaceArgs& a, long long q, int i, int j, int k) {
  if (!a.lvl) return true;
  const int L = a.lvl[q];
  if (L == 0) return true;
  if (L != 1) return false;
  return i == min((i & ~1) + 1, a.g.nx - 1) && j == min((j & ~1) + 1, a.g.ny - 1) &&
         k == min((k & ~1) + 1, a.g.nz - 1);
}

// the position of a sample point in the reference's visiting order: the linear cell
// id on the sparse grid (collectCellsWithMarker sorts them), the leafCells rank on a
// multires grid as (level, bx, by, bz, lz, ly, lx) packed into 64 bits
__device__ inline unsigned long long iface_order(const IfaceArgs& a, long long q, int i, int j, int k) {
  if (!a.lvl) return (unsigned long long)i + (unsigned long long)a.g.nx * (j + (unsigned long long)a.g.ny * k);
  const int L = a.lvl[q];
  const unsigned ci = (unsigned)i >> L, cj = (unsigned)j >> L, ck = (unsigned)k >> L;
  return ((unsigned long long)L << 60) | ((unsigned long long)(ci >> 2) << 42) |
         ((unsigned long long)(cj >> 2) << 24) | ((unsigned long long)(ck >> 2) << 6) |
         (unsigned long long)((ci & 3) + 4 * ((cj & 3) + 4 * (ck & 3)));
}

// sparseCellPhi (:137-148)
__global__ void __launch_bounds__(256) k_iface_phi(IfaceArgs a) {
#pragma clang fp contract(off)
  const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
  if (q >= a.dc.nblocks * kBlockVol) return;
  int i, j, k;
  cell_of(a.dc, q, i, j, k);
  double s = 0.0;
  if (i < a.g.nx && j < a.g.ny && k < a.g.nz) {
    s += phi_of_raw((double)a.mu[bm_index(a.du, i, j, k)], a.ph);
    s += phi_of_raw((double)a.mu[bm_index(a.du, i + 1, j, k)], a.ph);
    s += phi_of_raw((double)a.mv[bm_index(a.dv, i, j, k)], a.ph);
    s += phi_of_raw((double)a.mv[bm_index(a.dv, i, j + 1, k)], a.ph);
    s += phi_of_raw((double)a.mw[bm_index(a.dw, i, j, k)], a.ph);
    s += phi_of_raw((double)a.mw[bm_index(a.dw, i, j, k + 1)], a.ph);
    s = s / 6.0;
  }
  a.phi[q] = s;
}

__device__ inline double phi_at(const IfaceArgs& a, int i, int j, int k) {
  return a.phi[bm_index(a.dc, clampd(i, 0, a.g.nx - 1), clampd(j, 0, a.g.ny - 1), clampd(k, 0, a.g.nz - 1))];
}

// sparseGradient / sparseNormal (:158-186)
__device__ inline void gradient_at(const IfaceArgs& a, int i, int j, int k, double gr[3]) {
#pragma clang fp contract(off)
  const int im = clampd(i - 1, 0, a.g.nx - 1), ip = clampd(i + 1, 0, a.g.nx - 1);
  const int jm = clampd(j - 1, 0, a.g.ny - 1), jp = clampd(j + 1, 0, a.g.ny - 1);
  const int km = clampd(k - 1, 0, a.g.nz - 1), kp = clampd(k + 1, 0, a.g.nz - 1);
  const double dx = a.g.dx;
  gr[0] = (phi_at(a, ip, j, k) - phi_at(a, im, j, k)) / ((double)max(1, ip - im) * dx);
  gr[1] = (phi_at(a, i, jp, k) - phi_at(a, i, jm, k)) / ((double)max(1, jp - jm) * dx);
  gr[2] = (phi_at(a, i, j, kp) - phi_at(a, i, j, km)) / ((double)max(1, kp - km) * dx);
}

__device__ inline double vlen3(const double v[3]) {
#pragma clang fp contract(off)
  return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
}

__global__ void __launch_bounds__(256) k_iface_normal(IfaceArgs a) {
#pragma clang fp contract(off)
  const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
  if (q >= a.dc.nblocks * kBlockVol) return;
  int i, j, k;
  cell_of(a.dc, q, i, j, k);
  double gr[3] = {0.0, 0.0, 0.0}, n[3] = {0.0, 0.0, 0.0}, mag = 0.0;
  if (i < a.g.nx && j < a.g.ny && k < a.g.nz) {
    gradient_at(a, i, j, k, gr);
    mag = vlen3(gr);
    if (mag > kIfaceNormalEps && isfinite(mag)) {
      const double inv = 1.0 / mag;
      n[0] = gr[0] * inv;
      n[1] = gr[1] * inv;
      n[2] = gr[2] * inv;
    }
  }
  a.gmag[q] = mag;
  a.nrm[0][q] = n[0];
  a.nrm[1][q] = n[1];
  a.nrm[2][q] = n[2];
}

// sparseCurvature (:188-208), for the FLUID cells and every interface cell
__device__ inline double curvature_at(const IfaceArgs& a, int i, int j, int k) {
#pragma clang fp contract(off)
  const int im = clampd(i - 1, 0, a.g.nx - 1), ip = clampd(i + 1, 0, a.g.nx - 1);
  const int jm = clampd(j - 1, 0, a.g.ny - 1), jp = clampd(j + 1, 0, a.g.ny - 1);
  const int km = clampd(k - 1, 0, a.g.nz - 1), kp = clampd(k + 1, 0, a.g.nz - 1);
  const double dx = a.g.dx;
  const double xs = a.nrm[0][bm_index(a.dc, ip, j, k)] - a.nrm[0][bm_index(a.dc, im, j, k)];
  const double ys = a.nrm[1][bm_index(a.dc, i, jp, k)] - a.nrm[1][bm_index(a.dc, i, jm, k)];
  const double zs = a.nrm[2][bm_index(a.dc, i, j, kp)] - a.nrm[2][bm_index(a.dc, i, j, km)];
  return xs / ((double)max(1, ip - im) * dx) + ys / ((double)max(1, jp - jm) * dx) +
         zs / ((double)max(1, kp - km) * dx);
}

__global__ void __launch_bounds__(256) k_iface_curvature(IfaceArgs a) {
  const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
  if (q >= a.dc.nblocks * kBlockVol) return;
  int i, j, k;
  cell_of(a.dc, q, i, j, k);
  double kap = 0.0;
  if (i < a.g.nx && j < a.g.ny && k < a.g.nz && (a.marker[q] == CELL_FLUID || iface_phi(a.phi[q])))
    kap = curvature_at(a, i, j, k);
  a.kap[q] = kap;
}

// per-workgroup partials of the diagnostics (accumulate, :26-53): count, phi min,
// phi max, phi sum, interface count, grad sum, grad max, |kappa| sum, |kappa| max,
// non-finite count
__global__ void __launch_bounds__(256) k_iface_diag(IfaceArgs a, double* __restrict__ part) {
  __shared__ double s[4][kIfaceSlots];
  double v[kIfaceSlots] = {0.0, INFINITY, -INFINITY, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const long long n = a.dc.nblocks * kBlockVol;
  for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < n; q += (long long)gridDim.x * 256) {
    if (a.marker[q] != CELL_FLUID) continue;
    if (a.lvl) {
      int i, j, k;
      cell_of(a.dc, q, i, j, k);
      if (!iface_sample_point(a, q, i, j, k)) continue;
    }
    const double phi = a.phi[q], gm = a.gmag[q], kp = a.kap[q];
    if (!isfinite(phi) || !isfinite(gm) || !isfinite(kp)) {
      v[9] += 1.0;
      continue;
    }
    v[0] += 1.0;
    v[1] = fmin(v[1], phi);
    v[2] = fmax(v[2], phi);
    v[3] += phi;
    if (!iface_phi(phi)) continue;
    v[4] += 1.0;
    v[5] += gm;
    v[6] = fmax(v[6], gm);
    v[7] += fabs(kp);
    v[8] = fmax(v[8], fabs(kp));
  }
#pragma unroll
  for (int t = 0; t < kIfaceSlots; ++t) {
    double x = v[t];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double y = __shfl_xor(x, o, 64);
      x = (t == 1) ? fmin(x, y) : ((t == 2 || t == 6 || t == 8) ? fmax(x, y) : x + y);
    }
    v[t] = x;
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int t = 0; t < kIfaceSlots; ++t) s[w][t] = v[t];
  __syncthreads();
  if (threadIdx.x < kIfaceSlots) {
    const int t = threadIdx.x;
    double x = s[0][t];
    for (int q = 1; q < 4; ++q)
      x = (t == 1) ? fmin(x, s[q][t]) : ((t == 2 || t == 6 || t == 8) ? fmax(x, s[q][t]) : x + s[q][t]);
    part[(long long)blockIdx.x * kIfaceSlots + t] = x;
  }
}

// the combine of per-workgroup partials: n slots, op 0 = sum, 1 = min, 2 = max
struct CombineOps {
  int n;
  int op[16];
};
// Per-workgroup partials part[w * ops.n + t] combined per slot t by ops.op[t] (0 sum,
// 1 min, 2 max) in a fixed order: thread h folds workgroups h, h + 256, ... in turn, then
// a pairwise tree over the 256 threads.  The order does not depend on timing, so the
// result is the same from run to run.  fmin/fmax start from NaN, which they ignore, so
// a slot whose partials are all NaN stays NaN.
constexpr int kCombineThreads = 256;
__device__ inline double combine_op(int op, double x, double y) {
  return op == 1 ? fmin(x, y) : (op == 2 ? fmax(x, y) : x + y);
}
__global__ void __launch_bounds__(kCombineThreads) k_iface_combine(const double* __restrict__ part, int nwg,
                                                                   CombineOps ops, double* __restrict__ out) {
  __shared__ double s[16][kCombineThreads];
  const int h = threadIdx.x;
  double x[16];
#pragma unroll
  for (int t = 0; t < 16; ++t) x[t] = ops.op[t] == 0 ? 0.0 : NAN;
  for (int w = h; w < nwg; w += kCombineThreads) {
#pragma unroll
    for (int t = 0; t < 16; ++t)
      if (t < ops.n) x[t] = combine_op(ops.op[t], x[t], part[(long long)w * ops.n + t]);
  }
#pragma unroll
  for (int t = 0; t < 16; ++t) s[t][h] = x[t];
  __syncthreads();
  for (int half = kCombineThreads / 2; half > 0; half >>= 1) {
    if (h < half) {
#pragma unroll
      for (int t = 0; t < 16; ++t)
        if (t < ops.n) s[t][h] = combine_op(ops.op[t], s[t][h], s[t][h + half]);
    }
    __syncthreads();
  }
  if (h < ops.n) out[h] = s[h][0];
}

struct CsfArgs {
  IfaceArgs f;
  double dt, strength, maxd;
  int radius;
  double* delta[3];            // CSF increment per cell (NaN where none applies)
  unsigned long long* first_bad;  // least visiting-order key of a non-finite increment
};

constexpr int kCsfSlots = 8;   // applied, |d| sum, |d| max, |raw| sum, |raw| max, |sm| sum, |sm| max, non-finite

// the CSF increment of each FLUID interface cell (:413-431) and per-workgroup
// statistics (accumulateSurfaceStats, :104-122)
__global__ void __launch_bounds__(256) k_csf_delta(CsfArgs c, double* __restrict__ part) {
#pragma clang fp contract(off)
  const IfaceArgs& a = c.f;
  __shared__ double s[4][kCsfSlots];
  double v[kCsfSlots] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const long long n = a.dc.nblocks * kBlockVol;
  for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < n; q += (long long)gridDim.x * 256) {
    double d[3] = {NAN, NAN, NAN};
    int i = 0, j = 0, k = 0;
    if (a.marker[q] == CELL_FLUID && a.lvl) cell_of(a.dc, q, i, j, k);
    // a NaN gradient passes the threshold test (:425-426) and ends as a zero increment
    if (a.marker[q] == CELL_FLUID && iface_phi(a.phi[q]) && !(a.gmag[q] <= kIfaceGradEps) &&
        (!a.lvl || iface_sample_point(a, q, i, j, k))) {
      if (!a.lvl) cell_of(a.dc, q, i, j, k);
      double gr[3];
      gradient_at(a, i, j, k, gr);
      const double raw = a.kap[q];
      double sm = raw;
      if (c.radius > 0 && isfinite(raw)) {
        // sparseSmoothedCurvature (:210-235): clamped box, interface cells only
        double sum = 0.0;
        int cnt = 0;
        for (int dz = -c.radius; dz <= c.radius; ++dz)
          for (int dy = -c.radius; dy <= c.radius; ++dy)
            for (int dx = -c.radius; dx <= c.radius; ++dx) {
              const long long nq = bm_index(a.dc, clampd(i + dx, 0, a.g.nx - 1), clampd(j + dy, 0, a.g.ny - 1),
                                            clampd(k + dz, 0, a.g.nz - 1));
              if (!iface_phi(a.phi[nq])) continue;
              const double cv = a.kap[nq];
              if (!isfinite(cv)) continue;
              sum += cv;
              ++cnt;
            }
        sm = cnt > 0 ? sum / (double)cnt : raw;
      }
      // limitedDelta (:72-79)
      const double f = -c.strength * sm * c.dt;
      d[0] = gr[0] * f;
      d[1] = gr[1] * f;
      d[2] = gr[2] * f;
      const double m0 = vlen3(d);
      if (!isfinite(m0)) {
        d[0] = d[1] = d[2] = 0.0;
      } else if (c.maxd > 0.0 && m0 > c.maxd) {
        const double sc = c.maxd / m0;
        d[0] *= sc;
        d[1] *= sc;
        d[2] *= sc;
      }
      const double mag = vlen3(d);
      if (!isfinite(mag) || !isfinite(raw) || !isfinite(sm)) {
        v[7] += 1.0;
        atomicMin(c.first_bad, iface_order(a, q, i, j, k));
        d[0] = d[1] = d[2] = NAN;
      } else {
        v[0] += 1.0;
        v[1] += mag;
        v[2] = fmax(v[2], mag);
        v[3] += fabs(raw);
        v[4] = fmax(v[4], fabs(raw));
        v[5] += fabs(sm);
        v[6] = fmax(v[6], fabs(sm));
      }
    }
    c.delta[0][q] = d[0];
    c.delta[1][q] = d[1];
    c.delta[2][q] = d[2];
  }
#pragma unroll
  for (int t = 0; t < kCsfSlots; ++t) {
    double x = v[t];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double y = __shfl_xor(x, o, 64);
      x = (t == 2 || t == 4 || t == 6) ? fmax(x, y) : x + y;
    }
    v[t] = x;
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int t = 0; t < kCsfSlots; ++t) s[w][t] = v[t];
  __syncthreads();
  if (threadIdx.x < kCsfSlots) {
    const int t = threadIdx.x;
    double x = s[0][t];
    for (int q = 1; q < 4; ++q) x = (t == 2 || t == 4 || t == 6) ? fmax(x, s[q][t]) : x + s[q][t];
    part[(long long)blockIdx.x * kCsfSlots + t] = x;
  }
}

// the faces of one axis: u += (float)(0.5 d) from the two cells beside the face, in
// the reference's visiting order, where the face mass is positive, the cell's
// increment applies (not NaN) and the cell precedes the first non-finite increment
__global__ void __launch_bounds__(256) k_csf_faces(CsfArgs c, int axis, float* __restrict__ vel,
                                                   const float* __restrict__ mass) {
  const IfaceArgs& a = c.f;
  const Dims& fd = axis == 0 ? a.du : (axis == 1 ? a.dv : a.dw);
  const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
  if (q >= fd.nblocks * kBlockVol) return;
  int i, j, k;
  cell_of(fd, q, i, j, k);
  if (i >= fd.nx || j >= fd.ny || k >= fd.nz) return;
  if (!(mass[q] > 0.0f)) return;
  const unsigned long long bad = *c.first_bad;
  const int li = i - (axis == 0), lj = j - (axis == 1), lk = k - (axis == 2);
  double d0 = NAN, d1 = NAN;
  unsigned long long k0 = ~0ull, k1 = ~0ull;
  if (li >= 0 && lj >= 0 && lk >= 0) {
    const long long cq = bm_index(a.dc, li, lj, lk);
    k0 = iface_order(a, cq, li, lj, lk);
    if (k0 < bad) d0 = c.delta[axis][cq];
  }
  if (i < a.g.nx && j < a.g.ny && k < a.g.nz) {
    const long long cq = bm_index(a.dc, i, j, k);
    k1 = iface_order(a, cq, i, j, k);
    if (k1 < bad) d1 = c.delta[axis][cq];
  }
  if (k1 < k0) {
    const double t = d0;
    d0 = d1;
    d1 = t;
  }
  float u = vel[q];
  if (!isnan(d0)) u = u + (float)(0.5 * d0);
  if (!isnan(d1)) u = u + (float)(0.5 * d1);
  vel[q] = u;
}

}  // namespace pf
