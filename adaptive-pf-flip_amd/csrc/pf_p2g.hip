// pf_p2g.hip -- normalized-cubic P2G splat (Eq. 6) into MAC momentum and raw mass,
// and the single-phase trilinear splat (same launch structure).
//
// Reference: spP2G3D_tp / preactivateKernel / splatKernelColor
// (src/driver/sparse_ops3d_tp.cpp:15-75, 95-124).  Weight w = (1 - d^2/r^2)^3, r = 1.5,
// normalised per particle over the in-bounds nodes of its stencil; mass += w*m_p,
// momentum += w*m_p*v.
//
// MI355X mapping:
// * One workgroup = one 64-lane wavefront = one 4^3 source block of cells; lane l
//   owns cell l of the block and splats that cell's particles (they are contiguous
//   after binning).  For a particle in cell c the non-zero weights of a face field
//   lie in c-1..c+2 along the face axis and c-1..c+1 across it (the reference's
//   6-wide loop box adds only zero weights), so a lane keeps its 36 node sums per
//   axis in registers and a block's contributions stay inside a 7x6x6 node region.
// * Region sums are accumulated in LDS: for a fixed stencil offset the 64 lanes hit
//   64 distinct nodes, and one wave's LDS operations complete in program order, so
//   the sum order is fixed -- no global atomics, and the result does not change
//   from run to run.  A cell's particles are visited in ascending caller index.
// * Regions of two source blocks of equal color8 (src/grid/block_color.h:7) are at
//   least 8 cells apart and never overlap, so each of the 8 colour launches adds its
//   regions into HBM with plain read-modify-write.
// * The activation predicate kern(d^2) > 0 is d^2 < 2.25 (exact: d^2/2.25 < 1 in
//   correctly rounded FP64 iff d^2 < 2.25).  It is evaluated in FP32 and recomputed
//   in FP64 inside a 1e-5 band around 2.25, so active-block sets match the FP64
//   reference bit for bit while weights run in FP32.
#include "pf_internal.h"

namespace pf {

struct P2GArgs {
  Geom g;
  Dims dc;
  Dims df[3];
  float* mom[3];
  float* mass[3];
  uint8_t* act_vel[3];
  uint8_t* act_mass[3];
  const uint32_t* start;
  const uint32_t* list;   // blocks of this colour
  const uint32_t* count;  // number of blocks in list (device)
  PBuf P;
  bool has_vol;
  bool pow2;              // dx is a power of two: (x - o)/dx == (x - o)*inv_dx exactly
  double inv_dx;
  bool fast32;            // pow2, origin 0 and fewer than 2^22 cells a side (see splat_axis)
  float inv_dx_f;
  float mp_l, mp_g;       // rho * Vp per phase
  const uint8_t* valid[3];  // KT = 2 (multires): the enumerated fine faces per axis
};

constexpr int kOrdMax = 16;
constexpr float kGuard = 1e-5f;
constexpr float kGuardT = 1e-5f;  // on t = 1 - d^2/2.25: wider than kGuard / 2.25

template <int AX>
struct Ax {
  static constexpr int RX = AX == 0 ? 7 : 6;
  static constexpr int RY = AX == 1 ? 7 : 6;
  static constexpr int RZ = AX == 2 ? 7 : 6;
  static constexpr int NX = AX == 0 ? 4 : 3;
  static constexpr int NY = AX == 1 ? 4 : 3;
  static constexpr int NZ = AX == 2 ? 4 : 3;
  static constexpr int NN = NX * NY * NZ;  // 36
  static constexpr int RN = RX * RY * RZ;  // 252
};

__device__ inline double grid_coord(float x, double o, const P2GArgs& a) {
  double d = (double)x - o;
  return a.pow2 ? d * a.inv_dx : d / a.g.dx;
}

// exact FP64 activation test and weight, kern() of sparse_ops3d_tp.cpp:15-19
__device__ inline float exact_weight(double ddx, double ddy, double ddz, bool& active) {
  double d2 = __dadd_rn(__dadd_rn(__dmul_rn(ddx, ddx), __dmul_rn(ddy, ddy)), __dmul_rn(ddz, ddz));
  double q = __ddiv_rn(d2, 2.25);
  double t = __dsub_rn(1.0, q);
  active = t > 0.0;
  return active ? (float)(t * t * t) : 0.f;
}

// Single-phase trilinear splat, spP2G3D (src/driver/sparse_ops3d.cpp:22-68, 72-100):
// nodes floor(g)..floor(g)+1 per axis with weights (1-f, f), w = (float)(wx*wy*wz)
// in FP64, unit particle mass, no normalisation; every in-bounds stencil node
// activates its block, zero weights included.  For a particle in cell c the nodes
// lie in c..c+1 along the face axis and c-1..c+1 across it, inside the cubic
// kernel's register window, so both kernels share the region and flush logic.
template <int AX>
__device__ inline void trilinear_weights(double gx, double gy, double gz, int i, int j, int k,
                                         const bool* vx, const bool* vy, const bool* vz, float* wt,
                                         uint64_t& fbits) {
  using A = Ax<AX>;
  const int i0 = (int)floor(gx), j0 = (int)floor(gy), k0 = (int)floor(gz);
  const double fx = gx - i0, fy = gy - j0, fz = gz - k0;
  const double wx[2] = {1.0 - fx, fx}, wy[2] = {1.0 - fy, fy}, wz[2] = {1.0 - fz, fz};
#pragma unroll
  for (int n = 0; n < A::NN; ++n) wt[n] = 0.f;
#pragma unroll
  for (int dk = 0; dk < 2; ++dk)
#pragma unroll
    for (int dj = 0; dj < 2; ++dj)
#pragma unroll
      for (int di = 0; di < 2; ++di) {
        // window offsets (node - cell + 1); node indices of the window are in range
        const int ox = i0 + di - i + 1, oy = j0 + dj - j + 1, oz = k0 + dk - k + 1;
        if (ox < 0 || ox >= A::NX || oy < 0 || oy >= A::NY || oz < 0 || oz >= A::NZ) continue;
        if (!(vx[ox] && vy[oy] && vz[oz])) continue;
        const int n = ox + A::NX * (oy + A::NY * oz);
        wt[n] = (float)(__dmul_rn(__dmul_rn(wx[di], wy[dj]), wz[dk]));
        fbits |= (1ull << n);
      }
}

template <int AX, int KT>
__device__ inline void splat_axis(const P2GArgs& a, int i, int j, int k, uint32_t s0, uint32_t cnt,
                                  const uint8_t* ord, float* s_mom, float* s_mass, uint8_t* s_flag,
                                  int lx, int ly, int lz) {
  using A = Ax<AX>;
  const Dims& fd = a.df[AX];
  float am[A::NN], aq[A::NN];
  uint64_t fbits = 0;
#pragma unroll
  for (int n = 0; n < A::NN; ++n) { am[n] = 0.f; aq[n] = 0.f; }

  // node validity along each axis (field extents), fixed per lane
  bool vx[A::NX], vy[A::NY], vz[A::NZ];
#pragma unroll
  for (int o = 0; o < A::NX; ++o) { int c = i + o - 1; vx[o] = c >= 0 && c < fd.nx; }
#pragma unroll
  for (int o = 0; o < A::NY; ++o) { int c = j + o - 1; vy[o] = c >= 0 && c < fd.ny; }
#pragma unroll
  for (int o = 0; o < A::NZ; ++o) { int c = k + o - 1; vz[o] = c >= 0 && c < fd.nz; }
  // multires: only the enumerated faces take (and normalise) weights
  // (collectWeightedFaces, multires_ops3d_tp.cpp:62-75)
  uint64_t vbits = ~0ull;
  if (KT == 2) {
    vbits = 0;
#pragma unroll
    for (int oz = 0; oz < A::NZ; ++oz)
#pragma unroll
      for (int oy = 0; oy < A::NY; ++oy)
#pragma unroll
        for (int ox = 0; ox < A::NX; ++ox)
          if (vx[ox] && vy[oy] && vz[oz] &&
              a.valid[AX][bm_index(fd, i + ox - 1, j + oy - 1, k + oz - 1)])
            vbits |= 1ull << (ox + A::NX * (oy + A::NY * oz));
  }

  uint32_t last = 0;
  bool first = true;
  // the particle data of the next iteration are loaded while this one computes (for
  // cells of at most kOrdMax particles, whose order is the precomputed ranking)
  const bool piped = cnt <= (uint32_t)kOrdMax;
  const float* vp = AX == 0 ? a.P.u : (AX == 1 ? a.P.v : a.P.w);
  float nx_ = 0.f, ny_ = 0.f, nz_ = 0.f, nv_ = 0.f, nvol_ = 1.f;
  uint8_t nty_ = 0;
  if (piped && cnt > 0) {
    const uint32_t q0 = s0 + ord[0];
    nx_ = a.P.x[q0]; ny_ = a.P.y[q0]; nz_ = a.P.z[q0]; nv_ = vp[q0]; nty_ = a.P.type[q0];
    if (a.has_vol) nvol_ = a.P.vol[q0];
  }
  for (uint32_t it = 0; it < cnt; ++it) {
    uint32_t q;
    float px, py, pz, vel, pvol = 1.f;
    uint8_t ty;
    if (piped) {
      q = s0 + ord[it];
      px = nx_; py = ny_; pz = nz_; vel = nv_; ty = nty_; pvol = nvol_;
      if (it + 1 < cnt) {
        const uint32_t qn = s0 + ord[it + 1];
        nx_ = a.P.x[qn]; ny_ = a.P.y[qn]; nz_ = a.P.z[qn]; nv_ = vp[qn]; nty_ = a.P.type[qn];
        if (a.has_vol) nvol_ = a.P.vol[qn];
      }
    } else {  // selection by ascending caller index
      uint32_t best = 0xffffffffu, bi = 0;
      for (uint32_t t = 0; t < cnt; ++t) {
        uint32_t id = a.P.id[s0 + t];
        if ((first || id > last) && id <= best) { best = id; bi = t; }
      }
      last = best;
      first = false;
      q = s0 + bi;
      px = a.P.x[q]; py = a.P.y[q]; pz = a.P.z[q]; vel = vp[q]; ty = a.P.type[q];
      if (a.has_vol) pvol = a.P.vol[q];
    }
    // stencil coordinates of this axis' face grid (sparse_ops3d_tp.cpp:117-122) and
    // the particle's offsets from the lane's cell
    double gx, gy, gz;
    float fx, fy, fz;
    if (KT == 0 && a.fast32) {
      // x/dx is exact in FP32 here, and each offset is either exact or one correctly
      // rounded FP32 subtraction of exact values (sx - 0.5 is inexact only for
      // sx < 0.25, in cell 0): the FP64 path's single rounding, bit for bit.  The
      // FP64 coordinates (exact) are used only by the guard-band path.
      const float sx = px * a.inv_dx_f, sy = py * a.inv_dx_f, sz = pz * a.inv_dx_f;
      fx = (AX == 0 ? sx : sx - 0.5f) - (float)i;
      fy = (AX == 1 ? sy : sy - 0.5f) - (float)j;
      fz = (AX == 2 ? sz : sz - 0.5f) - (float)k;
      gx = AX == 0 ? (double)sx : (double)sx - 0.5;
      gy = AX == 1 ? (double)sy : (double)sy - 0.5;
      gz = AX == 2 ? (double)sz : (double)sz - 0.5;
    } else {
      double dpx = grid_coord(px, a.g.ox, a), dpy = grid_coord(py, a.g.oy, a), dpz = grid_coord(pz, a.g.oz, a);
      gx = AX == 0 ? dpx : dpx - 0.5;
      gy = AX == 1 ? dpy : dpy - 0.5;
      gz = AX == 2 ? dpz : dpz - 0.5;
      fx = (float)(gx - (double)i);
      fy = (float)(gy - (double)j);
      fz = (float)(gz - (double)k);
    }
    (void)q;
    if (KT == 1) {
      float wt[A::NN];
      trilinear_weights<AX>(gx, gy, gz, i, j, k, vx, vy, vz, wt, fbits);
      // mass += (float)(w * 1.0) = w; momentum += (float)(w * v) with v exact in FP64
#pragma unroll
      for (int n = 0; n < A::NN; ++n) {
        am[n] += wt[n];
        aq[n] += (float)((double)wt[n] * (double)vel);
      }
      continue;
    }
    float dx2[A::NX], dy2[A::NY], dz2[A::NZ];
#pragma unroll
    for (int o = 0; o < A::NX; ++o) { float d = fx - (float)(o - 1); dx2[o] = d * d; }
#pragma unroll
    for (int o = 0; o < A::NY; ++o) { float d = fy - (float)(o - 1); dy2[o] = d * d; }
#pragma unroll
    for (int o = 0; o < A::NZ; ++o) { float d = fz - (float)(o - 1); dz2[o] = d * d; }
    float wt[A::NN];
    float wsum = 0.f;
    // two-phase: the separable form t = (1 - (dx2 + dy2)/2.25) - dz2/2.25, w = t^3 for
    // t > 0, with rows and columns outside the domain made negative -- one
    // subtraction per node and no branch.  A node with |t| < kGuardT (its d^2 within
    // about 1e-5 of 2.25) sends the particle through the per-node loop below, which
    // decides those nodes with the FP64 predicate; every other node is far enough from
    // the edge that t > 0 and the FP64 d^2 < 2.25 agree.
    bool slow = KT != 0;
    if (KT == 0) {
      float Axy[A::NX * A::NY], Bz[A::NZ];
#pragma unroll
      for (int oy = 0; oy < A::NY; ++oy)
#pragma unroll
        for (int ox = 0; ox < A::NX; ++ox)
          Axy[ox + A::NX * oy] = (vx[ox] && vy[oy]) ? 1.f - (dx2[ox] + dy2[oy]) * (1.f / 2.25f) : -4.f;
#pragma unroll
      for (int oz = 0; oz < A::NZ; ++oz) Bz[oz] = vz[oz] ? dz2[oz] * (1.f / 2.25f) : 4.f;
      // (a fast-path node is activated by its positive weight: w = t^3 >= 1e-15 where
      // t > 0 here, so the lane's accumulated mass there is positive -- the
      // activation bits of these nodes are read from the sums after the loop)
      float tmin = 4.f;
#pragma unroll
      for (int oz = 0; oz < A::NZ; ++oz)
#pragma unroll
        for (int oy = 0; oy < A::NY; ++oy)
#pragma unroll
          for (int ox = 0; ox < A::NX; ++ox) {
            const int n = ox + A::NX * (oy + A::NY * oz);
            const float t = Axy[ox + A::NX * oy] - Bz[oz];
            tmin = fminf(tmin, fabsf(t));
            const float tp = fmaxf(t, 0.f);
            const float w = tp * tp * tp;
            wt[n] = w;
            wsum += w;
          }
      slow = tmin < kGuardT;
    }
    if (slow) {
      wsum = 0.f;
  #pragma unroll
      for (int oz = 0; oz < A::NZ; ++oz)
  #pragma unroll
        for (int oy = 0; oy < A::NY; ++oy)
  #pragma unroll
          for (int ox = 0; ox < A::NX; ++ox) {
            const int n = ox + A::NX * (oy + A::NY * oz);
            float d2 = dx2[ox] + dy2[oy] + dz2[oz];
            bool inb = vx[ox] && vy[oy] && vz[oz];
            float w;
            bool act;
            if (fabsf(d2 - 2.25f) < kGuard) {
              w = exact_weight(gx - (double)(i + ox - 1), gy - (double)(j + oy - 1), gz - (double)(k + oz - 1),
                               act);
            } else {
              act = d2 < 2.25f;
              float t = 1.f - d2 * (1.f / 2.25f);
              w = act ? t * t * t : 0.f;
            }
            act = act && inb && ((vbits >> n) & 1ull);
            w = act ? w : 0.f;
            wt[n] = w;
            wsum += w;
            if (act) fbits |= (1ull << n);
          }
    }
    float mp = (ty == 0) ? a.mp_l : a.mp_g;
    if (a.has_vol) mp *= pvol;
    float sm = wsum > 0.f ? mp / wsum : 0.f;
    float sq = sm * vel;
#pragma unroll
    for (int n = 0; n < A::NN; ++n) {
      am[n] = fmaf(wt[n], sm, am[n]);
      aq[n] = fmaf(wt[n], sq, aq[n]);
    }
  }
  if (KT == 0) {
#pragma unroll
    for (int n = 0; n < A::NN; ++n)
      if (am[n] > 0.f) fbits |= 1ull << n;
  }
  // deposit into the LDS region: for fixed n the 64 lanes address 64 distinct nodes
  const int base = lx + A::RX * (ly + A::RY * lz);
#pragma unroll
  for (int oz = 0; oz < A::NZ; ++oz)
#pragma unroll
    for (int oy = 0; oy < A::NY; ++oy)
#pragma unroll
      for (int ox = 0; ox < A::NX; ++ox) {
        const int n = ox + A::NX * (oy + A::NY * oz);
        if ((fbits >> n) & 1ull) {
          const int r = base + ox + A::RX * (oy + A::RY * oz);
          atomicAdd(&s_mom[r], aq[n]);
          atomicAdd(&s_mass[r], am[n]);
          s_flag[r] = 1;
        }
      }
}

template <int AX>
__device__ inline void flush_axis(const P2GArgs& a, int bx, int by, int bz, const float* s_mom,
                                  const float* s_mass, const uint8_t* s_flag, int lane) {
  using A = Ax<AX>;
  constexpr int kPer = (A::RN + 63) / 64;  // region nodes per lane
  const Dims& fd = a.df[AX];
  float* __restrict__ mom = a.mom[AX];
  float* __restrict__ mass = a.mass[AX];
  // all of the lane's nodes are read before any is written (the nodes are distinct),
  // so the loads of one lane overlap instead of waiting on each other's stores
  long long gi[kPer];
  float om[kPer], oq[kPer];
#pragma unroll
  for (int e = 0; e < kPer; ++e) {
    const int r = lane + 64 * e;
    gi[e] = -1;
    if (r >= A::RN || !s_flag[r]) continue;
    const int rx = r % A::RX, ry = (r / A::RX) % A::RY, rz = r / (A::RX * A::RY);
    const int ni = bx * 4 - 1 + rx, nj = by * 4 - 1 + ry, nk = bz * 4 - 1 + rz;
    if (ni < 0 || nj < 0 || nk < 0 || ni >= fd.nx || nj >= fd.ny || nk >= fd.nz) continue;
    gi[e] = bm_index(fd, ni, nj, nk);
    oq[e] = mom[gi[e]];
    om[e] = mass[gi[e]];
  }
#pragma unroll
  for (int e = 0; e < kPer; ++e) {
    if (gi[e] < 0) continue;
    const int r = lane + 64 * e;
    mom[gi[e]] = oq[e] + s_mom[r];
    mass[gi[e]] = om[e] + s_mass[r];
    const long long blk = gi[e] >> 6;
    a.act_vel[AX][blk] = 1;
    a.act_mass[AX][blk] = 1;
  }
}

// KT = 0: normalized cubic (two-phase); KT = 1: trilinear, unit mass (single-phase);
// KT = 2: normalized cubic over the enumerated multires faces
template <int KT>
__global__ void __launch_bounds__(64) k_p2g_color(P2GArgs a) {
  __shared__ float s_mom[3][256];
  __shared__ float s_mass[3][256];
  __shared__ uint8_t s_flag[3][256];
  __shared__ uint8_t s_ord[64 * kOrdMax];
  __shared__ uint32_t s_ids[kOrdMax * 64];
  if (blockIdx.x >= *a.count) return;
  const uint32_t b = a.list[blockIdx.x];
  const int lane = threadIdx.x;
  for (int t = lane; t < 3 * 256; t += 64) {
    (&s_mom[0][0])[t] = 0.f;
    (&s_mass[0][0])[t] = 0.f;
    (&s_flag[0][0])[t] = 0;
  }
  int bx, by, bz;
  block_coords(a.dc, b, bx, by, bz);
  const int lx = lane & 3, ly = (lane >> 2) & 3, lz = lane >> 4;
  const int i = bx * 4 + lx, j = by * 4 + ly, k = bz * 4 + lz;
  const long long c = (long long)b * kBlockVol + lane;
  const uint32_t s0 = a.start[c];
  const uint32_t cnt = a.start[c + 1] - s0;
  uint8_t* ord = &s_ord[lane * kOrdMax];
  if (cnt > 1 && cnt <= (uint32_t)kOrdMax) {
    // rank of each particle by caller index (ids are unique), from the cell's ids
    // loaded at once into the lane's LDS column
#pragma unroll
    for (int t = 0; t < kOrdMax; ++t)
      if ((uint32_t)t < cnt) s_ids[t * 64 + lane] = a.P.id[s0 + t];
    for (uint32_t t = 0; t < cnt; ++t) {
      const uint32_t id = s_ids[t * 64 + lane];
      uint32_t r = 0;
      for (uint32_t u = 0; u < cnt; ++u) r += s_ids[u * 64 + lane] < id;
      ord[r] = (uint8_t)t;
    }
  } else if (cnt == 1) {
    ord[0] = 0;
  }
  __syncthreads();
  splat_axis<0, KT>(a, i, j, k, s0, cnt, ord, s_mom[0], s_mass[0], s_flag[0], lx, ly, lz);
  splat_axis<1, KT>(a, i, j, k, s0, cnt, ord, s_mom[1], s_mass[1], s_flag[1], lx, ly, lz);
  splat_axis<2, KT>(a, i, j, k, s0, cnt, ord, s_mom[2], s_mass[2], s_flag[2], lx, ly, lz);
  __syncthreads();
  flush_axis<0>(a, bx, by, bz, s_mom[0], s_mass[0], s_flag[0], lane);
  flush_axis<1>(a, bx, by, bz, s_mom[1], s_mass[1], s_flag[1], lane);
  flush_axis<2>(a, bx, by, bz, s_mom[2], s_mass[2], s_flag[2], lane);
}

}  // namespace pf
