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
//   after binning) in ascending caller index.  For a particle in cell c the non-zero
//   weights of a face field lie in c-1..c+2 along the face axis and c-1..c+1 across
//   it (the reference's 6-wide loop box adds only zero weights), so a lane keeps its
//   36 node sums per axis in registers and a block's contributions stay inside a
//   7x6x6 node region.  The three face axes are splatted in turn in one launch.
// * A lane's node sums go into the region in LDS one stencil offset at a time: for a
//   fixed offset the 64 lanes hit 64 distinct nodes, and one wave's LDS operations
//   execute in program order, so a plain read-add-write per offset sums every node in
//   a fixed order -- no atomics there, and the result does not change from run to run.
// * Regions of two source blocks of equal color8 (src/grid/block_color.h:7) are at
//   least 8 cells apart and never overlap, so within one of the 8 colour launches a
//   node takes at most one addition from a region: a no-return L2 float atomic adds it
//   into HBM in the same order as a read-modify-write would, without the wave waiting.
// * The activation predicate kern(d^2) > 0 is d^2 < 2.25 (exact: d^2/2.25 < 1 in
//   correctly rounded FP64 iff d^2 < 2.25).  It is evaluated in FP32 and recomputed
//   in FP64 inside a 1e-5 band around 2.25, so active-block sets match the FP64
//   reference bit for bit while weights run in FP32.
#include <type_traits>

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

// The launch's arguments: what all three face axes share (grid, block list, particle
// arrays and constants) and, per face axis, its field's pointers -- each particle pointer
// once, so the kernel's scalar registers hold one copy
struct P2GShared {
  Geom g;
  Dims dc;
  const uint32_t* start;
  const uint32_t* list;
  const uint32_t* count;
  const float *x, *y, *z, *vol;
  const uint8_t* type;
  const uint32_t* id;
  double inv_dx;
  float inv_dx_f;
  float mp_l, mp_g;
  bool has_vol, pow2, fast32;
};

struct P2GField {
  Dims df;
  float* mom;
  float* mass;
  uint8_t* act_vel;
  uint8_t* act_mass;
  const float* vel;
  const uint8_t* valid;
};

inline P2GShared p2g_shared_args(const P2GArgs& a) {
  P2GShared r;
  r.g = a.g;
  r.dc = a.dc;
  r.start = a.start;
  r.list = a.list;
  r.count = a.count;
  r.x = a.P.x; r.y = a.P.y; r.z = a.P.z;
  r.vol = a.P.vol;
  r.type = a.P.type;
  r.id = a.P.id;
  r.inv_dx = a.inv_dx;
  r.inv_dx_f = a.inv_dx_f;
  r.mp_l = a.mp_l; r.mp_g = a.mp_g;
  r.has_vol = a.has_vol; r.pow2 = a.pow2; r.fast32 = a.fast32;
  return r;
}

inline P2GField p2g_field_args(const P2GArgs& a, int ax) {
  P2GField r;
  r.df = a.df[ax];
  r.mom = a.mom[ax];
  r.mass = a.mass[ax];
  r.act_vel = a.act_vel[ax];
  r.act_mass = a.act_mass[ax];
  r.vel = ax == 0 ? a.P.u : (ax == 1 ? a.P.v : a.P.w);
  r.valid = a.valid[ax];
  return r;
}

constexpr int kOrdMax = kRankMax;  // particles per cell ranked in LDS
constexpr uint32_t kFastMax = 64;  // cells the fast loop takes (its marks fit a 64-bit mask)
constexpr float kGuard = 1e-5f;
constexpr float kGuardT = 1e-5f;  // on t = 1 - d^2/2.25: wider than kGuard / 2.25
constexpr float kOutside = 1e4f;  // d^2 penalty of a stencil row outside the field
// the smallest particle mass of the fast path: w * m_p / wsum >= 1e-15 * 1e-20 / 64
// stays a normal FP32 number
constexpr float kMinFastMass = 1e-20f;

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

__device__ inline double grid_coord(float x, double o, const P2GShared& a) {
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

// One particle's weights by the per-node path: the FP32 d^2 with the FP64 predicate and
// weight inside the guard band (exact_weight), the multires face mask, and the
// explicit activation bits.  dx2/dy2/dz2 carry the out-of-field penalties.
template <int AX>
__device__ inline float node_weights(const float* dx2, const float* dy2, const float* dz2, double gx, double gy,
                                     double gz, int i, int j, int k, uint64_t vbits, float* wt, uint64_t& fbits) {
  using A = Ax<AX>;
  float wsum = 0.f;
  // rolled: the rare exact path must not raise the kernel's register count
#pragma unroll 1
  for (int oz = 0; oz < A::NZ; ++oz)
#pragma unroll 1
    for (int oy = 0; oy < A::NY; ++oy)
#pragma unroll 1
      for (int ox = 0; ox < A::NX; ++ox) {
        const int n = ox + A::NX * (oy + A::NY * oz);
        float d2 = dx2[ox] + dy2[oy] + dz2[oz];
        float w;
        bool act;
        if (fabsf(d2 - 2.25f) < kGuard) {
          w = exact_weight(gx - (double)(i + ox - 1), gy - (double)(j + oy - 1), gz - (double)(k + oz - 1), act);
        } else {
          act = d2 < 2.25f;
          float t = 1.f - d2 * (1.f / 2.25f);
          w = act ? t * t * t : 0.f;
        }
        // (a node outside the field has d2 >= kOutside: inactive, never in the band)
        act = act && ((vbits >> n) & 1ull);
        w = act ? w : 0.f;
        wt[n] = w;
        wsum += w;
        if (act) fbits |= (1ull << n);
      }
  return wsum;
}

// A lane splats its cell's particles (in ascending caller index) into its 36 node sums
// of one face axis, then adds them into the block's LDS region.
//
// KT = 0 runs two loops.  The first covers cells of at most kFastMax particles in
// caller-index order, with the next particle's data loaded while this one
// computes, and evaluates the cubic in the separable FP32 form
// t = (1 - (dx^2 + dy^2)/2.25) - dz^2/2.25, w = max(t, 0)^3 -- one subtraction per
// node and no branch.  A particle with a node at |t| < kGuardT (its d^2 within about
// 1e-5 of 2.25), or with a mass too small for its positive weights to leave a positive
// sum, is only marked there.  The second loop takes the marked particles and every
// particle of a fuller cell (in memory order up to kSortedMax particles, else selected
// by ascending caller index) through node_weights.
// Keeping the rare paths out of the first loop keeps its registers to what it needs.
// KT = 2 (multires, enumerated faces only) and KT = 1 (trilinear) use the second loop
// for every particle.
//
// F32: origin 0, a power-of-two dx and fewer than 2^22 cells a side, where x/dx is
// exact in FP32 and each offset from the lane's cell is either exact or one correctly
// rounded FP32 subtraction of exact values (sx - 0.5 is inexact only for sx < 0.25, in
// cell 0): the FP64 path's single rounding, bit for bit.  Otherwise the offsets come
// from FP64 coordinates.
template <int AX, int KT, bool F32>
__device__ inline void splat_axis(const P2GShared& a, const P2GField& f, int i, int j, int k, uint32_t s0,
                                  uint32_t cnt, const uint8_t* ord, float2* s_node, uint8_t* s_flag,
                                  int lx, int ly, int lz) {
  using A = Ax<AX>;
  const Dims& fd = f.df;
  float am[A::NN], aq[A::NN];
  uint64_t fbits = 0;  // activation bits of the second loop
#pragma unroll
  for (int n = 0; n < A::NN; ++n) { am[n] = 0.f; aq[n] = 0.f; }

  // node validity along each axis (field extents), fixed per lane, and the same as
  // additive d^2 penalties held in vector registers: a row or column outside the field
  // gets d^2 >= kOutside, so its t is far below zero and its weight zero without a
  // select (lane masks would sit in scalar registers across the node loops)
  bool vx[A::NX], vy[A::NY], vz[A::NZ];
#pragma unroll
  for (int o = 0; o < A::NX; ++o) { int c = i + o - 1; vx[o] = c >= 0 && c < fd.nx; }
#pragma unroll
  for (int o = 0; o < A::NY; ++o) { int c = j + o - 1; vy[o] = c >= 0 && c < fd.ny; }
#pragma unroll
  for (int o = 0; o < A::NZ; ++o) { int c = k + o - 1; vz[o] = c >= 0 && c < fd.nz; }
  float penx[A::NX], peny[A::NY], penz[A::NZ];
#pragma unroll
  for (int o = 0; o < A::NX; ++o) penx[o] = vx[o] ? 0.f : kOutside;
#pragma unroll
  for (int o = 0; o < A::NY; ++o) peny[o] = vy[o] ? 0.f : kOutside;
#pragma unroll
  for (int o = 0; o < A::NZ; ++o) penz[o] = vz[o] ? 0.f : kOutside;
  const float* vp = f.vel;

  // offsets of a particle from the lane's cell along this axis' face grid
  // (sparse_ops3d_tp.cpp:117-122), and the FP64 stencil coordinates when asked
  auto offsets = [&](float px, float py, float pz, float& fx, float& fy, float& fz) {
    if (F32) {
      const float sx = px * a.inv_dx_f, sy = py * a.inv_dx_f, sz = pz * a.inv_dx_f;
      fx = (AX == 0 ? sx : sx - 0.5f) - (float)i;
      fy = (AX == 1 ? sy : sy - 0.5f) - (float)j;
      fz = (AX == 2 ? sz : sz - 0.5f) - (float)k;
    } else {
      const double dpx = grid_coord(px, a.g.ox, a), dpy = grid_coord(py, a.g.oy, a), dpz = grid_coord(pz, a.g.oz, a);
      fx = (float)((AX == 0 ? dpx : dpx - 0.5) - (double)i);
      fy = (float)((AX == 1 ? dpy : dpy - 0.5) - (double)j);
      fz = (float)((AX == 2 ? dpz : dpz - 0.5) - (double)k);
    }
  };
  auto coords64 = [&](float px, float py, float pz, double& gx, double& gy, double& gz) {
    double dpx, dpy, dpz;
    if (F32) {
      dpx = (double)(px * a.inv_dx_f); dpy = (double)(py * a.inv_dx_f); dpz = (double)(pz * a.inv_dx_f);
    } else {
      dpx = grid_coord(px, a.g.ox, a); dpy = grid_coord(py, a.g.oy, a); dpz = grid_coord(pz, a.g.oz, a);
    }
    gx = AX == 0 ? dpx : dpx - 0.5;
    gy = AX == 1 ? dpy : dpy - 0.5;
    gz = AX == 2 ? dpz : dpz - 0.5;
  };
  auto particle_mass = [&](uint8_t ty, float pvol) {
    float mp = (ty == 0) ? a.mp_l : a.mp_g;
    if (a.has_vol) mp *= pvol;
    return mp;
  };
  auto accumulate = [&](const float* wt, float wsum, float mp, float vel) {
    const float sm = wsum > 0.f ? mp / wsum : 0.f;
    const float sq = sm * vel;
#pragma unroll
    for (int n = 0; n < A::NN; ++n) {
      am[n] = fmaf(wt[n], sm, am[n]);
      aq[n] = fmaf(wt[n], sq, aq[n]);
    }
  };

  // the order of a cell's particles: the LDS ranking (at most kOrdMax), memory order
  // (cells of up to kSortedMax particles were sorted by id after binning), or else
  // selection of the next caller index
  const bool ranked = cnt <= (uint32_t)kOrdMax;
  const bool sorted = !ranked && cnt <= (uint32_t)kSortedMax;
  const bool masked = cnt <= kFastMax;  // ranked or sorted
  auto slot = [&](uint32_t r) -> uint32_t { return ranked ? (uint32_t)ord[r] : r; };
  uint64_t marked = 0;  // ranks (bit it) left to the second loop
  if (KT == 0 && masked && cnt > 0) {
    // the particle of rank r (its data for the next iteration load while this one
    // computes)
    auto fetch = [&](uint32_t r, float& px, float& py, float& pz, float& vel, uint8_t& ty, float& pvol) {
      const uint32_t q = s0 + slot(r);
      px = a.x[q]; py = a.y[q]; pz = a.z[q]; vel = vp[q]; ty = a.type[q];
      pvol = a.has_vol ? a.vol[q] : 1.f;
    };
    float nx_, ny_, nz_, nv_, nvol_;
    uint8_t nty_;
    fetch(0, nx_, ny_, nz_, nv_, nty_, nvol_);
    for (uint32_t it = 0; it < cnt; ++it) {
      const float px = nx_, py = ny_, pz = nz_, vel = nv_, pvol = nvol_;
      const uint8_t ty = nty_;
      if (it + 1 < cnt) fetch(it + 1, nx_, ny_, nz_, nv_, nty_, nvol_);
      float fx, fy, fz;
      offsets(px, py, pz, fx, fy, fz);
      float dx2[A::NX], dy2[A::NY], dz2[A::NZ];
#pragma unroll
      for (int o = 0; o < A::NX; ++o) { float d = fx - (float)(o - 1); dx2[o] = fmaf(d, d, penx[o]); }
#pragma unroll
      for (int o = 0; o < A::NY; ++o) { float d = fy - (float)(o - 1); dy2[o] = fmaf(d, d, peny[o]); }
#pragma unroll
      for (int o = 0; o < A::NZ; ++o) { float d = fz - (float)(o - 1); dz2[o] = fmaf(d, d, penz[o]); }
      const float mp = particle_mass(ty, pvol);
      float Axy[A::NX * A::NY], Bz[A::NZ];
#pragma unroll
      for (int oy = 0; oy < A::NY; ++oy)
#pragma unroll
        for (int ox = 0; ox < A::NX; ++ox) Axy[ox + A::NX * oy] = 1.f - (dx2[ox] + dy2[oy]) * (1.f / 2.25f);
#pragma unroll
      for (int oz = 0; oz < A::NZ; ++oz) Bz[oz] = dz2[oz] * (1.f / 2.25f);
      float wt[A::NN];
      float wsum = 0.f, tmin = 4.f;
#pragma unroll
      for (int oz = 0; oz < A::NZ; ++oz)
#pragma unroll
        for (int oy = 0; oy < A::NY; ++oy)
#pragma unroll
          for (int ox = 0; ox < A::NX; ++ox) {
            const int n = ox + A::NX * (oy + A::NY * oz);
            const float t = Axy[ox + A::NX * oy] - Bz[oz];
            tmin = fminf(tmin, fabsf(t));
            // max(t, 0) as one median (t <= 1 here); fmaxf would first quiet a NaN t
            const float tp = __builtin_amdgcn_fmed3f(t, 0.f, 1.f);
            const float w = tp * tp * tp;
            wt[n] = w;
            wsum += w;
          }
      // (a node of a particle accumulated here is activated by its positive weight:
      // w = t^3 >= 1e-15 where t > 0 and m_p >= kMinFastMass, so the lane's mass sum
      // there is positive -- these activation bits are read from the sums below)
      if (tmin < kGuardT || !(mp >= kMinFastMass)) {
        marked |= 1ull << it;
        continue;
      }
      accumulate(wt, wsum, mp, vel);
    }
  } else if (masked && cnt > 0) {
    marked = (cnt >= 64u) ? ~0ull : ((1ull << cnt) - 1ull);
  }

  // the second loop: marked particles in rank order, then every particle of a cell of
  // more than kFastMax particles, in memory order or by selection
  uint64_t vbits = ~0ull;
  if (KT == 2) {
    // multires: only the enumerated faces take (and normalise) weights
    // (collectWeightedFaces, multires_ops3d_tp.cpp:62-75)
    vbits = 0;
#pragma unroll
    for (int oz = 0; oz < A::NZ; ++oz)
#pragma unroll
      for (int oy = 0; oy < A::NY; ++oy)
#pragma unroll
        for (int ox = 0; ox < A::NX; ++ox)
          if (vx[ox] && vy[oy] && vz[oz] && f.valid[bm_index(fd, i + ox - 1, j + oy - 1, k + oz - 1)])
            vbits |= 1ull << (ox + A::NX * (oy + A::NY * oz));
  }
  const uint32_t nsecond = masked ? (uint32_t)__popcll(marked) : cnt;
  uint32_t last = 0;
  bool first = true;
  for (uint32_t s2 = 0; s2 < nsecond; ++s2) {
    uint32_t q;
    if (masked) {
      const int it = __ffsll((unsigned long long)marked) - 1;
      marked &= marked - 1ull;
      q = s0 + slot((uint32_t)it);
    } else if (sorted) {
      q = s0 + s2;
    } else {
      uint32_t best = 0xffffffffu, bi = 0;
      for (uint32_t t = 0; t < cnt; ++t) {
        const uint32_t id = a.id[s0 + t];
        if ((first || id > last) && id <= best) { best = id; bi = t; }
      }
      last = best;
      first = false;
      q = s0 + bi;
    }
    const float px = a.x[q], py = a.y[q], pz = a.z[q], vel = vp[q];
    const float pvol = a.has_vol ? a.vol[q] : 1.f;
    const uint8_t ty = a.type[q];
    double gx, gy, gz;
    coords64(px, py, pz, gx, gy, gz);
    float wt[A::NN];
    if (KT == 1) {
      trilinear_weights<AX>(gx, gy, gz, i, j, k, vx, vy, vz, wt, fbits);
      // mass += (float)(w * 1.0) = w; momentum += (float)(w * v) with v exact in FP64
#pragma unroll
      for (int n = 0; n < A::NN; ++n) {
        am[n] += wt[n];
        aq[n] += (float)((double)wt[n] * (double)vel);
      }
      continue;
    }
    float fx, fy, fz;
    offsets(px, py, pz, fx, fy, fz);
    float dx2[A::NX], dy2[A::NY], dz2[A::NZ];
#pragma unroll
    for (int o = 0; o < A::NX; ++o) { float d = fx - (float)(o - 1); dx2[o] = fmaf(d, d, penx[o]); }
#pragma unroll
    for (int o = 0; o < A::NY; ++o) { float d = fy - (float)(o - 1); dy2[o] = fmaf(d, d, peny[o]); }
#pragma unroll
    for (int o = 0; o < A::NZ; ++o) { float d = fz - (float)(o - 1); dz2[o] = fmaf(d, d, penz[o]); }
    const float wsum = node_weights<AX>(dx2, dy2, dz2, gx, gy, gz, i, j, k, vbits, wt, fbits);
    accumulate(wt, wsum, particle_mass(ty, pvol), vel);
  }
  // deposit into the LDS region, one stencil offset n at a time: for fixed n the 64
  // lanes address 64 distinct nodes, and one wave's LDS operations execute in program
  // order, so a plain read-add-write per n adds a node's lane sums in n order with no
  // LDS atomics (float atomics in LDS were half of the kernel's time).  The empty asm
  // with a memory clobber after each n keeps the compiler from moving offset n+1's
  // read above offset n's write (to one lane they are different addresses; to the
  // wave they can be the same node).  A node's (momentum, mass) pair is one float2, so
  // each offset is one ds_read_b64 and one ds_write_b64 off the lane's base address.
  // (A volatile pointer kept the order too, but lost the LDS address space: flat loads
  // and stores with a wait on each.)
  // A node takes the lane's sums where they hold mass (every positive weight of the
  // first loop), or where the second loop set its activation bit (which it does for
  // every positive weight, and for the zero-weight nodes that still activate); a wave
  // with no second-loop bits skips that test.
  const int base = lx + A::RX * (ly + A::RY * lz);
  auto deposit = [&](auto with_bits) {
#pragma unroll
    for (int oz = 0; oz < A::NZ; ++oz)
#pragma unroll
      for (int oy = 0; oy < A::NY; ++oy)
#pragma unroll
        for (int ox = 0; ox < A::NX; ++ox) {
          const int n = ox + A::NX * (oy + A::NY * oz);
          const bool put = am[n] > 0.f || (decltype(with_bits)::value && ((fbits >> n) & 1ull));
          if (put) {
            const int r = base + ox + A::RX * (oy + A::RY * oz);
            float2 v = s_node[r];
            v.x = v.x + aq[n];
            v.y = v.y + am[n];
            s_node[r] = v;
            s_flag[r] = 1;
          }
          __asm__ volatile("" ::: "memory");
        }
  };
  if (__any(fbits != 0ull))
    deposit(std::true_type{});
  else
    deposit(std::false_type{});
}

// Adds the block's region sums into the face field and clears the region for the next
// axis.  Regions of same-colour blocks never overlap, so each node takes at most one
// addition per launch: a no-return L2 float atomic gives the same sum as a read-modify-
// write, in the same order (colour by colour), without the wave waiting for the read.
template <int AX>
__device__ inline void flush_axis(const P2GField& f, int bx, int by, int bz, float2* s_node,
                                  uint8_t* s_flag, int lane) {
  using A = Ax<AX>;
  constexpr int kPer = (A::RN + 63) / 64;  // region nodes per lane
  const Dims& fd = f.df;
#pragma unroll
  for (int e = 0; e < kPer; ++e) {
    const int r = lane + 64 * e;
    if (r >= A::RN) continue;
    if (s_flag[r]) {
      const int rx = r % A::RX, ry = (r / A::RX) % A::RY, rz = r / (A::RX * A::RY);
      const int ni = bx * 4 - 1 + rx, nj = by * 4 - 1 + ry, nk = bz * 4 - 1 + rz;
      if (ni >= 0 && nj >= 0 && nk >= 0 && ni < fd.nx && nj < fd.ny && nk < fd.nz) {
        const long long gi = bm_index(fd, ni, nj, nk);
        const float2 v = s_node[r];
        unsafeAtomicAdd(&f.mom[gi], v.x);
        unsafeAtomicAdd(&f.mass[gi], v.y);
        const long long blk = gi >> 6;
        f.act_vel[blk] = 1;
        f.act_mass[blk] = 1;
      }
    }
    s_node[r] = make_float2(0.f, 0.f);
    s_flag[r] = 0;
  }
}

// KT = 0: normalized cubic (two-phase); KT = 1: trilinear, unit mass (single-phase);
// KT = 2: normalized cubic over the enumerated multires faces.  One launch per colour:
// a block's particles are ranked once and splatted into the three face fields in turn
// through one LDS node region (12.3 KB of LDS per workgroup with the id ranking).
template <int KT, bool F32>
__global__ void __launch_bounds__(64) k_p2g_color(P2GShared a, P2GField fu, P2GField fv, P2GField fw) {
  __shared__ uint32_t s_ids[kOrdMax * 64];
  __shared__ uint8_t s_ord[64 * kOrdMax];
  __shared__ float2 s_node[256];  // the region's (momentum, mass) sums per node
  __shared__ uint8_t s_flag[256];
  if (blockIdx.x >= *a.count) return;
  const uint32_t b = a.list[blockIdx.x];
  const int lane = threadIdx.x;
  for (int t = lane; t < 256; t += 64) {
    s_node[t] = make_float2(0.f, 0.f);
    s_flag[t] = 0;
  }
  int bx, by, bz;
  block_coords(a.dc, b, bx, by, bz);
  const int lx = lane & 3, ly = (lane >> 2) & 3, lz = lane >> 4;
  const int i = bx * 4 + lx, j = by * 4 + ly, k = bz * 4 + lz;
  const long long c = (long long)b * kBlockVol + lane;
  const uint32_t s0 = a.start[c];
  const uint32_t cnt = a.start[c + 1] - s0;
  const bool rank = cnt > 1 && cnt <= (uint32_t)kOrdMax;
  // the cell's ids at once into the lane's LDS column
#pragma unroll
  for (int t = 0; t < kOrdMax; ++t)
    if (rank && (uint32_t)t < cnt) s_ids[t * 64 + lane] = a.id[s0 + t];
  uint8_t* ord = &s_ord[lane * kOrdMax];
  if (rank) {
    // rank of each particle by caller index (ids are unique)
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
  splat_axis<0, KT, F32>(a, fu, i, j, k, s0, cnt, ord, s_node, s_flag, lx, ly, lz);
  __syncthreads();
  flush_axis<0>(fu, bx, by, bz, s_node, s_flag, lane);
  __syncthreads();
  splat_axis<1, KT, F32>(a, fv, i, j, k, s0, cnt, ord, s_node, s_flag, lx, ly, lz);
  __syncthreads();
  flush_axis<1>(fv, bx, by, bz, s_node, s_flag, lane);
  __syncthreads();
  splat_axis<2, KT, F32>(a, fw, i, j, k, s0, cnt, ord, s_node, s_flag, lx, ly, lz);
  __syncthreads();
  flush_axis<2>(fw, bx, by, bz, s_node, s_flag, lane);
}

}  // namespace pf
