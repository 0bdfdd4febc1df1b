// pf_project.hip -- variable-coefficient pressure projection (Eq. 8).
//
// Reference: spProjectStepVC3D, src/driver/sparse_ops3d_tp.cpp:234-423.  Rows are the
// FLUID cells; the off-diagonal of a face is beta = 1/rho(phi(raw face mass))
// (Eq. 7, src/physics/phasefield.h:10-18) when both cells are FLUID and neither is
// the pinned cell, the diagonal sums beta over non-SOLID faces (AIR = Dirichlet 0),
// the RHS is -div(u)/dx + c_div with faces next to SOLID cells taken as 0, and the
// first FLUID cell (lowest linear id) is pinned when no FLUID cell touches AIR.
// Preconditioner: Jacobi (diagonal), as the reference's (:328-333); it has no
// ordering, so any sweep order reproduces it.
//
// MI355X form: the operator is matrix-free over block-major cells of the sorted
// list of blocks that hold FLUID cells; each cell stores its diagonal and the
// coefficient of its +x/+y/+z faces (the -x coefficient is the -x neighbour's +x
// one), in FP32 for the inner PCG and FP64 for the true residual.  The inner PCG
// runs FP32 stencils and vector updates, FP64 dot products, FP64 alpha/beta kept in
// device memory; an FP64 iterative-refinement loop recomputes b - A x in FP64 and
// restarts the inner solve on it until the stop rule holds on the FP64 residual.
#include "pf_internal.h"

namespace pf {

struct PhaseK {
  double rho_l, rho_g, rmin, den;  // rmin = ln(rho_l/rho_g)*rho_g*rho_tilde_0, den = alpha*rho_tilde_0*rho_l
};

// phiFromRawDensity, src/physics/phasefield.h:10-14
__device__ inline double phi_of_raw(double raw, const PhaseK& k) {
#pragma clang fp contract(off)
  if (raw < k.rmin) return 0.0;
  double d = raw - k.rmin;
  if (d < 0.0) d = 0.0;
  double v = sqrt(d / k.den);
  return 1.0 < v ? 1.0 : v;  // std::min(v, 1.0): a NaN mass gives a NaN phase
}

// betaOfRaw, sparse_ops3d_tp.cpp:89-91 with phiFromRawDensity / betaFromPhi
__device__ inline double beta_of_raw(double raw, const PhaseK& k) {
#pragma clang fp contract(off)
  const double phi = phi_of_raw(raw, k);
  return 1.0 / (phi * k.rho_l + (1.0 - phi) * k.rho_g);
}

// Face phase field phi(raw face mass) in FP64, dense linear order of the face field
// (the phi that the assembly's beta uses, phasefield.h:10-14)
__global__ void k_face_phi(Dims d, const float* __restrict__ mass, PhaseK ph, double* __restrict__ out) {
  const long long n = (long long)d.nx * d.ny * d.nz;
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int i = (int)(q % d.nx);
  const long long t = q / d.nx;
  const int j = (int)(t % d.ny), k = (int)(t / d.ny);
  out[q] = phi_of_raw((double)mass[bm_index(d, i, j, k)], ph);
}

struct CellNbr {
  long long xp, xm, yp, ym, zp, zm;
};

// block-major neighbour indices of cell c = b*64 + lane; a neighbour outside the
// block grid maps to the cell itself (its coefficient is zero: such cells are
// never FLUID), so every index stays inside the arrays
__device__ inline CellNbr neighbours(const Dims& dc, long long b, int lane) {
  const int lx = lane & 3, ly = (lane >> 2) & 3, lz = lane >> 4;
  int bx, by, bz;
  block_coords(dc, b, bx, by, bz);  // wave-uniform
  const long long c = b * kBlockVol + lane;
  const long long sy = (long long)dc.nbx * kBlockVol, sz = (long long)dc.nbx * dc.nby * kBlockVol;
  CellNbr n;
  n.xp = lx < 3 ? c + 1 : (bx + 1 < dc.nbx ? c + kBlockVol - 3 : c);
  n.xm = lx > 0 ? c - 1 : (bx > 0 ? c - kBlockVol + 3 : c);
  n.yp = ly < 3 ? c + 4 : (by + 1 < dc.nby ? c + sy - 12 : c);
  n.ym = ly > 0 ? c - 4 : (by > 0 ? c - sy + 12 : c);
  n.zp = lz < 3 ? c + 16 : (bz + 1 < dc.nbz ? c + sz - 48 : c);
  n.zm = lz > 0 ? c - 16 : (bz > 0 ? c - sz + 48 : c);
  return n;
}

__device__ inline uint8_t marker_at(const uint8_t* mk, const Dims& dc, const Geom& g, int i, int j, int k) {
  if (i < 0 || j < 0 || k < 0 || i >= g.nx || j >= g.ny || k >= g.nz) return CELL_SOLID;
  return mk[bm_index(dc, i, j, k)];
}

// ---------------------------------------------------------------- reductions

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// sum (or max) of NV values per thread over a 256-thread workgroup -> part[slot*stride + blockIdx]
template <int NV>
__device__ inline void block_reduce_store(double (&v)[NV], const bool (&is_max)[NV], double* part,
                                          long long stride) {
  __shared__ double s[4][NV];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NV; ++q) v[q] = is_max[q] ? wave_max(v[q]) : wave_sum(v[q]);
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < NV; ++q) s[wave][q] = v[q];
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      double r = s[0][q];
      for (int w = 1; w < 4; ++w) r = is_max[q] ? fmax(r, s[w][q]) : r + s[w][q];
      part[q * stride + blockIdx.x] = r;
    }
  }
}

// fixed-order sum of part[0..n) by one 256-thread workgroup: strided per-thread
// sums, then a wave butterfly and a 4-way combine (n <= 1024 in the solver loop)
__device__ inline double block_sum_fast(const double* part, int n) {
  __shared__ double s_w[4];
  double acc = 0.0;
  for (int t = threadIdx.x; t < n; t += 256) acc += part[t];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = acc;
  __syncthreads();
  const double r = (s_w[0] + s_w[1]) + (s_w[2] + s_w[3]);
  __syncthreads();
  return r;
}

// one workgroup sums part[slot*stride + 0..n) in a fixed order
__device__ inline double block_sum_partials(const double* part, int n, bool is_max) {
  __shared__ double s[256];
  double acc = is_max ? 0.0 : 0.0;
  for (int t = threadIdx.x; t < n; t += 256) acc = is_max ? fmax(acc, part[t]) : acc + part[t];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) s[threadIdx.x] = is_max ? fmax(s[threadIdx.x], s[threadIdx.x + w]) : s[threadIdx.x] + s[threadIdx.x + w];
    __syncthreads();
  }
  double r = s[0];
  __syncthreads();
  return r;
}

// red[q] = reduction of slot q (NV slots)
__global__ void __launch_bounds__(256) k_reduce_slots(const double* part, int n, int nslots,
                                                      unsigned max_mask, double* red) {
  for (int q = 0; q < nslots; ++q) {
    double r = block_sum_partials(part + (long long)q * kMaxPartials, n, (max_mask >> q) & 1u);
    if (threadIdx.x == 0) red[q] = r;
  }
}

// ------------------------------------------------------------ set-up kernels

// a value every lane of the wave holds, moved to a scalar register (the list
// position and block id below depend only on the wave index)
__device__ inline uint32_t wave_uniform_u32(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }

// list[t] for a wave-uniform t through the constant address space (the lists are not
// written while a kernel walks them): a scalar load, which leaves the vector-memory
// counter -- and the previous block's stores -- alone
__device__ inline uint32_t list_at(const uint32_t* list, uint32_t t) {
  typedef const __attribute__((address_space(4))) uint32_t* cptr;
  return ((cptr)(uintptr_t)list)[t];
}

// Iterate the block list with a fixed block -> workgroup assignment (so partial
// sums are deterministic) that is XCD-aware: workgroups are dispatched round-robin
// over the 8 XCDs, so workgroup w runs on XCD w % 8; each XCD gets one contiguous
// eighth of the (sorted) list, i.e. a slab of the domain whose stencil neighbours
// sit in that XCD's L2.  Requires gridDim.x to be a multiple of 8.
#define FOR_LIST_BLOCKS(A, b, lane)                                                                   \
  for (long long _c = ((A).nlist + 7) / 8, _x = blockIdx.x & 7, _g = gridDim.x >> 3,                  \
                 _e = min((long long)(A).nlist, (_x + 1) * _c),                                       \
                 _t = _x * _c + (long long)(blockIdx.x >> 3) * 4 + (threadIdx.x >> 6);                \
       _t < _e; _t += _g * 4)                                                                         \
    for (long long b = list_at((A).list, wave_uniform_u32((uint32_t)_t)), _once = 1; _once; _once = 0)

struct ProjArgs {
  Geom g;
  Dims dc, du, dv, dw;
  const uint8_t* marker;
  const uint32_t* list;
  long long nlist;
  PhaseK ph;
  double scale;   // dt / dx^2
  double cdiv;    // divergence correction
  uint32_t* pin;  // [0] touches-air flag, [1] min FLUID linear id, [4] FLUID cell count
  int single;     // single-phase spProjectStep3D (sparse_ops3d.cpp:170-232): unit
                  // coefficients, no pin, divergence without solid-face masking
};

__device__ inline bool list_block(const ProjArgs& a, long long t, long long& b) {
  if (t >= a.nlist) return false;
  b = a.list[t];
  return true;
}

// Pin detection, sparse_ops3d_tp.cpp:250-264: does any FLUID cell have an AIR
// 6-neighbour, the smallest FLUID linear id, and the FLUID cell count.  Grid-stride
// over the list (XCD-aware), combined per workgroup so each workgroup issues one
// atomic per quantity.
__global__ void __launch_bounds__(256) k_pin_detect(ProjArgs a) {
  __shared__ uint32_t s_air[4], s_min[4], s_cnt[4];
  const int lane = threadIdx.x & 63;
  bool anyair = false;
  uint32_t m = 0xffffffffu, nfl = 0;
  FOR_LIST_BLOCKS(a, b, lane) {
    int bx, by, bz;
    block_coords(a.dc, b, bx, by, bz);
    int i = bx * 4 + (lane & 3), j = by * 4 + ((lane >> 2) & 3), k = bz * 4 + (lane >> 4);
    bool fl = a.marker[b * kBlockVol + lane] == CELL_FLUID;
    bool air = false;
    uint32_t lin = 0xffffffffu;
    if (fl) {
      air = marker_at(a.marker, a.dc, a.g, i + 1, j, k) == CELL_AIR ||
            marker_at(a.marker, a.dc, a.g, i - 1, j, k) == CELL_AIR ||
            marker_at(a.marker, a.dc, a.g, i, j + 1, k) == CELL_AIR ||
            marker_at(a.marker, a.dc, a.g, i, j - 1, k) == CELL_AIR ||
            marker_at(a.marker, a.dc, a.g, i, j, k + 1) == CELL_AIR ||
            marker_at(a.marker, a.dc, a.g, i, j, k - 1) == CELL_AIR;
      lin = (uint32_t)(i + (long long)a.g.nx * (j + (long long)a.g.ny * k));
    }
    anyair = anyair || a.single || __any(air);  // the single-phase solve never pins
    nfl += (uint32_t)__popcll(__ballot(fl));
    m = min(m, lin);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = min(m, (uint32_t)__shfl_xor((int)m, o, 64));
  const int w = threadIdx.x >> 6;
  if (lane == 0) {
    s_air[w] = anyair ? 1u : 0u;
    s_min[w] = m;
    s_cnt[w] = nfl;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t air = s_air[0] | s_air[1] | s_air[2] | s_air[3];
    const uint32_t mn = min(min(s_min[0], s_min[1]), min(s_min[2], s_min[3]));
    const uint32_t cnt = s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
    if (air) atomicOr(&a.pin[0], 1u);
    if (mn != 0xffffffffu) atomicMin(&a.pin[1], mn);
    if (cnt) atomicAdd(&a.pin[4], cnt);
  }
}

struct AsmOut {
  double *b64, *cd64, *cx64, *cy64, *cz64;
  float *cd32, *cx32, *cy32, *cz32, *invd32;
  const float *u, *v, *w, *mu, *mv, *mw;
  double* part;   // slot 0: sum b^2, slot 1: max |b| (the FP64 residual of x = 0)
  float *fbeta, *fphi;  // optional learned-preconditioner features (mean face beta / phi)
  uint32_t* mx;   // optional: max of the FP32 cd, cx, cy, cz and max of inv (float bits,
                  // atomicMax; the values are >= 0), what the multigrid build scales by
};

// RHS and assembly, sparse_ops3d_tp.cpp:266-313
__global__ void __launch_bounds__(256) k_assemble(ProjArgs a, AsmOut o) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & 63;
  double bb = 0.0, bmax = 0.0;
  float amax = 0.f, imax = 0.f;
  FOR_LIST_BLOCKS(a, b, lane) {
    const long long c = b * kBlockVol + lane;
    int bx, by, bz;
    block_coords(a.dc, b, bx, by, bz);
    int i = bx * 4 + (lane & 3), j = by * 4 + ((lane >> 2) & 3), k = bz * 4 + (lane >> 4);
    bool fl = a.marker[c] == CELL_FLUID;
    double rhs = 0.0, cd = 0.0, cx = 0.0, cy = 0.0, cz = 0.0, inv = 0.0;
    if (fl) {
      const bool touches = a.pin[0] != 0;
      const uint32_t pc = touches ? 0xffffffffu : a.pin[1];
      const uint32_t lin = (uint32_t)(i + (long long)a.g.nx * (j + (long long)a.g.ny * k));
      const bool pinned = lin == pc;
      const int nb[6][3] = {{i + 1, j, k}, {i - 1, j, k}, {i, j + 1, k}, {i, j - 1, k}, {i, j, k + 1}, {i, j, k - 1}};
      uint8_t m[6];
#pragma unroll
      for (int q = 0; q < 6; ++q) m[q] = marker_at(a.marker, a.dc, a.g, nb[q][0], nb[q][1], nb[q][2]);
      double fb[6] = {1.0, 1.0, 1.0, 1.0, 1.0, 1.0};
      if (!a.single) {
        fb[0] = beta_of_raw(o.mu[bm_index(a.du, i + 1, j, k)], a.ph);
        fb[1] = beta_of_raw(o.mu[bm_index(a.du, i, j, k)], a.ph);
        fb[2] = beta_of_raw(o.mv[bm_index(a.dv, i, j + 1, k)], a.ph);
        fb[3] = beta_of_raw(o.mv[bm_index(a.dv, i, j, k)], a.ph);
        fb[4] = beta_of_raw(o.mw[bm_index(a.dw, i, j, k + 1)], a.ph);
        fb[5] = beta_of_raw(o.mw[bm_index(a.dw, i, j, k)], a.ph);
      }
      if (a.single) {
        // sparse_ops3d.cpp:180-187: the six face values summed left to right
        const double d = (double)o.u[bm_index(a.du, i + 1, j, k)] - (double)o.u[bm_index(a.du, i, j, k)] +
                         (double)o.v[bm_index(a.dv, i, j + 1, k)] - (double)o.v[bm_index(a.dv, i, j, k)] +
                         (double)o.w[bm_index(a.dw, i, j, k + 1)] - (double)o.w[bm_index(a.dw, i, j, k)];
        rhs = -(d / a.g.dx);
      } else if (!pinned) {
        double uR = m[0] == CELL_SOLID ? 0.0 : (double)o.u[bm_index(a.du, i + 1, j, k)];
        double uL = m[1] == CELL_SOLID ? 0.0 : (double)o.u[bm_index(a.du, i, j, k)];
        double vT = m[2] == CELL_SOLID ? 0.0 : (double)o.v[bm_index(a.dv, i, j + 1, k)];
        double vB = m[3] == CELL_SOLID ? 0.0 : (double)o.v[bm_index(a.dv, i, j, k)];
        double wU = m[4] == CELL_SOLID ? 0.0 : (double)o.w[bm_index(a.dw, i, j, k + 1)];
        double wD = m[5] == CELL_SOLID ? 0.0 : (double)o.w[bm_index(a.dw, i, j, k)];
        rhs = -((uR - uL) + (vT - vB) + (wU - wD)) / a.g.dx + a.cdiv;
      }
      double diag = 0.0;
#pragma unroll
      for (int q = 0; q < 6; ++q)
        if (m[q] != CELL_SOLID) diag += fb[q];
      // +x/+y/+z couplings: both cells FLUID and neither pinned
      auto lin_of = [&](int q) {
        return (uint32_t)(nb[q][0] + (long long)a.g.nx * (nb[q][1] + (long long)a.g.ny * nb[q][2]));
      };
      // conductance form: a face is coupled when both cells are FLUID and neither
      // is pinned; the other non-SOLID faces (AIR, or the pinned cell's column the
      // reference drops) are Dirichlet terms on the diagonal.  Rows are then
      // cdir*x_c + sum_f c_f (x_c - x_f), so rounding c_f and cdir to FP32 is a
      // relative perturbation of each conductance and keeps the near-null space.
      double dir = 0.0;
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        if (m[q] == CELL_SOLID) continue;
        const bool coupled = !pinned && m[q] == CELL_FLUID && lin_of(q) != pc;
        if (!coupled) dir += fb[q];
        if (coupled && q == 0) cx = a.scale * fb[0];
        if (coupled && q == 2) cy = a.scale * fb[2];
        if (coupled && q == 4) cz = a.scale * fb[4];
      }
      cd = pinned ? 1.0 : a.scale * dir;
      const double full = a.scale * diag;
      inv = pinned ? 1.0 : (full > 0.0 ? 1.0 / full : 0.0);
    }
    o.b64[c] = rhs;
    o.cd64[c] = cd; o.cx64[c] = cx; o.cy64[c] = cy; o.cz64[c] = cz;
    const float cd32 = (float)cd, cx32 = (float)cx, cy32 = (float)cy, cz32 = (float)cz, inv32 = (float)inv;
    o.cd32[c] = cd32; o.cx32[c] = cx32; o.cy32[c] = cy32; o.cz32[c] = cz32;
    o.invd32[c] = inv32;
    amax = fmaxf(amax, fmaxf(fmaxf(cd32, cx32), fmaxf(cy32, cz32)));
    imax = fmaxf(imax, inv32);
    if (o.fbeta) {
      // cell features of the learned preconditioner: mean beta and mean phi over the
      // six faces (the cell phi of interface_diagnostics3d.cpp:137-148)
      float fb_mean = 0.f, fp_mean = 0.f;
      if (fl) {
        if (a.single) {
          fb_mean = 1.f;
          fp_mean = 1.f;
        } else {
          const double raw[6] = {o.mu[bm_index(a.du, i + 1, j, k)], o.mu[bm_index(a.du, i, j, k)],
                                 o.mv[bm_index(a.dv, i, j + 1, k)], o.mv[bm_index(a.dv, i, j, k)],
                                 o.mw[bm_index(a.dw, i, j, k + 1)], o.mw[bm_index(a.dw, i, j, k)]};
          double sb = 0.0, sp = 0.0;
          for (int q = 0; q < 6; ++q) {
            sp += phi_of_raw(raw[q], a.ph);
            sb += beta_of_raw(raw[q], a.ph);
          }
          fb_mean = (float)(sb / 6.0);
          fp_mean = (float)(sp / 6.0);
        }
      }
      o.fbeta[c] = fb_mean;
      o.fphi[c] = fp_mean;
    }
    bb += rhs * rhs;
    bmax = fmax(bmax, fabs(rhs));
  }
  double v[2] = {bb, bmax};
  const bool mx[2] = {false, true};
  block_reduce_store<2>(v, mx, o.part, kMaxPartials);
  if (o.mx) {
    // (fmaxf drops a NaN, as the separate pass over the slab does)
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) {
      amax = fmaxf(amax, __shfl_xor(amax, s, 64));
      imax = fmaxf(imax, __shfl_xor(imax, s, 64));
    }
    if (lane == 0) {
      atomicMax(&o.mx[0], __float_as_uint(amax));
      atomicMax(&o.mx[1], __float_as_uint(imax));
    }
  }
}

// zero every per-cell array of the listed blocks (blocks that leave the fluid set
// must read as zero; see the reference's pf.clear(), sparse_ops3d_tp.cpp:240)
// The first nf32_all FP32 arrays are cleared on every listed block; the others (and
// every FP64 array) only on blocks that leave the system (keep[b] == 0, or every
// block when keep is null): the next solve rewrites them on all 64 cells of its blocks
// before reading them.
struct ZeroArgs {
  const uint32_t* list;
  long long nlist;
  float* f32[13];
  int nf32, nf32_all;
  double* f64[7];
  int nf64;
  const uint8_t* keep;   // per block: 1 if it holds FLUID cells of the next system
};
__global__ void __launch_bounds__(256) k_zero_blocks(ZeroArgs z) {
  long long t = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (t >= z.nlist) return;
  const uint32_t b = z.list[t];
  const long long c = (long long)b * kBlockVol + (threadIdx.x & 63);
  const bool stays = z.keep && z.keep[b];
  for (int q = 0; q < z.nf32; ++q)
    if (q < z.nf32_all || !stays) z.f32[q][c] = 0.f;
  if (!stays)
    for (int q = 0; q < z.nf64; ++q) z.f64[q][c] = 0.0;
}

// ------------------------------------------------------------- solve kernels
//
// Mixed-precision Jacobi PCG with reliable updates (residual replacement): the CG
// recursion runs in FP32 with FP64 dot products and FP64 alpha/beta; whenever the
// FP32 residual has dropped by a factor delta since the last replacement (or
// reached the tolerance), the host folds x32 into the FP64 solution, recomputes
// the true residual b - A x in FP64, and the recursion continues from it with the
// same search direction.  The two p buffers alternate by the device iteration
// count, so launches that return early (solver paused or done) never desynchronise
// them.

struct VecArgs {
  Dims dc;
  const uint32_t* list;
  long long nlist;
  SolveScalars* sc;
  double* part;
};

// FP64 residual r = b - A x (true residual), partial sum r^2, max |r|
__global__ void __launch_bounds__(256) k_residual64(VecArgs a, const double* __restrict__ b64,
                                                    const double* __restrict__ x, double* __restrict__ r,
                                                    const double* __restrict__ cd, const double* __restrict__ cx,
                                                    const double* __restrict__ cy, const double* __restrict__ cz) {
  const int lane = threadIdx.x & 63;
  double rr = 0.0, rm = 0.0;
  FOR_LIST_BLOCKS(a, b, lane) {
    long long c = b * kBlockVol + lane;
    // non-FLUID cells have zero coefficients and zero right-hand side
    CellNbr n = neighbours(a.dc, b, lane);
    const double xc = x[c];
    double ax = cd[c] * xc + cx[c] * (xc - x[n.xp]) + cx[n.xm] * (xc - x[n.xm]) + cy[c] * (xc - x[n.yp]) +
                cy[n.ym] * (xc - x[n.ym]) + cz[c] * (xc - x[n.zp]) + cz[n.zm] * (xc - x[n.zm]);
    double res = b64[c] - ax;
    r[c] = res;
    rr += res * res;
    rm = fmax(rm, fabs(res));
  }
  double v[2] = {rr, rm};
  const bool mx[2] = {false, true};
  block_reduce_store<2>(v, mx, a.part, kMaxPartials);
}

// r32 = (float)r64, x32 = 0, z = D^-1 r (and p_A = 0 on a fresh start);
// partial r.z, r.r.  keep_z (learned preconditioner): z is left for the network and
// slot 0 gets r_new . z_old (the flexible beta's correction term)
__global__ void __launch_bounds__(256) k_pcg_load(VecArgs a, const double* __restrict__ r64,
                                                  float* __restrict__ x32, float* __restrict__ r32,
                                                  float* __restrict__ z32, float* __restrict__ p,
                                                  const float* __restrict__ invd, int keep_z = 0) {
  const int lane = threadIdx.x & 63;
  double rz = 0.0, rr = 0.0;
  FOR_LIST_BLOCKS(a, b, lane) {
    long long c = b * kBlockVol + lane;
    float r = (float)r64[c];
    float z = keep_z ? z32[c] : invd[c] * r;
    x32[c] = 0.f; r32[c] = r;
    if (!keep_z) z32[c] = z;
    if (p) p[c] = 0.f;
    rz += (double)r * (double)z;
    rr += (double)r * (double)r;
  }
  double v[2] = {rz, rr};
  const bool mx[2] = {false, false};
  block_reduce_store<2>(v, mx, a.part, kMaxPartials);
}

// fresh start: scalars from the loaded residual
__global__ void __launch_bounds__(256) k_pcg_start(const double* part, int n, SolveScalars* sc, double delta2,
                                                   double tol2_abs, int max_iter, int between) {
  double rz = block_sum_partials(part, n, false);
  double rr = block_sum_partials(part + kMaxPartials, n, false);
  if (threadIdx.x == 0) {
    sc->rz = rz; sc->rr = rr; sc->rr0 = rr; sc->rr_best = rr; sc->tol2 = delta2;
    sc->rr_ref = rr; sc->tol2_abs = tol2_abs;
    sc->beta = 0.0; sc->alpha = 0.0; sc->pAp = 0.0;
    sc->iter = 0; sc->iter_best = 0; sc->max_iter = max_iter; sc->stall_window = between;
    sc->breakdown = 0; sc->stalled = 0; sc->upd = 0;
    sc->done = (rr == 0.0 || max_iter <= 0) ? 1 : 0;
  }
}

// after a residual replacement: beta from the replaced residual, same direction
__global__ void __launch_bounds__(256) k_pcg_resume(const double* part, int n, SolveScalars* sc) {
  double rz = block_sum_partials(part, n, false);
  double rr = block_sum_partials(part + kMaxPartials, n, false);
  if (threadIdx.x == 0) {
    sc->beta = sc->rz != 0.0 ? rz / sc->rz : 0.0;
    sc->rz = rz; sc->rr = rr; sc->rr_ref = rr;
    sc->iter_best = sc->iter;  // iteration of the last replacement
    sc->upd = 0;
    sc->done = (rr == 0.0 || sc->iter >= sc->max_iter) ? 1 : 0;
  }
}

// p_new = z + beta p_old (own cell and the six neighbours, recomputed on the fly),
// Ap = A p_new (FP32), partial p.Ap in FP64 -- for arrays of 2 GiB and more
__global__ void __launch_bounds__(256) k_pcg_spmv_gather(VecArgs a, const float* __restrict__ z,
                                                         float* __restrict__ pA, float* __restrict__ pB,
                                                         float* __restrict__ Ap, const float* __restrict__ cd,
                                                         const float* __restrict__ cx, const float* __restrict__ cy,
                                                         const float* __restrict__ cz) {
  if (a.sc->done) return;
  const int lane = threadIdx.x & 63;
  const bool odd = a.sc->iter & 1;
  const float* __restrict__ pold = odd ? pB : pA;
  float* __restrict__ pnew = odd ? pA : pB;
  const float beta = (float)a.sc->beta;
  double pap = 0.0;
  FOR_LIST_BLOCKS(a, b, lane) {
    long long c = b * kBlockVol + lane;
    // non-FLUID cells: z = p = 0 and zero coefficients give p_new = Ap = 0
    CellNbr n = neighbours(a.dc, b, lane);
    const float pc = fmaf(beta, pold[c], z[c]);
    float pxp = fmaf(beta, pold[n.xp], z[n.xp]), pxm = fmaf(beta, pold[n.xm], z[n.xm]);
    float pyp = fmaf(beta, pold[n.yp], z[n.yp]), pym = fmaf(beta, pold[n.ym], z[n.ym]);
    float pzp = fmaf(beta, pold[n.zp], z[n.zp]), pzm = fmaf(beta, pold[n.zm], z[n.zm]);
    float ap = cd[c] * pc;
    ap = fmaf(cx[c], pc - pxp, ap);
    ap = fmaf(cx[n.xm], pc - pxm, ap);
    ap = fmaf(cy[c], pc - pyp, ap);
    ap = fmaf(cy[n.ym], pc - pym, ap);
    ap = fmaf(cz[c], pc - pzp, ap);
    ap = fmaf(cz[n.zm], pc - pzm, ap);
    pnew[c] = pc;
    Ap[c] = ap;
    pap += (double)pc * (double)ap;
  }
  double v[1] = {pap};
  const bool mx[1] = {false};
  block_reduce_store<1>(v, mx, a.part, kMaxPartials);
}

// raw buffer access over a whole array of fewer than 2^31 bytes: 32-bit byte offsets
// from a scalar descriptor; an offset at or past the array's size reads 0 and touches
// no memory (the masked loads below use kNoLoad)
constexpr uint32_t kNoLoad = 0x80000000u;
__device__ inline __amdgpu_buffer_rsrc_t buf_rsrc(const void* p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}
__device__ inline float buf_ld(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  return __uint_as_float((uint32_t)__builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
__device__ inline void buf_st(__amdgpu_buffer_rsrc_t r, uint32_t off, float v) {
  __builtin_amdgcn_raw_buffer_store_b32((int)__float_as_uint(v), r, off, 0, 0);
}

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
struct F4 {
  float v[4];
};
__device__ inline F4 buf_ld4(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  const u32x4 q = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
  F4 f;
  f.v[0] = __uint_as_float(q.x);
  f.v[1] = __uint_as_float(q.y);
  f.v[2] = __uint_as_float(q.z);
  f.v[3] = __uint_as_float(q.w);
  return f;
}
__device__ inline void buf_st4(__amdgpu_buffer_rsrc_t r, uint32_t off, const float* v) {
  u32x4 q;
  q.x = __float_as_uint(v[0]);
  q.y = __float_as_uint(v[1]);
  q.z = __float_as_uint(v[2]);
  q.w = __float_as_uint(v[3]);
  __builtin_amdgcn_raw_buffer_store_b128(q, r, off, 0, 0);
}

// one lane's inputs for a cell row (below): own row, x-face cells, y/z-face rows
struct RowLoads {
  uint32_t o;
  bool fxp, fxm, fyp, fym, fzp, fzm;
  F4 po, zo, cd, cx, cy, cz;
  float poxp, zxp, poxm, zxm, cxxm;
  F4 poyp, zyp, poym, zym, cyym, pozp, zzp, pozm, zzm, czzm;
};

struct RowBufs {
  __amdgpu_buffer_rsrc_t z, po, pn, ap, cd, cx, cy, cz;
  uint32_t sy, sz;
};

__device__ inline void row_load(const RowBufs& B, const Dims& dc, uint32_t b, bool valid, int row, RowLoads& L) {
  const int ly = row & 3, lz = row >> 2;
  int bx, by, bz;
  block_coords(dc, b, bx, by, bz);
  const uint32_t o = valid ? (b * kBlockVol + 4 * row) * 4 : kNoLoad;
  L.o = o;
  L.fxp = valid && bx + 1 < dc.nbx;
  L.fxm = valid && bx > 0;
  L.fyp = valid && ly == 3 && by + 1 < dc.nby;
  L.fym = valid && ly == 0 && by > 0;
  L.fzp = valid && lz == 3 && bz + 1 < dc.nbz;
  L.fzm = valid && lz == 0 && bz > 0;
  const uint32_t oxp = L.fxp ? o + kBlockVol * 4 : kNoLoad, oxm = L.fxm ? o - kBlockVol * 4 + 12 : kNoLoad;
  const uint32_t oyp = L.fyp ? o + B.sy - 48 : kNoLoad, oym = L.fym ? o - B.sy + 48 : kNoLoad;
  const uint32_t ozp = L.fzp ? o + B.sz - 192 : kNoLoad, ozm = L.fzm ? o - B.sz + 192 : kNoLoad;
  L.po = buf_ld4(B.po, o);
  L.zo = buf_ld4(B.z, o);
  L.cd = buf_ld4(B.cd, o);
  L.cx = buf_ld4(B.cx, o);
  L.cy = buf_ld4(B.cy, o);
  L.cz = buf_ld4(B.cz, o);
  L.poxp = buf_ld(B.po, oxp);
  L.zxp = buf_ld(B.z, oxp);
  L.poxm = buf_ld(B.po, oxm);
  L.zxm = buf_ld(B.z, oxm);
  L.cxxm = buf_ld(B.cx, oxm);
  L.poyp = buf_ld4(B.po, oyp);
  L.zyp = buf_ld4(B.z, oyp);
  L.poym = buf_ld4(B.po, oym);
  L.zym = buf_ld4(B.z, oym);
  L.cyym = buf_ld4(B.cy, oym);
  L.pozp = buf_ld4(B.po, ozp);
  L.zzp = buf_ld4(B.z, ozp);
  L.pozm = buf_ld4(B.po, ozm);
  L.zzm = buf_ld4(B.z, ozm);
  L.czzm = buf_ld4(B.cz, ozm);
}

// the same product with arrays below 2 GiB, by cell rows: a lane holds one x-row of
// 4 cells (row = ly + 4 lz, 16 lanes per 4^3 block) and a wave four consecutive list
// blocks.  x-neighbours are in the lane's registers; y- and z-neighbour rows come
// from lanes +-1 and +-4 of the same block; only rows on a block face read the
// neighbouring block (other lanes issue no memory access there).  At the domain
// boundary the neighbour is the cell itself, whose term vanishes, as in
// neighbours().  Per cell the arithmetic is k_pcg_spmv_gather's, in the same order.
// Software-pipelined: the next four blocks' loads are issued before this four's
// stores, so waiting for them leaves the stores in flight.
__global__ void __launch_bounds__(256) k_pcg_spmv(VecArgs a, const float* __restrict__ z, float* __restrict__ pA,
                                                  float* __restrict__ pB, float* __restrict__ Ap,
                                                  const float* __restrict__ cd, const float* __restrict__ cx,
                                                  const float* __restrict__ cy, const float* __restrict__ cz) {
  if (a.sc->done) return;
  const int lane = threadIdx.x & 63;
  const int qb = lane >> 4, row = lane & 15, ly = row & 3, lz = row >> 2;
  const bool odd = a.sc->iter & 1;
  const uint32_t bytes = (uint32_t)(a.dc.nblocks * kBlockVol * 4);
  RowBufs B;
  B.z = buf_rsrc(z, bytes);
  B.po = buf_rsrc(odd ? pB : pA, bytes);
  B.pn = buf_rsrc(odd ? pA : pB, bytes);
  B.ap = buf_rsrc(Ap, bytes);
  B.cd = buf_rsrc(cd, bytes);
  B.cx = buf_rsrc(cx, bytes);
  B.cy = buf_rsrc(cy, bytes);
  B.cz = buf_rsrc(cz, bytes);
  B.sy = (uint32_t)a.dc.nbx * kBlockVol * 4;
  B.sz = (uint32_t)a.dc.nbx * a.dc.nby * kBlockVol * 4;
  const float beta = (float)a.sc->beta;
  double pap = 0.0;
  // the XCD segments of FOR_LIST_BLOCKS, walked in slots of four list entries
  const long long seg = (a.nlist + 7) / 8, xcd = blockIdx.x & 7;
  const long long s0 = xcd * seg, e = min((long long)a.nlist, (xcd + 1) * seg);
  const long long nslot = e > s0 ? (e - s0 + 3) / 4 : 0, step = (long long)(gridDim.x >> 3) * 4;
  const uint32_t last = e > 0 ? (uint32_t)(e - 1) : 0;
  // this lane's block of the slot starting at list position t0 (clamped in range)
  auto slot_block = [&](uint32_t t0, bool& valid) -> uint32_t {
    const uint32_t b0 = list_at(a.list, t0), b1 = list_at(a.list, min(t0 + 1, last));
    const uint32_t b2 = list_at(a.list, min(t0 + 2, last)), b3 = list_at(a.list, min(t0 + 3, last));
    valid = (long long)t0 + qb < e;
    return qb == 0 ? b0 : (qb == 1 ? b1 : (qb == 2 ? b2 : b3));
  };
  long long sl = (long long)(blockIdx.x >> 3) * 4 + (threadIdx.x >> 6);
  if (sl < nslot) {
    bool valid;
    uint32_t b = slot_block(wave_uniform_u32((uint32_t)(s0 + sl * 4)), valid);
    RowLoads cur, nxt;
    row_load(B, a.dc, b, valid, row, cur);
    for (;;) {
      const long long sn = sl + step;
      const bool more = sn < nslot;
      bool vn = false;
      const uint32_t bn = more ? slot_block(wave_uniform_u32((uint32_t)(s0 + sn * 4)), vn) : b;
      row_load(B, a.dc, bn, more && vn, row, nxt);
      // non-FLUID cells: z = p = 0 and zero coefficients give p_new = Ap = 0
      float pc[4], ap[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) pc[i] = fmaf(beta, cur.po.v[i], cur.zo.v[i]);
      const float xpn = fmaf(beta, cur.poxp, cur.zxp), xmn = fmaf(beta, cur.poxm, cur.zxm);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float syp = __shfl(pc[i], lane + 1, 64), sym = __shfl(pc[i], lane - 1, 64);
        const float szp = __shfl(pc[i], lane + 4, 64), szm = __shfl(pc[i], lane - 4, 64);
        const float scy = __shfl(cur.cy.v[i], lane - 1, 64), scz = __shfl(cur.cz.v[i], lane - 4, 64);
        const float pxp = i < 3 ? pc[i + 1] : (cur.fxp ? xpn : pc[3]);
        const float pxm = i > 0 ? pc[i - 1] : (cur.fxm ? xmn : pc[0]);
        const float cxm = i > 0 ? cur.cx.v[i - 1] : (cur.fxm ? cur.cxxm : 0.f);
        const float pyp = ly < 3 ? syp : (cur.fyp ? fmaf(beta, cur.poyp.v[i], cur.zyp.v[i]) : pc[i]);
        const float pym = ly > 0 ? sym : (cur.fym ? fmaf(beta, cur.poym.v[i], cur.zym.v[i]) : pc[i]);
        const float cym = ly > 0 ? scy : (cur.fym ? cur.cyym.v[i] : 0.f);
        const float pzp = lz < 3 ? szp : (cur.fzp ? fmaf(beta, cur.pozp.v[i], cur.zzp.v[i]) : pc[i]);
        const float pzm = lz > 0 ? szm : (cur.fzm ? fmaf(beta, cur.pozm.v[i], cur.zzm.v[i]) : pc[i]);
        const float czm = lz > 0 ? scz : (cur.fzm ? cur.czzm.v[i] : 0.f);
        float sacc = cur.cd.v[i] * pc[i];
        sacc = fmaf(cur.cx.v[i], pc[i] - pxp, sacc);
        sacc = fmaf(cxm, pc[i] - pxm, sacc);
        sacc = fmaf(cur.cy.v[i], pc[i] - pyp, sacc);
        sacc = fmaf(cym, pc[i] - pym, sacc);
        sacc = fmaf(cur.cz.v[i], pc[i] - pzp, sacc);
        sacc = fmaf(czm, pc[i] - pzm, sacc);
        ap[i] = sacc;
      }
      buf_st4(B.pn, cur.o, pc);
      buf_st4(B.ap, cur.o, ap);
#pragma unroll
      for (int i = 0; i < 4; ++i) pap += (double)pc[i] * (double)ap[i];
      if (!more) break;
      sl = sn;
      b = bn;
      cur = nxt;
    }
  }
  double v[1] = {pap};
  const bool mx[1] = {false};
  block_reduce_store<1>(v, mx, a.part, kMaxPartials);
}

// alpha = r.z / p.Ap from the SpMV partials (one workgroup)
__global__ void __launch_bounds__(256) k_pcg_alpha(const double* part, int n, SolveScalars* sc) {
  if (sc->done) return;
  const double pap = block_sum_fast(part, n);
  if (threadIdx.x == 0) {
    sc->pAp = pap;
    if (!(pap > 0.0) || !isfinite(pap)) {
      sc->breakdown = 1;
      sc->done = 1;
      sc->alpha = 0.0;
    } else {
      sc->alpha = sc->rz / pap;
    }
  }
}

// iteration end: count, decide on a residual replacement (pause) or continue with
// beta = r.z_new / r.z
__device__ inline void finish_iteration(SolveScalars* sc, double rz, double rr) {
  int it = sc->iter + 1;
  sc->iter = it;
  sc->rr = rr;
  if (!isfinite(rr) || !isfinite(rz)) {
    sc->breakdown = 1;
    sc->done = 1;
    return;
  }
  // residual replacement when the FP32 residual fell by delta below the largest
  // residual since the last replacement, reached the target, ran `stall_window`
  // iterations (the FP32 recursion drifts from b - A x with iteration count), or
  // the iteration budget is spent; beta is formed after it
  if (rr > sc->rr_ref) sc->rr_ref = rr;
  if (rr <= sc->tol2 * sc->rr_ref || rr <= sc->tol2_abs || it >= sc->max_iter ||
      it - sc->iter_best >= sc->stall_window) {
    sc->upd = 1;
    sc->done = 1;
    sc->rz_pending = rz;
    return;
  }
  sc->beta = sc->rz != 0.0 ? rz / sc->rz : 0.0;
  sc->rz = rz;
}

// x += alpha p, r -= alpha Ap, z = D^-1 r; partial r.z and r.r
__global__ void __launch_bounds__(256) k_pcg_update(VecArgs a, float* __restrict__ x, float* __restrict__ r,
                                                    float* __restrict__ z, const float* __restrict__ pA,
                                                    const float* __restrict__ pB, const float* __restrict__ Ap,
                                                    const float* __restrict__ invd) {
  if (a.sc->done) return;
  const int lane = threadIdx.x & 63;
  const float* __restrict__ p = (a.sc->iter & 1) ? pA : pB;  // the p_new of this iteration
  const float alpha = (float)a.sc->alpha;
  double rz = 0.0, rr = 0.0;
  // the blocks of FOR_LIST_BLOCKS in the same order, software-pipelined: the next
  // block's loads are issued before this block's stores, so waiting for them does
  // not wait for the stores (the vector-memory counter completes in issue order)
  const long long seg = (a.nlist + 7) / 8, xcd = blockIdx.x & 7, step = (long long)(gridDim.x >> 3) * 4;
  const long long end = min((long long)a.nlist, (xcd + 1) * seg);
  long long t = xcd * seg + (long long)(blockIdx.x >> 3) * 4 + (threadIdx.x >> 6);
  if (t < end) {
    long long c = (long long)list_at(a.list, wave_uniform_u32((uint32_t)t)) * kBlockVol + lane;
    float pv = p[c], xv = x[c], av = Ap[c], rv = r[c], dv = invd[c];
    for (;;) {
      const long long tn = t + step;
      const bool more = tn < end;
      const long long cn = more ? (long long)list_at(a.list, wave_uniform_u32((uint32_t)tn)) * kBlockVol + lane : c;
      const float pn = p[cn], xn0 = x[cn], an = Ap[cn], rn0 = r[cn], dn = invd[cn];
      const float xn = fmaf(alpha, pv, xv);
      const float rn = fmaf(-alpha, av, rv);
      const float zn = dv * rn;
      x[c] = xn; r[c] = rn; z[c] = zn;
      rz += (double)rn * (double)zn;
      rr += (double)rn * (double)rn;
      if (!more) break;
      t = tn; c = cn;
      pv = pn; xv = xn0; av = an; rv = rn0; dv = dn;
    }
  }
  double v[2] = {rz, rr};
  const bool mx[2] = {false, false};
  block_reduce_store<2>(v, mx, a.part, kMaxPartials);
}

__global__ void __launch_bounds__(256) k_pcg_beta(const double* part, int n, SolveScalars* sc) {
  if (sc->done) return;
  const double rz = block_sum_fast(part, n);
  const double rr = block_sum_fast(part + kMaxPartials, n);
  if (threadIdx.x == 0) finish_iteration(sc, rz, rr);
}

// Flexible PCG (Notay 2000): with a preconditioner that changes between iterations
// (the learned network), beta = r_new . (z_new - z_old) / (r_old . z_old); a
// negative or non-finite beta restarts the direction (beta = 0), as the reference's
// multires solver does (multires_pressure3d.cpp:1424-1434).
__device__ inline double flexible_beta(double rz_new, double rz_cross, double rz_old) {
  const double b = rz_old != 0.0 ? (rz_new - rz_cross) / rz_old : 0.0;
  return (isfinite(b) && b > 0.0) ? b : 0.0;
}

// update without z: x += alpha p, r -= alpha Ap; partial r_new . z_old (slot 0), r.r (slot 1)
__global__ void __launch_bounds__(256) k_pcg_update_flex(VecArgs a, float* __restrict__ x, float* __restrict__ r,
                                                         const float* __restrict__ z, const float* __restrict__ pA,
                                                         const float* __restrict__ pB, const float* __restrict__ Ap) {
  if (a.sc->done) return;
  const int lane = threadIdx.x & 63;
  const float* __restrict__ p = (a.sc->iter & 1) ? pA : pB;
  const float alpha = (float)a.sc->alpha;
  double rzo = 0.0, rr = 0.0;
  FOR_LIST_BLOCKS(a, b, lane) {
    long long c = b * kBlockVol + lane;
    float xn = fmaf(alpha, p[c], x[c]);
    float rn = fmaf(-alpha, Ap[c], r[c]);
    x[c] = xn; r[c] = rn;
    rzo += (double)rn * (double)z[c];
    rr += (double)rn * (double)rn;
  }
  double v[2] = {rzo, rr};
  const bool mx[2] = {false, false};
  block_reduce_store<2>(v, mx, a.part, kMaxPartials);
}

// iteration end with the flexible beta: slot 2 (nz partials) r.z_new, slots 0/1 (n)
__global__ void __launch_bounds__(256) k_pcg_beta_flex(const double* part, int n, int nz, SolveScalars* sc) {
  if (sc->done) return;
  const double rzo = block_sum_fast(part, n);
  const double rr = block_sum_fast(part + kMaxPartials, n);
  const double rz = block_sum_fast(part + 2 * kMaxPartials, nz);
  if (threadIdx.x == 0) {
    const double beta = flexible_beta(rz, rzo, sc->rz);
    sc->rz_cross = rzo;  // kept for the resume after a residual replacement
    finish_iteration(sc, rz, rr);
    if (!sc->done) sc->beta = beta;
  }
}

// fresh start / resume for the flexible solve (z from the network, slot 2)
__global__ void __launch_bounds__(256) k_pcg_start_flex(const double* part, int n, int nz, SolveScalars* sc,
                                                        double delta2, double tol2_abs, int max_iter, int between) {
  const double rr = block_sum_partials(part + kMaxPartials, n, false);
  const double rz = block_sum_partials(part + 2 * kMaxPartials, nz, false);
  if (threadIdx.x == 0) {
    sc->rz = rz; sc->rr = rr; sc->rr0 = rr; sc->rr_best = rr; sc->tol2 = delta2;
    sc->rr_ref = rr; sc->tol2_abs = tol2_abs;
    sc->beta = 0.0; sc->alpha = 0.0; sc->pAp = 0.0;
    sc->iter = 0; sc->iter_best = 0; sc->max_iter = max_iter; sc->stall_window = between;
    sc->breakdown = 0; sc->stalled = 0; sc->upd = 0;
    sc->done = (rr == 0.0 || max_iter <= 0) ? 1 : 0;
  }
}

// The network has already replaced z_old by M(r_recursive) in the paused iteration,
// so the cross term r_new . z_old is the one formed then with the recursive residual
// (the replaced residual differs from it by the FP32 drift only).
__global__ void __launch_bounds__(256) k_pcg_resume_flex(const double* part, int n, int nz, SolveScalars* sc) {
  const double rzo = sc->rz_cross;
  const double rr = block_sum_partials(part + kMaxPartials, n, false);
  const double rz = block_sum_partials(part + 2 * kMaxPartials, nz, false);
  if (threadIdx.x == 0) {
    sc->beta = flexible_beta(rz, rzo, sc->rz);
    sc->rz = rz; sc->rr = rr; sc->rr_ref = rr;
    sc->iter_best = sc->iter;
    sc->upd = 0;
    sc->done = (rr == 0.0 || sc->iter >= sc->max_iter) ? 1 : 0;
  }
}

// x64 += x32 (x32 is reset by the next k_pcg_load)
__global__ void __launch_bounds__(256) k_ir_accum(VecArgs a, double* __restrict__ x64,
                                                  const float* __restrict__ x32) {
  const int lane = threadIdx.x & 63;
  FOR_LIST_BLOCKS(a, b, lane) {
    long long c = b * kBlockVol + lane;
    x64[c] += (double)x32[c];
  }
}

// x64 start: 0, or the stored pressure (warm start) on FLUID cells
// (x64 holds the previous step's pressure for a warm start, saved before the
// per-solve zeroing; FLUID cells are the ones with a preconditioner entry)
__global__ void __launch_bounds__(256) k_ir_start(VecArgs a, double* __restrict__ x64,
                                                  const float* __restrict__ invd, int warm) {
  const int lane = threadIdx.x & 63;
  FOR_LIST_BLOCKS(a, b, lane) {
    long long c = b * kBlockVol + lane;
    x64[c] = (warm && invd[c] != 0.f) ? x64[c] : 0.0;
  }
}

// x64 = p over the previous solve's blocks (warm-start save)
__global__ void __launch_bounds__(256) k_warm_save(const uint32_t* __restrict__ list, long long nlist,
                                                   const float* __restrict__ p, double* __restrict__ x64) {
  long long t = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (t >= nlist) return;
  long long c = (long long)list[t] * kBlockVol + (threadIdx.x & 63);
  x64[c] = (double)p[c];
}

// ------------------------------------------------------- pressure and faces

// p = (float)x on FLUID cells, 0 elsewhere (sparse_ops3d_tp.cpp:366-372)
__global__ void __launch_bounds__(256) k_store_pressure(VecArgs a, const uint8_t* __restrict__ marker,
                                                        const double* __restrict__ x64, float* __restrict__ p,
                                                        uint8_t* __restrict__ act_p) {
  const int lane = threadIdx.x & 63;
  FOR_LIST_BLOCKS(a, b, lane) {
    long long c = b * kBlockVol + lane;
    p[c] = marker[c] == CELL_FLUID ? (float)x64[c] : 0.f;
    if (lane == 0) act_p[b] = 1;
  }
}

struct VelArgs {
  Geom g;
  Dims dc, df[3];
  const uint8_t* marker;
  const float* p;
  float* vel[3];
  const float* mass[3];
  uint8_t* act_vel[3];
  PhaseK ph;
  double s;  // dt / dx
  const uint32_t* list;
  long long nlist;
  int unit_beta;  // single-phase path: beta = 1 on every face (sparse_ops3d.cpp:270-316)
};

__device__ inline void update_face(const VelArgs& a, int ax, int fi, int fj, int fk, int li, int lj, int lk, int ri,
                                   int rj, int rk) {
  const Dims& fd = a.df[ax];
  uint8_t ml = marker_at(a.marker, a.dc, a.g, li, lj, lk), mr = marker_at(a.marker, a.dc, a.g, ri, rj, rk);
  bool lf = ml == CELL_FLUID, rf = mr == CELL_FLUID;
  if (!lf && !rf) return;
  long long gi = bm_index(fd, fi, fj, fk);
  a.act_vel[ax][gi >> 6] = 1;
  if (ml == CELL_SOLID || mr == CELL_SOLID) {
    a.vel[ax][gi] = 0.f;
    return;
  }
  double pl = lf ? (double)a.p[bm_index(a.dc, li, lj, lk)] : 0.0;
  double pr = rf ? (double)a.p[bm_index(a.dc, ri, rj, rk)] : 0.0;
  double dlt = a.unit_beta ? a.s * (pr - pl) : a.s * beta_of_raw((double)a.mass[ax][gi], a.ph) * (pr - pl);
  a.vel[ax][gi] = a.vel[ax][gi] - (float)dlt;
}

// u -= (dt/dx) beta_f (p_r - p_l) on faces of FLUID cells; faces touching SOLID := 0
// (sparse_ops3d_tp.cpp:374-422).  Each face next to a FLUID cell is visited once:
// as the lower face of a FLUID cell, or as the upper face when the upper
// neighbour is not FLUID.
__global__ void __launch_bounds__(256) k_velocity_update(VelArgs a) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & 63;
  long long t = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (t >= a.nlist) return;
  long long b = a.list[t];
  if (a.marker[b * kBlockVol + lane] != CELL_FLUID) return;
  int bx, by, bz;
  block_coords(a.dc, b, bx, by, bz);
  int i = bx * 4 + (lane & 3), j = by * 4 + ((lane >> 2) & 3), k = bz * 4 + (lane >> 4);
  // x faces
  if (i > 0) update_face(a, 0, i, j, k, i - 1, j, k, i, j, k);
  if (i + 1 < a.g.nx && marker_at(a.marker, a.dc, a.g, i + 1, j, k) != CELL_FLUID)
    update_face(a, 0, i + 1, j, k, i, j, k, i + 1, j, k);
  if (j > 0) update_face(a, 1, i, j, k, i, j - 1, k, i, j, k);
  if (j + 1 < a.g.ny && marker_at(a.marker, a.dc, a.g, i, j + 1, k) != CELL_FLUID)
    update_face(a, 1, i, j + 1, k, i, j, k, i, j + 1, k);
  if (k > 0) update_face(a, 2, i, j, k, i, j, k - 1, i, j, k);
  if (k + 1 < a.g.nz && marker_at(a.marker, a.dc, a.g, i, j, k + 1) != CELL_FLUID)
    update_face(a, 2, i, j, k + 1, i, j, k, i, j, k + 1);
}

}  // namespace pf
