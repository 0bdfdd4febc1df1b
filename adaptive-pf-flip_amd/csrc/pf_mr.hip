// pf_mr.hip -- the multires two-phase path on the fine grid (reference:
// src/grid/multires_layout3d.h, multires_scalar_grid3d.h, multires_mac_grid3d.h,
// src/driver/multires_ops3d_tp.cpp, src/pressure/multires_pressure3d.cpp,
// src/driver/multires_sim3d_tp.cpp).
//
// MI355X representation.  The reference keeps leaves in a sorted vector (linear scan
// per lookup) and faces in std::maps.  Here the 2-level layout becomes a level byte
// per fine cell (lvl, 255 = no leaf); a leaf cell of level L is the 2^L-aligned cube
// of fine cells holding it, and its scalars (marker, pressure) are stored at every
// fine cell of that cube, so a lookup is one load.  Face fields stay dense fine face
// arrays (the same block-major slabs as the sparse path) plus a validity byte per
// fine face: the faces the reference enumerates (1x1 patches on leaf-cell
// boundaries); the others always hold zero, which is what the reference's map
// lookups return for them.
//
// Pressure: leaf cell a (volume V_a = h_a^3) has the reference's row
//   sum_p c_p / V_a (x_a - x_nbr(p)) = rhs_a,  c_p = dt beta_p dx^2 / (h_a/2 + h_b/2),
// over its fine face patches p (multires_pressure3d.cpp:783-812).  Multiplying row a
// by V_a gives a symmetric conductance Laplacian S x = V rhs with the same solution,
// so the FP32 inner PCG runs on S (Jacobi on S equals Jacobi on the reference's
// rows).  The unknown of a leaf cell lives at its anchor (lowest) fine cell; c_p is
// stored per fine face (cf, zero unless the face couples two FLUID leaf cells, neither
// pinned) and the Dirichlet part of the diagonal per anchor (cd).  The FP64 residual
// reports |rhs - A x|_inf = |(V rhs - S x) / V|_inf for the reference's stop rule.
#include "pf_internal.h"

namespace pf {

struct MrGrid {
  Dims dc, du, dv, dw;
  Geom g;
  const uint8_t* lvl;      // per fine cell (block-major)
  const uint8_t* marker;   // per fine cell: its leaf cell's marker
  const uint8_t* valid[3]; // per fine face
};

__device__ inline int mr_level(const MrGrid& m, int i, int j, int k) {
  if (i < 0 || j < 0 || k < 0 || i >= m.g.nx || j >= m.g.ny || k >= m.g.nz) return -1;
  const int l = m.lvl[bm_index(m.dc, i, j, k)];
  return l == 255 ? -1 : l;
}

// markerAtFineCell, multires_pressure3d.cpp:81-88 (outside or without a leaf: SOLID)
__device__ inline int mr_marker(const MrGrid& m, int i, int j, int k) {
  if (mr_level(m, i, j, k) < 0) return CELL_SOLID;
  return m.marker[bm_index(m.dc, i, j, k)];
}

__device__ inline const Dims& mr_fdims(const MrGrid& m, int axis) { return axis == 0 ? m.du : (axis == 1 ? m.dv : m.dw); }

// leaf blocks (level, bx, by, bz) -> level byte per fine cell
__global__ void __launch_bounds__(256) k_mr_levels(const int32_t* __restrict__ leaves, int nleaves, Dims dc, Geom g,
                                                   uint8_t* __restrict__ lvl) {
  const int a = blockIdx.x;
  if (a >= nleaves) return;
  const int L = leaves[4 * a], s = 4 << L;
  const int x0 = leaves[4 * a + 1] * s, y0 = leaves[4 * a + 2] * s, z0 = leaves[4 * a + 3] * s;
  for (int q = threadIdx.x; q < s * s * s; q += 256) {
    const int i = x0 + q % s, j = y0 + (q / s) % s, k = z0 + q / (s * s);
    if (i < g.nx && j < g.ny && k < g.nz) lvl[bm_index(dc, i, j, k)] = (uint8_t)L;
  }
}

// enumerated faces (MRMacGrid3D::enumerateFaces, multires_mac_grid3d.h:107-149): a
// fine face is the low face of the leaf cell above it when its coordinate is a
// multiple of that cell's step, or the (domain-clipped) high face of the leaf cell
// below it
__global__ void __launch_bounds__(256) k_mr_valid(MrGrid m, int axis, uint8_t* __restrict__ valid) {
  const Dims& fd = axis == 0 ? m.du : (axis == 1 ? m.dv : m.dw);
  const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
  if (q >= fd.nblocks * kBlockVol) return;
  const long long b = q >> 6;
  const int lane = (int)(q & 63);
  int bx, by, bz;
  block_coords(fd, b, bx, by, bz);
  const int i = bx * 4 + (lane & 3), j = by * 4 + ((lane >> 2) & 3), k = bz * 4 + (lane >> 4);
  uint8_t v = 0;
  if (i < fd.nx && j < fd.ny && k < fd.nz) {
    const int c[3] = {i, j, k};
    const int n = axis == 0 ? m.g.nx : (axis == 1 ? m.g.ny : m.g.nz);
    const int lh = mr_level(m, i, j, k);
    const int ll = mr_level(m, i - (axis == 0), j - (axis == 1), k - (axis == 2));
    if (lh >= 0 && c[axis] % (1 << lh) == 0) v = 1;
    if (ll >= 0 && (c[axis] % (1 << ll) == 0 || c[axis] == n)) v = 1;
  }
  valid[q] = v;
}

// markCells (multires_sim3d_tp.cpp:39-70) per leaf cell: SOLID when one of its fine
// cells lies in the boundary layer, else FLUID when one holds a particle, else AIR;
// written to every fine cell of the leaf cell
__global__ void __launch_bounds__(256) k_mr_mark(MrGrid m, const uint32_t* __restrict__ start,
                                                 uint8_t* __restrict__ marker, uint8_t* __restrict__ act) {
  const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
  if (q >= m.dc.nblocks * kBlockVol) return;
  const long long b = q >> 6;
  const int lane = (int)(q & 63);
  int bx, by, bz;
  block_coords(m.dc, b, bx, by, bz);
  const int i = bx * 4 + (lane & 3), j = by * 4 + ((lane >> 2) & 3), k = bz * 4 + (lane >> 4);
  if (i >= m.g.nx || j >= m.g.ny || k >= m.g.nz) return;
  const int L = mr_level(m, i, j, k);
  uint8_t mk = CELL_AIR;
  if (L >= 0) {
    const int s = 1 << L;
    const int x0 = i - i % s, y0 = j - j % s, z0 = k - k % s;
    bool solid = false, fluid = false;
    for (int dz = 0; dz < s; ++dz)
      for (int dy = 0; dy < s; ++dy)
        for (int dx = 0; dx < s; ++dx) {
          const int x = x0 + dx, y = y0 + dy, z = z0 + dz;
          if (x >= m.g.nx || y >= m.g.ny || z >= m.g.nz) continue;
          solid = solid || x == 0 || y == 0 || z == 0 || x == m.g.nx - 1 || y == m.g.ny - 1 || z == m.g.nz - 1;
          const long long c = bm_index(m.dc, x, y, z);
          fluid = fluid || start[c + 1] > start[c];
        }
    mk = solid ? CELL_SOLID : (fluid ? CELL_FLUID : CELL_AIR);
  }
  marker[q] = mk;
  if (mk != CELL_AIR) act[b] = 1;
}

struct MrAsm {
  MrGrid m;
  const uint32_t* list;   // blocks holding FLUID anchors
  long long nlist;
  PhaseK ph;
  double dt, cdiv;
  uint32_t* pin;          // [0] touches AIR, [1] pinned cell linear id (all ones: none), [4] FLUID leaf cells
  const float *u, *v, *w, *mu, *mv, *mw;
  double *b64, *cd64, *cf64[3], *vol64;
  float *cd32, *cf32[3], *invd32;
  double* part;
  uint32_t* order_key;    // FLUID anchors: the reference's leafCells order (for the pin)
  // optional, the reference's own row quantities (for the reference-order solve):
  double* dref;           // diag = sum of cond / V over non-solid patches (1 on the pinned row)
  double* rref;           // rhs = -divergence + cdiv (0 on the pinned row)
  double* vtrue;          // h^3 of the leaf cell (the pinned one included)
  long long* pin_cell;    // the pinned anchor's block-major index
};

__device__ inline bool mr_is_anchor(const MrGrid& m, int i, int j, int k, int& L) {
  L = mr_level(m, i, j, k);
  if (L < 0) return false;
  const int s = 1 << L;
  return i % s == 0 && j % s == 0 && k % s == 0;
}

__device__ inline long long mr_anchor(const MrGrid& m, int i, int j, int k) {
  const int L = mr_level(m, i, j, k);
  const int s = 1 << (L < 0 ? 0 : L);
  return bm_index(m.dc, i - i % s, j - j % s, k - k % s);
}

__device__ inline float mr_face_mass(const MrAsm& a, int axis, long long fi) {
  return axis == 0 ? a.mu[fi] : (axis == 1 ? a.mv[fi] : a.mw[fi]);
}
__device__ inline float mr_face_vel(const MrAsm& a, int axis, long long fi) {
  return axis == 0 ? a.u[fi] : (axis == 1 ? a.v[fi] : a.w[fi]);
}

// visit the fine face patches of leaf cell (x0, y0, z0, step s) in the reference's
// order (visitCellFaces, multires_pressure3d.cpp:101-136): fn(axis, face i, j, k,
// neighbour i, j, k, sign)
template <typename Fn>
__device__ inline void mr_patches(const MrGrid& m, int x0, int y0, int z0, int s, Fn&& fn) {
  const int x1 = x0 + s, y1 = y0 + s, z1 = z0 + s;
  const int cx1 = min(x1, m.g.nx), cy1 = min(y1, m.g.ny), cz1 = min(z1, m.g.nz);
  for (int z = z0; z < cz1; ++z)
    for (int y = y0; y < cy1; ++y) {
      fn(0, x0, y, z, x0 - 1, y, z, -1.0);
      fn(0, min(x1, m.g.nx), y, z, x1, y, z, 1.0);
    }
  for (int z = z0; z < cz1; ++z)
    for (int x = x0; x < cx1; ++x) {
      fn(1, x, y0, z, x, y0 - 1, z, -1.0);
      fn(1, x, min(y1, m.g.ny), z, x, y1, z, 1.0);
    }
  for (int y = y0; y < cy1; ++y)
    for (int x = x0; x < cx1; ++x) {
      fn(2, x, y, z0, x, y, z0 - 1, -1.0);
      fn(2, x, y, min(z1, m.g.nz), x, y, z1, 1.0);
    }
}

// leafCells order of a leaf cell (leaf blocks sorted by (level, bx, by, bz), then
// lz, ly, lx): a 32-bit key for the pin (the first FLUID cell in that order)
__device__ inline uint32_t mr_order_key(const MrGrid& m, int i, int j, int k, int L) {
  const int s = 1 << L, bs = 4 * s;
  const int bx = i / bs, by = j / bs, bz = k / bs;
  const int lx = (i % bs) / s, ly = (j % bs) / s, lz = (k % bs) / s;
  const int nbx = (m.g.nx + 3) / 4, nby = (m.g.ny + 3) / 4, nbz = (m.g.nz + 3) / 4;
  // (level, bx, by, bz) lexicographic, then (lz, ly, lx): fits 32 bits for grids up to 4096 cells
  const uint32_t blk = ((uint32_t)bx * (uint32_t)nby + (uint32_t)by) * (uint32_t)nbz + (uint32_t)bz;
  return ((uint32_t)L << 31) | (blk << 6) | (uint32_t)(lx + 4 * (ly + 4 * lz));
}

// pin detection: does a FLUID leaf cell have an AIR neighbour; the first FLUID cell
__global__ void __launch_bounds__(256) k_mr_pin(MrAsm a) {
  __shared__ uint32_t s_air[4], s_min[4], s_cnt[4];
  const int lane = threadIdx.x & 63;
  bool anyair = false;
  uint32_t mn = 0xffffffffu, cnt = 0;
  FOR_LIST_BLOCKS(a, b, lane) {
    int bx, by, bz;
    block_coords(a.m.dc, b, bx, by, bz);
    const int i = bx * 4 + (lane & 3), j = by * 4 + ((lane >> 2) & 3), k = bz * 4 + (lane >> 4);
    int L;
    bool air = false;
    if (i < a.m.g.nx && j < a.m.g.ny && k < a.m.g.nz && mr_is_anchor(a.m, i, j, k, L) &&
        a.m.marker[b * kBlockVol + lane] == CELL_FLUID) {
      ++cnt;
      mn = min(mn, mr_order_key(a.m, i, j, k, L));
      mr_patches(a.m, i, j, k, 1 << L, [&](int, int, int, int, int ni, int nj, int nk, double) {
        air = air || (mr_level(a.m, ni, nj, nk) >= 0 && mr_marker(a.m, ni, nj, nk) == CELL_AIR);
      });
    }
    anyair = anyair || __any(air);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    mn = min(mn, (uint32_t)__shfl_xor((int)mn, o, 64));
    cnt += (uint32_t)__shfl_xor((int)cnt, o, 64);
  }
  const int w = threadIdx.x >> 6;
  if (lane == 0) { s_air[w] = anyair; s_min[w] = mn; s_cnt[w] = cnt; }
  __syncthreads();
  if (threadIdx.x == 0) {
    if (s_air[0] | s_air[1] | s_air[2] | s_air[3]) atomicOr(&a.pin[0], 1u);
    const uint32_t m4 = min(min(s_min[0], s_min[1]), min(s_min[2], s_min[3]));
    if (m4 != 0xffffffffu) atomicMin(&a.pin[1], m4);
    const uint32_t c4 = s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
    if (c4) atomicAdd(&a.pin[4], c4);
  }
}

// rows (multires_pressure3d.cpp:783-812) scaled by the cell volume: cd (Dirichlet
// diagonal), cf per fine face (coupling), 1/diag, V rhs; FP64 and FP32 copies
__global__ void __launch_bounds__(256) k_mr_assemble(MrAsm a) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & 63;
  double bb = 0.0;
  const double area = a.m.g.dx * a.m.g.dx;
  FOR_LIST_BLOCKS(a, b, lane) {
    const long long c = b * kBlockVol + lane;
    int bx, by, bz;
    block_coords(a.m.dc, b, bx, by, bz);
    const int i = bx * 4 + (lane & 3), j = by * 4 + ((lane >> 2) & 3), k = bz * 4 + (lane >> 4);
    int L = 0;
    double rhs = 0.0, cd = 0.0, diag = 0.0, V = 0.0, dref = 0.0, vtrue = 0.0;
    const bool anchor = i < a.m.g.nx && j < a.m.g.ny && k < a.m.g.nz && mr_is_anchor(a.m, i, j, k, L) &&
                        a.m.marker[c] == CELL_FLUID;
    bool pinned = false;
    if (anchor) {
      const int s = 1 << L;
      const double h = a.m.g.dx * (double)s;
      V = h * h * h;
      const uint32_t key = mr_order_key(a.m, i, j, k, L);
      pinned = a.pin[0] == 0 && key == a.pin[1];
      const uint32_t pin_key = a.pin[0] == 0 ? a.pin[1] : 0xffffffffu;
      double div = 0.0;
      mr_patches(a.m, i, j, k, s, [&](int axis, int fx, int fy, int fz, int ni, int nj, int nk, double sign) {
        const int mn = mr_marker(a.m, ni, nj, nk);
        const Dims& fd = mr_fdims(a.m, axis);
        const long long fi = bm_index(fd, fx, fy, fz);
        if (mn == CELL_SOLID) return;
        const int Ln = mr_level(a.m, ni, nj, nk);
        const double hn = Ln >= 0 ? a.m.g.dx * (double)(1 << Ln) : h;
        const double dist = 0.5 * h + 0.5 * hn;
        const double cond = (dist > 0.0) ? a.dt * beta_of_raw((double)mr_face_mass(a, axis, fi), a.ph) * area / dist : 0.0;
        diag += cond;
        dref += (V > 0.0) ? cond / V : 0.0;
        bool coupled = false;
        if (mn == CELL_FLUID) {
          const int sn = 1 << Ln;
          const uint32_t nkey = mr_order_key(a.m, ni - ni % sn, nj - nj % sn, nk - nk % sn, Ln);
          coupled = nkey != pin_key && !pinned;
        }
        if (!coupled) cd += cond;
        // a coupling face is written by both of its cells with the same value
        if (coupled) { a.cf64[axis][fi] = cond; a.cf32[axis][fi] = (float)cond; }
        div += sign * (double)mr_face_vel(a, axis, fi) * area / V;
      });
      rhs = pinned ? 0.0 : -div + a.cdiv;
      vtrue = V;
      if (pinned) { cd = 1.0; diag = 1.0; V = 1.0; dref = 1.0; }
      if (pinned && a.pin_cell) a.pin_cell[0] = c;
    }
    if (a.dref) { a.dref[c] = dref; a.rref[c] = anchor ? rhs : 0.0; a.vtrue[c] = vtrue; }
    // S = V A: rows of the symmetric conductance form
    a.b64[c] = V * rhs;
    a.vol64[c] = anchor ? V : 0.0;
    a.cd64[c] = cd;
    a.cd32[c] = (float)cd;
    a.invd32[c] = diag > 0.0 ? (float)(1.0 / diag) : 0.f;
    bb += V * rhs * V * rhs;
  }
  double v[1] = {bb};
  const bool mx[1] = {false};
  block_reduce_store<1>(v, mx, a.part, kMaxPartials);
}

struct MrOp {
  MrGrid m;
  const float *cd, *cf[3];
};
struct MrOp64 {
  MrGrid m;
  const double *cd, *cf[3], *vol;
};

// sum over the patches of anchor (i,j,k): c_p (x_a - x_anchor(nbr)) with the
// neighbour value produced by `val(anchor index)`
template <typename T, typename Val>
__device__ inline T mr_apply_row(const MrGrid& m, const T* const* cf, int i, int j, int k, int s, T xa, Val&& val) {
  T acc = (T)0;
  mr_patches(m, i, j, k, s, [&](int axis, int fx, int fy, int fz, int ni, int nj, int nk, double) {
    const long long fi = bm_index(mr_fdims(m, axis), fx, fy, fz);
    const T cfv = cf[axis][fi];
    if (cfv != (T)0) acc += cfv * (xa - val(mr_anchor(m, ni, nj, nk)));
  });
  return acc;
}

// The same row sum spread over the lanes of a 4^3 block whose cells all have level
// L <= 2 (a block lies inside one leaf): the s^3 lanes inside a leaf cell (s = 2^L)
// take its 6 s^2 patches in turn, q = r, r + s^3, ... (at most 6 per lane) for the
// lane's offset r inside the cell, so a level-2 cell's 96 patches cost each lane two
// instead of the anchor lane 96 in a row.  Patch q is axis q / 2s^2, then side and the
// in-face position.  The loads go in three waves: every patch's coefficient and
// neighbour level byte (independent), then the neighbour anchors' values, so a lane
// waits on three loads in a row instead of three per patch.  The xor shuffles over
// the offset bits (lane = x + 4 y + 16 z) leave every lane of the cell with the total.
// Wave-uniform L; lanes with live == false add nothing.  mr_row_slots is one lane's
// share: patches r0, r0 + stride, ... (at most 6) of the cell anchored at (ax, ay, az).
template <typename T, typename Val>
__device__ inline T mr_row_slots(const MrGrid& m, const T* const* cf, int ax, int ay, int az, int s, int r0, int stride,
                                 bool live, T xa, Val&& val) {
  const int s2 = s * s;
  T c[6];
  int n[6][3];
  int lv[6];
#pragma unroll
  for (int t = 0; t < 6; ++t) {
    const int q = r0 + t * stride;
    c[t] = (T)0;
    lv[t] = -1;
    if (!live || q >= 6 * s2) continue;
    const int axis = q / (2 * s2), rem = q - axis * 2 * s2;
    const int side = rem & 1, w = rem >> 1, u = w % s, v = w / s;
    int f[3] = {ax, ay, az};
    if (axis == 0) { f[1] += u; f[2] += v; }
    else if (axis == 1) { f[0] += u; f[2] += v; }
    else { f[0] += u; f[1] += v; }
    if (f[0] >= m.g.nx || f[1] >= m.g.ny || f[2] >= m.g.nz) continue;
    n[t][0] = f[0]; n[t][1] = f[1]; n[t][2] = f[2];
    const int lim = axis == 0 ? m.g.nx : (axis == 1 ? m.g.ny : m.g.nz);
    if (side) {
      n[t][axis] = f[axis] + s;
      f[axis] = min(n[t][axis], lim);
    } else {
      n[t][axis] = f[axis] - 1;
    }
    c[t] = cf[axis][bm_index(mr_fdims(m, axis), f[0], f[1], f[2])];
    lv[t] = mr_level(m, n[t][0], n[t][1], n[t][2]);
  }
  T acc = (T)0;
#pragma unroll
  for (int t = 0; t < 6; ++t) {
    if (c[t] != (T)0 && lv[t] >= 0) {
      const int sn = 1 << lv[t];
      acc += c[t] * (xa - val(bm_index(m.dc, n[t][0] & ~(sn - 1), n[t][1] & ~(sn - 1), n[t][2] & ~(sn - 1))));
    }
  }
  return acc;
}

template <typename T, typename Val>
__device__ inline T mr_apply_row_coop(const MrGrid& m, const T* const* cf, int i, int j, int k, int L, bool live,
                                      T xa, Val&& val) {
  const int s = 1 << L;
  const int ax = i & ~(s - 1), ay = j & ~(s - 1), az = k & ~(s - 1);
  T acc = mr_row_slots<T>(m, cf, ax, ay, az, s, (i - ax) + s * ((j - ay) + s * (k - az)), s * s * s, live, xa, val);
  for (int bit = 0; bit < L; ++bit) {
    acc += __shfl_xor(acc, 1 << bit, 64);
    acc += __shfl_xor(acc, 4 << bit, 64);
    acc += __shfl_xor(acc, 16 << bit, 64);
  }
  return acc;
}

// multires twin of k_pcg_spmv: p_new = z + beta p_old (recomputed at the neighbour
// anchors), Ap = S p_new, partial p.Ap
__global__ void __launch_bounds__(256) k_mr_spmv(VecArgs a, MrOp op, const float* __restrict__ z, float* __restrict__ pA,
                                                 float* __restrict__ pB, float* __restrict__ Ap) {
  if (a.sc->done) return;
  const int lane = threadIdx.x & 63;
  const bool odd = a.sc->iter & 1;
  const float* __restrict__ pold = odd ? pB : pA;
  float* __restrict__ pnew = odd ? pA : pB;
  const float beta = (float)a.sc->beta;
  double pap = 0.0;
  FOR_LIST_BLOCKS(a, b, lane) {
    const long long c = b * kBlockVol + lane;
    const float pc = fmaf(beta, pold[c], z[c]);
    float ap = 0.f;
    {
      // every anchor: a non-FLUID one has no couplings and cd = 0, so Ap = 0 there
      int bx, by, bz;
      block_coords(op.m.dc, b, bx, by, bz);
      const int i = bx * 4 + (lane & 3), j = by * 4 + ((lane >> 2) & 3), k = bz * 4 + (lane >> 4);
      const bool in = i < op.m.g.nx && j < op.m.g.ny && k < op.m.g.nz;
      auto pval = [&](long long n) { return fmaf(beta, pold[n], z[n]); };
      const int Lb = mr_level(op.m, bx * 4, by * 4, bz * 4);
      if (Lb >= 0 && Lb <= 2) {
        const int s = 1 << Lb;
        const long long anc = bm_index(op.m.dc, i & ~(s - 1), j & ~(s - 1), k & ~(s - 1));
        const float sum = mr_apply_row_coop<float>(op.m, op.cf, i, j, k, Lb, in, in ? pval(anc) : 0.f, pval);
        if (in && ((i | j | k) & (s - 1)) == 0) ap = op.cd[c] * pc + sum;
      } else {
        int L;
        if (in && mr_is_anchor(op.m, i, j, k, L))
          ap = op.cd[c] * pc + mr_apply_row<float>(op.m, op.cf, i, j, k, 1 << L, pc, pval);
      }
    }
    pnew[c] = pc;
    Ap[c] = ap;
    pap += (double)pc * (double)ap;
  }
  double v[1] = {pap};
  const bool mx[1] = {false};
  block_reduce_store<1>(v, mx, a.part, kMaxPartials);
}

// FP64 residual r = V rhs - S x; partial sums of r^2, max |r / V| (the reference's
// residual), and (slot 2) sum (V rhs)^2
__global__ void __launch_bounds__(256) k_mr_residual64(VecArgs a, MrOp64 op, const double* __restrict__ b64,
                                                       const double* __restrict__ x, double* __restrict__ r) {
  const int lane = threadIdx.x & 63;
  double rr = 0.0, rm = 0.0;
  FOR_LIST_BLOCKS(a, b, lane) {
    const long long c = b * kBlockVol + lane;
    double res = 0.0;
    int bx, by, bz;
    block_coords(op.m.dc, b, bx, by, bz);
    const int i = bx * 4 + (lane & 3), j = by * 4 + ((lane >> 2) & 3), k = bz * 4 + (lane >> 4);
    auto xval = [&](long long n) { return x[n]; };
    const int Lb = mr_level(op.m, bx * 4, by * 4, bz * 4);
    if (Lb >= 0 && Lb <= 2) {
      // the cell's lanes share its patches; vol is non-zero on FLUID anchors only
      const int s = 1 << Lb;
      const bool in = i < op.m.g.nx && j < op.m.g.ny && k < op.m.g.nz;
      const long long anc = bm_index(op.m.dc, i & ~(s - 1), j & ~(s - 1), k & ~(s - 1));
      const bool live = in && op.vol[anc] != 0.0;
      const double sum = mr_apply_row_coop<double>(op.m, op.cf, i, j, k, Lb, live, live ? x[anc] : 0.0, xval);
      if (live && anc == c) {
        res = b64[c] - (op.cd[c] * x[c] + sum);
        rm = fmax(rm, fabs(res / op.vol[c]));
      }
    } else if (op.vol[c] != 0.0) {
      int L;
      mr_is_anchor(op.m, i, j, k, L);
      const double xc = x[c];
      const double sx = op.cd[c] * xc + mr_apply_row<double>(op.m, op.cf, i, j, k, 1 << L, xc, xval);
      res = b64[c] - sx;
      rm = fmax(rm, fabs(res / op.vol[c]));
    }
    r[c] = res;
    rr += res * res;
  }
  double v[2] = {rr, rm};
  const bool mx[2] = {false, true};
  block_reduce_store<2>(v, mx, a.part, kMaxPartials);
}

// pressure per fine cell: the solution at the anchor of its leaf cell on FLUID cells
__global__ void __launch_bounds__(256) k_mr_store_pressure(MrGrid m, const double* __restrict__ x64, float* __restrict__ p,
                                                           double* __restrict__ pfull, uint8_t* __restrict__ act) {
  const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
  if (q >= m.dc.nblocks * kBlockVol) return;
  const long long b = q >> 6;
  const int lane = (int)(q & 63);
  int bx, by, bz;
  block_coords(m.dc, b, bx, by, bz);
  const int i = bx * 4 + (lane & 3), j = by * 4 + ((lane >> 2) & 3), k = bz * 4 + (lane >> 4);
  double v = 0.0;
  if (i < m.g.nx && j < m.g.ny && k < m.g.nz && m.marker[q] == CELL_FLUID && mr_level(m, i, j, k) >= 0)
    v = x64[mr_anchor(m, i, j, k)];
  p[q] = (float)v;
  pfull[q] = v;
  if (v != 0.0) act[b] = 1;
}

// face update over the enumerated faces (multires_pressure3d.cpp:1455-1524):
// u = (float)(u - dt beta (p_r - p_l) / distance), distance from the FLUID side's cell
__global__ void __launch_bounds__(256) k_mr_velocity(MrGrid m, int axis, const double* __restrict__ pfull,
                                                     float* __restrict__ vel, const float* __restrict__ mass,
                                                     PhaseK ph, double dt) {
#pragma clang fp contract(off)
  const Dims& fd = mr_fdims(m, axis);
  const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
  if (q >= fd.nblocks * kBlockVol) return;
  if (!m.valid[axis][q]) return;
  const long long b = q >> 6;
  const int lane = (int)(q & 63);
  int bx, by, bz;
  block_coords(fd, b, bx, by, bz);
  const int i = bx * 4 + (lane & 3), j = by * 4 + ((lane >> 2) & 3), k = bz * 4 + (lane >> 4);
  const int li = i - (axis == 0), lj = j - (axis == 1), lk = k - (axis == 2);
  const int lm = mr_marker(m, li, lj, lk), rm = mr_marker(m, i, j, k);
  const bool lf = lm == CELL_FLUID, rf = rm == CELL_FLUID;
  if (!lf && !rf) return;
  if (lm == CELL_SOLID || rm == CELL_SOLID) {
    vel[q] = 0.f;
    return;
  }
  const int Lr = lf ? mr_level(m, li, lj, lk) : mr_level(m, i, j, k);
  const int Lo = lf ? mr_level(m, i, j, k) : mr_level(m, li, lj, lk);
  const double h = m.g.dx * (double)(1 << Lr);
  const double ho = Lo >= 0 ? m.g.dx * (double)(1 << Lo) : h;
  const double dist = 0.5 * h + 0.5 * ho;
  const double pl = lf ? pfull[bm_index(m.dc, li, lj, lk)] : 0.0;
  const double pr = rf ? pfull[bm_index(m.dc, i, j, k)] : 0.0;
  const double beta = beta_of_raw((double)mass[q], ph);
  vel[q] = (float)((double)vel[q] - dt * beta * (pr - pl) / dist);
}

// the particle cell bounding boxes of updateDynamicRefinement (multires_sim3d_tp.cpp:
// 577-615): out[0..5] all particles, out[6..11] gas; min as x, max as -x (atomicMin)
__global__ void __launch_bounds__(256) k_mr_bbox(long long n, PBuf P, Geom g, int* __restrict__ out) {
  __shared__ int s[12];
  if (threadIdx.x < 12) s[threadIdx.x] = 0x7fffffff;
  __syncthreads();
  int v[12];
  for (int q = 0; q < 12; ++q) v[q] = 0x7fffffff;
  for (long long p = (long long)blockIdx.x * 256 + threadIdx.x; p < n; p += (long long)gridDim.x * 256) {
    const double x = P.x[p], y = P.y[p], z = P.z[p];
    const int c[3] = {(int)(x / g.dx), (int)(y / g.dx), (int)(z / g.dx)};
    if (c[0] < 0 || c[0] >= g.nx || c[1] < 0 || c[1] >= g.ny || c[2] < 0 || c[2] >= g.nz) continue;
    for (int a = 0; a < 3; ++a) {
      v[a] = min(v[a], c[a]);
      v[3 + a] = min(v[3 + a], -c[a]);
      if (P.type[p] == 1) {
        v[6 + a] = min(v[6 + a], c[a]);
        v[9 + a] = min(v[9 + a], -c[a]);
      }
    }
  }
  for (int q = 0; q < 12; ++q) {
    int w = v[q];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) w = min(w, __shfl_xor(w, o, 64));
    if ((threadIdx.x & 63) == 0) atomicMin(&s[q], w);
  }
  __syncthreads();
  if (threadIdx.x < 12) atomicMin(&out[threadIdx.x], s[threadIdx.x]);
}

// Bubbles of the multires multi-bubble tank, in domain units (centre, radius).
constexpr int kMaxSceneBubbles = 64;
struct MrScene {
  int nb;
  int water;  // fine cells j < water hold water
  float4 b[kMaxSceneBubbles];
};

// Seeds the multires multi-bubble tank: 8 particles uniform in every leaf cell's part of
// the interior water box [1, n-1) x [1, min(n-1, water)) x [1, n-1), one wave per leaf
// (lane = cell of the leaf's 4^3).  COUNT stores 8 x the seeded cells per leaf; the
// write pass places them from off[leaf] in lane order.  A particle inside a bubble is
// gas (type 1).  Positions are (float)((c0 + U[0,1) (c1 - c0)) dx).
template <bool COUNT>
__global__ void __launch_bounds__(256) k_mr_scene(const int32_t* __restrict__ leaves, int nleaves, Geom g, MrScene sc,
                                                  uint64_t seed, uint32_t* __restrict__ cnt,
                                                  const long long* __restrict__ off, PBuf P) {
#pragma clang fp contract(off)
  const int lf = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (lf >= nleaves) return;  // whole waves leave together
  const int L = leaves[4 * lf], h = 1 << L;
  const int x0 = leaves[4 * lf + 1] * 4 * h + (lane & 3) * h;
  const int y0 = leaves[4 * lf + 2] * 4 * h + ((lane >> 2) & 3) * h;
  const int z0 = leaves[4 * lf + 3] * 4 * h + (lane >> 4) * h;
  const int cx0 = max(x0, 1), cx1 = min(x0 + h, g.nx - 1);
  const int cy0 = max(y0, 1), cy1 = min(min(y0 + h, g.ny - 1), sc.water);
  const int cz0 = max(z0, 1), cz1 = min(z0 + h, g.nz - 1);
  const bool seeded = cx0 < cx1 && cy0 < cy1 && cz0 < cz1;
  const uint64_t m = __ballot(seeded);
  if (COUNT) {
    if (lane == 0) cnt[lf] = 8u * (uint32_t)__popcll(m);
    return;
  }
  if (!seeded) return;
  const long long base = off[lf] + 8ll * __popcll(m & ((1ull << lane) - 1ull));
  const uint64_t cid = (uint64_t)x0 + (uint64_t)g.nx * ((uint64_t)y0 + (uint64_t)g.ny * (uint64_t)z0);
  uint64_t s = seed * 0x9E3779B97F4A7C15ULL ^ (cid + 1) * 0xD1B54A32D192ED03ULL;
  const double dx = g.dx;
  for (int q = 0; q < 8; ++q) {
    const double ux = (double)(splitmix64(s) >> 11) * 0x1.0p-53;
    const double uy = (double)(splitmix64(s) >> 11) * 0x1.0p-53;
    const double uz = (double)(splitmix64(s) >> 11) * 0x1.0p-53;
    const float fx = (float)__dmul_rn(__dadd_rn((double)cx0, __dmul_rn(ux, (double)(cx1 - cx0))), dx);
    const float fy = (float)__dmul_rn(__dadd_rn((double)cy0, __dmul_rn(uy, (double)(cy1 - cy0))), dx);
    const float fz = (float)__dmul_rn(__dadd_rn((double)cz0, __dmul_rn(uz, (double)(cz1 - cz0))), dx);
    uint8_t t = 0;
    for (int b = 0; b < sc.nb; ++b) {
      const double ddx = (double)fx - sc.b[b].x, ddy = (double)fy - sc.b[b].y, ddz = (double)fz - sc.b[b].z;
      if (ddx * ddx + ddy * ddy + ddz * ddz < (double)sc.b[b].w * sc.b[b].w) t = 1;
    }
    const long long p = base + q;
    P.x[p] = fx; P.y[p] = fy; P.z[p] = fz;
    P.u[p] = 0.f; P.v[p] = 0.f; P.w[p] = 0.f;
    P.type[p] = t;
    P.id[p] = (uint32_t)p;
  }
}

}  // namespace pf
