// pf_mr_ref.hip -- the multires pressure solve in the reference's own order, in FP64,
// for the options of MRPressureSolveConfig3D that change the iteration itself
// (multires_pressure3d.h:137-166): no Jacobi, flexible beta, damped-Jacobi
// relaxation sweeps, the level-1 coarse correction and the additive coarse
// preconditioner (multires_pressure3d.cpp:39-70, 470-700, 846-1449).
//
// The host keeps the reference's control flow; every vector operation is a kernel
// over the fluid blocks (fine vectors live at the leaf-cell anchors, zero
// elsewhere) or over the coarse grid.  Rows are the reference's own: A x =
// diag x - sum coeff x_nbr with coeff = cf / V per fine face patch in
// visitCellFaces order, the pinned row the identity, so the operator rounds as the
// reference's applyA does.  Scalars (dots, norms) are fixed-order FP64 reductions.
//
// Coarse space (level1ParentTuple, :39-50): a leaf cell's aggregate is its fine
// anchor's 2x parent, so coarse vectors live on a dense (n/2)^3 grid; aggregate
// volumes, volume-weighted restriction and piecewise-constant prolongation gather
// over each aggregate's 2x2x2 fine footprint (no atomics).  Cells whose aggregate
// holds no FLUID leaf cell carry zero volume and drop out of every weighted dot.
#include "pf_internal.h"

namespace pf {

struct MrfArgs {
  VecArgs va;
  MrOp64 op;
  const double* dref;    // reference diagonal per anchor (1 on the pinned row)
  const double* vtrue;   // h^3 per FLUID anchor (the pinned one included)
  long long pin;         // pinned anchor (block-major index) or -1
};

struct CoarseGrid {
  int hx, hy, hz;        // (n + 1) / 2 per axis
  long long nc;
  Dims dc;
  Geom g;
};

__device__ inline long long mrf_parent(const CoarseGrid& cg, int i, int j, int k) {
  return (long long)(i >> 1) + (long long)cg.hx * ((long long)(j >> 1) + (long long)cg.hy * (long long)(k >> 1));
}

// Ax = A x (rows as above); with RESID also r = rhs - Ax and partial max |r|
// (+inf when an entry is not finite: residualFromPressure, :870-884)
template <bool RESID>
__global__ void __launch_bounds__(256) k_mrf_apply(MrfArgs a, const double* __restrict__ x, double* __restrict__ Ax,
                                                   const double* __restrict__ rhs, double* __restrict__ r) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & 63;
  double rm = 0.0;
  FOR_LIST_BLOCKS(a.va, b, lane) {
    const long long c = b * kBlockVol + lane;
    double out = 0.0;
    const bool anchor = a.op.vol[c] != 0.0;
    if (anchor) {
      if (c == a.pin) {
        out = x[c];
      } else {
        int bx, by, bz;
        block_coords(a.op.m.dc, b, bx, by, bz);
        const int i = bx * 4 + (lane & 3), j = by * 4 + ((lane >> 2) & 3), k = bz * 4 + (lane >> 4);
        int L;
        mr_is_anchor(a.op.m, i, j, k, L);
        const double V = a.vtrue[c];
        double off = 0.0;
        mr_patches(a.op.m, i, j, k, 1 << L, [&](int axis, int fx, int fy, int fz, int ni, int nj, int nk, double) {
          const long long fi = bm_index(mr_fdims(a.op.m, axis), fx, fy, fz);
          const double cf = a.op.cf[axis][fi];
          if (cf != 0.0) off += (cf / V) * x[mr_anchor(a.op.m, ni, nj, nk)];
        });
        out = a.dref[c] * x[c] - off;
      }
    }
    Ax[c] = out;
    if (RESID) {
      const double e = anchor ? rhs[c] - out : 0.0;
      r[c] = e;
      if (anchor) rm = isfinite(e) ? fmax(rm, fabs(e)) : INFINITY;
    }
  }
  if (RESID) {
    double v[1] = {rm};
    const bool mx[1] = {true};
    block_reduce_store<1>(v, mx, a.va.part, kMaxPartials);
  }
}

enum MrfVecOp {
  MRF_MAXABS = 0,      // slot 0: max |a|
  MRF_JACOBI = 1,      // z = (jacobi && d > 0) ? r / d : r; slot 0: r.z
  MRF_XR = 2,          // x += s0 p; r -= s0 Ap; slot 0: max |r| over finite r, slot 1: non-finite x or r count
  MRF_P = 3,           // p = z + s0 p
  MRF_AXPY = 4,        // out = a + s0 b
  MRF_DOT = 5,         // slot 0: a.b
  MRF_FLEX = 6,        // slot 0: r.(z - zp)
  MRF_COPY = 7         // out = a
};

struct MrfVec {
  int op, jacobi;
  double s0;
  const double *a, *b, *c;
  double *out, *out2;
};

// the reference's vector loops over the FLUID cells (non-anchors stay zero)
__global__ void __launch_bounds__(256) k_mrf_vec(MrfArgs a, MrfVec v) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & 63;
  double s0 = 0.0, s1 = 0.0;
  FOR_LIST_BLOCKS(a.va, b, lane) {
    const long long c = b * kBlockVol + lane;
    if (a.op.vol[c] == 0.0) continue;
    switch (v.op) {
      case MRF_MAXABS: s0 = fmax(s0, fabs(v.a[c])); break;
      case MRF_JACOBI: {
        const double d = a.dref[c];
        const double z = (v.jacobi && d > 0.0) ? v.a[c] / d : v.a[c];
        v.out[c] = z;
        s0 += v.a[c] * z;
        break;
      }
      case MRF_XR: {
        // v.out = x, v.out2 = r, v.a = p, v.b = Ap
        const double x = v.out[c] + v.s0 * v.a[c];
        const double r = v.out2[c] - v.s0 * v.b[c];
        v.out[c] = x;
        v.out2[c] = r;
        if (!isfinite(x) || !isfinite(r)) s1 += 1.0;
        if (isfinite(r)) s0 = fmax(s0, fabs(r));
        break;
      }
      case MRF_P: v.out[c] = v.a[c] + v.s0 * v.out[c]; break;
      case MRF_AXPY: v.out[c] = v.a[c] + v.s0 * v.b[c]; break;
      case MRF_DOT: s0 += v.a[c] * v.b[c]; break;
      case MRF_FLEX: s0 += v.a[c] * (v.b[c] - v.c[c]); break;
      case MRF_COPY: v.out[c] = v.a[c]; break;
    }
  }
  double r2[2] = {s0, s1};
  const bool mx[2] = {v.op == MRF_MAXABS || v.op == MRF_XR, false};
  block_reduce_store<2>(r2, mx, a.va.part, kMaxPartials);
}

// coarse loop with a fixed workgroup mapping (deterministic partials)
#define FOR_COARSE(cg, C) \
  for (long long C = (long long)blockIdx.x * 256 + threadIdx.x; C < (cg).nc; C += (long long)gridDim.x * 256)

// aggregate volumes (buildMRPressureAggregation3D, :470-505) and the aggregate count
__global__ void __launch_bounds__(256) k_mrf_coarse_volume(CoarseGrid cg, const double* __restrict__ vol64,
                                                           const double* __restrict__ vtrue, double* __restrict__ Vc,
                                                           double* __restrict__ part) {
#pragma clang fp contract(off)
  double cnt = 0.0;
  FOR_COARSE(cg, C) {
    const int I = (int)(C % cg.hx), J = (int)((C / cg.hx) % cg.hy), K = (int)(C / ((long long)cg.hx * cg.hy));
    double v = 0.0;
    for (int d = 0; d < 8; ++d) {
      const int i = 2 * I + (d & 1), j = 2 * J + ((d >> 1) & 1), k = 2 * K + (d >> 2);
      if (i >= cg.g.nx || j >= cg.g.ny || k >= cg.g.nz) continue;
      const long long f = bm_index(cg.dc, i, j, k);
      if (vol64[f] != 0.0) v += vtrue[f];
    }
    Vc[C] = v;
    if (v > 0.0) cnt += 1.0;
  }
  double r[1] = {cnt};
  const bool mx[1] = {false};
  block_reduce_store<1>(r, mx, part, kMaxPartials);
}

// restrictMRPressureVolumeWeighted3D (:518-539)
__global__ void __launch_bounds__(256) k_mrf_restrict(CoarseGrid cg, const double* __restrict__ vol64,
                                                      const double* __restrict__ vtrue, const double* __restrict__ Vc,
                                                      const double* __restrict__ fine, double* __restrict__ coarse) {
#pragma clang fp contract(off)
  FOR_COARSE(cg, C) {
    const int I = (int)(C % cg.hx), J = (int)((C / cg.hx) % cg.hy), K = (int)(C / ((long long)cg.hx * cg.hy));
    double s = 0.0;
    for (int d = 0; d < 8; ++d) {
      const int i = 2 * I + (d & 1), j = 2 * J + ((d >> 1) & 1), k = 2 * K + (d >> 2);
      if (i >= cg.g.nx || j >= cg.g.ny || k >= cg.g.nz) continue;
      const long long f = bm_index(cg.dc, i, j, k);
      if (vol64[f] != 0.0) s += vtrue[f] * fine[f];
    }
    coarse[C] = Vc[C] > 0.0 ? s / Vc[C] : 0.0;
  }
}

// prolongMRPressurePiecewiseConstant3D (:541-552) onto the FLUID anchors
__global__ void __launch_bounds__(256) k_mrf_prolong(MrfArgs a, CoarseGrid cg, const double* __restrict__ coarse,
                                                     double* __restrict__ fine) {
  const int lane = threadIdx.x & 63;
  FOR_LIST_BLOCKS(a.va, b, lane) {
    const long long c = b * kBlockVol + lane;
    double v = 0.0;
    if (a.op.vol[c] != 0.0) {
      int bx, by, bz;
      block_coords(a.op.m.dc, b, bx, by, bz);
      v = coarse[mrf_parent(cg, bx * 4 + (lane & 3), by * 4 + ((lane >> 2) & 3), bz * 4 + (lane >> 4))];
    }
    fine[c] = v;
  }
}

enum MrfCoarseOp {
  MRC_INIT = 0,    // x = 0, r = rhs, p = r, (pin) r = p = 0; slot 0: (r,r)_V
  MRC_XR = 1,      // x += s0 p, r -= s0 Ap, (pin) x = r = 0; slot 0: (r,r)_V
  MRC_P = 2,       // p = r + s0 p, (pin) p = 0
  MRC_DOT = 3,     // slot 0: (a,b)_V
  MRC_PINSET = 4   // out[pin] = a[pin] (the coarse pin row)
};

struct MrcVec {
  int op;
  long long pin;
  double s0;
  const double *Vc, *a, *b;
  double *x, *r, *p;
};

__global__ void __launch_bounds__(256) k_mrf_coarse(CoarseGrid cg, MrcVec v, double* __restrict__ part) {
#pragma clang fp contract(off)
  double s0 = 0.0;
  FOR_COARSE(cg, C) {
    const double w = v.Vc[C];
    switch (v.op) {
      case MRC_INIT: {
        const double rv = C == v.pin ? 0.0 : v.a[C];
        v.x[C] = 0.0;
        v.r[C] = rv;
        v.p[C] = rv;
        s0 += w * rv * rv;
        break;
      }
      case MRC_XR: {
        double x = v.x[C] + v.s0 * v.p[C];
        double r = v.r[C] - v.s0 * v.a[C];
        if (C == v.pin) { x = 0.0; r = 0.0; }
        v.x[C] = x;
        v.r[C] = r;
        s0 += w * r * r;
        break;
      }
      case MRC_P: v.p[C] = C == v.pin ? 0.0 : v.r[C] + v.s0 * v.p[C]; break;
      case MRC_DOT: s0 += w * v.a[C] * v.b[C]; break;
      case MRC_PINSET: if (C == v.pin) v.x[C] = v.a[C]; break;
    }
  }
  double r[1] = {s0};
  const bool mx[1] = {false};
  block_reduce_store<1>(r, mx, part, kMaxPartials);
}

}  // namespace pf
