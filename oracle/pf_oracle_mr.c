/*
 * pf_oracle_mr.c -- CPU restatement of the reference's 3D multires two-phase path
 * (TEST INFRASTRUCTURE ONLY; see pf_oracle.h for who may load this library).
 *
 * Follows: src/grid/multires_layout3d.h (2-level leaf layout of 4^3 blocks),
 * src/grid/multires_scalar_grid3d.h (per-leaf-cell scalars), src/grid/multires_mac_grid3d.h
 * (1x1 fine face patches on leaf-cell boundaries), src/driver/multires_ops3d_tp.cpp
 * (P2G / G2P / advection), src/pressure/multires_pressure3d.cpp (finite-volume
 * projection, default solve path: Jacobi PCG with plain dots, adaptive restart,
 * flexible beta, damped-Jacobi pre-relaxation; the coarse correction / coarse
 * preconditioner options are not restated), src/driver/multires_sim3d_tp.cpp
 * (markCells, gravity, walls, dynamic refinement, step).
 *
 * Representation: fine-resolution dense arrays.  A leaf cell (level L, step 2^L fine
 * cells) is identified by (leaf block index, lx, ly, lz); scalars (marker, pressure)
 * are stored per leaf cell; face fields per fine face with a validity flag for the
 * faces the reference enumerates.  faithful = 1 rounds every store through float
 * where the reference stores float.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "pf_oracle.h"

#define MRB 4

struct orc_mr {
  int nx, ny, nz;
  double dx;
  int faithful;
  int nleaves;
  int32_t* leaves;      /* [4 * nleaves]: level, bx, by, bz (sorted, unique) */
  int32_t* leaf_of;     /* per fine cell: leaf block index or -1 */
  double* mk;           /* per leaf cell id (leaf * 64 + lx + 4 (ly + 4 lz)) */
  double* p;
  double* f[6];         /* u v w mu mv mw, fine face arrays */
  uint8_t* valid[3];    /* enumerated faces per axis */
  int64_t fn[3];        /* face array sizes */
};

static inline double stf(const orc_mr* g, double v) { return g->faithful ? (double)(float)v : v; }

static inline int64_t cidx(const orc_mr* g, int i, int j, int k) {
  return i + (int64_t)g->nx * (j + (int64_t)g->ny * k);
}
static inline int64_t fidx(const orc_mr* g, int axis, int i, int j, int k) {
  const int fx = g->nx + (axis == 0), fy = g->ny + (axis == 1);
  return i + (int64_t)fx * (j + (int64_t)fy * k);
}

static int leaf_cmp(const void* a, const void* b) {
  const int32_t *x = (const int32_t*)a, *y = (const int32_t*)b;
  for (int q = 0; q < 4; ++q) {
    if (x[q] < y[q]) return -1;
    if (x[q] > y[q]) return 1;
  }
  return 0;
}

/* ---------------------------------------------------------------- layout */

static int block_intersects(int nx, int ny, int nz, const int32_t* key) {
  /* MRLayout3D::intersectsDomain, multires_layout3d.h:137-153 */
  int s = MRB * (1 << key[0]);
  int x0 = key[1] * s, y0 = key[2] * s, z0 = key[3] * s;
  int cx0 = x0 > 0 ? x0 : 0, cy0 = y0 > 0 ? y0 : 0, cz0 = z0 > 0 ? z0 : 0;
  int cx1 = x0 + s < nx ? x0 + s : nx, cy1 = y0 + s < ny ? y0 + s : ny, cz1 = z0 + s < nz ? z0 + s : nz;
  return cx0 < cx1 && cy0 < cy1 && cz0 < cz1;
}

static int sort_unique(int32_t* lv, int n) {
  qsort(lv, (size_t)n, 4 * sizeof(int32_t), leaf_cmp);
  int m = 0;
  for (int a = 0; a < n; ++a) {
    if (m > 0 && leaf_cmp(lv + 4 * (m - 1), lv + 4 * a) == 0) continue;
    memmove(lv + 4 * m, lv + 4 * a, 4 * sizeof(int32_t));
    ++m;
  }
  return m;
}

/* setCoarseEverywhere(level) then refineFineCellBox(box) (multires_layout3d.h:51-97),
 * i.e. layoutWithFineBox (multires_sim3d_tp.cpp:172-180) for a level-1 base.
 * Returns the leaf count; writes at most cap leaves. */
int64_t orc_mr_layout_fine_box(int nx, int ny, int nz, int level, int box_valid, const int32_t box[6],
                               int32_t* out, int64_t cap) {
  int s = MRB * (1 << level);
  int bxn = (nx + s - 1) / s, byn = (ny + s - 1) / s, bzn = (nz + s - 1) / s;
  int64_t n0 = (int64_t)bxn * byn * bzn;
  int32_t* cur = (int32_t*)malloc(sizeof(int32_t) * 4 * (size_t)(n0 > 0 ? n0 : 1));
  int64_t n = 0;
  for (int bz = 0; bz < bzn; ++bz)
    for (int by = 0; by < byn; ++by)
      for (int bx = 0; bx < bxn; ++bx) {
        cur[4 * n] = level; cur[4 * n + 1] = bx; cur[4 * n + 2] = by; cur[4 * n + 3] = bz;
        ++n;
      }
  int x0 = 0, y0 = 0, z0 = 0, x1 = 0, y1 = 0, z1 = 0;
  if (box_valid) {
    /* refineBox clamps to the domain and skips empty boxes (multires_sim3d_tp.cpp:158-170) */
    x0 = box[0] < 0 ? 0 : (box[0] > nx ? nx : box[0]);
    y0 = box[1] < 0 ? 0 : (box[1] > ny ? ny : box[1]);
    z0 = box[2] < 0 ? 0 : (box[2] > nz ? nz : box[2]);
    x1 = box[3] < 0 ? 0 : (box[3] > nx ? nx : box[3]);
    y1 = box[4] < 0 ? 0 : (box[4] > ny ? ny : box[4]);
    z1 = box[5] < 0 ? 0 : (box[5] > nz ? nz : box[5]);
    if (!(x0 < x1 && y0 < y1 && z0 < z1)) box_valid = 0;
  }
  if (box_valid) {
    int32_t* next = (int32_t*)malloc(sizeof(int32_t) * 4 * (size_t)(8 * n + 1));
    int64_t m = 0;
    for (int64_t a = 0; a < n; ++a) {
      const int32_t* key = cur + 4 * a;
      int ss = MRB * (1 << key[0]);
      int bx0 = key[1] * ss, by0 = key[2] * ss, bz0 = key[3] * ss;
      int bx1 = bx0 + ss < nx ? bx0 + ss : nx, by1 = by0 + ss < ny ? by0 + ss : ny;
      int bz1 = bz0 + ss < nz ? bz0 + ss : nz;
      int overlaps = bx0 < x1 && bx1 > x0 && by0 < y1 && by1 > y0 && bz0 < z1 && bz1 > z0;
      if (!overlaps || key[0] == 0) {
        memcpy(next + 4 * m, key, 4 * sizeof(int32_t));
        ++m;
        continue;
      }
      for (int cz = 0; cz < 2; ++cz)
        for (int cy = 0; cy < 2; ++cy)
          for (int cx = 0; cx < 2; ++cx) {
            int32_t child[4] = {key[0] - 1, key[1] * 2 + cx, key[2] * 2 + cy, key[3] * 2 + cz};
            if (block_intersects(nx, ny, nz, child)) {
              memcpy(next + 4 * m, child, 4 * sizeof(int32_t));
              ++m;
            }
          }
    }
    free(cur);
    cur = next;
    n = m;
  }
  n = sort_unique(cur, (int)n);
  if (out) memcpy(out, cur, sizeof(int32_t) * 4 * (size_t)(n < cap ? n : cap));
  free(cur);
  return n;
}

static void enumerate_faces(orc_mr* g);

orc_mr* orc_mr_create(int nx, int ny, int nz, double dx, int faithful) {
  orc_mr* g = (orc_mr*)calloc(1, sizeof(orc_mr));
  g->nx = nx; g->ny = ny; g->nz = nz; g->dx = dx; g->faithful = faithful;
  int64_t nc = (int64_t)nx * ny * nz;
  g->leaf_of = (int32_t*)malloc(sizeof(int32_t) * (size_t)nc);
  for (int64_t c = 0; c < nc; ++c) g->leaf_of[c] = -1;
  g->fn[0] = (int64_t)(nx + 1) * ny * nz;
  g->fn[1] = (int64_t)nx * (ny + 1) * nz;
  g->fn[2] = (int64_t)nx * ny * (nz + 1);
  for (int q = 0; q < 6; ++q) g->f[q] = (double*)calloc((size_t)g->fn[q % 3], sizeof(double));
  for (int a = 0; a < 3; ++a) g->valid[a] = (uint8_t*)calloc((size_t)g->fn[a], 1);
  g->leaves = (int32_t*)malloc(4 * sizeof(int32_t));
  g->mk = (double*)calloc(1, sizeof(double));
  g->p = (double*)calloc(1, sizeof(double));
  return g;
}

void orc_mr_destroy(orc_mr* g) {
  if (!g) return;
  free(g->leaves); free(g->leaf_of); free(g->mk); free(g->p);
  for (int q = 0; q < 6; ++q) free(g->f[q]);
  for (int a = 0; a < 3; ++a) free(g->valid[a]);
  free(g);
}

orc_mr* orc_mr_clone(const orc_mr* s) {
  orc_mr* g = orc_mr_create(s->nx, s->ny, s->nz, s->dx, s->faithful);
  int32_t* lv = (int32_t*)malloc(sizeof(int32_t) * 4 * (size_t)(s->nleaves + 1));
  memcpy(lv, s->leaves, sizeof(int32_t) * 4 * (size_t)s->nleaves);
  orc_mr_set_leaves(g, s->nleaves, lv);
  free(lv);
  memcpy(g->mk, s->mk, sizeof(double) * 64 * (size_t)s->nleaves);
  memcpy(g->p, s->p, sizeof(double) * 64 * (size_t)s->nleaves);
  for (int q = 0; q < 6; ++q) memcpy(g->f[q], s->f[q], sizeof(double) * (size_t)g->fn[q % 3]);
  return g;
}

/* a new layout: MRMacGrid3D<4>(layout) -- every field starts empty */
void orc_mr_set_leaves(orc_mr* g, int n, const int32_t* leaves) {
  free(g->leaves); free(g->mk); free(g->p);
  g->leaves = (int32_t*)malloc(sizeof(int32_t) * 4 * (size_t)(n > 0 ? n : 1));
  memcpy(g->leaves, leaves, sizeof(int32_t) * 4 * (size_t)n);
  g->nleaves = sort_unique(g->leaves, n);
  g->mk = (double*)calloc((size_t)64 * (g->nleaves > 0 ? g->nleaves : 1), sizeof(double));
  g->p = (double*)calloc((size_t)64 * (g->nleaves > 0 ? g->nleaves : 1), sizeof(double));
  int64_t nc = (int64_t)g->nx * g->ny * g->nz;
  for (int64_t c = 0; c < nc; ++c) g->leaf_of[c] = -1;
  /* leafAtFineCell returns the first leaf (in sorted order) containing the cell */
  for (int a = g->nleaves - 1; a >= 0; --a) {
    const int32_t* key = g->leaves + 4 * a;
    int s = MRB * (1 << key[0]);
    int x0 = key[1] * s, y0 = key[2] * s, z0 = key[3] * s;
    for (int k = z0; k < z0 + s && k < g->nz; ++k)
      for (int j = y0; j < y0 + s && j < g->ny; ++j)
        for (int i = x0; i < x0 + s && i < g->nx; ++i)
          if (i >= 0 && j >= 0 && k >= 0) g->leaf_of[cidx(g, i, j, k)] = a;
  }
  for (int q = 0; q < 6; ++q) memset(g->f[q], 0, sizeof(double) * (size_t)g->fn[q % 3]);
  enumerate_faces(g);
}

int64_t orc_mr_leaf_count(const orc_mr* g) { return g->nleaves; }
void orc_mr_get_leaves(const orc_mr* g, int32_t* out) { memcpy(out, g->leaves, sizeof(int32_t) * 4 * (size_t)g->nleaves); }

/* leaf cell of a fine cell (cellAtFineCell, multires_scalar_grid3d.h:41-51): id or -1 */
static inline int64_t leaf_cell(const orc_mr* g, int i, int j, int k) {
  if (i < 0 || j < 0 || k < 0 || i >= g->nx || j >= g->ny || k >= g->nz) return -1;
  int a = g->leaf_of[cidx(g, i, j, k)];
  if (a < 0) return -1;
  const int32_t* key = g->leaves + 4 * a;
  int step = 1 << key[0];
  int lx = (i - key[1] * MRB * step) / step, ly = (j - key[2] * MRB * step) / step;
  int lz = (k - key[3] * MRB * step) / step;
  return (int64_t)a * 64 + lx + MRB * (ly + MRB * lz);
}

static inline int cell_level(const orc_mr* g, int64_t id) { return g->leaves[4 * (id / 64)]; }

static inline void cell_box(const orc_mr* g, int64_t id, int* x0, int* y0, int* z0, int* step) {
  const int32_t* key = g->leaves + 4 * (id / 64);
  int s = 1 << key[0];
  int l = (int)(id % 64);
  *step = s;
  *x0 = key[1] * MRB * s + (l & 3) * s;
  *y0 = key[2] * MRB * s + ((l >> 2) & 3) * s;
  *z0 = key[3] * MRB * s + (l >> 4) * s;
}

/* MRMacGrid3D::enumerateFaces, multires_mac_grid3d.h:107-149 */
static void enumerate_faces(orc_mr* g) {
  for (int a = 0; a < 3; ++a) memset(g->valid[a], 0, (size_t)g->fn[a]);
  for (int a = 0; a < g->nleaves; ++a)
    for (int l = 0; l < 64; ++l) {
      int x0, y0, z0, s;
      cell_box(g, (int64_t)a * 64 + l, &x0, &y0, &z0, &s);
      int x1 = x0 + s, y1 = y0 + s, z1 = z0 + s;
      if (!(x0 < g->nx && x1 > 0 && y0 < g->ny && y1 > 0 && z0 < g->nz && z1 > 0)) continue;
      int sx0 = x0 > 0 ? x0 : 0, sy0 = y0 > 0 ? y0 : 0, sz0 = z0 > 0 ? z0 : 0;
      int sx1 = x1 < g->nx ? x1 : g->nx, sy1 = y1 < g->ny ? y1 : g->ny, sz1 = z1 < g->nz ? z1 : g->nz;
      int xs[2] = {x0, x1 < g->nx ? x1 : g->nx}, ys[2] = {y0, y1 < g->ny ? y1 : g->ny};
      int zs[2] = {z0, z1 < g->nz ? z1 : g->nz};
      for (int h = 0; h < 2; ++h) {
        if (xs[h] >= 0 && xs[h] <= g->nx)
          for (int z = sz0; z < sz1; ++z)
            for (int y = sy0; y < sy1; ++y) g->valid[0][fidx(g, 0, xs[h], y, z)] = 1;
        if (ys[h] >= 0 && ys[h] <= g->ny)
          for (int z = sz0; z < sz1; ++z)
            for (int x = sx0; x < sx1; ++x) g->valid[1][fidx(g, 1, x, ys[h], z)] = 1;
        if (zs[h] >= 0 && zs[h] <= g->nz)
          for (int y = sy0; y < sy1; ++y)
            for (int x = sx0; x < sx1; ++x) g->valid[2][fidx(g, 2, x, y, zs[h])] = 1;
      }
    }
}

void orc_mr_face_valid(const orc_mr* g, int axis, uint8_t* out) { memcpy(out, g->valid[axis], (size_t)g->fn[axis]); }

/* field access: faces (ORC_U..ORC_MW) dense fine arrays; ORC_P / ORC_MK per fine cell
 * (the value of its leaf cell) */
void orc_mr_field_dense(const orc_mr* g, int field, double* out) {
  if (field <= ORC_MW) {
    memcpy(out, g->f[field], sizeof(double) * (size_t)g->fn[field % 3]);
    return;
  }
  for (int k = 0; k < g->nz; ++k)
    for (int j = 0; j < g->ny; ++j)
      for (int i = 0; i < g->nx; ++i) {
        int64_t id = leaf_cell(g, i, j, k);
        out[cidx(g, i, j, k)] = id < 0 ? 0.0 : (field == ORC_P ? g->p[id] : g->mk[id]);
      }
}

/* test hooks: set a face value, or the marker of the leaf cell holding fine cell (i,j,k) */
void orc_mr_set(orc_mr* g, int field, int i, int j, int k, double v) {
  if (field <= ORC_MW) {
    g->f[field][fidx(g, field % 3, i, j, k)] = stf(g, v);
    return;
  }
  int64_t id = leaf_cell(g, i, j, k);
  if (id < 0) return;
  if (field == ORC_P) g->p[id] = stf(g, v);
  else g->mk[id] = stf(g, v);
}

static inline int marker_fine(const orc_mr* g, int i, int j, int k) {
  /* markerAtFineCell, multires_pressure3d.cpp:81-88 */
  int64_t id = leaf_cell(g, i, j, k);
  if (id < 0) return 2;
  return (int)(g->mk[id] + 0.5);
}

/* ---------------------------------------------------------------- stages */

/* markCells, multires_sim3d_tp.cpp:39-70 */
void orc_mr_mark_cells(orc_mr* g, int64_t n, const double* x, const double* y, const double* z) {
  memset(g->mk, 0, sizeof(double) * 64 * (size_t)(g->nleaves > 0 ? g->nleaves : 1));
  int nx = g->nx, ny = g->ny, nz = g->nz;
#define SETC(I, J, K, V) do { int64_t _id = leaf_cell(g, (I), (J), (K)); if (_id >= 0) g->mk[_id] = (V); } while (0)
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j) { SETC(0, j, k, 2.0); SETC(nx - 1, j, k, 2.0); }
  for (int k = 0; k < nz; ++k)
    for (int i = 0; i < nx; ++i) { SETC(i, 0, k, 2.0); SETC(i, ny - 1, k, 2.0); }
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) { SETC(i, j, 0, 2.0); SETC(i, j, nz - 1, 2.0); }
  for (int64_t p = 0; p < n; ++p) {
    int i = (int)(x[p] / g->dx), j = (int)(y[p] / g->dx), k = (int)(z[p] / g->dx);
    if (i >= 0 && i < nx && j >= 0 && j < ny && k >= 0 && k < nz) {
      int64_t id = leaf_cell(g, i, j, k);
      int cur = id < 0 ? 2 : (int)(g->mk[id] + 0.5);  /* cell(): absent leaf reads as 0 -> AIR */
      if (id < 0) cur = 0;
      if (cur != 2 && id >= 0) g->mk[id] = 1.0;
    }
  }
#undef SETC
}

static inline double kern15(double d2) {
  double q = d2 / (1.5 * 1.5);
  double t = 1.0 - q;
  return t > 0.0 ? t * t * t : 0.0;
}
static inline int lo_range(double c, int lo, int hi) {
  int v = (int)floor(c - 1.5);
  return v < lo ? lo : (v > hi ? hi : v);
}
static inline int hi_range(double c, int lo, int hi) {
  int v = (int)ceil(c + 1.5);
  return v < lo ? lo : (v > hi ? hi : v);
}

/* mrP2G3D_tp, multires_ops3d_tp.cpp:213-278 with collectWeightedFaces (:62-137):
 * normalisation over the enumerated faces with positive weight */
void orc_mr_p2g(orc_mr* g, int64_t n, const double* x, const double* y, const double* z, const double* u,
                const double* v, const double* w, const uint8_t* type, const double* vol,
                const orc_phase* pp, double Vp) {
  for (int q = 0; q < 6; ++q) memset(g->f[q], 0, sizeof(double) * (size_t)g->fn[q % 3]);
  int64_t* fl = (int64_t*)malloc(sizeof(int64_t) * 512);
  double* fw = (double*)malloc(sizeof(double) * 512);
  for (int64_t p = 0; p < n; ++p) {
    double rho = type[p] == 0 ? pp->rho_l : pp->rho_g;
    double mp = rho * Vp * (vol ? vol[p] : 1.0);
    double px = x[p] / g->dx, py = y[p] / g->dx, pz = z[p] / g->dx;
    const double vel[3] = {u[p], v[p], w[p]};
    for (int axis = 0; axis < 3; ++axis) {
      double sx = axis == 0 ? 0.0 : 0.5, sy = axis == 1 ? 0.0 : 0.5, sz = axis == 2 ? 0.0 : 0.5;
      int hx = g->nx - (axis == 0 ? 0 : 1), hy = g->ny - (axis == 1 ? 0 : 1), hz = g->nz - (axis == 2 ? 0 : 1);
      int ix0 = lo_range(px - sx, 0, hx), ix1 = hi_range(px - sx, 0, hx);
      int iy0 = lo_range(py - sy, 0, hy), iy1 = hi_range(py - sy, 0, hy);
      int iz0 = lo_range(pz - sz, 0, hz), iz1 = hi_range(pz - sz, 0, hz);
      int m = 0;
      double sum = 0.0;
      for (int k = iz0; k <= iz1; ++k)
        for (int j = iy0; j <= iy1; ++j)
          for (int i = ix0; i <= ix1; ++i) {
            double ddx = px - ((double)i + sx), ddy = py - ((double)j + sy), ddz = pz - ((double)k + sz);
            double wt = kern15(ddx * ddx + ddy * ddy + ddz * ddz);
            if (wt <= 0.0) continue;
            int64_t fi = fidx(g, axis, i, j, k);
            if (!g->valid[axis][fi]) continue;
            fl[m] = fi;
            fw[m] = wt;
            ++m;
            sum += wt;
          }
      if (!(sum > 0.0)) continue;
      double* V = g->f[axis];
      double* M = g->f[3 + axis];
      for (int q = 0; q < m; ++q) {
        double wn = fw[q] / sum;
        if (wn <= 0.0) continue;
        V[fl[q]] = g->faithful ? (double)((float)V[fl[q]] + (float)(wn * mp * vel[axis])) : V[fl[q]] + wn * mp * vel[axis];
        M[fl[q]] = g->faithful ? (double)((float)M[fl[q]] + (float)(wn * mp)) : M[fl[q]] + wn * mp;
      }
    }
  }
  free(fl);
  free(fw);
  for (int axis = 0; axis < 3; ++axis)
    for (int64_t fi = 0; fi < g->fn[axis]; ++fi) {
      double m = g->f[3 + axis][fi];
      if (m > 0.0)
        g->f[axis][fi] = g->faithful ? (double)((float)g->f[axis][fi] / (float)m) : g->f[axis][fi] / m;
    }
}

/* applyGravity / applyWallBoundary, multires_sim3d_tp.cpp:72-97 (over enumerated faces) */
void orc_mr_apply_gravity(orc_mr* g, double dt, double gravity) {
  double dv = dt * gravity;
  for (int64_t fi = 0; fi < g->fn[1]; ++fi)
    if (g->valid[1][fi] && g->f[ORC_MV][fi] > 0.0)
      g->f[ORC_V][fi] = g->faithful ? (double)((float)g->f[ORC_V][fi] + (float)dv) : g->f[ORC_V][fi] + dv;
}

void orc_mr_apply_walls(orc_mr* g) {
  for (int axis = 0; axis < 3; ++axis) {
    int n = axis == 0 ? g->nx : (axis == 1 ? g->ny : g->nz);
    int fx = g->nx + (axis == 0), fy = g->ny + (axis == 1), fz = g->nz + (axis == 2);
    for (int k = 0; k < fz; ++k)
      for (int j = 0; j < fy; ++j)
        for (int i = 0; i < fx; ++i) {
          int64_t fi = fidx(g, axis, i, j, k);
          if (!g->valid[axis][fi]) continue;
          int c = axis == 0 ? i : (axis == 1 ? j : k);
          if (c == 0 || c == 1 || c == n - 1 || c == n) g->f[axis][fi] = 0.0;
        }
  }
}

/* sampleU/V/W, multires_ops3d_tp.cpp:139-211: trilinear over fine faces, indices
 * clamped, absent faces read 0 */
static double mr_sample(const orc_mr* g, int axis, double x, double y, double z) {
  double px = x / g->dx, py = y / g->dx, pz = z / g->dx;
  double gx = axis == 0 ? px : px - 0.5, gy = axis == 1 ? py : py - 0.5, gz = axis == 2 ? pz : pz - 0.5;
  int i0 = (int)floor(gx), j0 = (int)floor(gy), k0 = (int)floor(gz);
  double fx = gx - i0, fy = gy - j0, fz = gz - k0;
  int hx = g->nx - (axis == 0 ? 0 : 1), hy = g->ny - (axis == 1 ? 0 : 1), hz = g->nz - (axis == 2 ? 0 : 1);
  double s = 0.0;
  for (int dz = 0; dz < 2; ++dz) {
    double wz = dz == 0 ? 1.0 - fz : fz;
    for (int dy = 0; dy < 2; ++dy) {
      double wy = dy == 0 ? 1.0 - fy : fy;
      for (int ddx = 0; ddx < 2; ++ddx) {
        double wx = ddx == 0 ? 1.0 - fx : fx;
        int i = i0 + ddx, j = j0 + dy, k = k0 + dz;
        i = i < 0 ? 0 : (i > hx ? hx : i);
        j = j < 0 ? 0 : (j > hy ? hy : j);
        k = k < 0 ? 0 : (k > hz ? hz : k);
        double val = (double)(float)g->f[axis][fidx(g, axis, i, j, k)];
        if (!g->faithful) val = g->f[axis][fidx(g, axis, i, j, k)];
        s += wx * wy * wz * val;
      }
    }
  }
  return s;
}

/* mrG2P3D_tp, multires_ops3d_tp.cpp:280-299 */
void orc_mr_g2p(const orc_mr* g, const orc_mr* saved, int64_t n, const double* x, const double* y,
                const double* z, double* u, double* v, double* w, const uint8_t* type, double aL, double aG) {
  for (int64_t p = 0; p < n; ++p) {
    double a = type[p] == 0 ? aL : aG;
    double un = mr_sample(g, 0, x[p], y[p], z[p]), vn = mr_sample(g, 1, x[p], y[p], z[p]);
    double wn = mr_sample(g, 2, x[p], y[p], z[p]);
    double du = un - mr_sample(saved, 0, x[p], y[p], z[p]);
    double dv = vn - mr_sample(saved, 1, x[p], y[p], z[p]);
    double dw = wn - mr_sample(saved, 2, x[p], y[p], z[p]);
    double fx = u[p] + du, fy = v[p] + dv, fz = w[p] + dw;
    u[p] = a * fx + (1.0 - a) * un;
    v[p] = a * fy + (1.0 - a) * vn;
    w[p] = a * fz + (1.0 - a) * wn;
  }
}

/* mrAdvect3D_tp, multires_ops3d_tp.cpp:303-355 (RK2 midpoint or RK3), clamp counters */
void orc_mr_advect(const orc_mr* g, int64_t n, double* x, double* y, double* z, const uint8_t* type, double dt,
                   int order, orc_escape_stats* es) {
  double lo[3] = {0.5 * g->dx, 0.5 * g->dx, 0.5 * g->dx};
  double hi[3] = {((double)g->nx - 0.5) * g->dx, ((double)g->ny - 0.5) * g->dx, ((double)g->nz - 0.5) * g->dx};
  if (es) memset(es, 0, sizeof(*es));
  for (int64_t p = 0; p < n; ++p) {
    double pos[3] = {x[p], y[p], z[p]}, nxt[3];
    double k1[3] = {mr_sample(g, 0, pos[0], pos[1], pos[2]), mr_sample(g, 1, pos[0], pos[1], pos[2]),
                    mr_sample(g, 2, pos[0], pos[1], pos[2])};
    if (order == 3) {
      double q[3], k2[3], k3[3];
      for (int c = 0; c < 3; ++c) q[c] = pos[c] + k1[c] * (0.5 * dt);
      for (int c = 0; c < 3; ++c) k2[c] = mr_sample(g, c, q[0], q[1], q[2]);
      for (int c = 0; c < 3; ++c) q[c] = pos[c] - k1[c] * dt + k2[c] * (2.0 * dt);
      for (int c = 0; c < 3; ++c) k3[c] = mr_sample(g, c, q[0], q[1], q[2]);
      for (int c = 0; c < 3; ++c) nxt[c] = pos[c] + (k1[c] + k2[c] * 4.0 + k3[c]) * (dt / 6.0);
    } else {
      double m[3], k2[3];
      for (int c = 0; c < 3; ++c) m[c] = pos[c] + k1[c] * (0.5 * dt);
      for (int c = 0; c < 3; ++c) k2[c] = mr_sample(g, c, m[0], m[1], m[2]);
      for (int c = 0; c < 3; ++c) nxt[c] = pos[c] + k2[c] * dt;
    }
    int any = 0;
    int32_t* cnt = es ? &es->x_lo : NULL;
    for (int c = 0; c < 3; ++c) {
      int l = nxt[c] < lo[c], h = nxt[c] > hi[c];
      if (cnt) { cnt[2 * c] += l; cnt[2 * c + 1] += h; }
      any |= l | h;
      nxt[c] = nxt[c] > hi[c] ? hi[c] : nxt[c];
      nxt[c] = nxt[c] < lo[c] ? lo[c] : nxt[c];
    }
    if (es && any) {
      if (type[p] == 0) es->clamped_liquid++;
      else if (type[p] == 1) es->clamped_gas++;
    }
    x[p] = nxt[0]; y[p] = nxt[1]; z[p] = nxt[2];
  }
}

/* ---------------------------------------------------------------- projection */

static double beta_raw(double raw, const orc_phase* pp) {
  /* betaFromRawMass, multires_pressure3d.cpp:94-99 */
  double rmin = log(pp->rho_l / pp->rho_g) * pp->rho_g * pp->rho_tilde_0;
  double invden = 1.0 / (pp->alpha_phi * pp->rho_tilde_0 * pp->rho_l);
  double phi = raw < rmin ? 0.0 : fmin(sqrt((raw - rmin) * invden), 1.0);
  return 1.0 / (phi * pp->rho_l + (1.0 - phi) * pp->rho_g);
}

typedef struct {
  int64_t fi;     /* face index */
  int axis;
  int nx, ny, nz; /* neighbour fine cell */
  double sign;
} mr_patch;

/* visitCellFaces, multires_pressure3d.cpp:101-136: the fine face patches of a leaf cell
 * in the reference's visiting order; returns the count */
static int cell_patches(const orc_mr* g, int64_t id, mr_patch* out) {
  int x0, y0, z0, s;
  cell_box(g, id, &x0, &y0, &z0, &s);
  int x1 = x0 + s, y1 = y0 + s, z1 = z0 + s;
  int cx0 = x0 > 0 ? x0 : 0, cx1 = x1 < g->nx ? x1 : g->nx;
  int cy0 = y0 > 0 ? y0 : 0, cy1 = y1 < g->ny ? y1 : g->ny;
  int cz0 = z0 > 0 ? z0 : 0, cz1 = z1 < g->nz ? z1 : g->nz;
  int m = 0;
  for (int z = cz0; z < cz1; ++z)
    for (int y = cy0; y < cy1; ++y) {
      out[m++] = (mr_patch){fidx(g, 0, x0, y, z), 0, x0 - 1, y, z, -1.0};
      out[m++] = (mr_patch){fidx(g, 0, x1 < g->nx ? x1 : g->nx, y, z), 0, x1, y, z, 1.0};
    }
  for (int z = cz0; z < cz1; ++z)
    for (int x = cx0; x < cx1; ++x) {
      out[m++] = (mr_patch){fidx(g, 1, x, y0, z), 1, x, y0 - 1, z, -1.0};
      out[m++] = (mr_patch){fidx(g, 1, x, y1 < g->ny ? y1 : g->ny, z), 1, x, y1, z, 1.0};
    }
  for (int y = cy0; y < cy1; ++y)
    for (int x = cx0; x < cx1; ++x) {
      out[m++] = (mr_patch){fidx(g, 2, x, y, z0), 2, x, y, z0 - 1, -1.0};
      out[m++] = (mr_patch){fidx(g, 2, x, y, z1 < g->nz ? z1 : g->nz), 2, x, y, z1, 1.0};
    }
  return m;
}

static inline double cell_h(const orc_mr* g, int64_t id) { return g->dx * (double)(1 << cell_level(g, id)); }

/* neighborDistance, multires_pressure3d.cpp:189-201 */
static double nbr_distance(const orc_mr* g, double h, int nx, int ny, int nz) {
  double other = h;
  int64_t o = leaf_cell(g, nx, ny, nz);
  if (o >= 0) other = cell_h(g, o);
  return 0.5 * h + 0.5 * other;
}

typedef struct { int64_t idx; double coeff; } mr_term;

/* ---- level-1 coarse machinery (multires_pressure3d.cpp:39-70, 470-700) */

typedef struct {           /* the assembled rows: A x = diag x - sum coeff x_nbr, pin row = identity */
  int64_t N, pin;
  const double* diag;
  const int* noff;
  const mr_term* off;
} mr_op;

/* applyA (rows of :783-812); the pinned row is the identity */
static void mr_apply(const mr_op* A, const double* x, double* out) {
  for (int64_t t = 0; t < A->N; ++t) {
    if (t == A->pin) { out[t] = x[t]; continue; }
    double o = 0.0;
    for (int q = 0; q < A->noff[t]; ++q) o += A->off[6 * 64 * t + q].coeff * x[A->off[6 * 64 * t + q].idx];
    out[t] = A->diag[t] * x[t] - o;
  }
}

typedef struct {           /* MRPressureAggregation3D, multires_pressure3d.h:29-36 */
  int64_t nf, nc;
  int64_t* f2c;
  double* fv;
  double* cv;
} mr_agg;

static void mr_agg_free(mr_agg* a) { free(a->f2c); free(a->fv); free(a->cv); }

/* buildLevel1AggregationFromCells (:52-70) with level1ParentTuple (:39-50): a cell's
 * coarse id is its 2x-parent's order of first appearance in the cell list; coarse
 * volumes sum the fine ones (buildMRPressureAggregation3D, :470-505).  Returns 0 when
 * a leaf has a level above 1 (the reference throws there). */
static int mr_level1_aggregation(const orc_mr* g, const int64_t* cells, int64_t N, mr_agg* a) {
  const int64_t hx = (g->nx + 1) / 2, hy = (g->ny + 1) / 2, hz = (g->nz + 1) / 2;
  int64_t* first = (int64_t*)malloc(sizeof(int64_t) * (size_t)(hx * hy * hz));
  for (int64_t q = 0; q < hx * hy * hz; ++q) first[q] = -1;
  a->nf = N;
  a->nc = 0;
  a->f2c = (int64_t*)malloc(sizeof(int64_t) * (size_t)(N > 0 ? N : 1));
  a->fv = (double*)malloc(sizeof(double) * (size_t)(N > 0 ? N : 1));
  for (int64_t t = 0; t < N; ++t) {
    if (cell_level(g, cells[t]) > 1) { free(first); return 0; }
    int x0, y0, z0, st;
    cell_box(g, cells[t], &x0, &y0, &z0, &st);
    const int64_t key = x0 / 2 + hx * (y0 / 2 + hy * (int64_t)(z0 / 2));
    if (first[key] < 0) first[key] = a->nc++;
    a->f2c[t] = first[key];
    const double h = cell_h(g, cells[t]);
    a->fv[t] = h * h * h;
  }
  free(first);
  a->cv = (double*)calloc((size_t)(a->nc > 0 ? a->nc : 1), sizeof(double));
  for (int64_t t = 0; t < N; ++t) a->cv[a->f2c[t]] += a->fv[t];
  return 1;
}

/* restrictMRPressureVolumeWeighted3D (:518-539) */
static void mr_restrict(const mr_agg* a, const double* fine, double* coarse) {
  for (int64_t c = 0; c < a->nc; ++c) coarse[c] = 0.0;
  for (int64_t i = 0; i < a->nf; ++i) coarse[a->f2c[i]] += a->fv[i] * fine[i];
  for (int64_t c = 0; c < a->nc; ++c) coarse[c] /= a->cv[c];
}

/* prolongMRPressurePiecewiseConstant3D (:541-552) */
static void mr_prolong(const mr_agg* a, const double* coarse, double* fine) {
  for (int64_t i = 0; i < a->nf; ++i) fine[i] = coarse[a->f2c[i]];
}

/* weightedDotVolumes (:369-380) */
static double mr_wdot(const mr_agg* a, const double* x, const double* y) {
  double s = 0.0;
  for (int64_t c = 0; c < a->nc; ++c) s += a->cv[c] * x[c] * y[c];
  return s;
}

static double mr_wnorm(const mr_agg* a, const double* v) {
  const double n2 = mr_wdot(a, v, v);
  return n2 > 0.0 ? sqrt(n2) : 0.0;
}

typedef struct {           /* scratch of the coarse solves */
  double *rhs, *x, *r, *p, *ap;    /* coarse */
  double *fp, *fap, *corr;         /* fine */
} mr_coarse_work;

/* the coarse operator R A P, applied matrix-free through the fine rows, with the
 * coarse pin as an identity row (applyCoarse, :922-929 and :1181-1190) */
static void mr_apply_coarse(const mr_op* A, const mr_agg* a, int64_t cpin, mr_coarse_work* w, const double* in,
                            double* out) {
  mr_prolong(a, in, w->fp);
  mr_apply(A, w->fp, w->fap);
  mr_restrict(a, w->fap, out);
  if (cpin >= 0) out[cpin] = in[cpin];
}

/* the coarse CG shared by the correction (:946-1015) and the preconditioner
 * (:1214-1281): x = 0, r = p = rhs (pin zeroed); stops when the volume-weighted norm
 * of r is at most `eff`; returns the iterations; *breakdown on a non-finite or
 * vanishing scalar */
static int mr_coarse_cg(const mr_op* A, const mr_agg* a, int64_t cpin, mr_coarse_work* w, int max_it, double eff,
                        int* converged, int* breakdown) {
  const int64_t nc = a->nc;
  for (int64_t c = 0; c < nc; ++c) { w->x[c] = 0.0; w->r[c] = w->rhs[c]; w->p[c] = w->r[c]; w->ap[c] = 0.0; }
  if (cpin >= 0) { w->r[cpin] = 0.0; w->p[cpin] = 0.0; }
  const double res0 = mr_wnorm(a, w->r);
  *converged = res0 <= eff;
  *breakdown = 0;
  double rr = mr_wdot(a, w->r, w->r);
  if (!isfinite(rr)) { *breakdown = 1; return 0; }
  int iters = 0;
  for (int it = 0; !*converged && !*breakdown && it < max_it; ++it) {
    mr_apply_coarse(A, a, cpin, w, w->p, w->ap);
    const double pAp = mr_wdot(a, w->p, w->ap);
    if (!isfinite(pAp) || fabs(pAp) < 1e-30) { *breakdown = 1; break; }
    const double alpha = rr / pAp;
    if (!isfinite(alpha)) { *breakdown = 1; break; }
    for (int64_t c = 0; c < nc; ++c) { w->x[c] += alpha * w->p[c]; w->r[c] -= alpha * w->ap[c]; }
    if (cpin >= 0) { w->x[cpin] = 0.0; w->r[cpin] = 0.0; }
    iters = it + 1;
    const double res = mr_wnorm(a, w->r);
    if (!isfinite(res)) { *breakdown = 1; break; }
    if (res <= eff) { *converged = 1; break; }
    const double rrn = mr_wdot(a, w->r, w->r);
    if (!isfinite(rrn) || fabs(rr) < 1e-30) { *breakdown = 1; break; }
    const double beta = rrn / rr;
    if (!isfinite(beta)) { *breakdown = 1; break; }
    rr = rrn;
    for (int64_t c = 0; c < nc; ++c) w->p[c] = w->r[c] + beta * w->p[c];
    if (cpin >= 0) w->p[cpin] = 0.0;
  }
  return iters;
}

/* residualFromPressure (:870-884): out = rhs - A x; max |out|, +inf if any entry is
 * not finite */
static double mr_residual_from(const mr_op* A, const double* rhs, const double* x, double* Ax, double* out) {
  mr_apply(A, x, Ax);
  double res = 0.0;
  for (int64_t i = 0; i < A->N; ++i) {
    out[i] = rhs[i] - Ax[i];
    if (!isfinite(out[i])) return INFINITY;
    res = fmax(res, fabs(out[i]));
  }
  return res;
}

typedef struct {           /* the coarse preconditioner's state across one solve (:1119-1160) */
  int ready, disabled, streak;
  int64_t cpin;
  mr_agg agg;
  mr_coarse_work w;
} mr_cprec;

/* applyPreconditioner with the additive coarse term (:1283-1340): Jacobi z, then
 * z + scale P x_c when a coarse solve ran and r.z grew by at least min_rz_gain */
static void mr_precondition(const mr_op* A, const orc_mr_solve_cfg* cf, orc_mr_solve_stats* st, mr_cprec* cp,
                            const double* r, double* z, double* cand) {
  const int64_t N = A->N;
  for (int64_t t = 0; t < N; ++t) z[t] = (cf->use_jacobi && A->diag[t] > 0.0) ? r[t] / A->diag[t] : r[t];
  if (!cp) return;
  double rzb = 0.0;
  for (int64_t t = 0; t < N; ++t) rzb += r[t] * z[t];
  const int base_ok = isfinite(rzb) && rzb > 0.0;
  if (cp->disabled) return;
  const int skipped_before = st->cp_skipped;
  int accepted = 0, wasted = 0, solved = 0;
  /* solveCoarsePreconditioner (:1192-1281) */
  if (cp->ready) {
    int max_c = cf->coarse_preconditioner_iterations;
    int budget_ok = 1;
    if (cf->coarse_preconditioner_max_work_ratio > 0.0) {
      const double allowed = cf->coarse_preconditioner_max_work_ratio * (double)(st->iterations + 1 > 1 ? st->iterations + 1 : 1);
      const int remaining = (int)floor(allowed + 1e-12) - st->cp_iterations;
      if (remaining < max_c) {
        st->cp_budget_limited++;
        st->cp_skipped++;
        st->cp_budget_exhausted = 1;
        budget_ok = 0;
      }
    }
    if (budget_ok) {
      st->cp_applications++;
      mr_restrict(&cp->agg, r, cp->w.rhs);
      double r0 = 0.0;
      {
        for (int64_t c = 0; c < cp->agg.nc; ++c) cp->w.r[c] = cp->w.rhs[c];
        if (cp->cpin >= 0) cp->w.r[cp->cpin] = 0.0;
        r0 = mr_wnorm(&cp->agg, cp->w.r);
      }
      const double eff = fmax(cf->coarse_preconditioner_absolute_tolerance,
                              cf->coarse_preconditioner_relative_tolerance * r0);
      st->cp_effective_tolerance = eff;
      int conv = 0, bd = 0;
      const int it = mr_coarse_cg(A, &cp->agg, cp->cpin, &cp->w, max_c, eff, &conv, &bd);
      if (bd) {
        st->cp_breakdown = 1;
      } else {
        st->cp_iterations += it;
        mr_prolong(&cp->agg, cp->w.x, cp->w.corr);
        solved = 1;
      }
    }
  }
  if (solved) {
    for (int64_t t = 0; t < N; ++t) cand[t] = z[t] + cf->coarse_preconditioner_scale * cp->w.corr[t];
    double rzc = 0.0;
    for (int64_t t = 0; t < N; ++t) rzc += r[t] * cand[t];
    if (isfinite(rzc) && base_ok) st->cp_last_rz_gain = rzc / rzb - 1.0;
    const double min_acc = base_ok ? rzb * (1.0 + cf->coarse_preconditioner_min_rz_gain) : 0.0;
    if (isfinite(rzc) && rzc > 0.0 && (!base_ok || rzc >= min_acc)) {
      memcpy(z, cand, sizeof(double) * (size_t)N);
      st->cp_accepted++;
      accepted = 1;
    } else {
      st->cp_rejected++;
      wasted = 1;
    }
  } else if (st->cp_skipped > skipped_before) {
    wasted = 1;
  }
  /* recordCoarsePreconditionerOutcome (:1157-1178) */
  if (accepted) {
    cp->streak = 0;
    st->cp_wasted_streak = 0;
    return;
  }
  if (!wasted) return;
  cp->streak++;
  st->cp_wasted_streak = cp->streak;
  if (cf->coarse_preconditioner_auto_disable && cf->coarse_preconditioner_auto_disable_after > 0 &&
      cp->streak >= cf->coarse_preconditioner_auto_disable_after) {
    cp->disabled = 1;
    st->cp_auto_disabled = 1;
    if (st->cp_auto_disable_iteration < 0) st->cp_auto_disable_iteration = st->iterations;
  }
}

static void mr_coarse_work_alloc(mr_coarse_work* w, int64_t nc, int64_t nf) {
  const size_t c = (size_t)(nc > 0 ? nc : 1), f = (size_t)(nf > 0 ? nf : 1);
  w->rhs = (double*)calloc(c, 8); w->x = (double*)calloc(c, 8); w->r = (double*)calloc(c, 8);
  w->p = (double*)calloc(c, 8); w->ap = (double*)calloc(c, 8);
  w->fp = (double*)calloc(f, 8); w->fap = (double*)calloc(f, 8); w->corr = (double*)calloc(f, 8);
}

static void mr_coarse_work_free(mr_coarse_work* w) {
  free(w->rhs); free(w->x); free(w->r); free(w->p); free(w->ap); free(w->fp); free(w->fap); free(w->corr);
}

/* projectMR3D (phase-aware), multires_pressure3d.cpp:747-1449 and the face update
 * :1451-1524, with the coarse correction (:886-1060) and the additive coarse
 * preconditioner (:1119-1341) when enabled. */
/* validateSolveConfig, multires_pressure3d.cpp:307-333 (the reference throws) */
int orc_mr_solve_cfg_valid(const orc_mr_solve_cfg* c) {
  return !(c->max_iterations < 0 || c->absolute_tolerance < 0.0 || c->relative_tolerance < 0.0 ||
           c->restart_growth_threshold < 0.0 || c->relaxation_sweeps < 0 || c->relaxation_omega < 0.0 ||
           c->relaxation_min_omega < 0.0 || c->coarse_correction_iterations < 0 ||
           c->coarse_correction_sweeps < 0 || c->coarse_correction_absolute_tolerance < 0.0 ||
           c->coarse_correction_relative_tolerance < 0.0 || c->coarse_correction_min_scale <= 0.0 ||
           c->coarse_correction_min_scale > 1.0 || c->coarse_preconditioner_iterations < 0 ||
           c->coarse_preconditioner_absolute_tolerance < 0.0 || c->coarse_preconditioner_relative_tolerance < 0.0 ||
           c->coarse_preconditioner_scale < 0.0 || c->coarse_preconditioner_min_rz_gain < 0.0 ||
           c->coarse_preconditioner_max_work_ratio < 0.0 || c->coarse_preconditioner_auto_disable_after < 0 ||
           !isfinite(c->divergence_correction));
}

/* the face update over the enumerated faces (multires_pressure3d.cpp:1455-1524) from the
 * row pressures x (rows numbered by idx) */
static void mr_face_update(orc_mr* g, const orc_phase* pp, double dt, const int64_t* idx, const double* x) {
  for (int axis = 0; axis < 3; ++axis) {
    int fx = g->nx + (axis == 0), fy = g->ny + (axis == 1), fz = g->nz + (axis == 2);
    for (int k = 0; k < fz; ++k)
      for (int j = 0; j < fy; ++j)
        for (int i = 0; i < fx; ++i) {
          int64_t fi = fidx(g, axis, i, j, k);
          if (!g->valid[axis][fi]) continue;
          int li = i - (axis == 0), lj = j - (axis == 1), lk = k - (axis == 2);
          int lm = marker_fine(g, li, lj, lk), rm = marker_fine(g, i, j, k);
          int lf = lm == 1, rf = rm == 1;
          if (!lf && !rf) continue;
          if (lm == 2 || rm == 2) { g->f[axis][fi] = 0.0; continue; }
          double dist = g->dx;
          int64_t ref = lf ? leaf_cell(g, li, lj, lk) : leaf_cell(g, i, j, k);
          if (ref >= 0 && idx[ref] >= 0)
            dist = lf ? nbr_distance(g, cell_h(g, ref), i, j, k) : nbr_distance(g, cell_h(g, ref), li, lj, lk);
          int64_t lc = leaf_cell(g, li, lj, lk), rc = leaf_cell(g, i, j, k);
          double pl = (lf && lc >= 0 && idx[lc] >= 0) ? x[idx[lc]] : 0.0;
          double pr = (rf && rc >= 0 && idx[rc] >= 0) ? x[idx[rc]] : 0.0;
          double beta = beta_raw(g->f[3 + axis][fi], pp);
          double cur = g->faithful ? (double)(float)g->f[axis][fi] : g->f[axis][fi];
          g->f[axis][fi] = stf(g, cur - dt * beta * (pr - pl) / dist);
        }
  }
}

/* The projection system (multires_pressure3d.cpp:166-175, 783-812): the FLUID leaf
 * cells in leafCells order (leaves sorted, then lz, ly, lx), the pin (the first of them
 * when no FLUID cell touches AIR), and the rows diag_t x_t - sum coeff x_nbr with the
 * divergence of the current face velocities. */
typedef struct {
  int64_t N, pin;
  int64_t* idx;   /* leaf cell -> row, or -1 */
  int64_t* cells; /* row -> leaf cell */
  double *diag, *divg;
  int* noff;
  mr_term* off;   /* 6 * 64 terms per row */
} mr_rows;

static void mr_rows_free(mr_rows* R) {
  free(R->idx); free(R->cells); free(R->diag); free(R->divg); free(R->noff); free(R->off);
  memset(R, 0, sizeof(*R));
}

static void mr_rows_build(orc_mr* g, const orc_phase* pp, double dt, mr_rows* R) {
  memset(R, 0, sizeof(*R));
  int64_t nl = (int64_t)g->nleaves * 64;
  int64_t* idx = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nl > 0 ? nl : 1));
  int64_t* cells = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nl > 0 ? nl : 1));
  int64_t N = 0;
  for (int64_t id = 0; id < nl; ++id) {
    idx[id] = -1;
    int x0, y0, z0, s;
    cell_box(g, id, &x0, &y0, &z0, &s);
    if (!(x0 < g->nx && x0 + s > 0 && y0 < g->ny && y0 + s > 0 && z0 < g->nz && z0 + s > 0)) continue;
    if ((int)(g->mk[id] + 0.5) != 1) continue;
    idx[id] = N;
    cells[N++] = id;
  }
  R->N = N;
  R->idx = idx;
  R->cells = cells;
  R->pin = -1;
  if (N == 0) return;
  mr_patch* pt = (mr_patch*)malloc(sizeof(mr_patch) * 6 * 64);
  /* findPinCell :166-175 */
  int has_air = 0;
  for (int64_t t = 0; t < N && !has_air; ++t) {
    int m = cell_patches(g, cells[t], pt);
    for (int q = 0; q < m; ++q)
      if (marker_fine(g, pt[q].nx, pt[q].ny, pt[q].nz) == 0) { has_air = 1; break; }
  }
  int64_t pin = has_air ? -1 : 0;
  R->pin = pin;
  /* rows :783-812 */
  double* diag = (double*)calloc((size_t)N, sizeof(double));
  double* divg = (double*)calloc((size_t)N, sizeof(double));
  int* noff = (int*)calloc((size_t)N, sizeof(int));
  mr_term* off = (mr_term*)malloc(sizeof(mr_term) * 6 * 64 * (size_t)N);
  double area = g->dx * g->dx;
  for (int64_t t = 0; t < N; ++t) {
    if (t == pin) { diag[t] = 1.0; continue; }
    double h = cell_h(g, cells[t]);
    double V = h * h * h;
    int m = cell_patches(g, cells[t], pt);
    for (int q = 0; q < m; ++q) {
      if (marker_fine(g, pt[q].nx, pt[q].ny, pt[q].nz) == 2) continue;
      double dist = nbr_distance(g, h, pt[q].nx, pt[q].ny, pt[q].nz);
      double coeff = 0.0;
      if (dist > 0.0 && V > 0.0)
        coeff = dt * beta_raw(g->f[3 + pt[q].axis][pt[q].fi], pp) * area / dist / V;
      diag[t] += coeff;
      int64_t o = leaf_cell(g, pt[q].nx, pt[q].ny, pt[q].nz);
      int64_t ni = o >= 0 ? idx[o] : -1;
      if (ni >= 0 && ni != pin) {
        off[6 * 64 * t + noff[t]].idx = ni;
        off[6 * 64 * t + noff[t]].coeff = coeff;
        noff[t]++;
      }
      double vel = g->f[pt[q].axis][pt[q].fi];
      divg[t] += pt[q].sign * vel * area / V;
    }
  }
  free(pt);
  R->diag = diag;
  R->divg = divg;
  R->noff = noff;
  R->off = off;
}

void orc_mr_project(orc_mr* g, const orc_phase* pp, double dt, const orc_mr_solve_cfg* cf, orc_mr_solve_stats* st) {
  orc_mr_solve_stats local;
  if (!st) st = &local;
  memset(st, 0, sizeof(*st));
  st->pinned_cell = -1;
  st->cp_auto_disable_iteration = -1;
  if (!orc_mr_solve_cfg_valid(cf)) {
    st->breakdown = 1;
    return;
  }
  mr_rows R;
  mr_rows_build(g, pp, dt, &R);
  const int64_t N = R.N, pin = R.pin;
  int64_t *idx = R.idx, *cells = R.cells;
  double *diag = R.diag, *divg = R.divg;
  int* noff = R.noff;
  mr_term* off = R.off;
  st->active_cells = N;
  st->pinned_cell = pin;
  memset(g->p, 0, sizeof(double) * (size_t)(g->nleaves > 0 ? (int64_t)g->nleaves * 64 : 1));
  if (N == 0) {
    st->converged = 1;
    mr_rows_free(&R);
    return;
  }
  double* x = (double*)calloc((size_t)N, sizeof(double));
  double* rhs = (double*)calloc((size_t)N, sizeof(double));
  double* r = (double*)calloc((size_t)N, sizeof(double));
  double* z = (double*)calloc((size_t)N, sizeof(double));
  double* zp = (double*)calloc((size_t)N, sizeof(double));
  double* p = (double*)calloc((size_t)N, sizeof(double));
  double* Ap = (double*)calloc((size_t)N, sizeof(double));
  double* cx = (double*)calloc((size_t)N, sizeof(double));
  double* cr = (double*)calloc((size_t)N, sizeof(double));
#define APPLY(XX, OUT)                                                              \
  for (int64_t t = 0; t < N; ++t) {                                                 \
    if (t == pin) { (OUT)[t] = (XX)[t]; continue; }                                 \
    double o_ = 0.0;                                                                \
    for (int q = 0; q < noff[t]; ++q) o_ += off[6 * 64 * t + q].coeff * (XX)[off[6 * 64 * t + q].idx]; \
    (OUT)[t] = diag[t] * (XX)[t] - o_;                                              \
  }
  double res0 = 0.0;
  for (int64_t t = 0; t < N; ++t) {
    rhs[t] = t == pin ? 0.0 : -divg[t] + cf->divergence_correction;
    r[t] = rhs[t];
    res0 = fmax(res0, fabs(r[t]));
  }
  st->initial_residual = res0;
  st->final_residual = res0;
  double tol = fmax(fmax(0.0, cf->absolute_tolerance), fmax(0.0, cf->relative_tolerance) * res0);
  st->effective_tolerance = tol;
  st->converged = res0 < tol;
#define PRECOND()                                                                   \
  for (int64_t t = 0; t < N; ++t) z[t] = (cf->use_jacobi && diag[t] > 0.0) ? r[t] / diag[t] : r[t];
  const mr_op A = {N, pin, diag, noff, off};
  if (res0 >= tol) {
    double cur = res0;
    /* Galerkin coarse correction sweeps with a halving scale search, :886-1060 */
    if (cf->use_coarse_correction && cur >= tol) {
      mr_agg agg;
      if (mr_level1_aggregation(g, cells, N, &agg)) {
        st->cc_cells = (int32_t)agg.nc;
        st->cc_initial_residual = cur;
        st->cc_final_residual = cur;
        if (agg.nc > 0 && agg.nc < N) {
          const int64_t cpin = pin >= 0 ? agg.f2c[pin] : -1;
          mr_coarse_work w;
          mr_coarse_work_alloc(&w, agg.nc, N);
          double* bx = (double*)malloc(8 * (size_t)N);
          double* br = (double*)malloc(8 * (size_t)N);
          for (int sw = 0; sw < cf->coarse_correction_sweeps && cur >= tol && !st->cc_breakdown; ++sw) {
            st->cc_sweeps++;
            mr_restrict(&agg, r, w.rhs);
            for (int64_t c = 0; c < agg.nc; ++c) w.r[c] = w.rhs[c];
            if (cpin >= 0) w.r[cpin] = 0.0;
            const double eff = fmax(cf->coarse_correction_absolute_tolerance,
                                    cf->coarse_correction_relative_tolerance * mr_wnorm(&agg, w.r));
            st->cc_effective_tolerance = eff;
            int conv = 0, bd = 0;
            const int its = mr_coarse_cg(&A, &agg, cpin, &w, cf->coarse_correction_iterations, eff, &conv, &bd);
            st->cc_converged = conv;
            st->cc_iterations += its;
            if (bd) { st->cc_breakdown = 1; break; }
            mr_prolong(&agg, w.x, w.corr);
            double best = cur, bscale = 0.0;
            int have = 0;
            for (double scale = 1.0; scale >= cf->coarse_correction_min_scale; scale *= 0.5) {
              for (int64_t t = 0; t < N; ++t) cx[t] = x[t] + scale * w.corr[t];
              const double cres = mr_residual_from(&A, rhs, cx, Ap, cr);
              if (isfinite(cres) && cres <= best) {
                best = cres;
                bscale = scale;
                memcpy(bx, cx, 8 * (size_t)N);
                memcpy(br, cr, 8 * (size_t)N);
                have = 1;
              }
            }
            st->cc_final_residual = best;
            if (best < cur && have) {
              memcpy(x, bx, 8 * (size_t)N);
              memcpy(r, br, 8 * (size_t)N);
              cur = best;
              st->final_residual = cur;
              st->cc_accepted = 1;
              st->cc_last_scale = bscale;
              st->cc_accepted_sweeps++;
            } else {
              st->cc_rejected_sweeps++;
              break;
            }
          }
          free(bx);
          free(br);
          mr_coarse_work_free(&w);
        }
        mr_agg_free(&agg);
      }
    }
    /* damped-Jacobi pre-relaxation with omega backoff, :1062-1110 */
    double om = fmax(0.0, cf->relaxation_omega), omin = fmax(0.0, cf->relaxation_min_omega);
    if (omin > om) omin = om;
    for (int sw = 0; sw < cf->relaxation_sweeps && cur >= tol; ++sw) {
      PRECOND();
      double trial = om;
      int accepted = 0;
      while (trial >= omin && trial > 0.0) {
        for (int64_t t = 0; t < N; ++t) cx[t] = x[t] + trial * z[t];
        APPLY(cx, Ap);
        double cres = 0.0;
        int fin = 1;
        for (int64_t t = 0; t < N; ++t) {
          cr[t] = rhs[t] - Ap[t];
          if (!isfinite(cr[t])) { fin = 0; break; }
          cres = fmax(cres, fabs(cr[t]));
        }
        if (!fin) cres = INFINITY;
        if (isfinite(cres) && cres <= cur) {
          memcpy(x, cx, 8 * (size_t)N);
          memcpy(r, cr, 8 * (size_t)N);
          cur = cres;
          st->final_residual = cur;
          om = trial;
          st->relaxation_accepted++;
          accepted = 1;
          break;
        }
        st->relaxation_rejected++;
        if (trial <= omin) break;
        trial = fmax(omin, trial * 0.5);
      }
      if (!accepted) break;
    }
    st->converged = cur < tol;
    if (st->converged) st->final_residual = cur;
  }
  if (!st->converged && res0 >= tol) {
    /* the additive coarse preconditioner's aggregation and pin, :1119-1143 */
    mr_cprec cpre;
    memset(&cpre, 0, sizeof(cpre));
    mr_cprec* cpp = NULL;
    if (cf->use_coarse_preconditioner && cf->coarse_preconditioner_iterations > 0 &&
        cf->coarse_preconditioner_scale > 0.0 && mr_level1_aggregation(g, cells, N, &cpre.agg)) {
      st->cp_cells = (int32_t)cpre.agg.nc;
      cpre.ready = cpre.agg.nc > 0 && cpre.agg.nc < N;
      cpre.cpin = (cpre.ready && pin >= 0) ? cpre.agg.f2c[pin] : -1;
      mr_coarse_work_alloc(&cpre.w, cpre.agg.nc, N);
      cpp = &cpre;
    }
#define PRECOND_PCG() mr_precondition(&A, cf, st, cpp, r, z, cx)
    PRECOND_PCG();
    memcpy(p, z, 8 * (size_t)N);
    double rz = 0.0;
    for (int64_t t = 0; t < N; ++t) rz += r[t] * z[t];
    if (!isfinite(rz)) st->breakdown = 1;
    double prev = st->final_residual;
    for (int it = 0; !st->breakdown && it < cf->max_iterations; ++it) {
      APPLY(p, Ap);
      double pAp = 0.0;
      for (int64_t t = 0; t < N; ++t) pAp += p[t] * Ap[t];
      if (!isfinite(pAp) || fabs(pAp) < 1e-30) { st->breakdown = 1; break; }
      double alpha = rz / pAp;
      if (!isfinite(alpha)) { st->breakdown = 1; break; }
      double res = 0.0;
      int fin = 1;
      for (int64_t t = 0; t < N; ++t) {
        x[t] += alpha * p[t];
        r[t] -= alpha * Ap[t];
        fin = fin && isfinite(x[t]) && isfinite(r[t]);
        if (isfinite(r[t])) res = fmax(res, fabs(r[t]));
      }
      st->iterations = it + 1;
      st->final_residual = res;
      if (!fin || !isfinite(res)) { st->breakdown = 1; break; }
      if (res < tol) { st->converged = 1; break; }
      if (cf->adaptive_restart && cf->restart_growth_threshold > 0.0 && res > prev * cf->restart_growth_threshold &&
          res > res0 * cf->restart_growth_threshold && it + 1 < cf->max_iterations) {
        PRECOND_PCG();
        rz = 0.0;
        for (int64_t t = 0; t < N; ++t) rz += r[t] * z[t];
        if (!isfinite(rz)) { st->breakdown = 1; break; }
        memcpy(p, z, 8 * (size_t)N);
        st->restarts++;
        prev = res;
        continue;
      }
      prev = res;
      memcpy(zp, z, 8 * (size_t)N);
      PRECOND_PCG();
      double rzn = 0.0;
      for (int64_t t = 0; t < N; ++t) rzn += r[t] * z[t];
      if (!isfinite(rzn)) { st->breakdown = 1; break; }
      double beta = 0.0;
      if (rz != 0.0) {
        if (cf->use_flexible_beta) {
          double num = 0.0;
          for (int64_t t = 0; t < N; ++t) num += r[t] * (z[t] - zp[t]);
          beta = num / rz;
        } else {
          beta = rzn / rz;
        }
      }
      if (!isfinite(beta)) { beta = 0.0; st->beta_resets++; }
      if (beta < 0.0) { beta = 0.0; st->beta_resets++; }
      rz = rzn;
      for (int64_t t = 0; t < N; ++t) p[t] = z[t] + beta * p[t];
    }
#undef PRECOND_PCG
    if (cpp) {
      mr_agg_free(&cpre.agg);
      mr_coarse_work_free(&cpre.w);
    }
  }
  /* true residual (FP64) for comparisons */
  APPLY(x, Ap);
  {
    double rm = 0.0, rr = 0.0, bb = 0.0;
    for (int64_t t = 0; t < N; ++t) {
      double e = rhs[t] - Ap[t];
      rm = fmax(rm, fabs(e));
      rr += e * e;
      bb += rhs[t] * rhs[t];
    }
    st->true_max_residual = rm;
    st->true_rel_residual = bb > 0.0 ? sqrt(rr / bb) : 0.0;
  }
#undef APPLY
#undef PRECOND
  for (int64_t t = 0; t < N; ++t) g->p[cells[t]] = stf(g, x[t]);
  mr_face_update(g, pp, dt, idx, x);
  mr_rows_free(&R);
  free(x); free(rhs); free(r); free(z); free(zp); free(p); free(Ap); free(cx); free(cr);
}

/* The projection system in CSR form, for a direct solve: rows in leafCells order, each
 * row's diagonal first, then its off-diagonal terms as assembled (the pin's column left
 * out), and rhs_t = -div_t + cdiv (0 at the pin).  Returns the row count and stores the
 * nonzero count and the pin row; the arrays are filled when the caps hold them. */
int64_t orc_mr_system(orc_mr* g, const orc_phase* pp, double dt, double cdiv, int64_t cap_rows, int64_t cap_nnz,
                      int64_t* cells, int64_t* row_ptr, int64_t* col, double* val, double* rhs, int64_t* nnz,
                      int64_t* pin) {
  mr_rows R;
  mr_rows_build(g, pp, dt, &R);
  int64_t m = 0;
  for (int64_t t = 0; t < R.N; ++t) m += 1 + R.noff[t];
  *nnz = m;
  *pin = R.pin;
  if (cap_rows >= R.N && cap_nnz >= m && row_ptr) {
    int64_t e = 0;
    for (int64_t t = 0; t < R.N; ++t) {
      cells[t] = R.cells[t];
      row_ptr[t] = e;
      col[e] = t;
      val[e++] = R.diag[t];
      for (int q = 0; q < R.noff[t]; ++q) {
        col[e] = R.off[6 * 64 * t + q].idx;
        val[e++] = -R.off[6 * 64 * t + q].coeff;
      }
      rhs[t] = t == R.pin ? 0.0 : -R.divg[t] + cdiv;
    }
    row_ptr[R.N] = e;
  }
  const int64_t N = R.N;
  mr_rows_free(&R);
  return N;
}

/* The end of a projection from given row pressures x (rows as orc_mr_system numbers
 * them): the leaf pressures and the face update. */
void orc_mr_apply_pressure(orc_mr* g, const orc_phase* pp, double dt, const double* x) {
  mr_rows R;
  mr_rows_build(g, pp, dt, &R);
  memset(g->p, 0, sizeof(double) * (size_t)(g->nleaves > 0 ? (int64_t)g->nleaves * 64 : 1));
  for (int64_t t = 0; t < R.N; ++t) g->p[R.cells[t]] = stf(g, x[t]);
  if (R.N > 0) mr_face_update(g, pp, dt, R.idx, x);
  mr_rows_free(&R);
}

/* ---------------------------------------------------------------- step */

/* the particle-occupancy box of updateDynamicRefinement, multires_sim3d_tp.cpp:577-669:
 * returns the target box (and updates the retained box) given the current layout */
static void dynamic_box(const orc_mr_sim* s, int64_t n, const double* x, const double* y, const double* z,
                        const uint8_t* type, int32_t box_out[6], int* valid_out, orc_mr_sim* st);

void orc_mr_dynamic_refinement(orc_mr* g, orc_mr_sim* sim, int64_t n, const double* x, const double* y,
                               const double* z, const uint8_t* type) {
  if (!sim->dynamic_refinement) return;
  int32_t box[6];
  int valid = 0;
  dynamic_box(sim, n, x, y, z, type, box, &valid, sim);
  int64_t cnt = orc_mr_layout_fine_box(g->nx, g->ny, g->nz, 1, valid, box, NULL, 0);
  int32_t* lv = (int32_t*)malloc(sizeof(int32_t) * 4 * (size_t)(cnt > 0 ? cnt : 1));
  orc_mr_layout_fine_box(g->nx, g->ny, g->nz, 1, valid, box, lv, cnt);
  int64_t fine = 0;
  for (int64_t a = 0; a < cnt; ++a) fine += lv[4 * a] == 0;
  sim->dynamic_last_fine_leaves = (int)fine;
  int same = cnt == g->nleaves && memcmp(lv, g->leaves, sizeof(int32_t) * 4 * (size_t)cnt) == 0;
  if (!same) orc_mr_set_leaves(g, (int)cnt, lv);
  free(lv);
}

static void dynamic_box(const orc_mr_sim* s, int64_t n, const double* x, const double* y, const double* z,
                        const uint8_t* type, int32_t box_out[6], int* valid_out, orc_mr_sim* st) {
  int nx = s->nx, ny = s->ny, nz = s->nz;
  int mn[3] = {nx, ny, nz}, mx[3] = {-1, -1, -1}, gmn[3] = {nx, ny, nz}, gmx[3] = {-1, -1, -1};
  for (int64_t p = 0; p < n; ++p) {
    if (!isfinite(x[p]) || !isfinite(y[p]) || !isfinite(z[p])) continue;
    int c[3] = {(int)(x[p] / s->dx), (int)(y[p] / s->dx), (int)(z[p] / s->dx)};
    if (c[0] < 0 || c[0] >= nx || c[1] < 0 || c[1] >= ny || c[2] < 0 || c[2] >= nz) continue;
    for (int a = 0; a < 3; ++a) {
      if (c[a] < mn[a]) mn[a] = c[a];
      if (c[a] > mx[a]) mx[a] = c[a];
      if (type[p] == 1) {
        if (c[a] < gmn[a]) gmn[a] = c[a];
        if (c[a] > gmx[a]) gmx[a] = c[a];
      }
    }
  }
  int dv = 0, d[6] = {0, 0, 0, 0, 0, 0};
#define INCLUDE(A0, A1, A2, B0, B1, B2)                                                              \
  do {                                                                                               \
    int _b[6] = {(A0), (A1), (A2), (B0), (B1), (B2)};                                                \
    if (_b[0] < _b[3] && _b[1] < _b[4] && _b[2] < _b[5]) {                                           \
      if (!dv) { memcpy(d, _b, sizeof(d)); dv = 1; }                                                 \
      else { for (int _a = 0; _a < 3; ++_a) { if (_b[_a] < d[_a]) d[_a] = _b[_a]; if (_b[3 + _a] > d[3 + _a]) d[3 + _a] = _b[3 + _a]; } } \
    }                                                                                                \
  } while (0)
  if (mx[0] >= 0) {
    int pad = s->particle_padding > 0 ? s->particle_padding : 0;
    INCLUDE(mn[0] - pad, mn[1] - pad, mn[2] - pad, mx[0] + pad + 1, mx[1] + pad + 1, mx[2] + pad + 1);
  }
  if (gmx[0] >= 0) {
    int pad = s->gas_padding > 0 ? s->gas_padding : 0;
    INCLUDE(gmn[0] - pad, gmn[1] - pad, gmn[2] - pad, gmx[0] + pad + 1, gmx[1] + pad + 1, gmx[2] + pad + 1);
  }
#undef INCLUDE
  const int lim[3] = {nx, ny, nz};
  /* clampBox */
  if (dv) {
    for (int a = 0; a < 3; ++a) {
      d[a] = d[a] < 0 ? 0 : (d[a] > lim[a] ? lim[a] : d[a]);
      d[3 + a] = d[3 + a] < 0 ? 0 : (d[3 + a] > lim[a] ? lim[a] : d[3 + a]);
    }
    dv = d[0] < d[3] && d[1] < d[4] && d[2] < d[5];
  }
  int t[6] = {0, 0, 0, 0, 0, 0}, tv = 0;
  if (dv) {
    const int* r = st->retained;
    int contains = st->retained_valid && r[0] <= d[0] && r[1] <= d[1] && r[2] <= d[2] && r[3] >= d[3] &&
                   r[4] >= d[4] && r[5] >= d[5];
    if (contains) {
      memcpy(t, r, sizeof(t));
      tv = 1;
    } else {
      int pad = s->hysteresis > 0 ? s->hysteresis : 0;
      for (int a = 0; a < 3; ++a) {
        t[a] = d[a] - pad; t[3 + a] = d[3 + a] + pad;
        t[a] = t[a] < 0 ? 0 : (t[a] > lim[a] ? lim[a] : t[a]);
        t[3 + a] = t[3 + a] < 0 ? 0 : (t[3 + a] > lim[a] ? lim[a] : t[3 + a]);
      }
      tv = t[0] < t[3] && t[1] < t[4] && t[2] < t[5];
    }
    /* fitBoxToFineLeafBudget (:194-210) */
    st->budget_limited = 0;
    if (tv && s->max_fine_leaves > 0) {
      int32_t b32[6];
      for (int a = 0; a < 6; ++a) b32[a] = t[a];
      int64_t fine = 0;
      {
        int64_t cnt = orc_mr_layout_fine_box(nx, ny, nz, 1, 1, b32, NULL, 0);
        int32_t* lv = (int32_t*)malloc(sizeof(int32_t) * 4 * (size_t)(cnt > 0 ? cnt : 1));
        orc_mr_layout_fine_box(nx, ny, nz, 1, 1, b32, lv, cnt);
        for (int64_t a = 0; a < cnt; ++a) fine += lv[4 * a] == 0;
        free(lv);
      }
      if (fine > s->max_fine_leaves) {
        st->budget_limited = 1;
        double c[3] = {0.5 * (t[0] + t[3]), 0.5 * (t[1] + t[4]), 0.5 * (t[2] + t[5])};
        for (;;) {
          for (int a = 0; a < 6; ++a) b32[a] = t[a];
          int64_t cnt = orc_mr_layout_fine_box(nx, ny, nz, 1, 1, b32, NULL, 0);
          int32_t* lv = (int32_t*)malloc(sizeof(int32_t) * 4 * (size_t)(cnt > 0 ? cnt : 1));
          orc_mr_layout_fine_box(nx, ny, nz, 1, 1, b32, lv, cnt);
          fine = 0;
          for (int64_t a = 0; a < cnt; ++a) fine += lv[4 * a] == 0;
          free(lv);
          if (fine <= s->max_fine_leaves) break;
          /* shrinkLargestAxis */
          int ax = 0;
          if (t[4] - t[1] > t[3] - t[0]) ax = 1;
          if (t[5] - t[2] > t[3 + ax] - t[ax]) ax = 2;
          if (t[3 + ax] - t[ax] <= 1) break;
          if ((t[3 + ax] - c[ax]) >= (c[ax] - t[ax])) --t[3 + ax];
          else ++t[ax];
          tv = t[0] < t[3] && t[1] < t[4] && t[2] < t[5];
          if (!tv) break;
        }
      }
    }
  } else {
    st->budget_limited = 0;
  }
  st->retained_valid = tv;
  if (tv) memcpy(st->retained, t, sizeof(t));
  memcpy(box_out, t, sizeof(t));
  *valid_out = tv;
}

/* ---------------------------------------------------------------- interface */
/* diagnoseMRInterface3D and applyMRSurfaceTension3D (interface_diagnostics3d.cpp:381-401,
 * :447-496): the sparse stencils over the fine 1x1 face patches (mrFacePhi, :238-251,
 * reads 0 for a patch the grid does not hold, as the array does), evaluated once per
 * FLUID leaf cell at its fine centre clamped into the grid (mrCellFineCenter,
 * :350-358), in leafCells order (leaves sorted, then lz, ly, lx; only cells that
 * intersect the domain, multires_scalar_grid3d.h:75-88, 126-137). */

static double mr_mass(const void* gp, int axis, int i, int j, int k) {
  const orc_mr* g = (const orc_mr*)gp;
  return g->f[ORC_MU + axis][fidx(g, axis, i, j, k)];
}

static orc_iface_src mr_src(const orc_mr* g) {
  orc_iface_src s = {g->nx, g->ny, g->nz, g->dx, g, mr_mass};
  return s;
}

static inline int clampi_mr(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* the clamped fine centre of FLUID leaf cell id, or 0 when it is skipped */
static int mr_fluid_centre(const orc_mr* g, int64_t id, int* i, int* j, int* k) {
  int x0, y0, z0, st;
  cell_box(g, id, &x0, &y0, &z0, &st);
  if (!(x0 < g->nx && x0 + st > 0 && y0 < g->ny && y0 + st > 0 && z0 < g->nz && z0 + st > 0)) return 0;
  if ((int)(g->mk[id] + 0.5) != 1) return 0;
  *i = clampi_mr(x0 + st / 2, 0, g->nx - 1);
  *j = clampi_mr(y0 + st / 2, 0, g->ny - 1);
  *k = clampi_mr(z0 + st / 2, 0, g->nz - 1);
  return 1;
}

void orc_mr_interface_diagnostics(const orc_mr* g, const orc_phase* pp, orc_iface_stats* S) {
  const orc_iface_src s = mr_src(g);
  memset(S, 0, sizeof(*S));
  S->finite = 1;
  for (int64_t id = 0; id < (int64_t)g->nleaves * 64; ++id) {
    int i, j, k;
    if (mr_fluid_centre(g, id, &i, &j, &k)) orc_iface_accumulate(&s, pp, i, j, k, S);
  }
  orc_iface_finish(S);
}

static void mr_face_add(orc_mr* g, int axis, int i, int j, int k, double add) {
  const int64_t fi = fidx(g, axis, i, j, k);
  if (!(g->f[ORC_MU + axis][fi] > 0.0)) return;
  double* r = &g->f[axis][fi];
  *r = g->faithful ? (double)((float)*r + (float)add) : *r + add;
}

void orc_mr_surface_tension(orc_mr* g, const orc_phase* pp, double dt, double strength, double maxd, int radius,
                            orc_st_stats* S) {
  const orc_iface_src s = mr_src(g);
  orc_st_init(S, pp, g->dx, dt, strength, maxd, radius);
  if (!S->enabled) return;
  for (int64_t id = 0; id < (int64_t)g->nleaves * 64; ++id) {
    int i, j, k;
    double d[3];
    if (!mr_fluid_centre(g, id, &i, &j, &k)) continue;
    if (orc_st_cell(&s, pp, i, j, k, dt, strength, maxd, S, d) != 1) continue;
    mr_face_add(g, 0, i, j, k, 0.5 * d[0]);
    mr_face_add(g, 0, i + 1, j, k, 0.5 * d[0]);
    mr_face_add(g, 1, i, j, k, 0.5 * d[1]);
    mr_face_add(g, 1, i, j + 1, k, 0.5 * d[1]);
    mr_face_add(g, 2, i, j, k, 0.5 * d[2]);
    mr_face_add(g, 2, i, j, k + 1, 0.5 * d[2]);
  }
  orc_st_finish(S);
}

/* mrPhaseCells (render_cache3d.h:590-613): each leaf cell in leafCells order whose
 * corner lies in the grid, sampled by mrCellPhi (:302-318) at its centre clamped to
 * the grid; kept unless AIR with phi <= 1e-8 */
int orc_render_cache_mr(const orc_mr* g, const orc_phase* pp, int64_t n, const double* x, const double* y,
                        const double* z, const double* u, const double* v, const double* w, const uint8_t* type,
                        const double* vol, const orc_render_meta* m, const char* path) {
  const orc_iface_src s = mr_src(g);
  const int64_t nl = (int64_t)g->nleaves * 64;
  int32_t* q = (int32_t*)malloc(sizeof(int32_t) * 5 * (size_t)(nl > 0 ? nl : 1));
  double* pv = (double*)malloc(sizeof(double) * 2 * (size_t)(nl > 0 ? nl : 1));
  double total = 0.0;
  int64_t nc = 0;
  for (int64_t id = 0; id < nl; ++id) {
    int x0, y0, z0, st;
    cell_box(g, id, &x0, &y0, &z0, &st);
    if (x0 >= g->nx || y0 >= g->ny || z0 >= g->nz) continue;
    const int ci = x0 + st / 2 < g->nx - 1 ? x0 + st / 2 : g->nx - 1;
    const int cj = y0 + st / 2 < g->ny - 1 ? y0 + st / 2 : g->ny - 1;
    const int ck = z0 + st / 2 < g->nz - 1 ? z0 + st / 2 : g->nz - 1;
    const int mk = (int)(g->mk[id] + 0.5);
    const double phi = orc_iface_cell_phi(&s, pp, ci, cj, ck);
    if (mk == 0 && phi <= 1e-8) continue;
    const double cd = g->dx * (double)st;
    const double wv = phi * cd * cd * cd;
    total += wv;
    int32_t* r = q + 5 * nc;
    r[0] = x0; r[1] = y0; r[2] = z0; r[3] = g->leaves[4 * (id / 64)]; r[4] = mk;
    pv[2 * nc] = phi;
    pv[2 * nc + 1] = wv;
    ++nc;
  }
  const int r = orc_render_frame_cells(path, "multires3d_tp", g->nx, g->ny, g->nz, g->dx, m, q, pv, nc, total, n, x,
                                       y, z, u, v, w, type, vol);
  free(pv);
  free(q);
  return r;
}

/* MRSim3DTP::step, multires_sim3d_tp.cpp:671-734 (no particle adaptivity or
 * secondary particles; interface diagnostics and optional surface tension) */
void orc_mr_step(orc_mr* g, orc_mr_sim* sim, int64_t n, double* x, double* y, double* z, double* u, double* v,
                 double* w, const uint8_t* type, const double* vol, orc_mr_step_stats* S) {
  orc_mr_step_stats local;
  if (!S) S = &local;
  memset(S, 0, sizeof(*S));
  orc_mr_dynamic_refinement(g, sim, n, x, y, z, type);
  orc_timestep(n, u, v, w, g->dx, sim->dt, sim->adaptive_timestep, sim->adaptive_cfl, sim->adaptive_min_dt, &S->ts);
  double dt = S->ts.effective_dt;
  double lv = 0.0;
  for (int64_t p = 0; p < n; ++p)
    if (type[p] == 0) lv += (vol ? vol[p] : 1.0) * sim->Vp;
  S->c_div = 0.0;
  if (sim->c_div_volume_correction && sim->c_div_strength > 0.0 && sim->liquid_volume_target > 0.0 && dt > 0.0)
    S->c_div = sim->c_div_strength * (sim->liquid_volume_target - lv) / (dt * sim->liquid_volume_target);
  orc_mr_mark_cells(g, n, x, y, z);
  orc_mr_p2g(g, n, x, y, z, u, v, w, type, vol, &sim->phase, sim->Vp);
  orc_mr_interface_diagnostics(g, &sim->phase, &S->iface);
  orc_mr* saved = orc_mr_clone(g);
  orc_mr_apply_gravity(g, dt, sim->gravity);
  if (sim->surface_tension)
    orc_mr_surface_tension(g, &sim->phase, dt, sim->surface_tension_strength, sim->surface_tension_max_delta_speed,
                           sim->surface_tension_smoothing_radius, &S->surface);
  else
    orc_st_init(&S->surface, &sim->phase, 0.0, 0.0, 0.0, 0.0, 0);
  orc_mr_apply_walls(g);
  orc_mr_solve_cfg cf = sim->solve;
  cf.divergence_correction = S->c_div;
  orc_mr_project(g, &sim->phase, dt, &cf, &S->solve);
  orc_mr_g2p(g, saved, n, x, y, z, u, v, w, type, sim->alpha_liquid, sim->alpha_gas);
  orc_mr_advect(g, n, x, y, z, type, dt, sim->advection_order, &S->escape);
  orc_mr_destroy(saved);
}

/* the reference's initBubbleTankInterfaceBand layout (refineBubbleStaticBand,
 * multires_sim3d_tp.cpp:212-228) on a level-1 base */
int64_t orc_mr_layout_bubble_band(int nx, int ny, int nz, int32_t* out, int64_t cap) {
  int water = ny / 2;
  int top = water + 2 < ny ? water + 2 : ny;
  int32_t b1[6] = {0, 0, 0, nx, top, nz};
  int64_t cnt = orc_mr_layout_fine_box(nx, ny, nz, 1, 1, b1, NULL, 0);
  int32_t* lv = (int32_t*)malloc(sizeof(int32_t) * 4 * (size_t)(8 * cnt + 8));
  orc_mr_layout_fine_box(nx, ny, nz, 1, 1, b1, lv, cnt);
  /* second refineBox on the same layout: only level-1 blocks split */
  int cx0 = nx / 2 - nx / 4 > 0 ? nx / 2 - nx / 4 : 0, cx1 = nx / 2 + nx / 4 < nx ? nx / 2 + nx / 4 : nx;
  int cy0 = water / 4 > 0 ? water / 4 : 0, cy1 = water < ny ? water : ny;
  int cz0 = nz / 2 - nz / 4 > 0 ? nz / 2 - nz / 4 : 0, cz1 = nz / 2 + nz / 4 < nz ? nz / 2 + nz / 4 : nz;
  int64_t m = cnt;
  if (cx0 < cx1 && cy0 < cy1 && cz0 < cz1) {
    int32_t* nxt = (int32_t*)malloc(sizeof(int32_t) * 4 * (size_t)(8 * cnt + 8));
    m = 0;
    for (int64_t a = 0; a < cnt; ++a) {
      const int32_t* key = lv + 4 * a;
      int s = MRB * (1 << key[0]);
      int bx0 = key[1] * s, by0 = key[2] * s, bz0 = key[3] * s;
      int bx1 = bx0 + s < nx ? bx0 + s : nx, by1 = by0 + s < ny ? by0 + s : ny, bz1 = bz0 + s < nz ? bz0 + s : nz;
      int ov = bx0 < cx1 && bx1 > cx0 && by0 < cy1 && by1 > cy0 && bz0 < cz1 && bz1 > cz0;
      if (!ov || key[0] == 0) { memcpy(nxt + 4 * m, key, 16); ++m; continue; }
      for (int cz = 0; cz < 2; ++cz)
        for (int cy = 0; cy < 2; ++cy)
          for (int cx = 0; cx < 2; ++cx) {
            int32_t ch[4] = {key[0] - 1, key[1] * 2 + cx, key[2] * 2 + cy, key[3] * 2 + cz};
            if (block_intersects(nx, ny, nz, ch)) { memcpy(nxt + 4 * m, ch, 16); ++m; }
          }
    }
    m = sort_unique(nxt, (int)m);
    memcpy(lv, nxt, 16 * (size_t)m);
    free(nxt);
  }
  if (out) memcpy(out, lv, 16 * (size_t)(m < cap ? m : cap));
  free(lv);
  return m;
}
