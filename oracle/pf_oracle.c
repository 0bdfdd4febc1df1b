/*
 * pf_oracle.c -- CPU restatement of the reference's sparse two-phase PF-FLIP step.
 *
 * TEST INFRASTRUCTURE ONLY (see pf_oracle.h): the parity checker, never part of
 * the product.  Compiled with -ffp-contract=off so that every double expression
 * rounds exactly as written, like the reference's own loops.
 *
 * Storage: each field keeps its values densely by block id (bid*64 + local
 * offset) plus the reference's block map (bid -> first-touch pool index or -1)
 * and the first-touch activation order.  Inactive blocks always hold zeros, which
 * is what SparseBlockGrid3D::get returns for them (sparse_block_grid3d.h:50-55).
 */
#include "pf_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define BLK 4
#define BVOL 64

typedef struct orc_field {
  int nx, ny, nz, nbx, nby, nbz;
  int64_t nblocks, nactive;
  int32_t* map;   /* bid -> pool index (first-touch rank) or -1 */
  int32_t* order; /* pool index -> bid */
  double* val;    /* nblocks * 64 */
} orc_field;

struct orc_grid {
  int nx, ny, nz;
  double dx, ox, oy, oz;
  int faithful;
  orc_field f[ORC_NFIELDS];
};

/* ------------------------------------------------------------------ fields */

static void field_init(orc_field* f, int nx, int ny, int nz) {
  f->nx = nx; f->ny = ny; f->nz = nz;
  /* sparse_block_grid3d.h:18-21 */
  f->nbx = (nx + BLK - 1) / BLK; f->nby = (ny + BLK - 1) / BLK; f->nbz = (nz + BLK - 1) / BLK;
  f->nblocks = (int64_t)f->nbx * f->nby * f->nbz;
  f->nactive = 0;
  f->map = (int32_t*)malloc(sizeof(int32_t) * f->nblocks);
  f->order = (int32_t*)malloc(sizeof(int32_t) * f->nblocks);
  f->val = (double*)calloc((size_t)f->nblocks * BVOL, sizeof(double));
  for (int64_t b = 0; b < f->nblocks; ++b) f->map[b] = -1;
}

static void field_free(orc_field* f) {
  free(f->map); free(f->order); free(f->val);
}

static void field_copy(orc_field* d, const orc_field* s) {
  *d = *s;
  d->map = (int32_t*)malloc(sizeof(int32_t) * s->nblocks);
  d->order = (int32_t*)malloc(sizeof(int32_t) * s->nblocks);
  d->val = (double*)malloc(sizeof(double) * (size_t)s->nblocks * BVOL);
  memcpy(d->map, s->map, sizeof(int32_t) * s->nblocks);
  memcpy(d->order, s->order, sizeof(int32_t) * s->nblocks);
  memcpy(d->val, s->val, sizeof(double) * (size_t)s->nblocks * BVOL);
}

static inline int64_t bid_of(const orc_field* f, int bx, int by, int bz) {
  return bx + (int64_t)f->nbx * (by + (int64_t)f->nby * bz); /* sparse_block_grid3d.h:24 */
}

static inline int local_of(int i, int j, int k) {
  return (i % BLK) + BLK * ((j % BLK) + BLK * (k % BLK)); /* sparse_block_grid3d.h:48 */
}

/* activateBlock, sparse_block_grid3d.h:34-44 (zero fill is the clear invariant) */
static inline void field_activate(orc_field* f, int bx, int by, int bz) {
  int64_t b = bid_of(f, bx, by, bz);
  if (f->map[b] < 0) {
    f->map[b] = (int32_t)f->nactive;
    f->order[f->nactive++] = (int32_t)b;
  }
}

/* ref(), sparse_block_grid3d.h:46-49: activates and returns the slot */
static inline double* field_ref(orc_field* f, int i, int j, int k) {
  field_activate(f, i / BLK, j / BLK, k / BLK);
  return &f->val[bid_of(f, i / BLK, j / BLK, k / BLK) * BVOL + local_of(i, j, k)];
}

/* get(), sparse_block_grid3d.h:50-55 */
static inline double field_get(const orc_field* f, int i, int j, int k) {
  if (i < 0 || i >= f->nx || j < 0 || j >= f->ny || k < 0 || k >= f->nz) return 0.0;
  int64_t b = bid_of(f, i / BLK, j / BLK, k / BLK);
  if (f->map[b] < 0) return 0.0;
  return f->val[b * BVOL + local_of(i, j, k)];
}

/* clear(), sparse_block_grid3d.h:64-67 */
static void field_clear(orc_field* f) {
  for (int64_t a = 0; a < f->nactive; ++a) {
    int64_t b = f->order[a];
    f->map[b] = -1;
    memset(&f->val[b * BVOL], 0, sizeof(double) * BVOL);
  }
  f->nactive = 0;
}

static inline void block_coords(const orc_field* f, int64_t b, int* bx, int* by, int* bz) {
  *bx = (int)(b % f->nbx); /* sparse_block_grid3d.h:58-63 */
  int64_t t = b / f->nbx;
  *by = (int)(t % f->nby);
  *bz = (int)(t / f->nby);
}

/* store rounding: the reference's pool is float */
static inline double st(const orc_grid* g, double v) { return g->faithful ? (double)(float)v : v; }
/* `float += (float)x` in faithful mode, double accumulate otherwise */
static inline double acc(const orc_grid* g, double cur, double add) {
  return g->faithful ? (double)((float)cur + (float)add) : cur + add;
}

orc_grid* orc_grid_create(int nx, int ny, int nz, double dx, double ox, double oy, double oz,
                          int faithful) {
  orc_grid* g = (orc_grid*)calloc(1, sizeof(orc_grid));
  g->nx = nx; g->ny = ny; g->nz = nz; g->dx = dx; g->ox = ox; g->oy = oy; g->oz = oz;
  g->faithful = faithful;
  /* sparse_mac_grid3d.h:12-17: face extents */
  field_init(&g->f[ORC_U], nx + 1, ny, nz);
  field_init(&g->f[ORC_V], nx, ny + 1, nz);
  field_init(&g->f[ORC_W], nx, ny, nz + 1);
  field_init(&g->f[ORC_MU], nx + 1, ny, nz);
  field_init(&g->f[ORC_MV], nx, ny + 1, nz);
  field_init(&g->f[ORC_MW], nx, ny, nz + 1);
  field_init(&g->f[ORC_P], nx, ny, nz);
  field_init(&g->f[ORC_MK], nx, ny, nz);
  return g;
}

/* switch the storage rounding of an existing grid (float-faithful or FP64) */
void orc_grid_set_faithful(orc_grid* g, int faithful) { g->faithful = faithful ? 1 : 0; }

void orc_grid_destroy(orc_grid* g) {
  if (!g) return;
  for (int f = 0; f < ORC_NFIELDS; ++f) field_free(&g->f[f]);
  free(g);
}

orc_grid* orc_grid_clone(const orc_grid* s) {
  orc_grid* g = (orc_grid*)calloc(1, sizeof(orc_grid));
  *g = *s;
  for (int f = 0; f < ORC_NFIELDS; ++f) field_copy(&g->f[f], &s->f[f]);
  return g;
}

void orc_field_dims(const orc_grid* g, int field, int32_t out[3]) {
  out[0] = g->f[field].nx; out[1] = g->f[field].ny; out[2] = g->f[field].nz;
}

double orc_get(const orc_grid* g, int field, int i, int j, int k) {
  return field_get(&g->f[field], i, j, k);
}

void orc_set(orc_grid* g, int field, int i, int j, int k, double v) {
  *field_ref(&g->f[field], i, j, k) = st(g, v);
}

void orc_field_dense(const orc_grid* g, int field, double* out) {
  const orc_field* f = &g->f[field];
  for (int k = 0; k < f->nz; ++k)
    for (int j = 0; j < f->ny; ++j)
      for (int i = 0; i < f->nx; ++i)
        out[i + (int64_t)f->nx * (j + (int64_t)f->ny * k)] = field_get(f, i, j, k);
}

int64_t orc_active_blocks(const orc_grid* g, int field, int32_t* out) {
  const orc_field* f = &g->f[field];
  if (out) memcpy(out, f->order, sizeof(int32_t) * f->nactive);
  return f->nactive;
}

void orc_clear_field(orc_grid* g, int field) { field_clear(&g->f[field]); }

static inline int cell_of(const orc_grid* g, int i, int j, int k) {
  /* SparseMacGrid3D::cell, sparse_mac_grid3d.h:36 */
  return (int)(field_get(&g->f[ORC_MK], i, j, k) + 0.5);
}
static inline int in_bounds(const orc_grid* g, int i, int j, int k) {
  return i >= 0 && i < g->nx && j >= 0 && j < g->ny && k >= 0 && k < g->nz;
}
static inline int is_fluid(const orc_grid* g, int i, int j, int k) { /* sparse_ops3d_tp.cpp:77-79 */
  return in_bounds(g, i, j, k) && cell_of(g, i, j, k) == ORC_FLUID;
}
static inline int is_solid(const orc_grid* g, int i, int j, int k) { /* :81-83 */
  return !in_bounds(g, i, j, k) || cell_of(g, i, j, k) == ORC_SOLID;
}
static inline int is_air(const orc_grid* g, int i, int j, int k) { /* :85-87 */
  return in_bounds(g, i, j, k) && cell_of(g, i, j, k) == ORC_AIR;
}

/* ----------------------------------------------------------------- physics */

/* phiFromRawDensity, src/physics/phasefield.h:10-16 (Eq. 7) */
double orc_phi(double rt, const orc_phase* pp) {
  double rmin = log(pp->rho_l / pp->rho_g) * pp->rho_g * pp->rho_tilde_0;
  if (rt < rmin) return 0.0;
  double d = rt - rmin;
  if (d < 0.0) d = 0.0;
  double v = sqrt(d / (pp->alpha_phi * pp->rho_tilde_0 * pp->rho_l));
  return 1.0 < v ? 1.0 : v;  /* std::min(v, 1.0): NaN stays NaN */
}

/* betaFromPhi, src/physics/phasefield.h:17-18 */
double orc_beta(double phi, const orc_phase* pp) {
  return 1.0 / (phi * pp->rho_l + (1.0 - phi) * pp->rho_g);
}

static inline double beta_of_raw(double raw, const orc_phase* pp) { /* sparse_ops3d_tp.cpp:89-91 */
  return orc_beta(orc_phi(raw, pp), pp);
}

/* kern, sparse_ops3d_tp.cpp:15-19 (Eq. 6 un-normalized weight) */
static inline double kern(double d2, double r) {
  double q = d2 / (r * r);
  double t = 1.0 - q;
  return (t > 0.0) ? t * t * t : 0.0;
}
#define KR 1.5

/* splatN, src/transfer/transfer3d_tp.cpp:9-27 (dense, double) */
static void dense_splat(double* field, double* mass, int sw, int W, int H, int D, double gx,
                        double gy, double gz, double mom, double m) {
  int rad = (int)ceil(KR);
  int i0 = (int)floor(gx), j0 = (int)floor(gy), k0 = (int)floor(gz);
  double wsum = 0.0;
  for (int dk = -rad; dk <= rad + 1; ++dk)
    for (int dj = -rad; dj <= rad + 1; ++dj)
      for (int di = -rad; di <= rad + 1; ++di) {
        int ii = i0 + di, jj = j0 + dj, kk = k0 + dk;
        if (ii < 0 || ii >= W || jj < 0 || jj >= H || kk < 0 || kk >= D) continue;
        double dx = gx - ii, dy = gy - jj, dz = gz - kk;
        wsum += kern(dx * dx + dy * dy + dz * dz, KR);
      }
  if (wsum <= 0.0) return;
  for (int dk = -rad; dk <= rad + 1; ++dk)
    for (int dj = -rad; dj <= rad + 1; ++dj)
      for (int di = -rad; di <= rad + 1; ++di) {
        int ii = i0 + di, jj = j0 + dj, kk = k0 + dk;
        if (ii < 0 || ii >= W || jj < 0 || jj >= H || kk < 0 || kk >= D) continue;
        double dx = gx - ii, dy = gy - jj, dz = gz - kk;
        double w = kern(dx * dx + dy * dy + dz * dz, KR) / wsum;
        if (w <= 0.0) continue;
        int64_t idx = ii + (int64_t)sw * (jj + (int64_t)H * kk);
        field[idx] += w * mom;
        mass[idx] += w * m;
      }
}

/* calibrateRhoTilde0, src/transfer/transfer3d_tp.cpp:29-36: raw u-face mass at
 * (4,4,4) of an 8^3 lattice of 8 liquid particles per cell, over rho_l. */
double orc_calibrate_rho_tilde0(const orc_phase* pp, double Vp) {
  const int n = 8;
  int64_t nu = (int64_t)(n + 1) * n * n;
  double* U = (double*)calloc(nu, sizeof(double));
  double* M = (double*)calloc(nu, sizeof(double));
  double mp = pp->rho_l * Vp;
  for (int k = 0; k < n; ++k)
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i)
        for (int s = 0; s < 8; ++s) {
          double x = i + 0.25 + 0.5 * (s & 1), y = j + 0.25 + 0.5 * ((s >> 1) & 1),
                 z = k + 0.25 + 0.5 * ((s >> 2) & 1);
          dense_splat(U, M, n + 1, n + 1, n, n, x, y - 0.5, z - 0.5, 0.0, mp);
        }
  double raw = M[4 + (int64_t)(n + 1) * (4 + (int64_t)n * 4)];
  free(U); free(M);
  return (raw > 0) ? raw / pp->rho_l : 1.0;
}

/* p2g_tp, src/transfer/transfer3d_tp.cpp:37-48 (dense two-phase splat + normalize) */
void orc_dense_p2g(int nx, int ny, int nz, double dx, int64_t n, const double* x,
                   const double* y, const double* z, const double* u, const double* v,
                   const double* w, const uint8_t* type, const double* vol,
                   const orc_phase* pp, double Vp, double* U, double* V, double* W, double* MU,
                   double* MV, double* MW) {
  int64_t nu = (int64_t)(nx + 1) * ny * nz, nv = (int64_t)nx * (ny + 1) * nz,
          nw = (int64_t)nx * ny * (nz + 1);
  memset(U, 0, nu * 8); memset(MU, 0, nu * 8);
  memset(V, 0, nv * 8); memset(MV, 0, nv * 8);
  memset(W, 0, nw * 8); memset(MW, 0, nw * 8);
  for (int64_t p = 0; p < n; ++p) {
    double rho = (type[p] == 0) ? pp->rho_l : pp->rho_g;
    double mp = rho * Vp * (vol ? vol[p] : 1.0);
    double px = x[p] / dx, py = y[p] / dx, pz = z[p] / dx;
    dense_splat(U, MU, nx + 1, nx + 1, ny, nz, px, py - 0.5, pz - 0.5, mp * u[p], mp);
    dense_splat(V, MV, nx, nx, ny + 1, nz, px - 0.5, py, pz - 0.5, mp * v[p], mp);
    dense_splat(W, MW, nx, nx, ny, nz + 1, px - 0.5, py - 0.5, pz, mp * w[p], mp);
  }
  for (int64_t i = 0; i < nu; ++i) if (MU[i] > 0) U[i] /= MU[i];
  for (int64_t i = 0; i < nv; ++i) if (MV[i] > 0) V[i] /= MV[i];
  for (int64_t i = 0; i < nw; ++i) if (MW[i] > 0) W[i] /= MW[i];
}

/* ------------------------------------------------------------------ stages */

/* markCells, src/driver/sparse_sim3d_tp.cpp:22-50 */
void orc_mark_cells(orc_grid* g, int64_t n, const double* x, const double* y, const double* z) {
  orc_field* mk = &g->f[ORC_MK];
  field_clear(mk);
  for (int k = 0; k < g->nz; ++k)
    for (int j = 0; j < g->ny; ++j) {
      *field_ref(mk, 0, j, k) = ORC_SOLID;
      *field_ref(mk, g->nx - 1, j, k) = ORC_SOLID;
    }
  for (int k = 0; k < g->nz; ++k)
    for (int i = 0; i < g->nx; ++i) {
      *field_ref(mk, i, 0, k) = ORC_SOLID;
      *field_ref(mk, i, g->ny - 1, k) = ORC_SOLID;
    }
  for (int j = 0; j < g->ny; ++j)
    for (int i = 0; i < g->nx; ++i) {
      *field_ref(mk, i, j, 0) = ORC_SOLID;
      *field_ref(mk, i, j, g->nz - 1) = ORC_SOLID;
    }
  for (int64_t p = 0; p < n; ++p) {
    int i = (int)((x[p] - g->ox) / g->dx);
    int j = (int)((y[p] - g->oy) / g->dx);
    int k = (int)((z[p] - g->oz) / g->dx);
    if (in_bounds(g, i, j, k) && cell_of(g, i, j, k) != ORC_SOLID) *field_ref(mk, i, j, k) = ORC_FLUID;
  }
}

/* preactivateKernel, sparse_ops3d_tp.cpp:23-40 */
static void preactivate(orc_field* f, double gx, double gy, double gz, int nx, int ny, int nz) {
  int rad = (int)ceil(KR);
  int i0 = (int)floor(gx), j0 = (int)floor(gy), k0 = (int)floor(gz);
  for (int dk = -rad; dk <= rad + 1; ++dk)
    for (int dj = -rad; dj <= rad + 1; ++dj)
      for (int di = -rad; di <= rad + 1; ++di) {
        int i = i0 + di, j = j0 + dj, k = k0 + dk;
        if (i < 0 || i >= nx || j < 0 || j >= ny || k < 0 || k >= nz) continue;
        double dx = gx - i, dy = gy - j, dz = gz - k;
        if (kern(dx * dx + dy * dy + dz * dz, KR) <= 0.0) continue;
        field_activate(f, i / BLK, j / BLK, k / BLK);
      }
}

/* splatKernelColor, sparse_ops3d_tp.cpp:42-75, with all eight colours merged: a
 * node receives its contributions in particle order either way, so the float
 * sums are identical to the reference's colour-parallel splat. */
static void splat(const orc_grid* g, orc_field* fld, orc_field* mass, double gx, double gy,
                  double gz, double mom, double m, int nx, int ny, int nz) {
  int rad = (int)ceil(KR);
  int i0 = (int)floor(gx), j0 = (int)floor(gy), k0 = (int)floor(gz);
  double wsum = 0.0;
  for (int dk = -rad; dk <= rad + 1; ++dk)
    for (int dj = -rad; dj <= rad + 1; ++dj)
      for (int di = -rad; di <= rad + 1; ++di) {
        int i = i0 + di, j = j0 + dj, k = k0 + dk;
        if (i < 0 || i >= nx || j < 0 || j >= ny || k < 0 || k >= nz) continue;
        double dx = gx - i, dy = gy - j, dz = gz - k;
        wsum += kern(dx * dx + dy * dy + dz * dz, KR);
      }
  if (wsum <= 0.0) return;
  for (int dk = -rad; dk <= rad + 1; ++dk)
    for (int dj = -rad; dj <= rad + 1; ++dj)
      for (int di = -rad; di <= rad + 1; ++di) {
        int i = i0 + di, j = j0 + dj, k = k0 + dk;
        if (i < 0 || i >= nx || j < 0 || j >= ny || k < 0 || k >= nz) continue;
        double dx = gx - i, dy = gy - j, dz = gz - k;
        double w = kern(dx * dx + dy * dy + dz * dz, KR) / wsum;
        if (w <= 0.0) continue;
        double* a = field_ref(fld, i, j, k);
        *a = acc(g, *a, w * mom);
        double* b = field_ref(mass, i, j, k);
        *b = acc(g, *b, w * m);
      }
}

static void normalize_axis(orc_grid* g, int fv, int fm, int lim_i, int lim_j, int lim_k) {
  orc_field* V = &g->f[fv];
  orc_field* M = &g->f[fm];
  for (int64_t a = 0; a < M->nactive; ++a) {
    int bx, by, bz;
    block_coords(M, M->order[a], &bx, &by, &bz);
    for (int lz = 0; lz < 4; ++lz)
      for (int ly = 0; ly < 4; ++ly)
        for (int lx = 0; lx < 4; ++lx) {
          int i = bx * 4 + lx, j = by * 4 + ly, k = bz * 4 + lz;
          if (i > lim_i || j > lim_j || k > lim_k) continue;
          double m = field_get(M, i, j, k);
          if (m > 0.0) {
            double q = g->faithful ? (double)((float)field_get(V, i, j, k) / (float)m)
                                   : field_get(V, i, j, k) / m;
            *field_ref(V, i, j, k) = q;
          }
        }
  }
}

/* ------------------------------------------------------------ host threads
 * orc_set_threads(n > 1) splits the particle loops of P2G, G2P and advection over n
 * POSIX threads so the checker finishes at 64^3-128^3 in seconds.  Every result is
 * bit-identical to the serial loops:
 *  - P2G: thread t owns a slab of block rows along z of each face field and walks all
 *    particles in order, writing only its slab's nodes, so every node still receives
 *    its contributions in particle order.  Activation records, per block, the first
 *    (particle, stencil position) that touches it; sorting those keys rebuilds the
 *    serial first-touch order exactly.
 *  - G2P / advection: per-particle work on disjoint particle ranges; the clamp
 *    counters are integer sums.
 * The FP64 pressure solve stays serial (its dot products are order-dependent). */
static int g_threads = 1;      /* P2G */
static int g_threads_adv = 1;  /* G2P and advection */

static int clamp_threads(int n) { return n < 1 ? 1 : (n > 256 ? 256 : n); }
void orc_set_threads(int n) { g_threads = g_threads_adv = clamp_threads(n); }
void orc_set_threads_split(int p2g, int particles) {
  g_threads = clamp_threads(p2g);
  g_threads_adv = clamp_threads(particles);
}
int orc_get_threads(void) { return g_threads; }

typedef void (*orc_task_fn)(void* arg, int t, int nt);
typedef struct { orc_task_fn fn; void* arg; int t, nt; } orc_task;

static void* orc_task_main(void* a) {
  orc_task* k = (orc_task*)a;
  k->fn(k->arg, k->t, k->nt);
  return NULL;
}

/* fn(arg, t, nt) for t = 0..nt-1 on nt threads (the caller runs t = 0) */
static void parallel_run(int nt, orc_task_fn fn, void* arg) {
  if (nt <= 1) { fn(arg, 0, 1); return; }
  pthread_t th[256];
  orc_task tk[256];
  int started[256];
  for (int t = 0; t < nt; ++t) {
    tk[t].fn = fn; tk[t].arg = arg; tk[t].t = t; tk[t].nt = nt;
    started[t] = 0;
  }
  for (int t = 1; t < nt; ++t) started[t] = pthread_create(&th[t], NULL, orc_task_main, &tk[t]) == 0;
  fn(arg, 0, nt);
  for (int t = 1; t < nt; ++t) {
    if (started[t]) pthread_join(th[t], NULL);
    else fn(arg, t, nt); /* thread creation failed: run that share here */
  }
}

typedef struct {
  orc_grid* g;
  int64_t n;
  const double *x, *y, *z, *u, *v, *w, *vol;
  const uint8_t* type;
  const orc_phase* pp;
  double Vp;
  int64_t* key[3]; /* first-touch key per block of the u, v, w face fields */
} p2g_job;

/* the slab of block rows [*b0, *b1) along z of a field with nbz rows, for thread t */
static void slab_rows(int nbz, int t, int nt, int* b0, int* b1) {
  *b0 = (int)((int64_t)nbz * t / nt);
  *b1 = (int)((int64_t)nbz * (t + 1) / nt);
}

static void p2g_offsets(const orc_grid* g, int axis, double px, double py, double pz, double* gx, double* gy,
                        double* gz) {
  *gx = axis == 0 ? px : px - 0.5;
  *gy = axis == 1 ? py : py - 0.5;
  *gz = axis == 2 ? pz : pz - 0.5;
  (void)g;
}

/* preactivateKernel restricted to node rows [klo, khi): the least (particle, stencil
 * position) touching each block, as p * 216 + position */
static void p2g_activate_task(void* a, int t, int nt) {
  p2g_job* J = (p2g_job*)a;
  orc_grid* g = J->g;
  const int rad = (int)ceil(KR);
  for (int axis = 0; axis < 3; ++axis) {
    orc_field* f = &g->f[ORC_U + axis];
    int b0, b1;
    slab_rows(f->nbz, t, nt, &b0, &b1);
    const int klo = 4 * b0, khi = 4 * b1;
    if (klo >= khi) continue;
    int64_t* key = J->key[axis];
    for (int64_t p = 0; p < J->n; ++p) {
      double px = (J->x[p] - g->ox) / g->dx, py = (J->y[p] - g->oy) / g->dx, pz = (J->z[p] - g->oz) / g->dx;
      double gx, gy, gz;
      p2g_offsets(g, axis, px, py, pz, &gx, &gy, &gz);
      int i0 = (int)floor(gx), j0 = (int)floor(gy), k0 = (int)floor(gz);
      if (k0 + rad + 1 < klo || k0 - rad >= khi) continue;
      int e = 0;
      for (int dk = -rad; dk <= rad + 1; ++dk)
        for (int dj = -rad; dj <= rad + 1; ++dj)
          for (int di = -rad; di <= rad + 1; ++di, ++e) {
            int i = i0 + di, j = j0 + dj, k = k0 + dk;
            if (k < klo || k >= khi) continue;
            if (i < 0 || i >= f->nx || j < 0 || j >= f->ny || k < 0 || k >= f->nz) continue;
            double ddx = gx - i, ddy = gy - j, ddz = gz - k;
            if (kern(ddx * ddx + ddy * ddy + ddz * ddz, KR) <= 0.0) continue;
            int64_t b = bid_of(f, i / BLK, j / BLK, k / BLK);
            if (key[b] == INT64_MAX) key[b] = p * 216 + e;
          }
    }
  }
}

static int cmp_key_bid(const void* a, const void* b) {
  const int64_t* x = (const int64_t*)a;
  const int64_t* y = (const int64_t*)b;
  return x[0] < y[0] ? -1 : (x[0] > y[0] ? 1 : 0);
}

/* activation order of field f from the first-touch keys */
static void activate_by_keys(orc_field* f, const int64_t* key) {
  int64_t cnt = 0;
  for (int64_t b = 0; b < f->nblocks; ++b) cnt += key[b] != INT64_MAX;
  int64_t* kb = (int64_t*)malloc(sizeof(int64_t) * 2 * (cnt > 0 ? cnt : 1));
  int64_t q = 0;
  for (int64_t b = 0; b < f->nblocks; ++b)
    if (key[b] != INT64_MAX) { kb[2 * q] = key[b]; kb[2 * q + 1] = b; ++q; }
  qsort(kb, (size_t)cnt, 2 * sizeof(int64_t), cmp_key_bid);
  for (int64_t a = 0; a < cnt; ++a) {
    int64_t b = kb[2 * a + 1];
    f->map[b] = (int32_t)f->nactive;
    f->order[f->nactive++] = (int32_t)b;
  }
  free(kb);
}

/* splatKernelColor restricted to node rows [klo, khi) of the thread's slab; every
 * written block is already active */
static void p2g_splat_task(void* a, int t, int nt) {
  p2g_job* J = (p2g_job*)a;
  orc_grid* g = J->g;
  const int rad = (int)ceil(KR);
  for (int axis = 0; axis < 3; ++axis) {
    orc_field* F = &g->f[ORC_U + axis];
    orc_field* M = &g->f[ORC_MU + axis];
    const double* vel = axis == 0 ? J->u : (axis == 1 ? J->v : J->w);
    int b0, b1;
    slab_rows(F->nbz, t, nt, &b0, &b1);
    const int klo = 4 * b0, khi = 4 * b1;
    if (klo >= khi) continue;
    const int nx = F->nx, ny = F->ny, nz = F->nz;
    for (int64_t p = 0; p < J->n; ++p) {
      double px = (J->x[p] - g->ox) / g->dx, py = (J->y[p] - g->oy) / g->dx, pz = (J->z[p] - g->oz) / g->dx;
      double gx, gy, gz;
      p2g_offsets(g, axis, px, py, pz, &gx, &gy, &gz);
      int i0 = (int)floor(gx), j0 = (int)floor(gy), k0 = (int)floor(gz);
      if (k0 + rad + 1 < klo || k0 - rad >= khi) continue;
      double rho = (J->type[p] == 0) ? J->pp->rho_l : J->pp->rho_g;
      double mp = rho * J->Vp * (J->vol ? J->vol[p] : 1.0);
      double mom = mp * vel[p];
      double wsum = 0.0;
      for (int dk = -rad; dk <= rad + 1; ++dk)
        for (int dj = -rad; dj <= rad + 1; ++dj)
          for (int di = -rad; di <= rad + 1; ++di) {
            int i = i0 + di, j = j0 + dj, k = k0 + dk;
            if (i < 0 || i >= nx || j < 0 || j >= ny || k < 0 || k >= nz) continue;
            double ddx = gx - i, ddy = gy - j, ddz = gz - k;
            wsum += kern(ddx * ddx + ddy * ddy + ddz * ddz, KR);
          }
      if (wsum <= 0.0) continue;
      for (int dk = -rad; dk <= rad + 1; ++dk)
        for (int dj = -rad; dj <= rad + 1; ++dj)
          for (int di = -rad; di <= rad + 1; ++di) {
            int i = i0 + di, j = j0 + dj, k = k0 + dk;
            if (k < klo || k >= khi) continue;
            if (i < 0 || i >= nx || j < 0 || j >= ny || k < 0 || k >= nz) continue;
            double ddx = gx - i, ddy = gy - j, ddz = gz - k;
            double wt = kern(ddx * ddx + ddy * ddy + ddz * ddz, KR) / wsum;
            if (wt <= 0.0) continue;
            int64_t off = bid_of(F, i / BLK, j / BLK, k / BLK) * BVOL + local_of(i, j, k);
            F->val[off] = acc(g, F->val[off], wt * mom);
            M->val[off] = acc(g, M->val[off], wt * mp);
          }
    }
  }
}

static void p2g_threaded(orc_grid* g, int64_t n, const double* x, const double* y, const double* z,
                         const double* u, const double* v, const double* w, const uint8_t* type,
                         const double* vol, const orc_phase* pp, double Vp) {
  p2g_job J = {g, n, x, y, z, u, v, w, vol, type, pp, Vp, {NULL, NULL, NULL}};
  for (int axis = 0; axis < 3; ++axis) {
    orc_field* f = &g->f[ORC_U + axis];
    J.key[axis] = (int64_t*)malloc(sizeof(int64_t) * f->nblocks);
    for (int64_t b = 0; b < f->nblocks; ++b) J.key[axis][b] = INT64_MAX;
  }
  parallel_run(g_threads, p2g_activate_task, &J);
  for (int axis = 0; axis < 3; ++axis) {
    activate_by_keys(&g->f[ORC_U + axis], J.key[axis]);  /* u, v, w and their masses: same sets */
    activate_by_keys(&g->f[ORC_MU + axis], J.key[axis]);
    free(J.key[axis]);
  }
  parallel_run(g_threads, p2g_splat_task, &J);
}

/* spP2G3D_tp, sparse_ops3d_tp.cpp:95-153 */
void orc_p2g(orc_grid* g, int64_t n, const double* x, const double* y, const double* z,
             const double* u, const double* v, const double* w, const uint8_t* type,
             const double* vol, const orc_phase* pp, double Vp) {
  for (int f = ORC_U; f <= ORC_MW; ++f) field_clear(&g->f[f]);
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  if (g_threads > 1) {
    p2g_threaded(g, n, x, y, z, u, v, w, type, vol, pp, Vp);
    n = 0; /* splat done: only the normalisation below remains */
  }
  for (int64_t p = 0; p < n; ++p) {
    double px = (x[p] - g->ox) / g->dx, py = (y[p] - g->oy) / g->dx, pz = (z[p] - g->oz) / g->dx;
    preactivate(&g->f[ORC_U], px, py - 0.5, pz - 0.5, nx + 1, ny, nz);
    preactivate(&g->f[ORC_MU], px, py - 0.5, pz - 0.5, nx + 1, ny, nz);
    preactivate(&g->f[ORC_V], px - 0.5, py, pz - 0.5, nx, ny + 1, nz);
    preactivate(&g->f[ORC_MV], px - 0.5, py, pz - 0.5, nx, ny + 1, nz);
    preactivate(&g->f[ORC_W], px - 0.5, py - 0.5, pz, nx, ny, nz + 1);
    preactivate(&g->f[ORC_MW], px - 0.5, py - 0.5, pz, nx, ny, nz + 1);
  }
  for (int64_t p = 0; p < n; ++p) {
    double rho = (type[p] == 0) ? pp->rho_l : pp->rho_g;
    double mp = rho * Vp * (vol ? vol[p] : 1.0);
    double px = (x[p] - g->ox) / g->dx, py = (y[p] - g->oy) / g->dx, pz = (z[p] - g->oz) / g->dx;
    splat(g, &g->f[ORC_U], &g->f[ORC_MU], px, py - 0.5, pz - 0.5, mp * u[p], mp, nx + 1, ny, nz);
    splat(g, &g->f[ORC_V], &g->f[ORC_MV], px - 0.5, py, pz - 0.5, mp * v[p], mp, nx, ny + 1, nz);
    splat(g, &g->f[ORC_W], &g->f[ORC_MW], px - 0.5, py - 0.5, pz, mp * w[p], mp, nx, ny, nz + 1);
  }
  normalize_axis(g, ORC_U, ORC_MU, nx, ny - 1, nz - 1);   /* :126-134 */
  normalize_axis(g, ORC_V, ORC_MV, nx - 1, ny, nz - 1);   /* :135-143 */
  normalize_axis(g, ORC_W, ORC_MW, nx - 1, ny - 1, nz);   /* :144-152 */
}

/* ------------------------------------------- single-phase path (sparse_ops3d.cpp) */

/* preactivateField, src/driver/sparse_ops3d.cpp:22-40: every in-bounds node of the
 * 2x2x2 trilinear stencil activates its block (zero weights included) */
static void preactivate_tri(orc_field* f, double gx, double gy, double gz, int nx, int ny, int nz) {
  int i0 = (int)floor(gx), j0 = (int)floor(gy), k0 = (int)floor(gz);
  for (int dk = 0; dk < 2; ++dk)
    for (int dj = 0; dj < 2; ++dj)
      for (int di = 0; di < 2; ++di) {
        int i = i0 + di, j = j0 + dj, k = k0 + dk;
        if (i < 0 || i >= nx || j < 0 || j >= ny || k < 0 || k >= nz) continue;
        field_activate(f, i / BLK, j / BLK, k / BLK);
      }
}

/* splatFieldColor, src/driver/sparse_ops3d.cpp:42-68, colours merged (particle order
 * per node is unchanged): w = (float)(wx*wy*wz); field += (float)(w*mom) */
static void splat_tri(const orc_grid* g, orc_field* fld, orc_field* mass, double gx, double gy,
                      double gz, double mom, double m, int nx, int ny, int nz) {
  int i0 = (int)floor(gx), j0 = (int)floor(gy), k0 = (int)floor(gz);
  double fx = gx - i0, fy = gy - j0, fz = gz - k0;
  double wx[2] = {1.0 - fx, fx}, wy[2] = {1.0 - fy, fy}, wz[2] = {1.0 - fz, fz};
  for (int dk = 0; dk < 2; ++dk)
    for (int dj = 0; dj < 2; ++dj)
      for (int di = 0; di < 2; ++di) {
        int i = i0 + di, j = j0 + dj, k = k0 + dk;
        if (i < 0 || i >= nx || j < 0 || j >= ny || k < 0 || k >= nz) continue;
        double w = st(g, wx[di] * wy[dj] * wz[dk]);
        double* a = field_ref(fld, i, j, k);
        *a = acc(g, *a, w * mom);
        double* b = field_ref(mass, i, j, k);
        *b = acc(g, *b, w * m);
      }
}

/* spP2G3D, src/driver/sparse_ops3d.cpp:72-125: trilinear splat, unit particle mass */
void orc_p2g_single(orc_grid* g, int64_t n, const double* x, const double* y, const double* z,
                    const double* u, const double* v, const double* w) {
  for (int f = ORC_U; f <= ORC_MW; ++f) field_clear(&g->f[f]);
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  const double mp = 1.0;
  for (int64_t p = 0; p < n; ++p) {
    double px = (x[p] - g->ox) / g->dx, py = (y[p] - g->oy) / g->dx, pz = (z[p] - g->oz) / g->dx;
    preactivate_tri(&g->f[ORC_U], px, py - 0.5, pz - 0.5, nx + 1, ny, nz);
    preactivate_tri(&g->f[ORC_MU], px, py - 0.5, pz - 0.5, nx + 1, ny, nz);
    preactivate_tri(&g->f[ORC_V], px - 0.5, py, pz - 0.5, nx, ny + 1, nz);
    preactivate_tri(&g->f[ORC_MV], px - 0.5, py, pz - 0.5, nx, ny + 1, nz);
    preactivate_tri(&g->f[ORC_W], px - 0.5, py - 0.5, pz, nx, ny, nz + 1);
    preactivate_tri(&g->f[ORC_MW], px - 0.5, py - 0.5, pz, nx, ny, nz + 1);
  }
  for (int64_t p = 0; p < n; ++p) {
    double px = (x[p] - g->ox) / g->dx, py = (y[p] - g->oy) / g->dx, pz = (z[p] - g->oz) / g->dx;
    splat_tri(g, &g->f[ORC_U], &g->f[ORC_MU], px, py - 0.5, pz - 0.5, mp * u[p], mp, nx + 1, ny, nz);
    splat_tri(g, &g->f[ORC_V], &g->f[ORC_MV], px - 0.5, py, pz - 0.5, mp * v[p], mp, nx, ny + 1, nz);
    splat_tri(g, &g->f[ORC_W], &g->f[ORC_MW], px - 0.5, py - 0.5, pz, mp * w[p], mp, nx, ny, nz + 1);
  }
  normalize_axis(g, ORC_U, ORC_MU, nx, ny - 1, nz - 1);   /* :101-110 */
  normalize_axis(g, ORC_V, ORC_MV, nx - 1, ny, nz - 1);   /* :111-119 */
  normalize_axis(g, ORC_W, ORC_MW, nx - 1, ny - 1, nz);   /* :120-124 */
}

/* applyGravity, src/driver/sparse_sim3d_tp.cpp:52-70 */
void orc_apply_gravity(orc_grid* g, double dt, double gravity) {
  orc_field* MV = &g->f[ORC_MV];
  orc_field* V = &g->f[ORC_V];
  double dv = dt * gravity;
  for (int64_t a = 0; a < MV->nactive; ++a) {
    int bx, by, bz;
    block_coords(MV, MV->order[a], &bx, &by, &bz);
    for (int lz = 0; lz < 4; ++lz)
      for (int ly = 0; ly < 4; ++ly)
        for (int lx = 0; lx < 4; ++lx) {
          int i = bx * 4 + lx, j = by * 4 + ly, k = bz * 4 + lz;
          if (i >= g->nx || j > g->ny || k >= g->nz) continue;
          if (field_get(MV, i, j, k) > 0.0) {
            double cur = field_get(V, i, j, k);
            *field_ref(V, i, j, k) = g->faithful ? (double)((float)cur + (float)dv) : cur + dv;
          }
        }
  }
}

/* applyWallBoundary, src/driver/sparse_sim3d_tp.cpp:72-97 */
void orc_apply_walls(orc_grid* g) {
  orc_field *U = &g->f[ORC_U], *V = &g->f[ORC_V], *W = &g->f[ORC_W];
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j) {
      *field_ref(U, 0, j, k) = 0.0; *field_ref(U, 1, j, k) = 0.0;
      *field_ref(U, nx - 1, j, k) = 0.0; *field_ref(U, nx, j, k) = 0.0;
    }
  for (int k = 0; k < nz; ++k)
    for (int i = 0; i < nx; ++i) {
      *field_ref(V, i, 0, k) = 0.0; *field_ref(V, i, 1, k) = 0.0;
      *field_ref(V, i, ny - 1, k) = 0.0; *field_ref(V, i, ny, k) = 0.0;
    }
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      *field_ref(W, i, j, 0) = 0.0; *field_ref(W, i, j, 1) = 0.0;
      *field_ref(W, i, j, nz - 1) = 0.0; *field_ref(W, i, j, nz) = 0.0;
    }
}

/* spProjectStepVC3D, sparse_ops3d_tp.cpp:234-423 (single = 0), and the single-phase
 * spProjectStep3D, sparse_ops3d.cpp:170-319 (single = 1: unit coefficients, no pin,
 * no c_div, the divergence summed left to right, :180-187) */
static void project_impl(orc_grid* g, const orc_phase* pp, double dt, int cg_iters, double cg_tol,
                         double div_correction, int stop_rule, int single, orc_solve_stats* stt) {
  orc_solve_stats dummy;
  orc_solve_stats* S = stt ? stt : &dummy;
  memset(S, 0, sizeof(*S));
  S->pinned_cell = -1;
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  const int64_t ncell = (int64_t)nx * ny * nz;
  orc_field* P = &g->f[ORC_P];
  orc_field* MK = &g->f[ORC_MK];
  field_clear(P);
  /* collectCellsWithMarker(g, 1), sparse_ops3d_common.h:65-85: sorted linear ids */
  char* isf = (char*)calloc(ncell, 1);
  for (int64_t a = 0; a < MK->nactive; ++a) {
    int bx, by, bz;
    block_coords(MK, MK->order[a], &bx, &by, &bz);
    for (int lz = 0; lz < 4; ++lz)
      for (int ly = 0; ly < 4; ++ly)
        for (int lx = 0; lx < 4; ++lx) {
          int i = bx * 4 + lx, j = by * 4 + ly, k = bz * 4 + lz;
          if (in_bounds(g, i, j, k) && cell_of(g, i, j, k) == ORC_FLUID)
            isf[i + (int64_t)nx * (j + (int64_t)ny * k)] = 1;
        }
  }
  int64_t N = 0;
  for (int64_t c = 0; c < ncell; ++c) N += isf[c];
  S->fluid_cells = N;
  if (N == 0) { free(isf); return; }
  int64_t* cells = (int64_t*)malloc(sizeof(int64_t) * N);
  int64_t* index = (int64_t*)malloc(sizeof(int64_t) * ncell); /* findSortedIndex by table */
  {
    int64_t t = 0;
    for (int64_t c = 0; c < ncell; ++c) {
      index[c] = -1;
      if (isf[c]) { index[c] = t; cells[t++] = c; }
    }
  }
  static const int di[6] = {1, -1, 0, 0, 0, 0};
  static const int dj[6] = {0, 0, 1, -1, 0, 0};
  static const int dk[6] = {0, 0, 0, 0, 1, -1};
  double scale = dt / (g->dx * g->dx);

  /* pin rule, :250-264 */
  int64_t pc = -1;
  int touches_air = 0;
  for (int64_t t = 0; t < N && !touches_air; ++t) {
    int i = (int)(cells[t] % nx);
    int64_t q = cells[t] / nx;
    int j = (int)(q % ny), k = (int)(q / ny);
    for (int nn = 0; nn < 6; ++nn)
      if (is_air(g, i + di[nn], j + dj[nn], k + dk[nn])) { touches_air = 1; break; }
  }
  if (!touches_air && !single) pc = cells[0];
  S->pinned = pc >= 0;
  S->pinned_cell = pc;

  double* x = (double*)calloc(N, 8);
  double* r = (double*)malloc(8 * N);
  double* z = (double*)malloc(8 * N);
  double* pd = (double*)malloc(8 * N);
  double* Ap = (double*)malloc(8 * N);
  double* b = (double*)malloc(8 * N);
  double* diag = (double*)calloc(N, 8);
  int64_t* nbr = (int64_t*)malloc(sizeof(int64_t) * 6 * N);
  double* coeff = (double*)calloc(6 * N, 8);
  orc_field *U = &g->f[ORC_U], *V = &g->f[ORC_V], *W = &g->f[ORC_W];
  orc_field *MU = &g->f[ORC_MU], *MV = &g->f[ORC_MV], *MW = &g->f[ORC_MW];

  /* RHS, :266-283 */
  for (int64_t t = 0; t < N; ++t) {
    if (cells[t] == pc) { r[t] = 0.0; continue; }
    int i = (int)(cells[t] % nx);
    int64_t q = cells[t] / nx;
    int j = (int)(q % ny), k = (int)(q / ny);
    if (single) {
      /* the reference sums the float face values in float arithmetic, then divides in double */
      double sum;
      if (g->faithful) {
        float fs = (float)field_get(U, i + 1, j, k) - (float)field_get(U, i, j, k) +
                   (float)field_get(V, i, j + 1, k) - (float)field_get(V, i, j, k) +
                   (float)field_get(W, i, j, k + 1) - (float)field_get(W, i, j, k);
        sum = fs;
      } else {
        sum = field_get(U, i + 1, j, k) - field_get(U, i, j, k) + field_get(V, i, j + 1, k) -
              field_get(V, i, j, k) + field_get(W, i, j, k + 1) - field_get(W, i, j, k);
      }
      r[t] = -(sum / g->dx);
      continue;
    }
    double uR = is_solid(g, i + 1, j, k) ? 0.0 : field_get(U, i + 1, j, k);
    double uL = is_solid(g, i - 1, j, k) ? 0.0 : field_get(U, i, j, k);
    double vT = is_solid(g, i, j + 1, k) ? 0.0 : field_get(V, i, j + 1, k);
    double vB = is_solid(g, i, j - 1, k) ? 0.0 : field_get(V, i, j, k);
    double wU = is_solid(g, i, j, k + 1) ? 0.0 : field_get(W, i, j, k + 1);
    double wD = is_solid(g, i, j, k - 1) ? 0.0 : field_get(W, i, j, k);
    r[t] = -((uR - uL) + (vT - vB) + (wU - wD)) / g->dx + div_correction;
  }
  /* assembly, :285-313 */
  for (int64_t t = 0; t < N; ++t) {
    int i = (int)(cells[t] % nx);
    int64_t q = cells[t] / nx;
    int j = (int)(q % ny), k = (int)(q / ny);
    double fb[6] = {1.0, 1.0, 1.0, 1.0, 1.0, 1.0};
    if (!single) {
      fb[0] = beta_of_raw(field_get(MU, i + 1, j, k), pp); fb[1] = beta_of_raw(field_get(MU, i, j, k), pp);
      fb[2] = beta_of_raw(field_get(MV, i, j + 1, k), pp); fb[3] = beta_of_raw(field_get(MV, i, j, k), pp);
      fb[4] = beta_of_raw(field_get(MW, i, j, k + 1), pp); fb[5] = beta_of_raw(field_get(MW, i, j, k), pp);
    }
    for (int nn = 0; nn < 6; ++nn) {
      nbr[6 * t + nn] = -1;
      int ni = i + di[nn], nj = j + dj[nn], nk = k + dk[nn];
      if (is_solid(g, ni, nj, nk)) continue;
      diag[t] += fb[nn];
      coeff[6 * t + nn] = fb[nn];
      int64_t nc = ni + (int64_t)nx * (nj + (int64_t)ny * nk);
      if (is_fluid(g, ni, nj, nk) && nc != pc) nbr[6 * t + nn] = index[nc];
    }
  }
#define APPLY_A(XX, OUT)                                                   \
  for (int64_t t = 0; t < N; ++t) {                                        \
    if (cells[t] == pc) { (OUT)[t] = (XX)[t]; continue; }                  \
    double off = 0.0;                                                      \
    for (int nn = 0; nn < 6; ++nn)                                         \
      if (nbr[6 * t + nn] >= 0) off += coeff[6 * t + nn] * (XX)[nbr[6 * t + nn]]; \
    (OUT)[t] = scale * (diag[t] * (XX)[t] - off);                          \
  }
#define PRECOND(IN, OUT)                                                   \
  for (int64_t t = 0; t < N; ++t) {                                        \
    double d = (cells[t] == pc) ? 1.0 : scale * diag[t];                   \
    (OUT)[t] = (d > 0.0) ? (IN)[t] / d : 0.0;                              \
  }
  memcpy(b, r, 8 * N);
  double bnorm2 = 0.0;
  for (int64_t t = 0; t < N; ++t) bnorm2 += b[t] * b[t];
  double bnorm = sqrt(bnorm2);

  /* Jacobi PCG, :315-364 */
  double res0 = 0.0;
  for (int64_t t = 0; t < N; ++t) res0 = fmax(res0, fabs(r[t]));
  S->initial_max_residual = res0;
  double res = res0;
  int done0 = (stop_rule == ORC_STOP_REL2) ? !(bnorm > 0.0) : !(res0 >= cg_tol);
  if (!done0) {
    PRECOND(r, z);
    memcpy(pd, z, 8 * N);
    double rz = 0.0;
    for (int64_t t = 0; t < N; ++t) rz += r[t] * z[t];
    for (int it = 0; it < cg_iters; ++it) {
      APPLY_A(pd, Ap);
      double pAp = 0.0;
      for (int64_t t = 0; t < N; ++t) pAp += pd[t] * Ap[t];
      if (fabs(pAp) < 1e-30) { S->breakdown = 1; break; }
      double alpha = rz / pAp;
      for (int64_t t = 0; t < N; ++t) { x[t] += alpha * pd[t]; r[t] -= alpha * Ap[t]; }
      S->iterations = it + 1;
      res = 0.0;
      for (int64_t t = 0; t < N; ++t) res = fmax(res, fabs(r[t]));
      if (stop_rule == ORC_STOP_REL2) {
        double rr = 0.0;
        for (int64_t t = 0; t < N; ++t) rr += r[t] * r[t];
        if (sqrt(rr) <= cg_tol * bnorm) { S->converged = 1; break; }
      } else if (res < cg_tol) { S->converged = 1; break; }
      PRECOND(r, z);
      double rzn = 0.0;
      for (int64_t t = 0; t < N; ++t) rzn += r[t] * z[t];
      double beta = rzn / rz;
      rz = rzn;
      for (int64_t t = 0; t < N; ++t) pd[t] = z[t] + beta * pd[t];
    }
  } else {
    S->converged = 1;
  }
  S->final_max_residual = res;
  /* true residual in FP64 */
  APPLY_A(x, Ap);
  {
    double rr = 0.0, rm = 0.0;
    for (int64_t t = 0; t < N; ++t) {
      double e = b[t] - Ap[t];
      rr += e * e;
      rm = fmax(rm, fabs(e));
    }
    S->true_rel_residual = bnorm > 0 ? sqrt(rr) / bnorm : 0.0;
    S->true_max_residual = rm;
  }
#undef APPLY_A
#undef PRECOND

  /* store pressure, :366-372 */
  for (int64_t t = 0; t < N; ++t) {
    int i = (int)(cells[t] % nx);
    int64_t q = cells[t] / nx;
    int j = (int)(q % ny), k = (int)(q / ny);
    *field_ref(P, i, j, k) = st(g, x[t]);
  }

  /* velocity update over the sorted unique face lists, :374-422 and
   * sparse_ops3d_common.h:92-138 */
  double s = dt / g->dx;
  for (int axis = 0; axis < 3; ++axis) {
    orc_field* F = axis == 0 ? U : axis == 1 ? V : W;
    orc_field* M = axis == 0 ? MU : axis == 1 ? MV : MW;
    int fnx = F->nx, fny = F->ny, fnz = F->nz;
    int64_t nf = (int64_t)fnx * fny * fnz;
    char* mark = (char*)calloc(nf, 1);
    for (int64_t t = 0; t < N; ++t) {
      int i = (int)(cells[t] % nx);
      int64_t q = cells[t] / nx;
      int j = (int)(q % ny), k = (int)(q / ny);
      if (axis == 0) {
        if (i > 0) mark[i + (int64_t)fnx * (j + (int64_t)fny * k)] = 1;
        if (i + 1 < nx) mark[(i + 1) + (int64_t)fnx * (j + (int64_t)fny * k)] = 1;
      } else if (axis == 1) {
        if (j > 0) mark[i + (int64_t)fnx * (j + (int64_t)fny * k)] = 1;
        if (j + 1 < ny) mark[i + (int64_t)fnx * ((j + 1) + (int64_t)fny * k)] = 1;
      } else {
        if (k > 0) mark[i + (int64_t)fnx * (j + (int64_t)fny * k)] = 1;
        if (k + 1 < nz) mark[i + (int64_t)fnx * (j + (int64_t)fny * (k + 1))] = 1;
      }
    }
    for (int64_t fid = 0; fid < nf; ++fid) {
      if (!mark[fid]) continue;
      int i = (int)(fid % fnx);
      int64_t q = fid / fnx;
      int j = (int)(q % fny), k = (int)(q / fny);
      int li = i - (axis == 0), lj = j - (axis == 1), lk = k - (axis == 2);
      int lf = is_fluid(g, li, lj, lk), rf = is_fluid(g, i, j, k);
      if (!lf && !rf) continue;
      if (is_solid(g, li, lj, lk) || is_solid(g, i, j, k)) {
        *field_ref(F, i, j, k) = 0.0;
        continue;
      }
      double pl = lf ? field_get(P, li, lj, lk) : 0.0;
      double pr = rf ? field_get(P, i, j, k) : 0.0;
      double cur = field_get(F, i, j, k);
      double dlt = single ? s * (pr - pl) : s * beta_of_raw(field_get(M, i, j, k), pp) * (pr - pl);
      *field_ref(F, i, j, k) = g->faithful ? (double)((float)cur - (float)dlt) : cur - dlt;
    }
    free(mark);
  }
  free(isf); free(cells); free(index); free(x); free(r); free(z); free(pd); free(Ap); free(b);
  free(diag); free(nbr); free(coeff);
}

void orc_project(orc_grid* g, const orc_phase* pp, double dt, int cg_iters, double cg_tol,
                 double div_correction, int stop_rule, orc_solve_stats* st) {
  project_impl(g, pp, dt, cg_iters, cg_tol, div_correction, stop_rule, 0, st);
}

/* spProjectStep3D, src/driver/sparse_ops3d.cpp:170-319 */
void orc_project_single(orc_grid* g, double dt, int cg_iters, double cg_tol, int stop_rule,
                        orc_solve_stats* st) {
  project_impl(g, NULL, dt, cg_iters, cg_tol, 0.0, stop_rule, 1, st);
}

/* sampleField, src/driver/sparse_ops3d_common.h:30-51 (returns float in the reference) */
static double sample_field(const orc_grid* g, const orc_field* f, double gx, double gy, double gz) {
  int i0 = (int)floor(gx), j0 = (int)floor(gy), k0 = (int)floor(gz);
  double fx = gx - i0, fy = gy - j0, fz = gz - k0;
  double wx[2] = {1.0 - fx, fx}, wy[2] = {1.0 - fy, fy}, wz[2] = {1.0 - fz, fz};
  double s = 0.0;
  for (int dk = 0; dk < 2; ++dk)
    for (int dj = 0; dj < 2; ++dj)
      for (int di = 0; di < 2; ++di) {
        int i = i0 + di, j = j0 + dj, k = k0 + dk;
        i = i < 0 ? 0 : (i > f->nx - 1 ? f->nx - 1 : i);
        j = j < 0 ? 0 : (j > f->ny - 1 ? f->ny - 1 : j);
        k = k < 0 ? 0 : (k > f->nz - 1 ? f->nz - 1 : k);
        s += wx[di] * wy[dj] * wz[dk] * field_get(f, i, j, k);
      }
  return g->faithful ? (double)(float)s : s;
}
/* sampleU/V/W, sparse_ops3d_common.h:53-63 */
static inline double sample_u(const orc_grid* g, double px, double py, double pz) {
  return sample_field(g, &g->f[ORC_U], px, py - 0.5, pz - 0.5);
}
static inline double sample_v(const orc_grid* g, double px, double py, double pz) {
  return sample_field(g, &g->f[ORC_V], px - 0.5, py, pz - 0.5);
}
static inline double sample_w(const orc_grid* g, double px, double py, double pz) {
  return sample_field(g, &g->f[ORC_W], px - 0.5, py - 0.5, pz);
}

/* spG2P3D_tp, sparse_ops3d_tp.cpp:155-172 */
static void g2p_range(const orc_grid* g, const orc_grid* saved, int64_t p0, int64_t p1, const double* x,
                      const double* y, const double* z, double* u, double* v, double* w,
                      const uint8_t* type, double aL, double aG) {
  for (int64_t p = p0; p < p1; ++p) {
    double a = (type[p] == 0) ? aL : aG;
    double px = (x[p] - g->ox) / g->dx, py = (y[p] - g->oy) / g->dx, pz = (z[p] - g->oz) / g->dx;
    double un = sample_u(g, px, py, pz), vn = sample_v(g, px, py, pz), wn = sample_w(g, px, py, pz);
    double du = un - sample_u(saved, px, py, pz);
    double dv = vn - sample_v(saved, px, py, pz);
    double dw = wn - sample_w(saved, px, py, pz);
    double fx = u[p] + du, fy = v[p] + dv, fz = w[p] + dw;
    u[p] = fx * a + un * (1.0 - a);
    v[p] = fy * a + vn * (1.0 - a);
    w[p] = fz * a + wn * (1.0 - a);
  }
}

typedef struct {
  const orc_grid *g, *saved;
  int64_t n;
  const double *x, *y, *z;
  double *u, *v, *w;
  const uint8_t* type;
  double aL, aG;
} g2p_job;

static void g2p_task(void* a, int t, int nt) {
  g2p_job* J = (g2p_job*)a;
  g2p_range(J->g, J->saved, J->n * t / nt, J->n * (t + 1) / nt, J->x, J->y, J->z, J->u, J->v, J->w, J->type,
            J->aL, J->aG);
}

void orc_g2p(const orc_grid* g, const orc_grid* saved, int64_t n, const double* x,
             const double* y, const double* z, double* u, double* v, double* w,
             const uint8_t* type, double aL, double aG) {
  g2p_job J = {g, saved, n, x, y, z, u, v, w, type, aL, aG};
  parallel_run(n >= 4096 ? g_threads_adv : 1, g2p_task, &J);
}

/* sampleVelocity, sparse_ops3d_tp.cpp:176-183 */
static void sample_vel(const orc_grid* g, double x, double y, double z, double out[3]) {
  double px = (x - g->ox) / g->dx, py = (y - g->oy) / g->dx, pz = (z - g->oz) / g->dx;
  out[0] = sample_u(g, px, py, pz);
  out[1] = sample_v(g, px, py, pz);
  out[2] = sample_w(g, px, py, pz);
}

/* spAdvect3D_tp, advectMidpoint, advectRK3: sparse_ops3d_tp.cpp:185-232 */
static void advect_range(const orc_grid* g, int64_t p0, int64_t p1, double* x, double* y, double* z,
                         const uint8_t* type, double dt, int order, orc_escape_stats* es,
                         int32_t* escaped_idx, int64_t* n_escaped);

typedef struct {
  const orc_grid* g;
  int64_t n;
  double *x, *y, *z;
  const uint8_t* type;
  double dt;
  int order;
  orc_escape_stats es[256];
  int64_t ne[256];
} advect_job;

static void advect_task(void* a, int t, int nt) {
  advect_job* J = (advect_job*)a;
  advect_range(J->g, J->n * t / nt, J->n * (t + 1) / nt, J->x, J->y, J->z, J->type, J->dt, J->order, &J->es[t], NULL,
               &J->ne[t]);
}

void orc_advect(const orc_grid* g, int64_t n, double* x, double* y, double* z,
                const uint8_t* type, double dt, int order, orc_escape_stats* es,
                int32_t* escaped_idx, int64_t* n_escaped) {
  if (g_threads_adv <= 1 || escaped_idx || n < 4096) {
    advect_range(g, 0, n, x, y, z, type, dt, order, es, escaped_idx, n_escaped);
    return;
  }
  advect_job* J = (advect_job*)calloc(1, sizeof(advect_job));
  J->g = g; J->n = n; J->x = x; J->y = y; J->z = z; J->type = type; J->dt = dt; J->order = order;
  parallel_run(g_threads_adv, advect_task, J);
  if (es) memset(es, 0, sizeof(*es));
  int64_t ne = 0;
  for (int t = 0; t < g_threads_adv; ++t) {
    ne += J->ne[t];
    if (!es) continue;
    es->clamped_liquid += J->es[t].clamped_liquid; es->clamped_gas += J->es[t].clamped_gas;
    es->x_lo += J->es[t].x_lo; es->x_hi += J->es[t].x_hi; es->y_lo += J->es[t].y_lo;
    es->y_hi += J->es[t].y_hi; es->z_lo += J->es[t].z_lo; es->z_hi += J->es[t].z_hi;
  }
  if (n_escaped) *n_escaped = ne;
  free(J);
}

static void advect_range(const orc_grid* g, int64_t p0, int64_t p1, double* x, double* y, double* z,
                         const uint8_t* type, double dt, int order, orc_escape_stats* es,
                         int32_t* escaped_idx, int64_t* n_escaped) {
  double lox = g->ox + 0.5 * g->dx, hix = g->ox + (g->nx - 0.5) * g->dx;
  double loy = g->oy + 0.5 * g->dx, hiy = g->oy + (g->ny - 0.5) * g->dx;
  double loz = g->oz + 0.5 * g->dx, hiz = g->oz + (g->nz - 0.5) * g->dx;
  if (es) memset(es, 0, sizeof(*es));
  int64_t ne = 0;
  for (int64_t p = p0; p < p1; ++p) {
    double P[3] = {x[p], y[p], z[p]}, k1[3], k2[3], k3[3], nx_[3];
    sample_vel(g, P[0], P[1], P[2], k1);
    if (order == 3) {
      double a[3], b[3];
      for (int c = 0; c < 3; ++c) a[c] = P[c] + k1[c] * (0.5 * dt);
      sample_vel(g, a[0], a[1], a[2], k2);
      for (int c = 0; c < 3; ++c) b[c] = (P[c] - k1[c] * dt) + k2[c] * (2.0 * dt);
      sample_vel(g, b[0], b[1], b[2], k3);
      for (int c = 0; c < 3; ++c) nx_[c] = P[c] + ((k1[c] + k2[c] * 4.0) + k3[c]) * (dt / 6.0);
    } else {
      double mid[3];
      for (int c = 0; c < 3; ++c) mid[c] = P[c] + k1[c] * (0.5 * dt);
      sample_vel(g, mid[0], mid[1], mid[2], k2);
      for (int c = 0; c < 3; ++c) nx_[c] = P[c] + k2[c] * dt;
    }
    int xLo = nx_[0] < lox, xHi = nx_[0] > hix, yLo = nx_[1] < loy, yHi = nx_[1] > hiy;
    int zLo = nx_[2] < loz, zHi = nx_[2] > hiz;
    int clamped = xLo || xHi || yLo || yHi || zLo || zHi;
    x[p] = fmax(lox, fmin(hix, nx_[0]));
    y[p] = fmax(loy, fmin(hiy, nx_[1]));
    z[p] = fmax(loz, fmin(hiz, nx_[2]));
    if (es && clamped) { /* ParticleEscapeStats3D::recordClamp, particle_escape3d.h:60-79 */
      if (type[p] == 0) es->clamped_liquid++;
      else if (type[p] == 1) es->clamped_gas++;
      es->x_lo += xLo; es->x_hi += xHi; es->y_lo += yLo; es->y_hi += yHi;
      es->z_lo += zLo; es->z_hi += zHi;
    }
    if (clamped) {
      if (escaped_idx) escaped_idx[ne] = (int32_t)p;
      ++ne;
    }
  }
  if (n_escaped) *n_escaped = ne;
}

/* computeAdaptiveParticleTimestep3D, src/driver/timestep3d.h:16-50 */
void orc_timestep(int64_t n, const double* u, const double* v, const double* w, double dx,
                  double requested_dt, int enabled, double cfl, double min_dt,
                  orc_timestep_stats* o) {
  o->requested_dt = requested_dt;
  o->effective_dt = requested_dt;
  o->cfl_limit_dt = requested_dt;
  o->max_particle_speed = 0.0;
  o->limited = 0;
  for (int64_t i = 0; i < n; ++i) {
    if (!isfinite(u[i]) || !isfinite(v[i]) || !isfinite(w[i])) continue;
    o->max_particle_speed = fmax(o->max_particle_speed, sqrt(u[i] * u[i] + v[i] * v[i] + w[i] * w[i]));
  }
  if (!enabled || requested_dt <= 0.0 || dx <= 0.0 || cfl <= 0.0 || o->max_particle_speed <= 0.0)
    return;
  o->cfl_limit_dt = cfl * dx / o->max_particle_speed;
  double lim = fmin(requested_dt, o->cfl_limit_dt);
  if (min_dt > 0.0) lim = fmax(min_dt, lim);
  o->effective_dt = fmin(requested_dt, lim);
  o->limited = o->effective_dt < requested_dt ? 1 : 0;
}

static double now_ms(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

/* SparseSim3DTP::step, src/driver/sparse_sim3d_tp.cpp:854-884 (default flags) */
void orc_step(orc_grid* g, int64_t n, double* x, double* y, double* z, double* u, double* v,
              double* w, const uint8_t* type, const double* vol, const orc_step_cfg* c,
              orc_step_stats* S) {
  orc_step_stats local;
  if (!S) S = &local;
  memset(S, 0, sizeof(*S));
  double t0 = now_ms();
  orc_timestep(n, u, v, w, g->dx, c->dt, c->adaptive_timestep, c->adaptive_cfl,
               c->adaptive_min_dt, &S->ts);
  double stepDt = S->ts.effective_dt;
  /* updateVolumeCorrectionStats, sparse_sim3d_tp.cpp:368-379 */
  double lv = 0.0;
  for (int64_t p = 0; p < n; ++p)
    if (type[p] == 0) lv += (vol ? vol[p] : 1.0) * c->Vp;
  S->liquid_volume = lv;
  S->c_div = 0.0;
  if (c->c_div_volume_correction && c->c_div_strength > 0.0 && c->liquid_volume_target > 0.0 &&
      stepDt > 0.0)
    S->c_div = c->c_div_strength * (c->liquid_volume_target - lv) / (stepDt * c->liquid_volume_target);
  double t1 = now_ms();
  orc_mark_cells(g, n, x, y, z);
  double t2 = now_ms();
  orc_p2g(g, n, x, y, z, u, v, w, type, vol, &c->phase, c->Vp);
  orc_interface_diagnostics(g, &c->phase, &S->iface);
  double t3 = now_ms();
  orc_grid* saved = orc_grid_clone(g);
  orc_apply_gravity(g, stepDt, c->gravity);
  if (c->surface_tension)
    orc_surface_tension(g, &c->phase, stepDt, c->surface_tension_strength, c->surface_tension_max_delta_speed,
                        c->surface_tension_smoothing_radius, &S->surface);
  else  /* SurfaceTensionStats3D() */
    orc_st_init(&S->surface, &c->phase, 0.0, 0.0, 0.0, 0.0, 0);
  orc_apply_walls(g);
  double t4 = now_ms();
  orc_project(g, &c->phase, stepDt, c->cg_iters, c->cg_tol, S->c_div, c->stop_rule, &S->solve);
  double t5 = now_ms();
  orc_g2p(g, saved, n, x, y, z, u, v, w, type, c->alpha_liquid, c->alpha_gas);
  double t6 = now_ms();
  orc_advect(g, n, x, y, z, type, stepDt, c->advection_order, &S->escape, 0, 0);
  double t7 = now_ms();
  orc_grid_destroy(saved);
  S->ms_mark = t2 - t1; S->ms_p2g = t3 - t2; S->ms_save_gravity_walls = t4 - t3;
  S->ms_project = t5 - t4; S->ms_g2p = t6 - t5; S->ms_advect = t7 - t6; S->ms_total = t7 - t0;
}

/* SparseSim3D::step, src/driver/sparse_sim3d.cpp:106-115: mark, trilinear P2G, saved
 * copy, gravity, walls, unit-density projection, FLIP blend with one alpha, RK2.
 * markCells / applyGravity / applyWallBoundary (:26-104) are the two-phase ones. */
void orc_step_single(orc_grid* g, int64_t n, double* x, double* y, double* z, double* u,
                     double* v, double* w, double dt, double alpha, double gravity, int cg_iters,
                     double cg_tol, int stop_rule, orc_step_stats* S) {
  orc_step_stats local;
  if (!S) S = &local;
  memset(S, 0, sizeof(*S));
  uint8_t* type = (uint8_t*)calloc(n > 0 ? (size_t)n : 1, 1);
  double t0 = now_ms();
  orc_mark_cells(g, n, x, y, z);
  double t1 = now_ms();
  orc_p2g_single(g, n, x, y, z, u, v, w);
  double t2 = now_ms();
  orc_grid* saved = orc_grid_clone(g);
  orc_apply_gravity(g, dt, gravity);
  orc_apply_walls(g);
  double t3 = now_ms();
  orc_project_single(g, dt, cg_iters, cg_tol, stop_rule, &S->solve);
  double t4 = now_ms();
  orc_g2p(g, saved, n, x, y, z, u, v, w, type, alpha, alpha);
  double t5 = now_ms();
  orc_advect(g, n, x, y, z, type, dt, 2, &S->escape, 0, 0);
  double t6 = now_ms();
  orc_grid_destroy(saved);
  free(type);
  S->ts.requested_dt = S->ts.effective_dt = dt;
  S->ms_mark = t1 - t0; S->ms_p2g = t2 - t1; S->ms_save_gravity_walls = t3 - t2;
  S->ms_project = t4 - t3; S->ms_g2p = t5 - t4; S->ms_advect = t6 - t5; S->ms_total = t6 - t0;
}

/* ------------------------------------------------------------------ scenes */

/* splitmix64 (Steele, Lea, Flood 2014) */
static inline uint64_t splitmix64(uint64_t* s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

static inline int scene_cell_filled(int kind, int i, int j, int k, int nx, int ny, int nz) {
  (void)k; (void)nz;
  double cx = (i + 0.5) / nx, cy = (j + 0.5) / ny;
  if (kind == ORC_SCENE_BUBBLE) return cy < 0.6;
  if (kind == ORC_SCENE_DAMBREAK) return cx < 0.4 && cy < 0.6;
  return 1;
}

int64_t orc_scene_count(int kind, int nx, int ny, int nz) {
  int64_t c = 0;
  for (int k = 1; k < nz - 1; ++k)
    for (int j = 1; j < ny - 1; ++j)
      for (int i = 1; i < nx - 1; ++i) c += scene_cell_filled(kind, i, j, k, nx, ny, nz);
  return 8 * c;
}

/* 8 particles per filled interior cell, jittered uniformly by splitmix64 seeded from
 * (seed, cell id); FP32 positions are canonical.  Types: RT heavy (0) above
 * y = 0.5 Ly + 0.05 Lx cos(2 pi x/Lx) cos(2 pi z/Lz); bubble gas (1) inside
 * r = 0.12 Lx around (0.5, 0.25, 0.5) L; dam-break all liquid. */
int64_t orc_scene_fill(int kind, int nx, int ny, int nz, double dx, uint64_t seed, float* X,
                       float* Y, float* Z, uint8_t* T) {
  const double Lx = nx * dx, Ly = ny * dx, Lz = nz * dx;
  const double two_pi = 6.283185307179586;
  int64_t p = 0;
  for (int k = 1; k < nz - 1; ++k)
    for (int j = 1; j < ny - 1; ++j)
      for (int i = 1; i < nx - 1; ++i) {
        if (!scene_cell_filled(kind, i, j, k, nx, ny, nz)) continue;
        uint64_t cid = (uint64_t)i + (uint64_t)nx * ((uint64_t)j + (uint64_t)ny * (uint64_t)k);
        uint64_t s = seed * 0x9E3779B97F4A7C15ULL ^ (cid + 1) * 0xD1B54A32D192ED03ULL;
        for (int q = 0; q < 8; ++q) {
          double ux = (double)(splitmix64(&s) >> 11) * 0x1.0p-53;
          double uy = (double)(splitmix64(&s) >> 11) * 0x1.0p-53;
          double uz = (double)(splitmix64(&s) >> 11) * 0x1.0p-53;
          float fx = (float)((i + ux) * dx), fy = (float)((j + uy) * dx), fz = (float)((k + uz) * dx);
          uint8_t t = 0;
          if (kind == ORC_SCENE_RT) {
            double yi = 0.5 * Ly + 0.05 * Lx * cos(two_pi * fx / Lx) * cos(two_pi * fz / Lz);
            t = ((double)fy > yi) ? 0 : 1;
          } else if (kind == ORC_SCENE_BUBBLE) {
            double ddx = fx - 0.5 * Lx, ddy = fy - 0.25 * Ly, ddz = fz - 0.5 * Lz, r = 0.12 * Lx;
            t = (ddx * ddx + ddy * ddy + ddz * ddz < r * r) ? 1 : 0;
          }
          if (X) { X[p] = fx; Y[p] = fy; Z[p] = fz; T[p] = t; }
          ++p;
        }
      }
  return p;
}

/* ------------------------------------------------ interface diagnostics, CSF */
/* src/driver/interface_diagnostics3d.cpp: cell phi as the mean of the six face phi
 * (:137-148), central-difference gradient with clamped indices (:158-175), normal
 * grad / |grad| (zero below 1e-12, :177-181), curvature as the divergence of the
 * normals (:188-208), smoothing over a clamped (2r+1)^3 box of interface cells
 * (:210-235), CSF velocity increment grad * (-strength * kappa * dt) limited to
 * max_delta_speed and added half to each face of the cell (:404-440).  The stencils
 * read face masses through an orc_iface_src, so the multires variants (mrFacePhi
 * and the rest, :237-340, in pf_oracle_mr.c) evaluate the same fine-grid formulas. */

#define IFACE_PHI_LO 0.05
#define IFACE_PHI_HI 0.95
#define IFACE_GRAD_EPS 1e-5
#define IFACE_NORMAL_EPS 1e-12

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

double orc_iface_cell_phi(const orc_iface_src* s, const orc_phase* pp, int i, int j, int k) {
  if (i < 0 || j < 0 || k < 0 || i >= s->nx || j >= s->ny || k >= s->nz) return 0.0;
  double sum = 0.0;
  sum += orc_phi(s->mass(s->g, 0, i, j, k), pp);
  sum += orc_phi(s->mass(s->g, 0, i + 1, j, k), pp);
  sum += orc_phi(s->mass(s->g, 1, i, j, k), pp);
  sum += orc_phi(s->mass(s->g, 1, i, j + 1, k), pp);
  sum += orc_phi(s->mass(s->g, 2, i, j, k), pp);
  sum += orc_phi(s->mass(s->g, 2, i, j, k + 1), pp);
  return sum / 6.0;
}

static double phi_clamped(const orc_iface_src* s, const orc_phase* pp, int i, int j, int k) {
  return orc_iface_cell_phi(s, pp, clampi(i, 0, s->nx - 1), clampi(j, 0, s->ny - 1), clampi(k, 0, s->nz - 1));
}

void orc_iface_gradient(const orc_iface_src* s, const orc_phase* pp, int i, int j, int k, double out[3]) {
  const int im = clampi(i - 1, 0, s->nx - 1), ip = clampi(i + 1, 0, s->nx - 1);
  const int jm = clampi(j - 1, 0, s->ny - 1), jp = clampi(j + 1, 0, s->ny - 1);
  const int km = clampi(k - 1, 0, s->nz - 1), kp = clampi(k + 1, 0, s->nz - 1);
  const double dx = s->dx;
  out[0] = (phi_clamped(s, pp, ip, j, k) - phi_clamped(s, pp, im, j, k)) / ((ip - im > 1 ? ip - im : 1) * dx);
  out[1] = (phi_clamped(s, pp, i, jp, k) - phi_clamped(s, pp, i, jm, k)) / ((jp - jm > 1 ? jp - jm : 1) * dx);
  out[2] = (phi_clamped(s, pp, i, j, kp) - phi_clamped(s, pp, i, j, km)) / ((kp - km > 1 ? kp - km : 1) * dx);
}

static double vlen(const double v[3]) { return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

static void iface_normal(const orc_iface_src* s, const orc_phase* pp, int i, int j, int k, double n[3]) {
  double gr[3];
  orc_iface_gradient(s, pp, i, j, k, gr);
  const double mag = vlen(gr);
  if (mag <= IFACE_NORMAL_EPS || !isfinite(mag)) { n[0] = n[1] = n[2] = 0.0; return; }
  const double inv = 1.0 / mag;
  n[0] = gr[0] * inv; n[1] = gr[1] * inv; n[2] = gr[2] * inv;
}

double orc_iface_curvature(const orc_iface_src* s, const orc_phase* pp, int i, int j, int k) {
  const int im = clampi(i - 1, 0, s->nx - 1), ip = clampi(i + 1, 0, s->nx - 1);
  const int jm = clampi(j - 1, 0, s->ny - 1), jp = clampi(j + 1, 0, s->ny - 1);
  const int km = clampi(k - 1, 0, s->nz - 1), kp = clampi(k + 1, 0, s->nz - 1);
  const double dx = s->dx;
  double nxm[3], nxp[3], nym[3], nyp[3], nzm[3], nzp[3];
  iface_normal(s, pp, im, j, k, nxm); iface_normal(s, pp, ip, j, k, nxp);
  iface_normal(s, pp, i, jm, k, nym); iface_normal(s, pp, i, jp, k, nyp);
  iface_normal(s, pp, i, j, km, nzm); iface_normal(s, pp, i, j, kp, nzp);
  return (nxp[0] - nxm[0]) / ((ip - im > 1 ? ip - im : 1) * dx) +
         (nyp[1] - nym[1]) / ((jp - jm > 1 ? jp - jm : 1) * dx) +
         (nzp[2] - nzm[2]) / ((kp - km > 1 ? kp - km : 1) * dx);
}

static int iface_phi(double phi) { return phi > IFACE_PHI_LO && phi < IFACE_PHI_HI; }

/* sparseSmoothedCurvature / mrSmoothedCurvature, :210-235 and :322-348 */
static double smoothed_curvature(const orc_iface_src* s, const orc_phase* pp, int i, int j, int k, double raw,
                                 int r) {
  if (r <= 0 || !isfinite(raw)) return raw;
  double sum = 0.0;
  int cnt = 0;
  for (int dz = -r; dz <= r; ++dz)
    for (int dy = -r; dy <= r; ++dy)
      for (int dx = -r; dx <= r; ++dx) {
        const int ni = clampi(i + dx, 0, s->nx - 1), nj = clampi(j + dy, 0, s->ny - 1);
        const int nk = clampi(k + dz, 0, s->nz - 1);
        if (!iface_phi(orc_iface_cell_phi(s, pp, ni, nj, nk))) continue;
        const double c = orc_iface_curvature(s, pp, ni, nj, nk);
        if (!isfinite(c)) continue;
        sum += c;
        ++cnt;
      }
  return cnt > 0 ? sum / (double)cnt : raw;
}

/* accumulate, :26-53 */
void orc_iface_accumulate(const orc_iface_src* s, const orc_phase* pp, int i, int j, int k, orc_iface_stats* S) {
  const double phi = orc_iface_cell_phi(s, pp, i, j, k);
  double gr[3];
  orc_iface_gradient(s, pp, i, j, k, gr);
  const double gm = vlen(gr), kap = orc_iface_curvature(s, pp, i, j, k);
  if (!isfinite(phi) || !isfinite(gm) || !isfinite(kap)) { S->finite = 0; return; }
  if (S->sample_cells == 0) { S->phi_min = phi; S->phi_max = phi; }
  else { S->phi_min = fmin(S->phi_min, phi); S->phi_max = fmax(S->phi_max, phi); }
  S->sample_cells++;
  S->phi_mean += phi;
  if (!iface_phi(phi)) return;
  S->interface_cells++;
  S->grad_mean += gm;
  S->grad_max = fmax(S->grad_max, gm);
  S->curvature_abs_mean += fabs(kap);
  S->curvature_abs_max = fmax(S->curvature_abs_max, fabs(kap));
}

/* finishStats, :56-66 */
void orc_iface_finish(orc_iface_stats* S) {
  if (S->sample_cells > 0) S->phi_mean /= (double)S->sample_cells;
  if (S->interface_cells > 0) {
    S->grad_mean /= (double)S->interface_cells;
    S->curvature_abs_mean /= (double)S->interface_cells;
  }
  S->surface_tension_candidate = S->finite && S->interface_cells > 0 && S->grad_max > IFACE_GRAD_EPS;
}

/* initSurfaceStats and capillaryDtLimit, :81-102 */
void orc_st_init(orc_st_stats* S, const orc_phase* pp, double dx, double dt, double strength, double maxd,
                 int radius) {
  memset(S, 0, sizeof(*S));
  S->finite = 1;
  S->enabled = strength > 0.0 && dt > 0.0;
  S->strength = strength;
  S->max_delta_speed_limit = maxd;
  S->curvature_smoothing_radius = clampi(radius, 0, 3);
  const double rmin = pp->rho_l < pp->rho_g ? pp->rho_l : pp->rho_g;
  const double pi = 3.14159265358979323846;
  S->capillary_dt_limit = (dx <= 0.0 || strength <= 0.0 || rmin <= 0.0 || !isfinite(rmin))
                              ? 0.0 : sqrt((rmin * dx * dx * dx) / (2.0 * pi * strength));
  S->capillary_stable = !S->enabled || S->capillary_dt_limit <= 0.0 || dt <= S->capillary_dt_limit + 1e-12;
}

/* one FLUID cell of applySparseSurfaceTension3D / applyMRSurfaceTension3D (:415-431,
 * :462-477): -1 when it is no interface cell or its gradient is below the threshold,
 * else accumulateSurfaceStats (:104-122) and 1 when the increment d applies (the
 * stats are still finite), 0 when it does not */
int orc_st_cell(const orc_iface_src* s, const orc_phase* pp, int i, int j, int k, double dt, double strength,
                double maxd, orc_st_stats* S, double d[3]) {
  const double phi = orc_iface_cell_phi(s, pp, i, j, k);
  if (!iface_phi(phi)) return -1;
  double gr[3];
  orc_iface_gradient(s, pp, i, j, k, gr);
  if (vlen(gr) <= IFACE_GRAD_EPS) return -1;
  const double raw = orc_iface_curvature(s, pp, i, j, k);
  const double sm = smoothed_curvature(s, pp, i, j, k, raw, S->curvature_smoothing_radius);
  /* limitedDelta, :72-79 */
  const double f = -strength * sm * dt;
  d[0] = gr[0] * f; d[1] = gr[1] * f; d[2] = gr[2] * f;
  const double mag0 = vlen(d);
  if (!isfinite(mag0)) { d[0] = d[1] = d[2] = 0.0; }
  else if (maxd > 0.0 && mag0 > maxd) {
    const double sc = maxd / mag0;
    d[0] *= sc; d[1] *= sc; d[2] *= sc;
  }
  const double mag = vlen(d);
  if (!isfinite(mag) || !isfinite(raw) || !isfinite(sm)) { S->finite = 0; return 0; }
  S->applied_cells++;
  S->mean_delta_speed += mag;
  S->max_delta_speed = fmax(S->max_delta_speed, mag);
  S->raw_curvature_abs_mean += fabs(raw);
  S->raw_curvature_abs_max = fmax(S->raw_curvature_abs_max, fabs(raw));
  S->smoothed_curvature_abs_mean += fabs(sm);
  S->smoothed_curvature_abs_max = fmax(S->smoothed_curvature_abs_max, fabs(sm));
  return S->finite ? 1 : 0;
}

/* finishSurfaceStats, :124-131 */
void orc_st_finish(orc_st_stats* S) {
  if (S->applied_cells > 0) {
    const double n = (double)S->applied_cells;
    S->mean_delta_speed /= n;
    S->raw_curvature_abs_mean /= n;
    S->smoothed_curvature_abs_mean /= n;
  }
}

static double sparse_mass(const void* gp, int axis, int i, int j, int k) {
  const orc_grid* g = (const orc_grid*)gp;
  return field_get(&g->f[ORC_MU + axis], i, j, k);
}

static orc_iface_src sparse_src(const orc_grid* g) {
  orc_iface_src s = {g->nx, g->ny, g->nz, g->dx, g, sparse_mass};
  return s;
}

void orc_interface_fields(const orc_grid* g, const orc_phase* pp, double* phi, double* grad, double* curvature) {
  const orc_iface_src s = sparse_src(g);
  for (int k = 0; k < g->nz; ++k)
    for (int j = 0; j < g->ny; ++j)
      for (int i = 0; i < g->nx; ++i) {
        const int64_t c = i + (int64_t)g->nx * (j + (int64_t)g->ny * k);
        double gr[3];
        orc_iface_gradient(&s, pp, i, j, k, gr);
        phi[c] = orc_iface_cell_phi(&s, pp, i, j, k);
        grad[c] = vlen(gr);
        curvature[c] = orc_iface_curvature(&s, pp, i, j, k);
      }
}

/* diagnoseSparseInterface3D, :362-379, over the FLUID cells in linear order */
void orc_interface_diagnostics(const orc_grid* g, const orc_phase* pp, orc_iface_stats* S) {
  const orc_iface_src s = sparse_src(g);
  memset(S, 0, sizeof(*S));
  S->finite = 1;
  for (int k = 0; k < g->nz; ++k)
    for (int j = 0; j < g->ny; ++j)
      for (int i = 0; i < g->nx; ++i)
        if ((int)(field_get(&g->f[ORC_MK], i, j, k) + 0.5) == ORC_FLUID) orc_iface_accumulate(&s, pp, i, j, k, S);
  orc_iface_finish(S);
}

/* `float += (float)x` (faithful) or a double add on face field f */
static void face_add(orc_grid* g, int f, int i, int j, int k, double add) {
  double* r = field_ref(&g->f[f], i, j, k);
  *r = acc(g, *r, add);
}

/* applySparseSurfaceTension3D, :404-440 */
void orc_surface_tension(orc_grid* g, const orc_phase* pp, double dt, double strength, double maxd, int radius,
                         orc_st_stats* S) {
  const orc_iface_src s = sparse_src(g);
  orc_st_init(S, pp, g->dx, dt, strength, maxd, radius);
  if (!S->enabled) return;
  for (int k = 0; k < g->nz; ++k)
    for (int j = 0; j < g->ny; ++j)
      for (int i = 0; i < g->nx; ++i) {
        if ((int)(field_get(&g->f[ORC_MK], i, j, k) + 0.5) != ORC_FLUID) continue;
        double d[3];
        if (orc_st_cell(&s, pp, i, j, k, dt, strength, maxd, S, d) != 1) continue;
        if (field_get(&g->f[ORC_MU], i, j, k) > 0.0) face_add(g, ORC_U, i, j, k, 0.5 * d[0]);
        if (field_get(&g->f[ORC_MU], i + 1, j, k) > 0.0) face_add(g, ORC_U, i + 1, j, k, 0.5 * d[0]);
        if (field_get(&g->f[ORC_MV], i, j, k) > 0.0) face_add(g, ORC_V, i, j, k, 0.5 * d[1]);
        if (field_get(&g->f[ORC_MV], i, j + 1, k) > 0.0) face_add(g, ORC_V, i, j + 1, k, 0.5 * d[1]);
        if (field_get(&g->f[ORC_MW], i, j, k) > 0.0) face_add(g, ORC_W, i, j, k, 0.5 * d[2]);
        if (field_get(&g->f[ORC_MW], i, j, k + 1) > 0.0) face_add(g, ORC_W, i, j, k + 1, 0.5 * d[2]);
      }
  orc_st_finish(S);
}

/* ------------------------------------------------ render-cache cells */
/* sparsePhaseCells (:557-588) needs sparseCellPhi (:279-300) over the marker's active
 * blocks in allocation order (the pressure field's when the marker has none) */
int orc_render_cache_sparse(const orc_grid* g, const orc_phase* pp, int64_t n, const double* x, const double* y,
                            const double* z, const double* u, const double* v, const double* w,
                            const uint8_t* type, const double* vol, const orc_render_meta* m, const char* path) {
  int64_t nb = orc_active_blocks(g, ORC_MK, NULL);
  int field = ORC_MK;
  if (nb == 0) { field = ORC_P; nb = orc_active_blocks(g, ORC_P, NULL); }
  int32_t* order = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nb > 0 ? nb : 1));
  orc_active_blocks(g, field, order);
  int32_t* q = (int32_t*)malloc(sizeof(int32_t) * 5 * 64 * (size_t)(nb > 0 ? nb : 1));
  double* pvv = (double*)malloc(sizeof(double) * 2 * 64 * (size_t)(nb > 0 ? nb : 1));
  const int nbx = (g->nx + 3) / 4, nby = (g->ny + 3) / 4;
  const double cv = g->dx * g->dx * g->dx;
  double pv = 0.0;
  int64_t nc = 0;
  for (int64_t a = 0; a < nb; ++a) {
    const int b = order[a], bx = b % nbx, by = (b / nbx) % nby, bz = b / (nbx * nby);
    for (int lz = 0; lz < 4; ++lz)
      for (int ly = 0; ly < 4; ++ly)
        for (int lx = 0; lx < 4; ++lx) {
          const int i = bx * 4 + lx, j = by * 4 + ly, k = bz * 4 + lz;
          if (i >= g->nx || j >= g->ny || k >= g->nz) continue;
          const int mk = (int)(orc_get(g, ORC_MK, i, j, k) + 0.5);
          double s = 0.0;
          s += orc_phi(orc_get(g, ORC_MU, i, j, k), pp);
          s += orc_phi(orc_get(g, ORC_MU, i + 1, j, k), pp);
          s += orc_phi(orc_get(g, ORC_MV, i, j, k), pp);
          s += orc_phi(orc_get(g, ORC_MV, i, j + 1, k), pp);
          s += orc_phi(orc_get(g, ORC_MW, i, j, k), pp);
          s += orc_phi(orc_get(g, ORC_MW, i, j, k + 1), pp);
          const double phi = s / 6.0;
          if (mk == 0 && phi <= 1e-8) continue;
          const double wv = phi * cv;
          pv += wv;
          int32_t* r = q + 5 * nc;
          r[0] = i; r[1] = j; r[2] = k; r[3] = 0; r[4] = mk;
          pvv[2 * nc] = phi;
          pvv[2 * nc + 1] = wv;
          ++nc;
        }
  }
  const int r = orc_render_frame_cells(path, "sparse3d_tp", g->nx, g->ny, g->nz, g->dx, m, q, pvv, nc, pv, n, x, y, z,
                                       u, v, w, type, vol);
  free(pvv);
  free(q);
  free(order);
  return r;
}
