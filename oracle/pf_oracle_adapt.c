/* pf_oracle_adapt.c -- oracle (test infrastructure, not the product) for particle
 * adaptivity: narrow-band air, typed particle coarsening and liquid refill
 * (src/driver/particle_adaptivity3d.h:75-376), applied in SparseSim3DTP /
 * MRSim3DTP::applyParticleAdaptivity order (sparse_sim3d_tp.cpp:847-852,
 * multires_sim3d_tp.cpp:570-575).
 *
 * Particles are structure-of-arrays in FP64 (the reference's Vec3 and volume);
 * removal keeps the survivors' order (Particles3DTP::eraseIf, particles3d_tp.h:12-33)
 * and refill appends (Particles3DTP::add).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "pf_oracle.h"

/* mix32, particle_adaptivity3d.h:75-82 */
uint32_t orc_mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

/* ParticleCellDomain3D::particleCell, :35-47 */
static int particle_cell(const orc_pa_domain* d, double x, double y, double z, int* i, int* j, int* k) {
  if (!isfinite(x) || !isfinite(y) || !isfinite(z)) return 0;
  *i = (int)((x - d->ox) / d->dx);
  *j = (int)((y - d->oy) / d->dx);
  *k = (int)((z - d->oz) / d->dx);
  return *i >= 0 && *i < d->nx && *j >= 0 && *j < d->ny && *k >= 0 && *k < d->nz;
}

static int64_t cell_index(const orc_pa_domain* d, int i, int j, int k) {
  return i + (int64_t)d->nx * (j + (int64_t)d->ny * k);
}

/* particleScore, :84-103 */
uint32_t orc_particle_score(const orc_pa_domain* d, double x, double y, double z, int i, int j, int k,
                            uint32_t seed) {
  const int qx = (int)((x - (d->ox + i * d->dx)) / d->dx * 1024.0);
  const int qy = (int)((y - (d->oy + j * d->dx)) / d->dx * 1024.0);
  const int qz = (int)((z - (d->oz + k * d->dx)) / d->dx * 1024.0);
  uint32_t h = seed;
  h ^= orc_mix32((uint32_t)i + 0x9e3779b9u);
  h ^= orc_mix32((uint32_t)j + 0x85ebca6bu);
  h ^= orc_mix32((uint32_t)k + 0xc2b2ae35u);
  h ^= orc_mix32((uint32_t)(qx & 2047) + 0x27d4eb2fu);
  h ^= orc_mix32((uint32_t)(qy & 2047) + 0x165667b1u);
  h ^= orc_mix32((uint32_t)(qz & 2047) + 0xd3a2646cu);
  return orc_mix32(h);
}

/* slotScore, :105-110 */
uint32_t orc_slot_score(uint64_t cell, int slot, uint32_t seed) {
  uint32_t h = seed;
  h ^= orc_mix32((uint32_t)cell + 0x68bc21ebu);
  h ^= orc_mix32((uint32_t)slot + 0x02e5be93u);
  return orc_mix32(h);
}

typedef struct {
  double *x, *y, *z, *u, *v, *w, *vol;
  uint8_t* t;
} parts;

/* eraseIf with a precomputed flag per particle; returns the new count */
static int64_t erase_flagged(parts* P, int64_t n, const uint8_t* erase) {
  int64_t wr = 0;
  for (int64_t r = 0; r < n; ++r) {
    if (erase[r]) continue;
    if (wr != r) {
      P->x[wr] = P->x[r]; P->y[wr] = P->y[r]; P->z[wr] = P->z[r];
      P->u[wr] = P->u[r]; P->v[wr] = P->v[r]; P->w[wr] = P->w[r];
      P->t[wr] = P->t[r]; P->vol[wr] = P->vol[r];
    }
    ++wr;
  }
  return wr;
}

/* true when a flagged cell lies in the clamped box of the given radius around (i,j,k) */
static int box_any(const orc_pa_domain* d, const uint8_t* flag, int i, int j, int k, int r) {
  for (int dk = -r; dk <= r; ++dk) {
    const int kk = k + dk;
    if (kk < 0 || kk >= d->nz) continue;
    for (int dj = -r; dj <= r; ++dj) {
      const int jj = j + dj;
      if (jj < 0 || jj >= d->ny) continue;
      for (int di = -r; di <= r; ++di) {
        const int ii = i + di;
        if (ii < 0 || ii >= d->nx) continue;
        if (flag[cell_index(d, ii, jj, kk)]) return 1;
      }
    }
  }
  return 0;
}

/* applyNarrowBandAir, :112-167 */
static int64_t narrow_band_air(const orc_pa_domain* d, int radius, parts* P, int64_t n, orc_adapt_stats* S) {
  const int64_t nc = (int64_t)d->nx * d->ny * d->nz;
  uint8_t* liquid = (uint8_t*)calloc((size_t)nc, 1);
  for (int64_t p = 0; p < n; ++p) {
    if (P->t[p] != 0) continue;
    int i, j, k;
    if (!particle_cell(d, P->x[p], P->y[p], P->z[p], &i, &j, &k)) continue;
    uint8_t* o = &liquid[cell_index(d, i, j, k)];
    if (!*o) { *o = 1; S->narrow_band_air_liquid_cells++; }
  }
  if (S->narrow_band_air_liquid_cells == 0) { free(liquid); return n; }
  radius = radius > 0 ? radius : 0;
  uint8_t* erase = (uint8_t*)calloc((size_t)(n > 0 ? n : 1), 1);
  for (int64_t p = 0; p < n; ++p) {
    if (P->t[p] != 1) continue;
    S->narrow_band_air_gas_before++;
    int i, j, k;
    const int keep = particle_cell(d, P->x[p], P->y[p], P->z[p], &i, &j, &k) && box_any(d, liquid, i, j, k, radius);
    if (keep) S->narrow_band_air_gas_after++;
    erase[p] = !keep;
  }
  const int64_t m = erase_flagged(P, n, erase);
  S->narrow_band_air_removed = (int32_t)(n - m);
  free(erase);
  free(liquid);
  return m;
}

typedef struct {
  int64_t index, cell;
  uint32_t score;
  double volume;
} entry;

static int entry_cmp(const void* a, const void* b) {
  const entry *x = (const entry*)a, *y = (const entry*)b;
  if (x->cell != y->cell) return x->cell < y->cell ? -1 : 1;
  if (x->score != y->score) return x->score < y->score ? -1 : 1;
  return x->index < y->index ? -1 : (x->index > y->index);
}

/* applyTypedParticleCoarsening, :169-241; out = {removed, cells, overfull, before, after} */
static int64_t typed_coarsening(const orc_pa_domain* d, uint8_t type, int target, uint32_t seed, parts* P,
                                int64_t n, int32_t out[5]) {
  target = target > 1 ? target : 1;
  entry* e = (entry*)malloc(sizeof(entry) * (size_t)(n > 0 ? n : 1));
  uint8_t* keep = (uint8_t*)calloc((size_t)(n > 0 ? n : 1), 1);
  int64_t ne = 0;
  for (int64_t p = 0; p < n; ++p) {
    if (P->t[p] != type) continue;
    out[3]++;
    int i, j, k;
    if (!particle_cell(d, P->x[p], P->y[p], P->z[p], &i, &j, &k)) continue;
    e[ne].index = p;
    e[ne].cell = cell_index(d, i, j, k);
    e[ne].score = orc_particle_score(d, P->x[p], P->y[p], P->z[p], i, j, k, seed);
    e[ne].volume = P->vol[p];
    ++ne;
  }
  qsort(e, (size_t)ne, sizeof(entry), entry_cmp);
  int64_t g0 = 0;
  while (g0 < ne) {
    int64_t g1 = g0 + 1;
    while (g1 < ne && e[g1].cell == e[g0].cell) ++g1;
    out[1]++;
    const int64_t cnt = g1 - g0;
    if (cnt > target) out[2]++;
    const int64_t kc = cnt < target ? cnt : target;
    double vs = 0.0;
    for (int64_t t = g0; t < g1; ++t) vs += e[t].volume;
    const double kv = kc > 0 ? vs / (double)kc : 0.0;
    for (int64_t t = 0; t < kc; ++t) {
      const int64_t p = e[g0 + t].index;
      keep[p] = 1;
      P->vol[p] = kv;
      out[4]++;
    }
    g0 = g1;
  }
  uint8_t* erase = (uint8_t*)calloc((size_t)(n > 0 ? n : 1), 1);
  for (int64_t p = 0; p < n; ++p) erase[p] = P->t[p] == type && !keep[p];
  const int64_t m = erase_flagged(P, n, erase);
  out[0] = (int32_t)(n - m);
  free(erase);
  free(keep);
  free(e);
  return m;
}

/* applyLiquidParticleRefill, :243-374; returns the new count or -1 past cap */
static int64_t liquid_refill(const orc_pa_domain* d, const orc_adapt_cfg* c, parts* P, int64_t n, int64_t cap,
                             orc_adapt_stats* S) {
  const int target = c->liquid_refill_particles_per_cell_target > 1 ? c->liquid_refill_particles_per_cell_target : 1;
  const int64_t nc = (int64_t)d->nx * d->ny * d->nz;
  int32_t* count = (int32_t*)calloc((size_t)nc, sizeof(int32_t));
  double* vsum = (double*)calloc((size_t)nc, sizeof(double));
  double* velsum = (double*)calloc((size_t)nc * 3, sizeof(double));
  uint8_t* gas = (uint8_t*)calloc((size_t)nc, 1);
  int64_t* cell_of = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
  for (int64_t p = 0; p < n; ++p) {
    cell_of[p] = -1;
    int i, j, k;
    if (!particle_cell(d, P->x[p], P->y[p], P->z[p], &i, &j, &k)) continue;
    const int64_t q = cell_index(d, i, j, k);
    if (P->t[p] == 0) {
      cell_of[p] = q;
      count[q]++;
      vsum[q] += P->vol[p];
      velsum[3 * q] += P->u[p]; velsum[3 * q + 1] += P->v[p]; velsum[3 * q + 2] += P->w[p];
    } else if (P->t[p] == 1) {
      gas[q] = 1;
    }
  }
  /* the liquid particles of each cell (CellInfo::indices) */
  int64_t* first = (int64_t*)calloc((size_t)nc + 1, sizeof(int64_t));
  int64_t* members = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
  for (int64_t q = 0; q < nc; ++q) first[q + 1] = first[q] + count[q];
  int64_t* fill = (int64_t*)calloc((size_t)nc, sizeof(int64_t));
  for (int64_t p = 0; p < n; ++p)
    if (cell_of[p] >= 0) members[first[cell_of[p]] + fill[cell_of[p]]++] = p;
  free(fill);
  const int ir = c->liquid_particle_refill_interface_radius > 0 ? c->liquid_particle_refill_interface_radius : 0;
  int64_t budget = c->liquid_particle_refill_max_added_per_step > 0 ? c->liquid_particle_refill_max_added_per_step
                                                                    : (int64_t)2147483647;
  int64_t m = n;
  for (int64_t q = 0; q < nc; ++q) {
    const int cnt = count[q];
    if (cnt <= 0) continue;
    S->refill_cells++;
    S->refill_before += cnt;
    S->refill_after += cnt;
    const int i = (int)(q % d->nx), j = (int)((q / d->nx) % d->ny), k = (int)(q / ((int64_t)d->nx * d->ny));
    const int near = box_any(d, gas, i, j, k, ir);
    if (near) S->refill_interface_cells++;
    if (c->liquid_particle_refill_interface_only && !near) continue;
    if (cnt >= target) continue;
    S->refill_underfull_cells++;
    if (budget <= 0) { S->refill_budget_limited = 1; continue; }
    const double inv = 1.0 / (double)cnt;
    const double au = velsum[3 * q] * inv, av = velsum[3 * q + 1] * inv, aw = velsum[3 * q + 2] * inv;
    int slots[8] = {0, 1, 2, 3, 4, 5, 6, 7};
    for (int a = 1; a < 8; ++a)   /* insertion sort by (slotScore, slot) */
      for (int b = a; b > 0; --b) {
        const uint32_t sa = orc_slot_score((uint64_t)q, slots[b - 1], c->liquid_particle_refill_seed);
        const uint32_t sb = orc_slot_score((uint64_t)q, slots[b], c->liquid_particle_refill_seed);
        if (sa < sb || (sa == sb && slots[b - 1] < slots[b])) break;
        const int t = slots[b]; slots[b] = slots[b - 1]; slots[b - 1] = t;
      }
    const int wanted = target - cnt;
    const int add = (int64_t)wanted < budget ? wanted : (int)budget;
    if (add < wanted) S->refill_budget_limited = 1;
    const double split = vsum[q] / (double)(cnt + add);
    for (int64_t t = first[q]; t < first[q + 1]; ++t) P->vol[members[t]] = split;
    for (int a = 0; a < add; ++a) {
      if (m >= cap) { m = -1; goto done; }
      const int s = slots[a & 7];
      const double sx = 0.25 + 0.5 * (double)(s & 1), sy = 0.25 + 0.5 * (double)((s >> 1) & 1);
      const double sz = 0.25 + 0.5 * (double)((s >> 2) & 1);
      P->x[m] = d->ox + ((double)i + sx) * d->dx;
      P->y[m] = d->oy + ((double)j + sy) * d->dx;
      P->z[m] = d->oz + ((double)k + sz) * d->dx;
      P->u[m] = au; P->v[m] = av; P->w[m] = aw;
      P->t[m] = 0;
      P->vol[m] = split;
      ++m;
      S->refill_added++;
      S->refill_after++;
    }
    budget -= add;
  }
done:
  free(members); free(first);
  free(cell_of); free(gas); free(velsum); free(vsum); free(count);
  return m;
}

int64_t orc_particle_adaptivity(const orc_pa_domain* d, const orc_adapt_cfg* c, int64_t n, int64_t cap, double* x,
                                double* y, double* z, double* u, double* v, double* w, uint8_t* type, double* vol,
                                orc_adapt_stats* S) {
  memset(S, 0, sizeof(*S));
  parts P = {x, y, z, u, v, w, vol, type};
  if (c->narrow_band_air) n = narrow_band_air(d, c->narrow_band_air_radius, &P, n, S);
  if (c->gas_particle_coarsening) {
    int32_t o[5] = {0, 0, 0, 0, 0};
    n = typed_coarsening(d, 1, c->gas_particles_per_cell_target, c->gas_particle_coarsening_seed, &P, n, o);
    S->gas_removed = o[0]; S->gas_cells = o[1]; S->gas_overfull_cells = o[2];
    S->gas_before = o[3]; S->gas_after = o[4];
  }
  if (c->liquid_particle_coarsening) {
    int32_t o[5] = {0, 0, 0, 0, 0};
    n = typed_coarsening(d, 0, c->liquid_particles_per_cell_target, c->liquid_particle_coarsening_seed, &P, n, o);
    S->liquid_removed = o[0]; S->liquid_cells = o[1]; S->liquid_overfull_cells = o[2];
    S->liquid_before = o[3]; S->liquid_after = o[4];
  }
  if (c->liquid_particle_refill) n = liquid_refill(d, c, &P, n, cap, S);
  return n;
}
