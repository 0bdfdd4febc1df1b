// pf_adapt.hip -- particle adaptivity on the binned particles
// (src/driver/particle_adaptivity3d.h:112-374, applied in applyParticleAdaptivity
// order, sparse_sim3d_tp.cpp:847-852 / multires_sim3d_tp.cpp:570-575).
//
// The particles arrive sorted by cell (block-major cell index), so every per-cell
// decision is made by one thread over its cell's contiguous range:
//  * narrow-band air: cell flags (liquid / gas present), a separable box dilation of
//    the liquid flag, then gas particles outside the band die;
//  * typed coarsening: each particle's (mix32 score, id) key once, then per cell the
//    rank of every entry by counting smaller keys, the volume sum in rank order, and
//    the `target` lowest ranks kept with the mean volume;
//  * liquid refill: cell flags again, the gas flag dilated by the interface radius,
//    the wanted count per cell in LINEAR cell order (the reference's loop order),
//    one scan of it fixes every cell's share of the budget and the slot of its new
//    particles; sums over a cell's particles run in id order, the reference's array
//    order.
// Removal keeps the survivors' order: particles are written out in id order with
// ids renumbered by a scan of the survivors (Particles3DTP::eraseIf), and refilled
// particles follow them (Particles3DTP::add).  Integer decisions are exact; FP64
// sums of the FP32 volumes and velocities follow the reference's order and round to
// FP32 when stored.
#include "pf_internal.h"

namespace pf {

enum {
  PA_NB_REMOVED = 0, PA_NB_LIQUID_CELLS, PA_NB_GAS_BEFORE, PA_NB_GAS_AFTER,
  PA_GAS_REMOVED, PA_GAS_CELLS, PA_GAS_OVERFULL, PA_GAS_BEFORE, PA_GAS_AFTER,
  PA_LIQ_REMOVED, PA_LIQ_CELLS, PA_LIQ_OVERFULL, PA_LIQ_BEFORE, PA_LIQ_AFTER,
  PA_RF_ADDED, PA_RF_CELLS, PA_RF_INTERFACE, PA_RF_UNDERFULL, PA_RF_BUDGET, PA_RF_BEFORE, PA_RF_AFTER,
  PA_NCOUNTERS
};

// mix32, particle_adaptivity3d.h:75-82
__host__ __device__ inline uint32_t pa_mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

// particleScore, :84-103 (positions are the FP32 particle values, widened exactly)
__device__ inline uint32_t pa_particle_score(const Geom& g, double x, double y, double z, int i, int j, int k,
                                             uint32_t seed) {
#pragma clang fp contract(off)
  const int qx = (int)((x - (g.ox + i * g.dx)) / g.dx * 1024.0);
  const int qy = (int)((y - (g.oy + j * g.dx)) / g.dx * 1024.0);
  const int qz = (int)((z - (g.oz + k * g.dx)) / g.dx * 1024.0);
  uint32_t h = seed;
  h ^= pa_mix32((uint32_t)i + 0x9e3779b9u);
  h ^= pa_mix32((uint32_t)j + 0x85ebca6bu);
  h ^= pa_mix32((uint32_t)k + 0xc2b2ae35u);
  h ^= pa_mix32((uint32_t)(qx & 2047) + 0x27d4eb2fu);
  h ^= pa_mix32((uint32_t)(qy & 2047) + 0x165667b1u);
  h ^= pa_mix32((uint32_t)(qz & 2047) + 0xd3a2646cu);
  return pa_mix32(h);
}

// slotScore, :105-110
__device__ inline uint32_t pa_slot_score(uint32_t cell, int slot, uint32_t seed) {
  uint32_t h = seed;
  h ^= pa_mix32(cell + 0x68bc21ebu);
  h ^= pa_mix32((uint32_t)slot + 0x02e5be93u);
  return pa_mix32(h);
}

// ParticleCellDomain3D::particleCell, :35-47: false for a non-finite position or one
// outside the grid
__device__ inline bool pa_cell(const Geom& g, float fx, float fy, float fz, int& i, int& j, int& k) {
  const double x = fx, y = fy, z = fz;
  if (!isfinite(x) || !isfinite(y) || !isfinite(z)) return false;
  i = (int)((x - g.ox) / g.dx);
  j = (int)((y - g.oy) / g.dx);
  k = (int)((z - g.oz) / g.dx);
  return i >= 0 && i < g.nx && j >= 0 && j < g.ny && k >= 0 && k < g.nz;
}

struct AdaptArgs {
  Geom g;
  Dims dc;
  PBuf P;                  // binned particles (sorted by cell)
  long long np;
  const uint32_t* start;   // cell_start, block-major cells
  uint8_t* alive;          // per particle
  uint64_t* key;           // per particle: (score << 32) | id
  uint8_t* cflag;          // per cell: bit 0 liquid present, bit 1 gas present
  uint32_t* cnt;           // PA_NCOUNTERS
};

// block-wide sum of a per-thread count, one atomic per workgroup
__device__ inline void pa_count(uint32_t* cnt, int slot, uint32_t v) {
  __shared__ uint32_t s[4];
  uint32_t x = v;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t t = s[0] + s[1] + s[2] + s[3];
    if (t) atomicAdd(&cnt[slot], t);
  }
  __syncthreads();
}

__device__ inline void pa_cell_coords(const Dims& dc, long long q, int& i, int& j, int& k) {
  int bx, by, bz;
  block_coords(dc, q >> 6, bx, by, bz);
  const int l = (int)(q & 63);
  i = bx * 4 + (l & 3);
  j = by * 4 + ((l >> 2) & 3);
  k = bz * 4 + (l >> 4);
}

// per cell: which live particle types it holds; counts cells holding liquid into slot
// (negative: no count)
__global__ void __launch_bounds__(256) k_pa_flags(AdaptArgs a, int liquid_slot) {
  const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long ncell = a.dc.nblocks * kBlockVol;
  uint8_t f = 0;
  if (q < ncell) {
    const uint32_t s = a.start[q], e = a.start[q + 1];
    for (uint32_t p = s; p < e; ++p) {
      if (!a.alive[p]) continue;
      int i, j, k;
      if (!pa_cell(a.g, a.P.x[p], a.P.y[p], a.P.z[p], i, j, k)) continue;
      f |= a.P.type[p] == 0 ? 1 : (a.P.type[p] == 1 ? 2 : 0);
    }
    a.cflag[q] = f;
  }
  if (liquid_slot >= 0) pa_count(a.cnt, liquid_slot, (f & 1) ? 1u : 0u);
}

// one axis of a box dilation: out = any(in within r cells along axis, in range);
// mask selects the bits of in that count
__global__ void __launch_bounds__(256) k_pa_dilate(Dims dc, Geom g, int axis, int r, uint8_t mask,
                                                   const uint8_t* __restrict__ in, uint8_t* __restrict__ out) {
  const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
  if (q >= dc.nblocks * kBlockVol) return;
  int i, j, k;
  pa_cell_coords(dc, q, i, j, k);
  if (i >= g.nx || j >= g.ny || k >= g.nz) {
    out[q] = 0;
    return;
  }
  const int c = axis == 0 ? i : (axis == 1 ? j : k);
  const int n = axis == 0 ? g.nx : (axis == 1 ? g.ny : g.nz);
  uint8_t any = 0;
  for (int o = max(c - r, 0); o <= min(c + r, n - 1) && !any; ++o) {
    const long long nq = axis == 0 ? bm_index(dc, o, j, k) : (axis == 1 ? bm_index(dc, i, o, k) : bm_index(dc, i, j, o));
    any = (in[nq] & mask) ? 1 : 0;
  }
  out[q] = any;
}

// applyNarrowBandAir's erase predicate (:150-157): a gas particle lives iff its cell
// is in the liquid band
__global__ void __launch_bounds__(256) k_pa_narrow_band(AdaptArgs a, const uint8_t* __restrict__ band) {
  const long long p = (long long)blockIdx.x * 256 + threadIdx.x;
  uint32_t before = 0, after = 0;
  if (p < a.np && a.alive[p] && a.P.type[p] == 1) {
    before = 1;
    int i, j, k;
    const bool keep = pa_cell(a.g, a.P.x[p], a.P.y[p], a.P.z[p], i, j, k) && band[bm_index(a.dc, i, j, k)];
    if (keep) after = 1;
    else a.alive[p] = 0;
  }
  pa_count(a.cnt, PA_NB_GAS_BEFORE, before);
  pa_count(a.cnt, PA_NB_GAS_AFTER, after);
}

// the coarsening key of every live particle of one type: (score << 32) | id;
// counters at slot0 (removed) and slot0 + 3 (before)
__global__ void __launch_bounds__(256) k_pa_keys(AdaptArgs a, int type, uint32_t seed, int slot0) {
  const long long p = (long long)blockIdx.x * 256 + threadIdx.x;
  uint32_t before = 0, lost = 0;
  if (p < a.np) {
    uint64_t key = ~0ull;
    if (a.alive[p] && a.P.type[p] == type) {
      before = 1;
      int i, j, k;
      const float x = a.P.x[p], y = a.P.y[p], z = a.P.z[p];
      if (pa_cell(a.g, x, y, z, i, j, k)) {
        key = ((uint64_t)pa_particle_score(a.g, x, y, z, i, j, k, seed) << 32) | a.P.id[p];
      } else {
        a.alive[p] = 0;   // no cell: never an entry, so erased (:199-202, 236-239)
        lost = 1;
      }
    }
    a.key[p] = key;
  }
  pa_count(a.cnt, slot0 + 3, before);
  pa_count(a.cnt, slot0, lost);
}

// applyTypedParticleCoarsening's per-cell groups (:210-233), one thread per cell
__global__ void __launch_bounds__(256) k_pa_coarsen(AdaptArgs a, int type, int target, int slot0) {
  const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
  uint32_t cells = 0, overfull = 0, after = 0, removed = 0;
  if (q < a.dc.nblocks * kBlockVol) {
    const uint32_t s = a.start[q], e = a.start[q + 1];
    uint32_t m = 0;
    for (uint32_t p = s; p < e; ++p) m += a.key[p] != ~0ull && a.P.type[p] == type;
    if (m > 0) {
      cells = 1;
      overfull = m > (uint32_t)target;
      const uint32_t keep = min(m, (uint32_t)target);
      after = keep;
      removed = m - keep;
      // volumes summed in (score, id) order: walk the keys upwards
      double vs = 0.0;
      uint64_t last = 0;
      bool first = true;
      for (uint32_t r = 0; r < m; ++r) {
        uint64_t best = ~0ull;
        uint32_t bp = s;
        for (uint32_t p = s; p < e; ++p) {
          const uint64_t kk = a.key[p];
          if (kk == ~0ull || a.P.type[p] != type) continue;
          if ((first || kk > last) && kk < best) {
            best = kk;
            bp = p;
          }
        }
        vs += a.P.vol ? (double)a.P.vol[bp] : 1.0;
        last = best;
        first = false;
      }
      const double kv = vs / (double)keep;
      for (uint32_t p = s; p < e; ++p) {
        const uint64_t kk = a.key[p];
        if (kk == ~0ull || a.P.type[p] != type) continue;
        uint32_t rank = 0;
        for (uint32_t o = s; o < e; ++o) rank += a.key[o] != ~0ull && a.P.type[o] == type && a.key[o] < kk;
        if (rank < keep) a.P.vol[p] = (float)kv;
        else a.alive[p] = 0;
      }
    }
  }
  pa_count(a.cnt, slot0 + 1, cells);
  pa_count(a.cnt, slot0 + 2, overfull);
  pa_count(a.cnt, slot0 + 4, after);
  pa_count(a.cnt, slot0, removed);
}

struct RefillArgs {
  int target, interface_only;
  uint32_t seed;
  long long budget;           // remaining budget at the start (INT_MAX when unlimited)
  const uint8_t* near_gas;    // per cell (block-major)
  uint32_t* wanted;           // per cell in linear order
  const uint32_t* wscan;      // exclusive scan of wanted (linear order)
  long long n_kept;           // particles that survive (the first new id)
  PBuf out;                   // destination of the added particles
};

// live liquid particles of cell q: count
__device__ inline uint32_t pa_liquid_count(const AdaptArgs& a, uint32_t s, uint32_t e) {
  uint32_t m = 0;
  for (uint32_t p = s; p < e; ++p) m += a.alive[p] && a.P.type[p] == 0;
  return m;
}

// per cell: the refill counters and the wanted count (:305-327)
__global__ void __launch_bounds__(256) k_pa_refill_count(AdaptArgs a, RefillArgs r) {
  const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
  uint32_t cells = 0, before = 0, iface = 0, under = 0;
  if (q < a.dc.nblocks * kBlockVol) {
    int i, j, k;
    pa_cell_coords(a.dc, q, i, j, k);
    if (i < a.g.nx && j < a.g.ny && k < a.g.nz) {
      const uint32_t m = pa_liquid_count(a, a.start[q], a.start[q + 1]);
      uint32_t want = 0;
      if (m > 0) {
        cells = 1;
        before = m;
        const bool near = r.near_gas[q] != 0;
        iface = near;
        if (!(r.interface_only && !near) && m < (uint32_t)r.target) {
          under = 1;
          want = (uint32_t)r.target - m;
        }
      }
      r.wanted[i + (long long)a.g.nx * (j + (long long)a.g.ny * k)] = want;
    }
  }
  pa_count(a.cnt, PA_RF_CELLS, cells);
  pa_count(a.cnt, PA_RF_BEFORE, before);
  pa_count(a.cnt, PA_RF_INTERFACE, iface);
  pa_count(a.cnt, PA_RF_UNDERFULL, under);
}

// per underfull cell: its share of the budget, the volume split and the new
// particles (:329-371); a cell's share is min(wanted, budget - wanted before it)
__global__ void __launch_bounds__(256) k_pa_refill_apply(AdaptArgs a, RefillArgs r) {
#pragma clang fp contract(off)
  const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
  uint32_t limited = 0;
  if (q < a.dc.nblocks * kBlockVol) {
    int i, j, k;
    pa_cell_coords(a.dc, q, i, j, k);
    if (i < a.g.nx && j < a.g.ny && k < a.g.nz) {
      const long long lin = i + (long long)a.g.nx * (j + (long long)a.g.ny * k);
      const long long want = r.wanted[lin];
      if (want > 0) {
        const long long wb = r.wscan[lin];
        const long long remaining = wb < r.budget ? r.budget - wb : 0;
        const long long add = want < remaining ? want : remaining;
        limited = add < want;
        if (add > 0) {
          const uint32_t s = a.start[q], e = a.start[q + 1];
          const uint32_t m = pa_liquid_count(a, s, e);
          // sums over the cell's liquid particles in id (array) order
          double vs = 0.0, su = 0.0, sv = 0.0, sw = 0.0;
          uint32_t last = 0;
          for (uint32_t t = 0; t < m; ++t) {
            uint32_t best = 0xffffffffu, bp = s;
            for (uint32_t p = s; p < e; ++p) {
              if (!a.alive[p] || a.P.type[p] != 0) continue;
              const uint32_t id = a.P.id[p];
              if ((t == 0 || id > last) && id < best) {
                best = id;
                bp = p;
              }
            }
            vs += a.P.vol ? (double)a.P.vol[bp] : 1.0;
            su += (double)a.P.u[bp];
            sv += (double)a.P.v[bp];
            sw += (double)a.P.w[bp];
            last = best;
          }
          const double split = vs / (double)(m + add);
          for (uint32_t p = s; p < e; ++p)
            if (a.alive[p] && a.P.type[p] == 0) a.P.vol[p] = (float)split;
          const double inv = 1.0 / (double)m;
          const float au = (float)(su * inv), av = (float)(sv * inv), aw = (float)(sw * inv);
          // slots ordered by (slotScore, slot)
          int slots[8];
          uint32_t sc[8];
#pragma unroll
          for (int t = 0; t < 8; ++t) {
            slots[t] = t;
            sc[t] = pa_slot_score((uint32_t)lin, t, r.seed);
          }
          for (int t = 1; t < 8; ++t)
            for (int b = t; b > 0; --b) {
              const uint32_t s0 = sc[slots[b - 1]], s1 = sc[slots[b]];
              if (s0 < s1 || (s0 == s1 && slots[b - 1] < slots[b])) break;
              const int tmp = slots[b];
              slots[b] = slots[b - 1];
              slots[b - 1] = tmp;
            }
          const long long base = r.n_kept + (wb < r.budget ? wb : r.budget);
          for (long long t = 0; t < add; ++t) {
            const int sl = slots[t & 7];
            const double sx = 0.25 + 0.5 * (double)(sl & 1), sy = 0.25 + 0.5 * (double)((sl >> 1) & 1);
            const double sz = 0.25 + 0.5 * (double)((sl >> 2) & 1);
            const long long d = base + t;
            r.out.x[d] = (float)(a.g.ox + ((double)i + sx) * a.g.dx);
            r.out.y[d] = (float)(a.g.oy + ((double)j + sy) * a.g.dx);
            r.out.z[d] = (float)(a.g.oz + ((double)k + sz) * a.g.dx);
            r.out.u[d] = au;
            r.out.v[d] = av;
            r.out.w[d] = aw;
            r.out.type[d] = 0;
            r.out.vol[d] = (float)split;
            r.out.id[d] = (uint32_t)d;
          }
        }
      }
    }
  }
  pa_count(a.cnt, PA_RF_BUDGET, limited);
}

// survivors by id, for the renumbering scan
__global__ void __launch_bounds__(256) k_pa_alive_by_id(AdaptArgs a, uint32_t* __restrict__ by_id) {
  const long long p = (long long)blockIdx.x * 256 + threadIdx.x;
  if (p < a.np) by_id[a.P.id[p]] = a.alive[p];
}

// survivors to their new index (the rank of their id among the survivors), which is
// also their new id; the output is in id (caller) order
__global__ void __launch_bounds__(256) k_pa_compact(AdaptArgs a, const uint32_t* __restrict__ new_id, PBuf out,
                                                    bool has_vol) {
  const long long p = (long long)blockIdx.x * 256 + threadIdx.x;
  if (p >= a.np || !a.alive[p]) return;
  const uint32_t d = new_id[a.P.id[p]];
  out.x[d] = a.P.x[p]; out.y[d] = a.P.y[p]; out.z[d] = a.P.z[p];
  out.u[d] = a.P.u[p]; out.v[d] = a.P.v[p]; out.w[d] = a.P.w[p];
  out.type[d] = a.P.type[p];
  out.id[d] = d;
  if (has_vol) out.vol[d] = a.P.vol[p];
}

// per-workgroup FP64 partial sums of the liquid particle volume (1.0 without volumes)
__global__ void __launch_bounds__(256) k_pa_liquid_volume(PBuf P, long long n, bool has_vol, double* __restrict__ part) {
  __shared__ double s[4];
  double v = 0.0;
  for (long long p = (long long)blockIdx.x * 256 + threadIdx.x; p < n; p += (long long)gridDim.x * 256)
    if (P.type[p] == 0) v += has_vol ? (double)P.vol[p] : 1.0;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = s[0] + s[1] + s[2] + s[3];
}

}  // namespace pf

namespace pf {

// types and volumes back to the caller's order (out[id[p]] = in[p])
__global__ void k_unsort_ext(long long n, PBuf S, bool has_vol, uint8_t* __restrict__ type,
                             float* __restrict__ vol) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint32_t d = S.id[p];
  if (type) type[d] = S.type[p];
  if (vol) vol[d] = has_vol ? S.vol[p] : 1.0f;
}

}  // namespace pf
