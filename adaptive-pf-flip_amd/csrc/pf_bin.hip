// pf_bin.hip -- particle binning by cell (counting sort), per-colour block lists,
// caller-order download and the seeded scene generator.
//
// Binning replaces the reference's per-particle block activation bookkeeping with a
// sort: the key of a particle is the block-major index of its cell
// (int)((x - o)/dx) evaluated in FP64 exactly as markCells does
// (src/driver/sparse_sim3d_tp.cpp:43-46).  A cell's particles end up contiguous,
// blocks in bid order, so every later kernel reads particles coalesced and finds a
// cell's range from the scanned counts.
#include "pf_internal.h"

namespace pf {

__device__ inline int cell_coord(float x, double o, double dx) {
  return (int)(((double)x - o) / dx);
}

// Rank of this lane among the lanes of its wave with the same key, plus one
// atomicAdd per distinct key of the wave (a wave of cell-sorted particles covers
// about eight cells, so this is ~8x fewer same-address atomics).  The groups are
// found first (ballots only); then every group leader issues its atomic in the same
// instruction, so the wave waits for one atomic round trip instead of one per key.
__device__ inline uint32_t wave_aggregated_increment(uint32_t* cnt, uint32_t key) {
  const int lane = threadIdx.x & 63;
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
  uint64_t todo = __ballot(1);
  int my_leader = lane;
  uint32_t my_off = 0, my_count = 0;
  while (todo) {
    const int leader = __ffsll((unsigned long long)todo) - 1;
    const uint32_t k = __shfl(key, leader, 64);
    const uint64_t grp = __ballot(key == k) & todo;
    if ((grp >> lane) & 1ull) {
      my_leader = leader;
      my_off = (uint32_t)__popcll(grp & below);
    }
    if (lane == leader) my_count = (uint32_t)__popcll(grp);
    todo &= ~grp;
  }
  uint32_t base = 0;
  if (my_count) base = atomicAdd(&cnt[key], my_count);
  return __shfl(base, my_leader, 64) + my_off;
}

// the block-major index of a particle's cell.  fscale > 0 (origin 0 and a power-of-two
// dx, fscale = 1/dx): x * fscale is exact in FP32 and equals the FP64 (x - 0) / dx, so
// the index is the same without three FP64 divisions (which made the count pass
// VALU-bound)
__device__ inline uint32_t cell_key(float x, float y, float z, const Geom& g, const Dims& dc, float fscale) {
  int i, j, k;
  if (fscale > 0.f) {
    i = (int)(x * fscale);
    j = (int)(y * fscale);
    k = (int)(z * fscale);
  } else {
    i = cell_coord(x, g.ox, g.dx);
    j = cell_coord(y, g.oy, g.dx);
    k = cell_coord(z, g.oz, g.dx);
  }
  // positions are validated into the cell range on entry and clamped by advection;
  // the clamp only guards memory
  i = min(max(i, 0), g.nx - 1);
  j = min(max(j, 0), g.ny - 1);
  k = min(max(k, 0), g.nz - 1);
  return (uint32_t)bm_index(dc, i, j, k);
}

__global__ void k_bin_count(long long n, PBuf P, Geom g, Dims dc, float fscale, uint32_t* __restrict__ rank,
                            uint32_t* __restrict__ cnt) {
  long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  rank[p] = wave_aggregated_increment(cnt, cell_key(P.x[p], P.y[p], P.z[p], g, dc, fscale));
}

// the scatter reads the positions anyway, so it recomputes the key instead of the
// count pass storing it (8 bytes less traffic per particle)
__global__ void k_bin_scatter(long long n, PBuf S, PBuf D, bool has_vol, Geom g, Dims dc, float fscale,
                              const uint32_t* __restrict__ rank, const uint32_t* __restrict__ start) {
  long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const float x = S.x[p], y = S.y[p], z = S.z[p];
  uint32_t d = start[cell_key(x, y, z, g, dc, fscale)] + rank[p];
  D.x[d] = x; D.y[d] = y; D.z[d] = z;
  D.u[d] = S.u[p]; D.v[d] = S.v[p]; D.w[d] = S.w[p];
  D.type[d] = S.type[p];
  D.id[d] = S.id[p];
  if (has_vol) D.vol[d] = S.vol[p];
}

// Per-colour lists of cell blocks that hold particles (the P2G source blocks).
// Append order is free: same-colour blocks write disjoint regions.  One atomic per
// colour present in the wave (lanes take consecutive slots by their rank among the
// wave's blocks of that colour): a wave of consecutive block ids holds few colours,
// where one atomic per block would queue on 8 counters.
__global__ void k_color_lists(Dims dc, const uint32_t* __restrict__ start, uint32_t* __restrict__ cnt,
                              uint32_t* __restrict__ lists, long long cap) {
  const long long b = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const bool has = b < dc.nblocks && start[(b + 1) * kBlockVol] != start[b * kBlockVol];
  int col = -1;
  if (has) {
    int bx, by, bz;
    block_coords(dc, b, bx, by, bz);
    col = color8(bx, by, bz);
  }
  for (int k = 0; k < 8; ++k) {
    const unsigned long long m = __ballot(col == k);
    if (m == 0) continue;
    const int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(&cnt[k], (uint32_t)__popcll(m));
    base = __shfl(base, leader, 64);
    if (col == k) {
      const uint32_t rank = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
      lists[k * cap + base + rank] = (uint32_t)b;
    }
  }
}

// Cells of lo < count <= hi particles, appended in any order (each is sorted on its
// own); one atomic per wave.  cap >= np / (lo + 1) bounds the list.
__global__ void k_crowded_cells(long long ncell, const uint32_t* __restrict__ start, uint32_t lo, uint32_t hi,
                                uint32_t* __restrict__ list, uint32_t* __restrict__ count, long long cap) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  uint32_t cnt = 0;
  if (c < ncell) cnt = start[c + 1] - start[c];
  const bool crowd = cnt > lo && cnt <= hi;
  const unsigned long long m = __ballot(crowd);
  if (m == 0) return;
  const int leader = __ffsll(m) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(count, (uint32_t)__popcll(m));
  base = __shfl(base, leader, 64);
  if (crowd) {
    const long long q = (long long)base + __popcll(m & ((1ull << lane) - 1ull));
    if (q < cap) list[q] = (uint32_t)c;
  }
}

// Sorts the particles of each listed cell by caller id, in place: a bitonic sort of
// (id, slot) pairs in LDS, then every particle array permuted through LDS.  A
// workgroup takes list entries w, w + gridDim.x, ...
template <typename T>
__device__ inline void permute_cell(T* a, uint32_t s0, uint32_t cnt, const uint16_t* pos, uint32_t* tmp) {
  for (uint32_t t = threadIdx.x; t < cnt; t += blockDim.x) {
    T v = a[s0 + t];
    uint32_t bits = 0;
    __builtin_memcpy(&bits, &v, sizeof(T));
    tmp[t] = bits;
  }
  __syncthreads();
  for (uint32_t t = threadIdx.x; t < cnt; t += blockDim.x) {
    const uint32_t bits = tmp[pos[t]];
    T v;
    __builtin_memcpy(&v, &bits, sizeof(T));
    a[s0 + t] = v;
  }
  __syncthreads();
}

__global__ void __launch_bounds__(256) k_sort_crowded(const uint32_t* __restrict__ list,
                                                      const uint32_t* __restrict__ count,
                                                      const uint32_t* __restrict__ start, PBuf P, bool has_vol) {
  __shared__ uint32_t key[kSortedMax];
  __shared__ uint16_t pos[kSortedMax];
  __shared__ uint32_t tmp[kSortedMax];
  const uint32_t n = *count;
  for (uint32_t e = blockIdx.x; e < n; e += gridDim.x) {
    const uint32_t c = list[e];
    const uint32_t s0 = start[c], cnt = start[c + 1] - s0;  // kRankMax < cnt <= kSortedMax
    uint32_t m = 1;
    while (m < cnt) m <<= 1;
    for (uint32_t t = threadIdx.x; t < m; t += blockDim.x) {
      key[t] = t < cnt ? P.id[s0 + t] : 0xffffffffu;
      pos[t] = (uint16_t)t;
    }
    __syncthreads();
    for (uint32_t k = 2; k <= m; k <<= 1) {
      for (uint32_t j = k >> 1; j > 0; j >>= 1) {
        for (uint32_t t = threadIdx.x; t < m; t += blockDim.x) {
          const uint32_t u = t ^ j;
          if (u > t) {
            const bool up = (t & k) == 0;
            const uint32_t a = key[t], b = key[u];
            if ((a > b) == up) {
              key[t] = b; key[u] = a;
              const uint16_t pa = pos[t];
              pos[t] = pos[u]; pos[u] = pa;
            }
          }
        }
        __syncthreads();
      }
    }
    permute_cell(P.x, s0, cnt, pos, tmp);
    permute_cell(P.y, s0, cnt, pos, tmp);
    permute_cell(P.z, s0, cnt, pos, tmp);
    permute_cell(P.u, s0, cnt, pos, tmp);
    permute_cell(P.v, s0, cnt, pos, tmp);
    permute_cell(P.w, s0, cnt, pos, tmp);
    permute_cell(P.type, s0, cnt, pos, tmp);
    permute_cell(P.id, s0, cnt, pos, tmp);
    if (has_vol) permute_cell(P.vol, s0, cnt, pos, tmp);
  }
}

// Per-cell particle counts of the binning in dense linear order i + nx*(j + ny*k).
__global__ void k_counts_dense(Dims dc, const uint32_t* __restrict__ cnt, uint32_t* __restrict__ out) {
  const long long n = (long long)dc.nx * dc.ny * dc.nz;
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int i = (int)(q % dc.nx);
  const long long t = q / dc.nx;
  const int j = (int)(t % dc.ny), k = (int)(t / dc.ny);
  out[q] = cnt[bm_index(dc, i, j, k)];
}

// Return particles to the caller's order (out[id[p]] = in[p]).
__global__ void k_unsort(long long n, PBuf S, float* __restrict__ x, float* __restrict__ y,
                         float* __restrict__ z, float* __restrict__ u, float* __restrict__ v,
                         float* __restrict__ w) {
  long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  uint32_t d = S.id[p];
  x[d] = S.x[p]; y[d] = S.y[p]; z[d] = S.z[p];
  u[d] = S.u[p]; v[d] = S.v[p]; w[d] = S.w[p];
}

// Validates that positions lie in the cell range (sets *bad otherwise).
__global__ void k_validate(long long n, PBuf P, Geom g, uint32_t* bad) {
  long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  double x = P.x[p], y = P.y[p], z = P.z[p];
  bool ok = x >= g.ox && y >= g.oy && z >= g.oz && x < g.ox + g.nx * g.dx && y < g.oy + g.ny * g.dx &&
            z < g.oz + g.nz * g.dx;
  if (ok) {
    int i = cell_coord(P.x[p], g.ox, g.dx), j = cell_coord(P.y[p], g.oy, g.dx),
        k = cell_coord(P.z[p], g.oz, g.dx);
    ok = i >= 0 && i < g.nx && j >= 0 && j < g.ny && k >= 0 && k < g.nz;
  }
  if (!ok) atomicOr(bad, 1u);
}

// splitmix64 (Steele, Lea, Flood 2014)
__host__ __device__ inline uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

// Seeded synthetic scene: 8 jittered particles per filled interior cell of the box
// [1, ie) x [1, je) x [1, nz-1), in (k, j, i, particle) order.  Positions are
// (float)((i + U[0,1)) * dx); types as in the construction of the three workloads.
__global__ void k_scene(int kind, Geom g, int ie, int je, uint64_t seed, PBuf P) {
#pragma clang fp contract(off)
  long long ni = ie - 1, nj = je - 1, nk = g.nz - 2;
  long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ni * nj * nk) return;
  int i = (int)(c % ni) + 1;
  int j = (int)((c / ni) % nj) + 1;
  int k = (int)(c / (ni * nj)) + 1;
  const double dx = g.dx;
  const double Lx = g.nx * dx, Ly = g.ny * dx, Lz = g.nz * dx;
  const double two_pi = 6.283185307179586;
  uint64_t cid = (uint64_t)i + (uint64_t)g.nx * ((uint64_t)j + (uint64_t)g.ny * (uint64_t)k);
  uint64_t s = seed * 0x9E3779B97F4A7C15ULL ^ (cid + 1) * 0xD1B54A32D192ED03ULL;
  long long base = c * 8;
  for (int q = 0; q < 8; ++q) {
    double ux = (double)(splitmix64(s) >> 11) * 0x1.0p-53;
    double uy = (double)(splitmix64(s) >> 11) * 0x1.0p-53;
    double uz = (double)(splitmix64(s) >> 11) * 0x1.0p-53;
    float fx = (float)__dmul_rn(__dadd_rn((double)i, ux), dx);
    float fy = (float)__dmul_rn(__dadd_rn((double)j, uy), dx);
    float fz = (float)__dmul_rn(__dadd_rn((double)k, uz), dx);
    uint8_t t = 0;
    if (kind == 0) {
      double yi = 0.5 * Ly + 0.05 * Lx * cos(two_pi * fx / Lx) * cos(two_pi * fz / Lz);
      t = ((double)fy > yi) ? 0 : 1;
    } else if (kind == 1) {
      double ddx = fx - 0.5 * Lx, ddy = fy - 0.25 * Ly, ddz = fz - 0.5 * Lz, r = 0.12 * Lx;
      t = (ddx * ddx + ddy * ddy + ddz * ddz < r * r) ? 1 : 0;
    }
    long long p = base + q;
    P.x[p] = fx; P.y[p] = fy; P.z[p] = fz;
    P.u[p] = 0.f; P.v[p] = 0.f; P.w[p] = 0.f;
    P.type[p] = t;
    P.id[p] = (uint32_t)p;
  }
}

}  // namespace pf
