// pf_grid.hip -- cell markers, the face-field post-P2G pass (normalize, save,
// gravity, walls) and dense <-> block-major field transfers.
//
// All kernels map one 64-lane wavefront to one 4^3 block (lane = in-block offset),
// four blocks per 256-thread workgroup, so every block is one coalesced 256-byte
// access and block-level predicates are a single wave vote.
#include "pf_internal.h"

namespace pf {

// markCells, src/driver/sparse_sim3d_tp.cpp:22-50: walls SOLID, cells holding a
// particle FLUID unless solid, AIR otherwise.  Particle counts come from binning.
__global__ void __launch_bounds__(256) k_mark(Dims dc, Geom g, const uint32_t* __restrict__ start,
                                              uint8_t* __restrict__ marker, uint8_t* __restrict__ act_mk) {
  long long b = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  int lane = threadIdx.x & 63;
  if (b >= dc.nblocks) return;
  int bx, by, bz;
  block_coords(dc, b, bx, by, bz);
  int i = bx * 4 + (lane & 3), j = by * 4 + ((lane >> 2) & 3), k = bz * 4 + (lane >> 4);
  bool in = i < g.nx && j < g.ny && k < g.nz;
  long long c = b * kBlockVol + lane;
  uint32_t cnt = start[c + 1] - start[c];
  bool wall = in && (i == 0 || i == g.nx - 1 || j == 0 || j == g.ny - 1 || k == 0 || k == g.nz - 1);
  uint8_t m = !in ? CELL_AIR : (wall ? CELL_SOLID : (cnt > 0 ? CELL_FLUID : CELL_AIR));
  marker[c] = m;
  bool any = __any(m != CELL_AIR);
  if (lane == 0) act_mk[b] = any ? 1 : 0;
}

// Per-block "holds a FLUID cell" flag (the pressure field's active set).
__global__ void __launch_bounds__(256) k_fluid_flags(Dims dc, const uint8_t* __restrict__ marker,
                                                     uint8_t* __restrict__ flag) {
  long long b = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  int lane = threadIdx.x & 63;
  if (b >= dc.nblocks) return;
  bool f = marker[b * kBlockVol + lane] == CELL_FLUID;
  bool any = __any(f);
  if (lane == 0) flag[b] = any ? 1 : 0;
}

enum { FO_NORMALIZE = 1, FO_SAVE = 2, FO_GRAVITY = 4, FO_WALLS = 8 };

// One face axis: u = mom/mass where mass > 0 (sparse_ops3d_tp.cpp:126-152),
// saved = u (sparse_sim3d_tp.cpp:866), v += (float)(dt*g) where mass > 0 (:52-70),
// u = 0 on the two face layers next to each wall (:72-97).
template <int AX>
__global__ void __launch_bounds__(256) k_face_ops(Dims d, Geom g, float* __restrict__ vel,
                                                  const float* __restrict__ mass, float* __restrict__ saved,
                                                  uint8_t* __restrict__ act_vel, uint8_t* __restrict__ act_saved,
                                                  int ops, float dtg) {
  long long b = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  int lane = threadIdx.x & 63;
  if (b >= d.nblocks) return;
  int bx, by, bz;
  block_coords(d, b, bx, by, bz);
  int i = bx * 4 + (lane & 3), j = by * 4 + ((lane >> 2) & 3), k = bz * 4 + (lane >> 4);
  bool in = i < d.nx && j < d.ny && k < d.nz;
  long long gi = b * kBlockVol + lane;
  float val = vel[gi];
  float m = (ops & (FO_NORMALIZE | FO_GRAVITY)) ? mass[gi] : 0.f;
  if ((ops & FO_NORMALIZE) && m > 0.f) val = val / m;
  if (ops & FO_SAVE) {
    saved[gi] = val;
    if (lane == 0) act_saved[b] = act_vel[b];
  }
  if ((ops & FO_GRAVITY) && AX == 1 && m > 0.f) val = val + dtg;
  if (ops & FO_WALLS) {
    int a = AX == 0 ? i : (AX == 1 ? j : k);
    int n = AX == 0 ? g.nx : (AX == 1 ? g.ny : g.nz);
    bool wall = in && (a == 0 || a == 1 || a == n - 1 || a == n);
    if (wall) val = 0.f;
    bool any = __any(wall);
    if (lane == 0 && any) act_vel[b] = 1;
  }
  if (in) vel[gi] = val;
}

// dense (i + nx*(j + ny*k)) -> block-major; activity = given flags or "any non-zero"
__global__ void __launch_bounds__(256) k_from_dense(Dims d, const float* __restrict__ dense,
                                                    float* __restrict__ out, uint8_t* __restrict__ out_u8,
                                                    uint8_t* __restrict__ act, const uint8_t* __restrict__ act_in) {
  long long b = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  int lane = threadIdx.x & 63;
  if (b >= d.nblocks) return;
  int bx, by, bz;
  block_coords(d, b, bx, by, bz);
  int i = bx * 4 + (lane & 3), j = by * 4 + ((lane >> 2) & 3), k = bz * 4 + (lane >> 4);
  bool in = i < d.nx && j < d.ny && k < d.nz;
  float v = in ? dense[i + (long long)d.nx * (j + (long long)d.ny * k)] : 0.f;
  long long gi = b * kBlockVol + lane;
  if (out) out[gi] = v;
  if (out_u8) out_u8[gi] = (uint8_t)(v + 0.5f);
  bool any = __any(v != 0.f);
  if (lane == 0) act[b] = act_in ? (act_in[b] ? 1 : 0) : (any ? 1 : 0);
}

__global__ void k_to_dense(Dims d, const float* __restrict__ f, const uint8_t* __restrict__ f_u8,
                           float* __restrict__ dense) {
  long long n = (long long)d.nx * d.ny * d.nz;
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  int i = (int)(t % d.nx);
  long long q = t / d.nx;
  int j = (int)(q % d.ny), k = (int)(q / d.ny);
  long long gi = bm_index(d, i, j, k);
  dense[t] = f ? f[gi] : (float)f_u8[gi];
}

}  // namespace pf
