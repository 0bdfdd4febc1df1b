// pf_internal.h -- device data model of the MI355X sparse two-phase PF-FLIP step.
//
// Grid storage: every field is a block-major slab of 4^3 blocks (64 values = 256 B
// for FP32), indexed directly by the reference's block id
// bid = bx + nbx*(by + nby*bz) (src/grid/sparse_block_grid3d.h:24) and the in-block
// offset (i&3) + 4*((j&3) + 4*(k&3)) (:48).  A block's pool slot is its bid, so the
// reference's block-pointer map becomes a one-byte activity flag per block plus
// sorted compact lists built by scan; inactive blocks always hold zeros, which is
// what SparseBlockGrid3D::get returns for them (:50-55).  This spends HBM (cheap at
// 288 GB) to remove the per-tap indirection of the block map.
//
// Particles: FP32 structure-of-arrays, kept sorted by cell (block-major cell index)
// by a counting sort each step; `id` is the caller's index, used to return
// particles in the caller's order and to fix the within-cell splat order.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "pf_gpu.h"

struct pf_lpc;

namespace pf {

constexpr int kBlock = 4;
constexpr int kBlockVol = 64;

// One field family: node extents and block counts (faces have one more node on
// their own axis, src/grid/sparse_mac_grid3d.h:14-17).
struct Dims {
  int nx, ny, nz;
  int nbx, nby, nbz;
  long long nblocks;
  uint64_t inv_nbx, inv_nby;  // division constants of nbx and nby (fastdiv_u32)
};

// ceil(2^64 / d) for d >= 2, 0 for d = 1 (Lemire, Kaser, Kurz, "Faster remainder by
// direct computation", 2019: floor(n / d) = floor(c n / 2^64) for every 32-bit n)
inline uint64_t fastdiv_constant(uint32_t d) { return d <= 1 ? 0 : ~0ull / d + 1; }

// floor(n / d) from c = fastdiv_constant(d): the high 64 bits of the 96-bit product
// c n, formed from 32-bit halves (no 64-bit division on the device)
__host__ __device__ inline uint32_t fastdiv_u32(uint32_t n, uint64_t c) {
  if (c == 0) return n;
  const uint64_t lo = (c & 0xffffffffull) * n, hi = (c >> 32) * n;
  return (uint32_t)((hi + (lo >> 32)) >> 32);
}

inline Dims make_dims(int nx, int ny, int nz) {
  Dims d;
  d.nx = nx; d.ny = ny; d.nz = nz;
  d.nbx = (nx + kBlock - 1) / kBlock;
  d.nby = (ny + kBlock - 1) / kBlock;
  d.nbz = (nz + kBlock - 1) / kBlock;
  d.nblocks = (long long)d.nbx * d.nby * d.nbz;
  d.inv_nbx = fastdiv_constant((uint32_t)d.nbx);
  d.inv_nby = fastdiv_constant((uint32_t)d.nby);
  return d;
}

__host__ __device__ inline long long bid_of(const Dims& d, int bx, int by, int bz) {
  return bx + (long long)d.nbx * (by + (long long)d.nby * bz);
}

__host__ __device__ inline long long bm_index(const Dims& d, int i, int j, int k) {
  return bid_of(d, i >> 2, j >> 2, k >> 2) * kBlockVol + (i & 3) + 4 * (j & 3) + 16 * (k & 3);
}

// block ids are below 2^26 (cells below 2^32), so 32-bit arithmetic is exact
__host__ __device__ inline void block_coords(const Dims& d, long long b, int& bx, int& by, int& bz) {
  const uint32_t n = (uint32_t)b;
  const uint32_t t = fastdiv_u32(n, d.inv_nbx);
  const uint32_t z = fastdiv_u32(t, d.inv_nby);
  bx = (int)(n - t * (uint32_t)d.nbx);
  by = (int)(t - z * (uint32_t)d.nby);
  bz = (int)z;
}

// color8, src/grid/block_color.h:7
__host__ __device__ inline int color8(int bx, int by, int bz) {
  return (bx & 1) + 2 * (by & 1) + 4 * (bz & 1);
}

struct Geom {
  int nx, ny, nz;
  double dx, ox, oy, oz;
};

// Particle SoA (one of two ping-pong buffers).
struct PBuf {
  float *x, *y, *z, *u, *v, *w;
  uint8_t* type;
  float* vol;     // may be null (all 1.0)
  uint32_t* id;
};

enum { F_U = 0, F_V, F_W, F_MU, F_MV, F_MW, F_P, F_MK, F_SU, F_SV, F_SW, F_N };

// Particle order inside a cell, which fixes P2G's summation order: a cell of at most
// kRankMax particles is ranked by caller id in P2G itself; after binning, a cell of
// kRankMax < count <= kSortedMax particles is sorted by caller id in memory
// (k_sort_crowded); a fuller cell is walked by selection of the next id.
constexpr int kRankMax = 32;
constexpr int kSortedMax = 2048;
enum { CELL_AIR = 0, CELL_FLUID = 1, CELL_SOLID = 2 };

// FP64 scalars of one inner PCG solve, kept on the device so the iteration
// needs no host round trip.
struct SolveScalars {
  double rz, pAp, alpha, beta;
  double rr, rr0, rr_best, tol2;     // tol2 = delta^2 (replacement trigger)
  double rr_ref, tol2_abs, rz_pending, rz_cross;  // rz_cross: r_new . z_old at a pause (flexible beta)
  int iter, done, breakdown, stalled;
  int iter_best, max_iter, stall_window, upd;
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  Geom geom{};
  Dims du{}, dv{}, dw{}, dc{};   // u-faces, v-faces, w-faces, cells
  std::string err;

  // grid fields (block-major)
  float *f[F_N] = {};            // F_MK slot unused (marker is u8)
  uint8_t* marker = nullptr;
  uint8_t* act[F_N] = {};        // activity flag per block
  uint32_t* cell_count = nullptr;   // per cell (dense block-major)
  uint32_t* cell_start = nullptr;   // exclusive scan, ncells + 1

  // block lists
  uint32_t* color_list = nullptr;   // 8 * color_cap
  uint32_t* color_cnt = nullptr;    // 8
  long long color_cap = 0;
  uint32_t* crowd_list = nullptr;   // cells of kRankMax < count <= kSortedMax particles (sorted by id)
  long long crowd_cap = 0;
  uint32_t* pblk = nullptr;         // sorted list of fluid cell blocks
  uint32_t* pblk_prev = nullptr;    // previous projection's list (zeroed on reuse)
  long long npblk = 0, npblk_prev = 0;
  uint32_t* blk_flag32 = nullptr;   // scratch: per-block u32 flags for compaction
  uint32_t* blk_scan = nullptr;     // scratch: scan of block flags (>= max(blocks, Morton codes) + 1)
  uint8_t* morton_flags = nullptr;  // scratch: block flags in Morton order
  long long morton_cap = 0;

  // particles
  long long np = 0, pcap = 0;
  PBuf pa{}, pb{};                  // pa is current
  bool has_vol = false;
  bool binned = false;
  bool mark_order_valid = false;    // the particle order is the binning the last marking used
  uint32_t* rank = nullptr;
  double liquid_volume = 0.0;       // sum of volume over liquid particles

  // pressure solve (cells, block-major)
  float *x32 = nullptr, *r32 = nullptr, *z32 = nullptr, *pA = nullptr, *pB = nullptr, *Ap = nullptr;
  float *cd32 = nullptr, *cx32 = nullptr, *cy32 = nullptr, *cz32 = nullptr, *invd32 = nullptr;
  double *b64 = nullptr, *x64 = nullptr, *r64 = nullptr;
  double *cd64 = nullptr, *cx64 = nullptr, *cy64 = nullptr, *cz64 = nullptr;
  SolveScalars* sc = nullptr;       // device
  double* partial = nullptr;        // per-workgroup partial sums (kMaxPartials * 4)
  double* red = nullptr;            // reduced values (device, 16 doubles)
  double* red_host = nullptr;       // pinned host mirror
  int* done_host = nullptr;         // pinned host mirror of sc->done
  SolveScalars* sc_host = nullptr;  // pinned host mirror of the whole struct
  uint32_t* u32_dev = nullptr;      // small device scratch (pin detection, counters), 64 words
  uint32_t* u32_host = nullptr;     // pinned mirror, 64 words
  int32_t* esc_dev = nullptr;       // 8 clamp counters
  float* vmax_dev = nullptr;        // max particle speed (as float bits via atomicMax)
  float4* g2p_quad[3] = {nullptr, nullptr, nullptr};  // per face (vel, saved) of it and of its +x neighbour (G2P)
  bool vmax_valid = false;          // vmax_dev holds the max speed of the current velocities

  // scan scratch
  uint32_t* scan_tiles = nullptr;
  long long scan_tile_cap = 0;

  // staging for field transfers
  float* stage_f = nullptr;
  long long stage_cap = 0;

  hipEvent_t ev[PF_NSTAGES + 1] = {};

  // learned preconditioner (attached, not owned) and its per-system data
  ::pf_lpc* lpc = nullptr;
  float* fbeta = nullptr;           // cell beta feature (mean face beta) on FLUID cells
  float* fphi = nullptr;            // cell phi feature (mean face phi) on FLUID cells
  uint32_t* tiles = nullptr;        // 16^3 tiles of level-1 aggregates holding FLUID cells (sorted ids)
  uint8_t* tile_flags = nullptr;
  long long ntiles = 0, ntiles_all = 0;
  int ntx = 0, nty = 0, ntz = 0;
  bool tiles_valid = false;         // tiles/features describe the last assembled system
  // residual snapshots recorded during solves (training data)
  float* snaps = nullptr;
  int snap_cap = 0, nsnap = 0, snap_every = 0, snap_last_iter = -1;

  // multires layout (2-level leaves of 4^3 blocks over the fine grid)
  bool mr = false;
  std::vector<int32_t> mr_leaves;   // sorted (level, bx, by, bz)
  int32_t* mr_leaves_dev = nullptr;
  long long mr_leaves_cap = 0;
  uint8_t* mr_lvl = nullptr;        // level byte per fine cell
  uint8_t* mr_valid[3] = {};        // enumerated fine faces
  float* mr_cf32[3] = {};           // coupling conductance per fine face
  double* mr_cf64[3] = {};
  double* mr_vol64 = nullptr;       // leaf-cell volume at FLUID anchors
  double* mr_pfull = nullptr;       // FP64 pressure per fine cell
  int* mr_bbox = nullptr;           // particle cell boxes (device)
  int* mr_bbox_host = nullptr;
  bool mr_retained_valid = false;   // dynamic refinement state (MRSim3DTP::dynamic_retained_*)
  int mr_retained[6] = {0, 0, 0, 0, 0, 0};
  int mr_budget_limited = 0, mr_last_fine_leaves = 0;
  // the reference-order FP64 multires solve (options of MRPressureSolveConfig3D)
  const pf_mr_solve_cfg* mr_cfg = nullptr;  // set by pf_mr_project / pf_mr_step for one solve
  pf_mr_solve_detail mr_detail{};
  double* mrf = nullptr;            // 15 fine FP64 vectors (ncell each)
  double* mrc = nullptr;            // 6 coarse FP64 vectors
  // the multires cycle's block lists: levels 0-1 (a wave per block) and level 2 (four
  // blocks per wave), split from the fluid list per solve
  uint8_t *mr_f01 = nullptr, *mr_f2 = nullptr;
  uint32_t *mr_list01 = nullptr, *mr_list2 = nullptr;
  long long mr_n01 = 0, mr_n2 = 0;
  long long mrc_cap = 0;
  long long* mr_pin_dev = nullptr;  // pinned anchor (block-major index) or -1
  long long* mr_pin_host = nullptr;

  // interface fields and CSF increments (FP64 per cell of the slab: phi, |grad phi|,
  // three normal components, curvature, three increments), allocated on first use
  double* iface = nullptr;
  unsigned long long* iface_bad = nullptr;  // least visiting-order key of a non-finite increment

  // particle adaptivity scratch (per particle of the input set, per cell)
  uint8_t* pa_alive = nullptr;
  uint64_t* pa_key = nullptr;
  uint32_t* pa_by_id = nullptr;     // survivors by id, then its exclusive scan (np + 1)
  uint32_t* pa_new_id = nullptr;
  long long pa_cap = 0;
  uint8_t *pa_cflag = nullptr, *pa_d1 = nullptr, *pa_d2 = nullptr;  // per cell
  uint32_t *pa_wanted = nullptr, *pa_wscan = nullptr;              // per cell, linear order (+1)
  uint32_t* pa_cnt = nullptr;       // counters

  // aggregation multigrid preconditioner: the coarse levels (level 0 is the assembled
  // system) and a level-0 work array
  struct MgBuf {
    Dims d;
    int nx = 0, ny = 0, nz = 0;
    float *cd = nullptr, *cx = nullptr, *cy = nullptr, *cz = nullptr, *inv = nullptr;
    float *x = nullptr, *r = nullptr, *t = nullptr;
    // the level's operator as scaled FP16 for the row kernels (cd, cx, cy, cz, inv),
    // its maxima (float bits) and scales
    uint16_t* h[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    uint32_t* mx = nullptr;
    float* unscale = nullptr;
  };
  std::vector<MgBuf> mg;
  float* mg_t0 = nullptr;
  // FP16 copies of the level-0 cd, cx, cy, cz and diagonal inverse (scaled), their
  // maxima (float bits) and scales
  uint16_t *mg_h[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  uint32_t* mg_mx = nullptr;
  bool mg_mx_valid = false;   // mg_mx holds the maxima of the current system (set by the assembly)
  float* mg_unscale = nullptr;
  // per-aggregate scaling s = sqrt(w) of level 1's restricted residual and correction
  // (P S B_1 S P^T in place of P B_1 P^T), or null: w = 1.  mg_cs holds the caller's
  // weights, mg_cs_net the learned ones of the last assembled system; the cycle reads
  // mg_cs_active
  float* mg_cs = nullptr;
  float* mg_cs_net = nullptr;
  const float* mg_cs_active = nullptr;
  pf_mg_cfg mg_cfg{};   // the cycle (set at creation, pf_set_multigrid)

  // per-kernel event timing (measurement hook)
  uint32_t ktime = 0;                // kernels (bit PF_K_*) bracketed with events
  std::vector<hipEvent_t> kev;
  size_t kev_used = 0;
  std::vector<std::pair<int, size_t>> krec;  // (kernel id, index of its start event)
};

constexpr int kMaxPartials = 4096;

// ---- launch helpers shared by the kernel files (implemented in pf_util.hip etc.)
hipError_t fill_u8(hipStream_t s, uint8_t* p, uint8_t v, long long n);
hipError_t fill_u32(hipStream_t s, uint32_t* p, uint32_t v, long long n);
hipError_t fill_f32(hipStream_t s, float* p, float v, long long n);
hipError_t fill_f64(hipStream_t s, double* p, double v, long long n);
// out[0..n] = exclusive scan of in[0..n), out[n] = total
hipError_t scan_u32(Ctx& c, const uint32_t* in, uint32_t* out, long long n);
// list = sorted ids b with flag[b] != 0; count written to *count_dev (device)
hipError_t compact_flags(Ctx& c, const uint8_t* flags, long long n, uint32_t* list,
                         uint32_t* count_dev);
// the same set of flagged blocks (flags indexed by bid), listed in Morton order
hipError_t compact_flags_morton(Ctx& c, const Dims& d, const uint8_t* flags, uint32_t* list,
                                uint32_t* count_dev);
// size of the Morton code space covering d's block grid
long long morton_codes(const Dims& d);

}  // namespace pf

struct pf_ctx : public pf::Ctx {};
