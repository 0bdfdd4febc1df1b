// pf_api.hip -- the C ABI of include/pf_gpu.h: context, memory, stage sequencing.
//
// Built as one translation unit with the kernel files so every launch is a direct
// call (no relocatable device code).  All work is stream-ordered on the context's
// stream; the only host synchronisations are the ones a result needs: the
// fluid-block count before the solve, the refinement loop's residual check, the
// convergence flag every `check_every` inner iterations, and downloads.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "pf_internal.h"
#include "pf_util.hip"
#include "pf_bin.hip"
#include "pf_grid.hip"
#include "pf_p2g.hip"
#include "pf_project.hip"
#include "pf_g2p.hip"
#include "pf_lpc.hip"
#include "pf_mr.hip"
#include "pf_mr_ref.hip"
#include "pf_iface.hip"
#include "pf_adapt.hip"
#include "pf_render.hip"
#include "pf_mg.hip"

using namespace pf;

// Learned-preconditioner handle: FP32 master weights with Adam moments, their BF16
// layer matrices, and activation buffers grown to the largest tile count seen.
struct pf_lpc {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  float *w = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr;
  LpcW bf{};
  bool prepped = false;
  int adam_t = 0;
  long long cap = 0, cap_train = 0;
  __bf16 *X = nullptr, *h1 = nullptr, *h2 = nullptr, *h3 = nullptr, *d3 = nullptr, *d2 = nullptr, *d1 = nullptr;
  float *o = nullptr, *go = nullptr, *scale = nullptr;
  float* part = nullptr;
  long long part_cap = 0;
  float *zbuf = nullptr, *qbuf = nullptr, *gz = nullptr;
  long long cell_cap = 0;
  float* stage = nullptr;
  long long stage_cap = 0;
};

#define HIP_TRY(c, expr)                                                         \
  do {                                                                           \
    hipError_t _e = (expr);                                                      \
    if (_e != hipSuccess) {                                                      \
      (c)->err = std::string(#expr) + ": " + hipGetErrorString(_e);              \
      return PF_ERR_HIP;                                                         \
    }                                                                            \
  } while (0)

#define LAUNCH_CHECK(c) HIP_TRY(c, hipGetLastError())

namespace {

int fail(pf_ctx* c, int code, const std::string& m) {
  if (c) c->err = m;
  return code;
}

// per-kernel event bracket (active only while pf_kernel_timing is enabled)
long long kbegin(pf_ctx* c, int kid) {
  if (!((c->ktime >> kid) & 1u)) return -1;
  while (c->kev.size() < c->kev_used + 2) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return -1;
    c->kev.push_back(e);
  }
  size_t i = c->kev_used;
  c->kev_used += 2;
  (void)hipEventRecord(c->kev[i], c->stream);
  c->krec.push_back({kid, i});
  return (long long)i;
}
void kend(pf_ctx* c, long long i) {
  if (i >= 0) (void)hipEventRecord(c->kev[(size_t)i + 1], c->stream);
}
#define TIMED(c, kid, ...)              \
  do {                                  \
    long long _ki = kbegin((c), (kid)); \
    __VA_ARGS__;                        \
    kend((c), _ki);                     \
  } while (0)

template <typename T>
hipError_t dalloc(T** p, long long n) {
  *p = nullptr;
  if (n <= 0) n = 1;
  hipError_t e = hipMalloc((void**)p, sizeof(T) * (size_t)n);
  if (e != hipSuccess) return e;
  return hipMemset(*p, 0, sizeof(T) * (size_t)n);
}

unsigned grid_for(long long work, long long per_block) {
  long long g = (work + per_block - 1) / per_block;
  if (g < 1) g = 1;
  return (unsigned)g;
}

// workgroups for a particle loop split into 8 contiguous XCD shares: a multiple
// of 8, at most cap
unsigned xcd_grid(long long work, long long per_block, unsigned cap) {
  long long g = (work + per_block - 1) / per_block;
  g = ((g + 7) / 8) * 8;
  if (g < 8) g = 8;
  if (g > cap) g = cap;
  return (unsigned)g;
}

// workgroups for the list kernels: a multiple of 8 (one share per XCD), at most
// kMaxPartials (one partial sum each)
unsigned list_grid(long long nlist) {
  long long g = (nlist + 3) / 4;
  g = ((g + 7) / 8) * 8;
  if (g < 8) g = 8;
  if (g > kMaxPartials) g = kMaxPartials;
  return (unsigned)g;
}

const Dims& dims_of(const pf_ctx* c, int field) {
  switch (field) {
    case PF_FIELD_U: case PF_FIELD_MU: case PF_FIELD_SAVED_U: return c->du;
    case PF_FIELD_V: case PF_FIELD_MV: case PF_FIELD_SAVED_V: return c->dv;
    case PF_FIELD_W: case PF_FIELD_MW: case PF_FIELD_SAVED_W: return c->dw;
    default: return c->dc;
  }
}

int field_slot(int field) {
  switch (field) {
    case PF_FIELD_U: return F_U;
    case PF_FIELD_V: return F_V;
    case PF_FIELD_W: return F_W;
    case PF_FIELD_MU: return F_MU;
    case PF_FIELD_MV: return F_MV;
    case PF_FIELD_MW: return F_MW;
    case PF_FIELD_P: return F_P;
    case PF_FIELD_MARKER: return F_MK;
    case PF_FIELD_SAVED_U: return F_SU;
    case PF_FIELD_SAVED_V: return F_SV;
    case PF_FIELD_SAVED_W: return F_SW;
    default: return -1;
  }
}

bool is_pow2(double dx) {
  int e;
  double m = std::frexp(dx, &e);
  return m == 0.5;
}

PhaseK phase_k(const pf_phase* ph) {
  PhaseK k;
  k.rho_l = ph->rho_l;
  k.rho_g = ph->rho_g;
  // etaPhi(pp) * pp.rho_g * pp.rho_tilde_0, phasefield.h:10-11
  k.rmin = std::log(ph->rho_l / ph->rho_g) * ph->rho_g * ph->rho_tilde_0;
  k.den = ph->alpha_phi * ph->rho_tilde_0 * ph->rho_l;
  return k;
}

int grow_particles(pf_ctx* c, long long n, bool vol) {
  if (n <= c->pcap && (!vol || c->pa.vol)) return PF_OK;
  long long cap = std::max(n, c->pcap);
  PBuf* bufs[2] = {&c->pa, &c->pb};
  for (PBuf* b : bufs) {
    if (n > c->pcap) {
      hipFree(b->x); hipFree(b->y); hipFree(b->z); hipFree(b->u); hipFree(b->v); hipFree(b->w);
      hipFree(b->type); hipFree(b->id); hipFree(b->vol);
      *b = PBuf{};
      HIP_TRY(c, dalloc(&b->x, cap)); HIP_TRY(c, dalloc(&b->y, cap)); HIP_TRY(c, dalloc(&b->z, cap));
      HIP_TRY(c, dalloc(&b->u, cap)); HIP_TRY(c, dalloc(&b->v, cap)); HIP_TRY(c, dalloc(&b->w, cap));
      HIP_TRY(c, dalloc(&b->type, cap)); HIP_TRY(c, dalloc(&b->id, cap));
    }
    if (vol && !b->vol) HIP_TRY(c, dalloc(&b->vol, cap));
  }
  if (n > c->pcap) {
    hipFree(c->key); hipFree(c->rank);
    HIP_TRY(c, dalloc(&c->key, cap));
    HIP_TRY(c, dalloc(&c->rank, cap));
    c->pcap = cap;
  }
  return PF_OK;
}

// counting sort of particles by block-major cell index + per-colour source lists
int ensure_binned(pf_ctx* c) {
  if (c->binned) return PF_OK;
  c->mark_order_valid = false;
  const long long ncell = c->dc.nblocks * kBlockVol;
  HIP_TRY(c, fill_u32(c->stream, c->cell_count, 0u, ncell));
  if (c->np > 0) {
    TIMED(c, PF_K_BIN_COUNT,
          hipLaunchKernelGGL(k_bin_count, dim3(grid_for(c->np, 256)), dim3(256), 0, c->stream, c->np, c->pa,
                             c->geom, c->dc, c->key, c->rank, c->cell_count));
    LAUNCH_CHECK(c);
  }
  HIP_TRY(c, scan_u32(*c, c->cell_count, c->cell_start, ncell));
  if (c->np > 0) {
    TIMED(c, PF_K_BIN_SCATTER,
          hipLaunchKernelGGL(k_bin_scatter, dim3(grid_for(c->np, 256)), dim3(256), 0, c->stream, c->np, c->pa,
                             c->pb, c->has_vol, c->key, c->rank, c->cell_start));
    LAUNCH_CHECK(c);
    std::swap(c->pa, c->pb);
  }
  HIP_TRY(c, fill_u32(c->stream, c->color_cnt, 0u, 8));
  hipLaunchKernelGGL(k_color_lists, dim3(grid_for(c->dc.nblocks, 256)), dim3(256), 0, c->stream, c->dc,
                     c->cell_start, c->color_cnt, c->color_list, c->color_cap);
  LAUNCH_CHECK(c);
  c->binned = true;
  return PF_OK;
}

int run_mark(pf_ctx* c) {
  int st = ensure_binned(c);
  if (st) return st;
  hipLaunchKernelGGL(k_mark, dim3(grid_for(c->dc.nblocks, 4)), dim3(256), 0, c->stream, c->dc, c->geom,
                     c->cell_start, c->marker, c->act[F_MK]);
  LAUNCH_CHECK(c);
  c->mark_order_valid = true;
  return PF_OK;
}

int run_face_ops(pf_ctx* c, int ops, float dtg) {
  const Dims* d[3] = {&c->du, &c->dv, &c->dw};
  for (int ax = 0; ax < 3; ++ax) {
    dim3 grid(grid_for(d[ax]->nblocks, 4));
    float* vel = c->f[F_U + ax];
    const float* mass = c->f[F_MU + ax];
    float* sv = c->f[F_SU + ax];
    uint8_t* av = c->act[F_U + ax];
    uint8_t* as = c->act[F_SU + ax];
    long long ki = kbegin(c, PF_K_FACE_OPS);
    if (ax == 0)
      hipLaunchKernelGGL(k_face_ops<0>, grid, dim3(256), 0, c->stream, *d[ax], c->geom, vel, mass, sv, av, as, ops, dtg);
    else if (ax == 1)
      hipLaunchKernelGGL(k_face_ops<1>, grid, dim3(256), 0, c->stream, *d[ax], c->geom, vel, mass, sv, av, as, ops, dtg);
    else
      hipLaunchKernelGGL(k_face_ops<2>, grid, dim3(256), 0, c->stream, *d[ax], c->geom, vel, mass, sv, av, as, ops, dtg);
    kend(c, ki);
    LAUNCH_CHECK(c);
  }
  return PF_OK;
}

// single: the single-phase trilinear unit-mass splat (ph and Vp unused)
// mr: the cubic splat over the enumerated multires faces
int run_p2g_splat(pf_ctx* c, const pf_phase* ph, double Vp, bool single = false, bool mr = false) {
  int st = ensure_binned(c);
  if (st) return st;
  const Dims* d[3] = {&c->du, &c->dv, &c->dw};
  for (int ax = 0; ax < 3; ++ax) {
    HIP_TRY(c, fill_f32(c->stream, c->f[F_U + ax], 0.f, d[ax]->nblocks * kBlockVol));
    HIP_TRY(c, fill_f32(c->stream, c->f[F_MU + ax], 0.f, d[ax]->nblocks * kBlockVol));
    HIP_TRY(c, fill_u8(c->stream, c->act[F_U + ax], 0, d[ax]->nblocks));
    HIP_TRY(c, fill_u8(c->stream, c->act[F_MU + ax], 0, d[ax]->nblocks));
  }
  P2GArgs a;
  a.g = c->geom;
  a.dc = c->dc;
  for (int ax = 0; ax < 3; ++ax) {
    a.df[ax] = *d[ax];
    a.mom[ax] = c->f[F_U + ax];
    a.mass[ax] = c->f[F_MU + ax];
    a.act_vel[ax] = c->act[F_U + ax];
    a.act_mass[ax] = c->act[F_MU + ax];
  }
  a.start = c->cell_start;
  a.P = c->pa;
  a.has_vol = c->has_vol;
  a.pow2 = is_pow2(c->geom.dx);
  a.inv_dx = 1.0 / c->geom.dx;
  a.fast32 = a.pow2 && c->geom.ox == 0.0 && c->geom.oy == 0.0 && c->geom.oz == 0.0 &&
             std::max({c->geom.nx, c->geom.ny, c->geom.nz}) < (1 << 22);
  a.inv_dx_f = (float)a.inv_dx;
  for (int ax = 0; ax < 3; ++ax) a.valid[ax] = mr ? c->mr_valid[ax] : nullptr;
  a.mp_l = single ? 1.f : (float)(ph->rho_l * Vp);
  a.mp_g = single ? 1.f : (float)(ph->rho_g * Vp);
  if (c->np == 0) return PF_OK;
  for (int col = 0; col < 8; ++col) {
    a.list = c->color_list + (long long)col * c->color_cap;
    a.count = c->color_cnt + col;
    const dim3 grid((unsigned)std::max<long long>(c->color_cap, 1));
    if (mr)
      TIMED(c, PF_K_P2G, hipLaunchKernelGGL(k_p2g_color<2>, grid, dim3(64), 0, c->stream, a));
    else if (single)
      TIMED(c, PF_K_P2G, hipLaunchKernelGGL(k_p2g_color<1>, grid, dim3(64), 0, c->stream, a));
    else
      TIMED(c, PF_K_P2G, hipLaunchKernelGGL(k_p2g_color<0>, grid, dim3(64), 0, c->stream, a));
    LAUNCH_CHECK(c);
  }
  return PF_OK;
}

// ---- learned preconditioner helpers

template <typename T>
void dfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

constexpr int kLpcWgradWGs = 512;

int lpc_grow(pf_lpc* L, long long nt, bool train) {
  if (nt > L->cap) {
    dfree(L->X); dfree(L->h1); dfree(L->h2); dfree(L->o); dfree(L->scale);
    HIP_TRY(L, dalloc(&L->X, nt * kHaloVol * 4));
    HIP_TRY(L, dalloc(&L->h1, nt * kTileVol * 16));
    HIP_TRY(L, dalloc(&L->h2, nt * kTileVol * 32));
    HIP_TRY(L, dalloc(&L->o, nt * kTileVol));
    HIP_TRY(L, dalloc(&L->scale, nt));
    L->cap = nt;
  }
  if (train && nt > L->cap_train) {
    dfree(L->h3); dfree(L->d3); dfree(L->d2); dfree(L->d1); dfree(L->go);
    HIP_TRY(L, dalloc(&L->h3, nt * kTileVol * 16));
    HIP_TRY(L, dalloc(&L->d3, nt * kTileVol * 16));
    HIP_TRY(L, dalloc(&L->d2, nt * kTileVol * 32));
    HIP_TRY(L, dalloc(&L->d1, nt * kTileVol * 16));
    HIP_TRY(L, dalloc(&L->go, nt * kTileVol));
    L->cap_train = nt;
  }
  const long long need = std::max<long long>((long long)kLpcWgradWGs * kNW2, 1024LL * 16);
  if (need > L->part_cap) {
    dfree(L->part);
    HIP_TRY(L, dalloc(&L->part, need));
    L->part_cap = need;
  }
  return PF_OK;
}

int lpc_prep(pf_lpc* L, hipStream_t s) {
  if (L->prepped) return PF_OK;
  hipLaunchKernelGGL(k_lpc_prep, dim3((32 * kpad(16) + 255) / 256 + 1), dim3(256), 0, s, L->w, L->bf);
  HIP_TRY(L, hipGetLastError());
  L->prepped = true;
  return PF_OK;
}

// o (and h1, h2, optionally h3) from L->X for nt tiles
int lpc_forward(pf_lpc* L, hipStream_t s, long long nt, bool store_h3, const int* done) {
  int st = lpc_prep(L, s);
  if (st) return st;
  const dim3 grid((unsigned)(nt * 4));
  ConvArgs a{};
  a.ntiles = nt;
  a.done = done;
  a.in = L->X; a.wt = L->bf.w1f; a.out = L->h1;
  hipLaunchKernelGGL((k_lpc_conv<4, 16, true, EPI_RELU>), grid, dim3(256), 0, s, a);
  a.in = L->h1; a.wt = L->bf.w2f; a.out = L->h2;
  hipLaunchKernelGGL((k_lpc_conv<16, 32, false, EPI_RELU>), grid, dim3(256), 0, s, a);
  a.in = L->h2; a.wt = L->bf.w3f; a.out = store_h3 ? L->h3 : nullptr; a.wh = L->bf.wh; a.o = L->o;
  hipLaunchKernelGGL((k_lpc_conv<32, 16, false, EPI_HEAD>), grid, dim3(256), 0, s, a);
  HIP_TRY(L, hipGetLastError());
  return PF_OK;
}

template <int CIN, int COUT, bool PADDED>
int lpc_wgrad(pf_lpc* L, hipStream_t s, const __bf16* in, const __bf16* g, long long nt, float* dst) {
  const int nwg = (int)std::min<long long>(kLpcWgradWGs, nt * 16);
  hipLaunchKernelGGL((k_lpc_wgrad<CIN, COUT, PADDED>), dim3(nwg), dim3(256), 0, s, in, g, nt, L->part);
  const int n = 27 * CIN * COUT;
  hipLaunchKernelGGL(k_lpc_sum_parts, dim3((n + 255) / 256), dim3(256), 0, s, L->part, nwg, n, dst);
  HIP_TRY(L, hipGetLastError());
  return PF_OK;
}

// L->grad = d(sum o * go)/dw, from the activations of the last lpc_forward(store_h3)
int lpc_backward(pf_lpc* L, hipStream_t s, long long nt) {
  const long long ncells = nt * kTileVol;
  const int hg = (int)std::min<long long>(1024, (ncells + 15) / 16);
  hipLaunchKernelGGL(k_lpc_head_bwd, dim3(hg), dim3(256), 0, s, L->h3, L->go, L->bf.wh, ncells, L->d3, L->part);
  hipLaunchKernelGGL(k_lpc_sum_parts, dim3(1), dim3(256), 0, s, L->part, hg, kNWH, L->grad + kOffWH);
  const dim3 grid((unsigned)(nt * 4));
  ConvArgs a{};
  a.ntiles = nt;
  a.in = L->d3; a.wt = L->bf.w3b; a.out = L->d2; a.mask = L->h2;
  hipLaunchKernelGGL((k_lpc_conv<16, 32, false, EPI_MASK>), grid, dim3(256), 0, s, a);
  a.in = L->d2; a.wt = L->bf.w2b; a.out = L->d1; a.mask = L->h1;
  hipLaunchKernelGGL((k_lpc_conv<32, 16, false, EPI_MASK>), grid, dim3(256), 0, s, a);
  HIP_TRY(L, hipGetLastError());
  int st = lpc_wgrad<32, 16, false>(L, s, L->h2, L->d3, nt, L->grad + kOffW3);
  if (!st) st = lpc_wgrad<16, 32, false>(L, s, L->h1, L->d2, nt, L->grad + kOffW2);
  if (!st) st = lpc_wgrad<4, 16, true>(L, s, L->X, L->d1, nt, L->grad + kOffW1);
  return st;
}

LpcTiles lpc_tiles_args(pf_ctx* c, const float* r) {
  LpcTiles T{};
  T.dc = c->dc; T.g = c->geom; T.ntx = c->ntx; T.nty = c->nty; T.ntz = c->ntz;
  T.tiles = c->tiles; T.ntiles = c->ntiles;
  T.marker = c->marker; T.r = r; T.invd = c->invd32; T.fbeta = c->fbeta; T.fphi = c->fphi;
  T.scale = c->lpc->scale; T.X = c->lpc->X;
  return T;
}

unsigned lpc_scatter_grid(long long nt) { return (unsigned)std::max<long long>(1, std::min<long long>(nt, 1024)); }

// z = M(r) on the context's stream for the last assembled system; partial r.z in
// part[slot] (gridDim = lpc_scatter_grid); skipped while *done
int lpc_apply(pf_ctx* c, const float* r, float* z, const int* done, double* part, int slot) {
  pf_lpc* L = c->lpc;
  int st = lpc_grow(L, c->ntiles, false);
  if (st) return fail(c, st, "learned preconditioner: " + L->err);
  LpcTiles T = lpc_tiles_args(c, r);
  T.done = done;
  hipLaunchKernelGGL(k_lpc_gather, dim3((unsigned)c->ntiles), dim3(256), 0, c->stream, T);
  st = lpc_forward(L, c->stream, c->ntiles, false, done);
  if (st) return fail(c, st, "learned preconditioner: " + L->err);
  hipLaunchKernelGGL(k_lpc_scatter, dim3(lpc_scatter_grid(c->ntiles)), dim3(256), 0, c->stream, T, L->o, z, part, slot);
  LAUNCH_CHECK(c);
  return PF_OK;
}

// 16^3 tiles holding the current fluid-block list (c->pblk, npblk blocks)
int lpc_build_tiles(pf_ctx* c) {
  c->ntx = (c->dc.nbx + 3) / 4; c->nty = (c->dc.nby + 3) / 4; c->ntz = (c->dc.nbz + 3) / 4;
  const long long nall = (long long)c->ntx * c->nty * c->ntz;
  if (nall > c->ntiles_all) {
    dfree(c->tiles); dfree(c->tile_flags);
    HIP_TRY(c, dalloc(&c->tiles, nall));
    HIP_TRY(c, dalloc(&c->tile_flags, nall));
    c->ntiles_all = nall;
  }
  HIP_TRY(c, fill_u8(c->stream, c->tile_flags, 0, nall));
  hipLaunchKernelGGL(k_lpc_tile_flags, dim3(grid_for(c->npblk, 256)), dim3(256), 0, c->stream, c->pblk, c->npblk, c->dc,
                     c->ntx, c->nty, c->tile_flags);
  LAUNCH_CHECK(c);
  HIP_TRY(c, compact_flags(*c, c->tile_flags, nall, c->tiles, c->u32_dev + 8));
  HIP_TRY(c, hipMemcpyAsync(c->u32_host + 8, c->u32_dev + 8, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  c->ntiles = c->u32_host[8];
  return PF_OK;
}

int lpc_features_alloc(pf_ctx* c) {
  if (c->fbeta) return PF_OK;
  const long long ncell = c->dc.nblocks * kBlockVol;
  HIP_TRY(c, dalloc(&c->fbeta, ncell));
  HIP_TRY(c, dalloc(&c->fphi, ncell));
  return PF_OK;
}

// ---- multires layout helpers (host)

MrGrid mr_grid(const pf_ctx* c) {
  MrGrid m;
  m.dc = c->dc; m.du = c->du; m.dv = c->dv; m.dw = c->dw; m.g = c->geom;
  m.lvl = c->mr_lvl; m.marker = c->marker;
  for (int a = 0; a < 3; ++a) m.valid[a] = c->mr_valid[a];
  return m;
}


using Leaf = std::array<int32_t, 4>;

// MRLayout3D::intersectsDomain (multires_layout3d.h:137-153)
bool mr_block_in_domain(int nx, int ny, int nz, const Leaf& k) {
  const int s = 4 << k[0];
  const int x0 = k[1] * s, y0 = k[2] * s, z0 = k[3] * s;
  return std::max(0, x0) < std::min(nx, x0 + s) && std::max(0, y0) < std::min(ny, y0 + s) &&
         std::max(0, z0) < std::min(nz, z0 + s);
}

// refineFineCellBox (multires_layout3d.h:62-97): leaves overlapping the box split one level
void mr_refine(std::vector<Leaf>& leaves, int nx, int ny, int nz, int x0, int y0, int z0, int x1, int y1, int z1) {
  std::vector<Leaf> next;
  next.reserve(leaves.size() * 2);
  for (const Leaf& k : leaves) {
    const int s = 4 << k[0];
    const int bx0 = k[1] * s, by0 = k[2] * s, bz0 = k[3] * s;
    const int bx1 = std::min(nx, bx0 + s), by1 = std::min(ny, by0 + s), bz1 = std::min(nz, bz0 + s);
    const bool overlaps = bx0 < x1 && bx1 > x0 && by0 < y1 && by1 > y0 && bz0 < z1 && bz1 > z0;
    if (!overlaps || k[0] == 0) {
      next.push_back(k);
      continue;
    }
    for (int cz = 0; cz < 2; ++cz)
      for (int cy = 0; cy < 2; ++cy)
        for (int cx = 0; cx < 2; ++cx) {
          const Leaf ch = {k[0] - 1, k[1] * 2 + cx, k[2] * 2 + cy, k[3] * 2 + cz};
          if (mr_block_in_domain(nx, ny, nz, ch)) next.push_back(ch);
        }
  }
  std::sort(next.begin(), next.end());
  next.erase(std::unique(next.begin(), next.end()), next.end());
  leaves.swap(next);
}

// layoutWithFineBox (multires_sim3d_tp.cpp:172-180) on a `level` base
std::vector<Leaf> mr_layout_box(int nx, int ny, int nz, int level, bool valid, const int32_t* box) {
  std::vector<Leaf> lv;
  const int s = 4 << level;
  for (int bz = 0; bz < (nz + s - 1) / s; ++bz)
    for (int by = 0; by < (ny + s - 1) / s; ++by)
      for (int bx = 0; bx < (nx + s - 1) / s; ++bx) lv.push_back({level, bx, by, bz});
  std::sort(lv.begin(), lv.end());
  if (valid) {
    const int x0 = std::clamp(box[0], 0, nx), y0 = std::clamp(box[1], 0, ny), z0 = std::clamp(box[2], 0, nz);
    const int x1 = std::clamp(box[3], 0, nx), y1 = std::clamp(box[4], 0, ny), z1 = std::clamp(box[5], 0, nz);
    if (x0 < x1 && y0 < y1 && z0 < z1) mr_refine(lv, nx, ny, nz, x0, y0, z0, x1, y1, z1);
  }
  return lv;
}

long long mr_fine_count(const std::vector<Leaf>& lv) {
  long long n = 0;
  for (const Leaf& k : lv) n += k[0] == 0;
  return n;
}

int mr_apply_layout(pf_ctx* c, std::vector<Leaf> lv) {
  std::sort(lv.begin(), lv.end());
  lv.erase(std::unique(lv.begin(), lv.end()), lv.end());
  const Geom& g = c->geom;
  // the leaves must tile the domain: every fine cell in exactly one leaf
  long long covered = 0;
  for (const Leaf& k : lv) {
    if (k[0] < 0 || k[0] > 6 || k[1] < 0 || k[2] < 0 || k[3] < 0 || !mr_block_in_domain(g.nx, g.ny, g.nz, k))
      return fail(c, PF_ERR_ARG, "multires leaf outside the domain or with a level outside 0..6");
    const int s = 4 << k[0];
    covered += (long long)(std::min(g.nx, (k[1] + 1) * s) - k[1] * s) * (std::min(g.ny, (k[2] + 1) * s) - k[2] * s) *
               (std::min(g.nz, (k[3] + 1) * s) - k[3] * s);
  }
  if (covered != (long long)g.nx * g.ny * g.nz) return fail(c, PF_ERR_ARG, "multires leaves must tile the domain once");
  const long long ncell = c->dc.nblocks * kBlockVol;
  const Dims* fd[3] = {&c->du, &c->dv, &c->dw};
  if (!c->mr_lvl) {
    HIP_TRY(c, dalloc(&c->mr_lvl, ncell));
    for (int a = 0; a < 3; ++a) {
      HIP_TRY(c, dalloc(&c->mr_valid[a], fd[a]->nblocks * kBlockVol));
      HIP_TRY(c, dalloc(&c->mr_cf32[a], fd[a]->nblocks * kBlockVol));
      HIP_TRY(c, dalloc(&c->mr_cf64[a], fd[a]->nblocks * kBlockVol));
    }
    HIP_TRY(c, dalloc(&c->mr_vol64, ncell));
    HIP_TRY(c, dalloc(&c->mr_pfull, ncell));
  }
  if ((long long)lv.size() > c->mr_leaves_cap) {
    dfree(c->mr_leaves_dev);
    HIP_TRY(c, dalloc(&c->mr_leaves_dev, 4 * (long long)lv.size()));
    c->mr_leaves_cap = (long long)lv.size();
  }
  c->mr_leaves.assign(4 * lv.size(), 0);
  for (size_t a = 0; a < lv.size(); ++a)
    for (int q = 0; q < 4; ++q) c->mr_leaves[4 * a + q] = lv[a][q];
  HIP_TRY(c, hipMemcpyAsync(c->mr_leaves_dev, c->mr_leaves.data(), sizeof(int32_t) * c->mr_leaves.size(),
                            hipMemcpyHostToDevice, c->stream));
  HIP_TRY(c, fill_u8(c->stream, c->mr_lvl, 255, ncell));
  hipLaunchKernelGGL(k_mr_levels, dim3((unsigned)lv.size()), dim3(256), 0, c->stream, c->mr_leaves_dev, (int)lv.size(),
                     c->dc, c->geom, c->mr_lvl);
  LAUNCH_CHECK(c);
  c->mr = true;
  for (int a = 0; a < 3; ++a) {
    hipLaunchKernelGGL(k_mr_valid, dim3(grid_for(fd[a]->nblocks * kBlockVol, 256)), dim3(256), 0, c->stream, mr_grid(c),
                       a, c->mr_valid[a]);
    LAUNCH_CHECK(c);
  }
  // MRMacGrid3D<4>(layout): every field starts empty
  for (int f = 0; f < F_N; ++f) {
    if (f == F_MK) continue;
    const Dims& d = f == F_P ? c->dc : *fd[(f == F_SU || f == F_SV || f == F_SW) ? f - F_SU : f % 3];
    HIP_TRY(c, fill_f32(c->stream, c->f[f], 0.f, d.nblocks * kBlockVol));
    HIP_TRY(c, fill_u8(c->stream, c->act[f], 0, d.nblocks));
  }
  HIP_TRY(c, fill_u8(c->stream, c->marker, 0, ncell));
  HIP_TRY(c, fill_f64(c->stream, c->mr_pfull, 0.0, ncell));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return PF_OK;
}

// updateDynamicRefinement (multires_sim3d_tp.cpp:577-669): particle and gas cell
// boxes (padded), hysteresis, fine-leaf budget; the layout changes only when its
// leaves differ
int mr_dynamic(pf_ctx* c, const pf_mr_refine_cfg* rc) {
  if (!rc->dynamic_refinement) return PF_OK;
  const Geom& g = c->geom;
  const int lim[3] = {g.nx, g.ny, g.nz};
  // the box buffers belong to refinement, which may run before any layout exists
  if (!c->mr_bbox) HIP_TRY(c, dalloc(&c->mr_bbox, 16));
  if (!c->mr_bbox_host) HIP_TRY(c, hipHostMalloc((void**)&c->mr_bbox_host, sizeof(int) * 16));
  if (c->np > 0 && !c->pa.x) return fail(c, PF_ERR_STATE, "particle buffers missing");
  HIP_TRY(c, fill_u32(c->stream, reinterpret_cast<uint32_t*>(c->mr_bbox), 0x7fffffffu, 12));
  if (c->np > 0)
    hipLaunchKernelGGL(k_mr_bbox, dim3(std::min(grid_for(c->np, 256), 1024u)), dim3(256), 0, c->stream, c->np, c->pa,
                       c->geom, c->mr_bbox);
  LAUNCH_CHECK(c);
  HIP_TRY(c, hipMemcpyAsync(c->mr_bbox_host, c->mr_bbox, sizeof(int) * 12, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  const int* bb = c->mr_bbox_host;
  bool dv = false;
  int d[6] = {0, 0, 0, 0, 0, 0};
  auto include = [&](int x0, int y0, int z0, int x1, int y1, int z1) {
    if (x0 >= x1 || y0 >= y1 || z0 >= z1) return;
    const int b[6] = {x0, y0, z0, x1, y1, z1};
    if (!dv) {
      std::copy(b, b + 6, d);
      dv = true;
      return;
    }
    for (int a = 0; a < 3; ++a) { d[a] = std::min(d[a], b[a]); d[3 + a] = std::max(d[3 + a], b[3 + a]); }
  };
  if (bb[0] != 0x7fffffff) {
    const int p = std::max(0, rc->particle_padding);
    include(bb[0] - p, bb[1] - p, bb[2] - p, -bb[3] + p + 1, -bb[4] + p + 1, -bb[5] + p + 1);
  }
  if (bb[6] != 0x7fffffff) {
    const int p = std::max(0, rc->gas_padding);
    include(bb[6] - p, bb[7] - p, bb[8] - p, -bb[9] + p + 1, -bb[10] + p + 1, -bb[11] + p + 1);
  }
  auto clampbox = [&](int* b, bool& valid) {
    if (!valid) return;
    for (int a = 0; a < 3; ++a) { b[a] = std::clamp(b[a], 0, lim[a]); b[3 + a] = std::clamp(b[3 + a], 0, lim[a]); }
    valid = b[0] < b[3] && b[1] < b[4] && b[2] < b[5];
  };
  clampbox(d, dv);
  int t[6] = {0, 0, 0, 0, 0, 0};
  bool tv = false;
  c->mr_budget_limited = 0;
  if (dv) {
    const int* r = c->mr_retained;
    const bool contains = c->mr_retained_valid && r[0] <= d[0] && r[1] <= d[1] && r[2] <= d[2] && r[3] >= d[3] &&
                          r[4] >= d[4] && r[5] >= d[5];
    if (contains) {
      std::copy(r, r + 6, t);
      tv = true;
    } else {
      const int pad = std::max(0, rc->hysteresis_cells);
      for (int a = 0; a < 3; ++a) { t[a] = d[a] - pad; t[3 + a] = d[3 + a] + pad; }
      tv = true;
      clampbox(t, tv);
    }
    // fitBoxToFineLeafBudget (:194-210): shrink the longest axis away from the centre
    if (tv && rc->max_fine_leaves > 0 &&
        mr_fine_count(mr_layout_box(g.nx, g.ny, g.nz, 1, true, t)) > rc->max_fine_leaves) {
      c->mr_budget_limited = 1;
      const double cc[3] = {0.5 * (t[0] + t[3]), 0.5 * (t[1] + t[4]), 0.5 * (t[2] + t[5])};
      while (mr_fine_count(mr_layout_box(g.nx, g.ny, g.nz, 1, true, t)) > rc->max_fine_leaves) {
        int ax = 0;
        if (t[4] - t[1] > t[3] - t[0]) ax = 1;
        if (t[5] - t[2] > t[3 + ax] - t[ax]) ax = 2;
        if (t[3 + ax] - t[ax] <= 1) break;
        if ((t[3 + ax] - cc[ax]) >= (cc[ax] - t[ax])) --t[3 + ax];
        else ++t[ax];
        tv = t[0] < t[3] && t[1] < t[4] && t[2] < t[5];
        if (!tv) break;
      }
    }
  }
  c->mr_retained_valid = tv;
  if (tv) std::copy(t, t + 6, c->mr_retained);
  std::vector<Leaf> next = mr_layout_box(g.nx, g.ny, g.nz, 1, tv, t);
  c->mr_last_fine_leaves = (int)mr_fine_count(next);
  bool same = c->mr && next.size() * 4 == c->mr_leaves.size();
  for (size_t a = 0; same && a < next.size(); ++a)
    for (int q = 0; q < 4; ++q) same = same && next[a][q] == c->mr_leaves[4 * a + q];
  if (!same) return mr_apply_layout(c, next);
  return PF_OK;
}

int read_small(pf_ctx* c) {
  HIP_TRY(c, hipMemcpyAsync(c->red_host, c->red, sizeof(double) * 16, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipMemcpyAsync(c->u32_host, c->u32_dev, sizeof(uint32_t) * 16, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return PF_OK;
}

void record(pf_ctx* c, bool timing, int slot) {
  if (timing) hipEventRecord(c->ev[slot], c->stream);
}

// ---------------------------------------------------------------- reference-order multires solve

// validateSolveConfig, multires_pressure3d.cpp:307-333
bool mr_cfg_valid(const pf_mr_solve_cfg& c) {
  return !(c.max_iterations < 0 || c.absolute_tolerance < 0.0 || c.relative_tolerance < 0.0 ||
           c.restart_growth_threshold < 0.0 || c.relaxation_sweeps < 0 || c.relaxation_omega < 0.0 ||
           c.relaxation_min_omega < 0.0 || c.coarse_correction_iterations < 0 || c.coarse_correction_sweeps < 0 ||
           c.coarse_correction_absolute_tolerance < 0.0 || c.coarse_correction_relative_tolerance < 0.0 ||
           c.coarse_correction_min_scale <= 0.0 || c.coarse_correction_min_scale > 1.0 ||
           c.coarse_preconditioner_iterations < 0 || c.coarse_preconditioner_absolute_tolerance < 0.0 ||
           c.coarse_preconditioner_relative_tolerance < 0.0 || c.coarse_preconditioner_scale < 0.0 ||
           c.coarse_preconditioner_min_rz_gain < 0.0 || c.coarse_preconditioner_max_work_ratio < 0.0 ||
           c.coarse_preconditioner_auto_disable_after < 0);
}

constexpr int kMrfVectors = 15;

// projectMR3D's iteration (multires_pressure3d.cpp:846-1449) on FP64 device vectors:
// the coarse correction, the damped-Jacobi relaxation, then PCG with adaptive
// restart, (flexible) beta with resets and the additive coarse preconditioner.  x is
// c->x64 (zero on entry); the rows, the pin and the reference rhs come from the
// assembly.  Fills c->mr_detail and the iteration outcome.
int mr_reference_solve(pf_ctx* c, const pf_mr_solve_cfg& cf, unsigned G, const VecArgs& va, const MrOp64& op,
                       long long nfluid, int& iters, bool& conv, bool& brk) {
  pf_mr_solve_detail& D = c->mr_detail;
  const long long ncell = c->dc.nblocks * kBlockVol;
  double* F[kMrfVectors];
  for (int q = 0; q < kMrfVectors; ++q) F[q] = c->mrf + (long long)q * ncell;
  double *r = F[0], *z = F[1], *zp = F[2], *p = F[3], *Ap = F[4], *cand_x = F[5], *cand_r = F[6];
  double *best_x = F[7], *best_r = F[8], *fp = F[9], *fap = F[10], *corr = F[11];
  const double *rhs = F[12], *dref = F[13], *vtrue = F[14];
  double* x = c->x64;
  MrfArgs A;
  A.va = va;
  A.op = op;
  A.dref = dref;
  A.vtrue = vtrue;
  A.pin = *c->mr_pin_host;
  const long long N = nfluid;

  auto reduce = [&](unsigned g, int nslots, unsigned mask) -> int {
    hipLaunchKernelGGL(k_reduce_slots, dim3(1), dim3(256), 0, c->stream, c->partial, (int)g, nslots, mask, c->red);
    LAUNCH_CHECK(c);
    return read_small(c);
  };
  // one vector operation; its reductions land in red_host[0..1]
  auto vec = [&](int opk, double s0, const double* a, const double* b, const double* cc, double* out, double* out2,
                 bool want) -> int {
    MrfVec v{opk, cf.use_jacobi_preconditioner ? 1 : 0, s0, a, b, cc, out, out2};
    hipLaunchKernelGGL(k_mrf_vec, dim3(G), dim3(256), 0, c->stream, A, v);
    LAUNCH_CHECK(c);
    const bool maxslot = opk == MRF_MAXABS || opk == MRF_XR;
    return want ? reduce(G, 2, maxslot ? 1u : 0u) : PF_OK;
  };
  auto apply = [&](const double* in, double* out) -> int {
    hipLaunchKernelGGL(k_mrf_apply<false>, dim3(G), dim3(256), 0, c->stream, A, in, out, nullptr, nullptr);
    LAUNCH_CHECK(c);
    return PF_OK;
  };
  // residualFromPressure (:870-884): out = rhs - A xin, max |out| (+inf if not finite)
  auto residual_from = [&](const double* xin, double* out, double& res) -> int {
    hipLaunchKernelGGL(k_mrf_apply<true>, dim3(G), dim3(256), 0, c->stream, A, xin, Ap, rhs, out);
    LAUNCH_CHECK(c);
    int st = reduce(G, 1, 1u);
    res = c->red_host[0];
    return st;
  };
  auto copy = [&](double* dst, const double* src) -> int {
    HIP_TRY(c, hipMemcpyAsync(dst, src, sizeof(double) * (size_t)ncell, hipMemcpyDeviceToDevice, c->stream));
    return PF_OK;
  };
  auto max_abs_r = [&](double& out) -> int {
    int s2 = vec(MRF_MAXABS, 0.0, r, nullptr, nullptr, nullptr, nullptr, true);
    out = c->red_host[0];
    return s2;
  };

  // ---- the level-1 coarse space (:39-70, 470-552) on a dense (n/2)^3 grid
  const Geom& g = c->geom;
  CoarseGrid cg;
  cg.hx = (g.nx + 1) / 2;
  cg.hy = (g.ny + 1) / 2;
  cg.hz = (g.nz + 1) / 2;
  cg.nc = (long long)cg.hx * cg.hy * cg.hz;
  cg.dc = c->dc;
  cg.g = g;
  const unsigned Gc = (unsigned)std::min<long long>(kMaxPartials, std::max<long long>(1, (cg.nc + 255) / 256));
  bool coarse_built = false;
  long long n_agg = 0, cpin = -1;
  double *Vc = nullptr, *crhs = nullptr, *cx = nullptr, *cr = nullptr, *cp = nullptr, *cap = nullptr;
  auto build_coarse = [&]() -> int {
    if (coarse_built) return PF_OK;
    if (cg.nc * 6 > c->mrc_cap) {
      dfree(c->mrc);
      c->mrc = nullptr;
      c->mrc_cap = 0;
      HIP_TRY(c, dalloc(&c->mrc, cg.nc * 6));
      c->mrc_cap = cg.nc * 6;
    }
    Vc = c->mrc;
    crhs = Vc + cg.nc;
    cx = crhs + cg.nc;
    cr = cx + cg.nc;
    cp = cr + cg.nc;
    cap = cp + cg.nc;
    hipLaunchKernelGGL(k_mrf_coarse_volume, dim3(Gc), dim3(256), 0, c->stream, cg, op.vol, vtrue, Vc, c->partial);
    LAUNCH_CHECK(c);
    int s2 = reduce(Gc, 1, 0u);
    if (s2) return s2;
    n_agg = (long long)c->red_host[0];
    cpin = -1;
    if (A.pin >= 0) {
      const long long b = A.pin >> 6;
      const int lane = (int)(A.pin & 63);
      int bx, by, bz;
      block_coords(c->dc, b, bx, by, bz);
      const int i = bx * 4 + (lane & 3), j = by * 4 + ((lane >> 2) & 3), k = bz * 4 + (lane >> 4);
      cpin = (long long)(i >> 1) + (long long)cg.hx * ((long long)(j >> 1) + (long long)cg.hy * (long long)(k >> 1));
    }
    coarse_built = true;
    return PF_OK;
  };
  auto coarse_op = [&](int opk, double s0, const double* a, const double* b, bool want) -> int {
    MrcVec v{opk, cpin, s0, Vc, a, b, cx, cr, cp};
    if (opk == MRC_PINSET) v.x = cap;
    hipLaunchKernelGGL(k_mrf_coarse, dim3(Gc), dim3(256), 0, c->stream, cg, v, c->partial);
    LAUNCH_CHECK(c);
    return want ? reduce(Gc, 1, 0u) : PF_OK;
  };
  auto restrict_to = [&](const double* fine, double* coarse) -> int {
    hipLaunchKernelGGL(k_mrf_restrict, dim3(Gc), dim3(256), 0, c->stream, cg, op.vol, vtrue, Vc, fine, coarse);
    LAUNCH_CHECK(c);
    return PF_OK;
  };
  auto prolong_to = [&](const double* coarse, double* fine) -> int {
    hipLaunchKernelGGL(k_mrf_prolong, dim3(G), dim3(256), 0, c->stream, A, cg, coarse, fine);
    LAUNCH_CHECK(c);
    return PF_OK;
  };
  // the pinned coarse CG on R A P (:946-1015 and :1214-1281), right-hand side in crhs
  auto coarse_cg = [&](int max_it, double abs_t, double rel_t, double& eff, int& its, bool& cconv,
                       bool& cbd) -> int {
    int s2 = coarse_op(MRC_INIT, 0.0, crhs, nullptr, true);
    if (s2) return s2;
    double rr = c->red_host[0];
    const double res0 = rr > 0.0 ? std::sqrt(rr) : 0.0;
    eff = std::max(abs_t, rel_t * res0);
    cconv = res0 <= eff;
    cbd = false;
    its = 0;
    if (!std::isfinite(rr)) {
      cbd = true;
      return PF_OK;
    }
    for (int it = 0; !cconv && !cbd && it < max_it; ++it) {
      if ((s2 = prolong_to(cp, fp)) || (s2 = apply(fp, fap)) || (s2 = restrict_to(fap, cap)) ||
          (s2 = coarse_op(MRC_PINSET, 0.0, cp, nullptr, false)))
        return s2;
      if ((s2 = coarse_op(MRC_DOT, 0.0, cp, cap, true))) return s2;
      const double pAp = c->red_host[0];
      if (!std::isfinite(pAp) || std::abs(pAp) < 1e-30) { cbd = true; break; }
      const double alpha = rr / pAp;
      if (!std::isfinite(alpha)) { cbd = true; break; }
      if ((s2 = coarse_op(MRC_XR, alpha, cap, nullptr, true))) return s2;
      const double rrn = c->red_host[0];
      its = it + 1;
      const double res = rrn > 0.0 ? std::sqrt(rrn) : 0.0;
      if (!std::isfinite(res)) { cbd = true; break; }
      if (res <= eff) { cconv = true; break; }
      if (!std::isfinite(rrn) || std::abs(rr) < 1e-30) { cbd = true; break; }
      const double beta = rrn / rr;
      if (!std::isfinite(beta)) { cbd = true; break; }
      rr = rrn;
      if ((s2 = coarse_op(MRC_P, beta, nullptr, nullptr, false))) return s2;
    }
    return PF_OK;
  };

  int st = PF_OK;
  iters = 0;
  conv = false;
  brk = false;
  // rhs and the initial residual (:846-863); x = 0
  if ((st = copy(r, rhs))) return st;
  double res0 = 0.0;
  if ((st = max_abs_r(res0))) return st;
  D.initial_residual = res0;
  D.final_residual = res0;
  const double tol = std::max(std::max(0.0, cf.absolute_tolerance), std::max(0.0, cf.relative_tolerance) * res0);
  D.effective_tolerance = tol;
  conv = res0 < tol;
  double cur = res0;
  if (res0 >= tol) {
    // the coarse correction sweeps with a halving scale search (:886-1060)
    if (cf.use_coarse_correction && cur >= tol) {
      if ((st = build_coarse())) return st;
      D.coarse_correction_cells = (int32_t)n_agg;
      D.coarse_correction_initial_residual = cur;
      D.coarse_correction_final_residual = cur;
      if (n_agg > 0 && n_agg < N) {
        for (int sw = 0; sw < cf.coarse_correction_sweeps && cur >= tol && !D.coarse_correction_breakdown; ++sw) {
          D.coarse_correction_sweeps++;
          if ((st = restrict_to(r, crhs))) return st;
          double eff = 0.0;
          int its = 0;
          bool cc = false, cb = false;
          if ((st = coarse_cg(cf.coarse_correction_iterations, cf.coarse_correction_absolute_tolerance,
                              cf.coarse_correction_relative_tolerance, eff, its, cc, cb)))
            return st;
          D.coarse_correction_effective_tolerance = eff;
          D.coarse_correction_converged = cc ? 1 : 0;
          D.coarse_correction_iterations += its;
          if (cb) {
            D.coarse_correction_breakdown = 1;
            break;
          }
          if ((st = prolong_to(cx, corr))) return st;
          double best = cur, bscale = 0.0;
          bool have = false;
          for (double scale = 1.0; scale >= cf.coarse_correction_min_scale; scale *= 0.5) {
            if ((st = vec(MRF_AXPY, scale, x, corr, nullptr, cand_x, nullptr, false))) return st;
            double cres = 0.0;
            if ((st = residual_from(cand_x, cand_r, cres))) return st;
            if (std::isfinite(cres) && cres <= best) {
              best = cres;
              bscale = scale;
              std::swap(cand_x, best_x);
              std::swap(cand_r, best_r);
              have = true;
            }
          }
          D.coarse_correction_final_residual = best;
          if (best < cur && have) {
            if ((st = copy(x, best_x)) || (st = copy(r, best_r))) return st;
            cur = best;
            D.final_residual = cur;
            D.coarse_correction_accepted = 1;
            D.coarse_correction_last_scale = bscale;
            D.coarse_correction_accepted_sweeps++;
          } else {
            D.coarse_correction_rejected_sweeps++;
            break;
          }
        }
      }
    }
    // damped-Jacobi relaxation with omega backoff (:1062-1110)
    double om = std::max(0.0, cf.relaxation_omega), omin = std::max(0.0, cf.relaxation_min_omega);
    if (omin > om) omin = om;
    D.relaxation_final_omega = om;
    for (int sw = 0; sw < cf.relaxation_sweeps && cur >= tol; ++sw) {
      if ((st = vec(MRF_JACOBI, 0.0, r, nullptr, nullptr, z, nullptr, false))) return st;
      double trial = om;
      bool accepted = false;
      while (trial >= omin && trial > 0.0) {
        if ((st = vec(MRF_AXPY, trial, x, z, nullptr, cand_x, nullptr, false))) return st;
        double cres = 0.0;
        if ((st = residual_from(cand_x, cand_r, cres))) return st;
        if (std::isfinite(cres) && cres <= cur) {
          if ((st = copy(x, cand_x)) || (st = copy(r, cand_r))) return st;
          cur = cres;
          D.final_residual = cur;
          om = trial;
          D.relaxation_final_omega = om;
          D.relaxation_accepted++;
          accepted = true;
          break;
        }
        D.relaxation_rejected++;
        if (trial <= omin) break;
        trial = std::max(omin, trial * 0.5);
      }
      if (!accepted) break;
    }
    conv = cur < tol;
    if (conv) D.final_residual = cur;
  }
  if (conv || res0 < tol) return PF_OK;

  // PCG with the additive coarse preconditioner (:1112-1449)
  const bool use_cp = cf.use_coarse_preconditioner && cf.coarse_preconditioner_iterations > 0 &&
                      cf.coarse_preconditioner_scale > 0.0;
  bool cp_ready = false, cp_disabled = false;
  int cp_streak = 0;
  if (use_cp) {
    if ((st = build_coarse())) return st;
    D.coarse_preconditioner_cells = (int32_t)n_agg;
    cp_ready = n_agg > 0 && n_agg < N;
  }
  // applyPreconditioner (:1283-1340): Jacobi, then z + scale P x_c when a coarse
  // solve ran and r.z grew by at least min_rz_gain
  auto precondition = [&]() -> int {
    int s2 = vec(MRF_JACOBI, 0.0, r, nullptr, nullptr, z, nullptr, use_cp);
    if (s2 || !use_cp) return s2;
    const double rzb = c->red_host[0];
    const bool base_ok = std::isfinite(rzb) && rzb > 0.0;
    if (cp_disabled) return PF_OK;
    const int skipped_before = D.coarse_preconditioner_skipped_applications;
    bool accepted = false, wasted = false, solved = false;
    if (cp_ready) {
      const int max_c = cf.coarse_preconditioner_iterations;
      bool budget_ok = true;
      if (cf.coarse_preconditioner_max_work_ratio > 0.0) {
        const double allowed = cf.coarse_preconditioner_max_work_ratio * (double)std::max(1, iters + 1);
        const int remaining = (int)std::floor(allowed + 1e-12) - D.coarse_preconditioner_iterations;
        if (remaining < max_c) {
          D.coarse_preconditioner_budget_limited_applications++;
          D.coarse_preconditioner_skipped_applications++;
          D.coarse_preconditioner_budget_exhausted = 1;
          budget_ok = false;
        }
      }
      if (budget_ok) {
        D.coarse_preconditioner_applications++;
        if ((s2 = restrict_to(r, crhs))) return s2;
        double eff = 0.0;
        int its = 0;
        bool cc = false, cb = false;
        if ((s2 = coarse_cg(max_c, cf.coarse_preconditioner_absolute_tolerance,
                            cf.coarse_preconditioner_relative_tolerance, eff, its, cc, cb)))
          return s2;
        D.coarse_preconditioner_effective_tolerance = eff;
        if (cb) {
          D.coarse_preconditioner_breakdown = 1;
        } else {
          D.coarse_preconditioner_iterations += its;
          if ((s2 = prolong_to(cx, corr))) return s2;
          solved = true;
        }
      }
    }
    if (solved) {
      if ((s2 = vec(MRF_AXPY, cf.coarse_preconditioner_scale, z, corr, nullptr, cand_x, nullptr, false))) return s2;
      if ((s2 = vec(MRF_DOT, 0.0, r, cand_x, nullptr, nullptr, nullptr, true))) return s2;
      const double rzc = c->red_host[0];
      if (std::isfinite(rzc) && base_ok) D.coarse_preconditioner_last_rz_gain = rzc / rzb - 1.0;
      const double min_acc = base_ok ? rzb * (1.0 + cf.coarse_preconditioner_min_rz_gain) : 0.0;
      if (std::isfinite(rzc) && rzc > 0.0 && (!base_ok || rzc >= min_acc)) {
        if ((s2 = copy(z, cand_x))) return s2;
        D.coarse_preconditioner_accepted_applications++;
        accepted = true;
      } else {
        D.coarse_preconditioner_rejected_applications++;
        wasted = true;
      }
    } else if (D.coarse_preconditioner_skipped_applications > skipped_before) {
      wasted = true;
    }
    // recordCoarsePreconditionerOutcome (:1157-1178)
    if (accepted) {
      cp_streak = 0;
      return PF_OK;
    }
    if (!wasted) return PF_OK;
    ++cp_streak;
    if (cf.coarse_preconditioner_auto_disable && cf.coarse_preconditioner_auto_disable_after > 0 &&
        cp_streak >= cf.coarse_preconditioner_auto_disable_after) {
      cp_disabled = true;
      D.coarse_preconditioner_auto_disabled = 1;
      if (D.coarse_preconditioner_auto_disable_iteration < 0) D.coarse_preconditioner_auto_disable_iteration = iters;
    }
    return PF_OK;
  };
  auto dot = [&](const double* a, const double* b, double& out) -> int {
    int s2 = vec(MRF_DOT, 0.0, a, b, nullptr, nullptr, nullptr, true);
    out = c->red_host[0];
    return s2;
  };
  auto stop_on_breakdown = [&]() -> int {
    brk = true;
    return max_abs_r(D.final_residual);
  };

  if ((st = precondition()) || (st = copy(p, z))) return st;
  double rz = 0.0;
  if ((st = dot(r, z, rz))) return st;
  if (!std::isfinite(rz) && (st = stop_on_breakdown())) return st;
  double prev = D.final_residual;
  for (int it = 0; !brk && it < cf.max_iterations; ++it) {
    if ((st = apply(p, Ap))) return st;
    double pAp = 0.0;
    if ((st = dot(p, Ap, pAp))) return st;
    if (!std::isfinite(pAp) || std::abs(pAp) < 1e-30) {
      if ((st = stop_on_breakdown())) return st;
      break;
    }
    const double alpha = rz / pAp;
    if (!std::isfinite(alpha)) {
      if ((st = stop_on_breakdown())) return st;
      break;
    }
    if ((st = vec(MRF_XR, alpha, p, Ap, nullptr, x, r, true))) return st;
    const double res = c->red_host[0];
    const bool finite_state = c->red_host[1] == 0.0;
    iters = it + 1;
    D.final_residual = res;
    if (!finite_state || !std::isfinite(res)) {
      if ((st = stop_on_breakdown())) return st;
      break;
    }
    if (res < tol) {
      conv = true;
      break;
    }
    if (cf.adaptive_restart && cf.restart_growth_threshold > 0.0 && res > prev * cf.restart_growth_threshold &&
        res > res0 * cf.restart_growth_threshold && it + 1 < cf.max_iterations) {
      if ((st = precondition()) || (st = dot(r, z, rz))) return st;
      if (!std::isfinite(rz)) {
        if ((st = stop_on_breakdown())) return st;
        break;
      }
      if ((st = copy(p, z))) return st;
      D.restarts++;
      prev = res;
      continue;
    }
    prev = res;
    if ((st = copy(zp, z)) || (st = precondition())) return st;
    double rzn = 0.0;
    if ((st = dot(r, z, rzn))) return st;
    if (!std::isfinite(rzn)) {
      if ((st = stop_on_breakdown())) return st;
      break;
    }
    double beta = 0.0;
    if (rz != 0.0) {
      if (cf.use_flexible_cg_beta) {
        if ((st = vec(MRF_FLEX, 0.0, r, z, zp, nullptr, nullptr, true))) return st;
        beta = c->red_host[0] / rz;
      } else {
        beta = rzn / rz;
      }
    }
    if (!std::isfinite(beta)) {
      beta = 0.0;
      D.beta_resets++;
    }
    if (beta < 0.0) {
      beta = 0.0;
      D.beta_resets++;
    }
    rz = rzn;
    if ((st = vec(MRF_P, beta, z, nullptr, nullptr, p, nullptr, false))) return st;
  }
  return PF_OK;
}

// single: the single-phase unit-density projection (ph unused, cdiv ignored);
// c->mr: the multires finite-volume projection, stopping when
// |rhs - A x|_inf < max(cfg->tol, mr_rel_tol * |rhs|_inf)
// ---- aggregation multigrid preconditioner (pf_mg.hip)

bool mg_cfg_valid(const pf_mg_cfg& q) {
  return q.sweeps >= 1 && q.sweeps <= 16 && q.coarsest_sweeps >= 1 && q.coarsest_sweeps <= 1024 &&
         q.omega > 0.0f && q.omega < 1.0f && q.correction > 0.0f && q.correction < 2.0f &&
         (q.chebyshev_ratio == 0.0f || (q.chebyshev_ratio > 1.0f && q.chebyshev_ratio <= 1000.0f));
}

// the built-in cycle, with the PF_MG_* environment overrides (read once)
const pf_mg_cfg& mg_default_cfg() {
  static const pf_mg_cfg p = [] {
    pf_mg_cfg q{2, 32, 0.9f, 1.8f, 6.0f};
    pf_mg_cfg e = q;
    if (const char* v = std::getenv("PF_MG_SWEEPS")) e.sweeps = std::atoi(v);
    if (const char* v = std::getenv("PF_MG_COARSEST_SWEEPS")) e.coarsest_sweeps = std::atoi(v);
    if (const char* v = std::getenv("PF_MG_OMEGA")) e.omega = (float)std::atof(v);
    if (const char* v = std::getenv("PF_MG_CORRECTION")) e.correction = (float)std::atof(v);
    if (const char* v = std::getenv("PF_MG_CHEB")) e.chebyshev_ratio = (float)std::atof(v);
    return mg_cfg_valid(e) ? e : q;
  }();
  return p;
}

MgLevel mg_level0(const pf_ctx* c) {
  MgLevel L;
  L.d = c->dc;
  L.nx = c->geom.nx; L.ny = c->geom.ny; L.nz = c->geom.nz;
  L.cd = c->cd32; L.cx = c->cx32; L.cy = c->cy32; L.cz = c->cz32; L.inv = c->invd32;
  return L;
}

MgLevel mg_level(const Ctx::MgBuf& b) {
  MgLevel L;
  L.d = b.d;
  L.nx = b.nx; L.ny = b.ny; L.nz = b.nz;
  L.cd = b.cd; L.cx = b.cx; L.cy = b.cy; L.cz = b.cz; L.inv = b.inv;
  return L;
}

// the coarse operators of the assembled system (levels allocated on first use:
// halving down to at most 4 cells a side; a grid of at most 4 cells a side has none)
void mg_free_level(Ctx::MgBuf& b) {
  float** arrs[8] = {&b.cd, &b.cx, &b.cy, &b.cz, &b.inv, &b.x, &b.r, &b.t};
  for (float** a : arrs) dfree(*a);
}

// every multigrid buffer, or none: a partial allocation is released and the context
// keeps no levels, so a later call starts again from nothing
int mg_alloc(pf_ctx* c) {
  std::vector<Ctx::MgBuf> levels;
  float* t0 = nullptr;
  uint16_t* h[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  uint32_t* mx = nullptr;
  float* unscale = nullptr;
  auto release = [&]() {
    for (auto& b : levels) mg_free_level(b);
    dfree(t0);
    for (auto& p : h) dfree(p);
    dfree(mx);
    dfree(unscale);
  };
  hipError_t e = hipSuccess;
  int nx = c->geom.nx, ny = c->geom.ny, nz = c->geom.nz;
  while (e == hipSuccess && std::max({nx, ny, nz}) > 4) {
    nx = (nx + 1) / 2; ny = (ny + 1) / 2; nz = (nz + 1) / 2;
    Ctx::MgBuf b;
    b.d = make_dims(nx, ny, nz);
    b.nx = nx; b.ny = ny; b.nz = nz;
    const long long n = b.d.nblocks * kBlockVol;
    float** arrs[8] = {&b.cd, &b.cx, &b.cy, &b.cz, &b.inv, &b.x, &b.r, &b.t};
    for (float** a : arrs)
      if (e == hipSuccess) e = dalloc(a, n);
    levels.push_back(b);
  }
  const long long n0 = c->dc.nblocks * kBlockVol;
  if (e == hipSuccess) e = dalloc(&t0, n0);
  for (auto& p : h)
    if (e == hipSuccess) e = dalloc(&p, n0);
  if (e == hipSuccess) e = dalloc(&mx, 2);
  if (e == hipSuccess) e = dalloc(&unscale, 2);
  if (e != hipSuccess) {
    release();
    return fail(c, PF_ERR_NOMEM, std::string("multigrid levels: ") + hipGetErrorString(e));
  }
  c->mg = std::move(levels);
  c->mg_t0 = t0;
  for (int q = 0; q < 5; ++q) c->mg_h[q] = h[q];
  c->mg_mx = mx;
  c->mg_unscale = unscale;
  return PF_OK;
}

int mg_build(pf_ctx* c) {
  if (!c->mg_t0) {
    int st = mg_alloc(c);
    if (st) return st;
  }
  // level 1's residual is restricted from the FLUID blocks only: the aggregates of the
  // other blocks stay zero through the solve
  if (!c->mg.empty()) HIP_TRY(c, fill_f32(c->stream, c->mg[0].r, 0.f, c->mg[0].d.nblocks * kBlockVol));
  // the FP16 level-0 operator of the row kernels
  {
    const long long n0 = c->dc.nblocks * kBlockVol;
    HIP_TRY(c, fill_u32(c->stream, c->mg_mx, 0u, 2));
    hipLaunchKernelGGL(k_mg_absmax, dim3(std::min<unsigned>(grid_for(n0, 256), 2048u)), dim3(256), 0, c->stream, n0,
                       c->cd32, c->cx32, c->cy32, c->cz32, c->invd32, c->mg_mx);
    _Float16* h[5];
    for (int q = 0; q < 5; ++q) h[q] = reinterpret_cast<_Float16*>(c->mg_h[q]);
    hipLaunchKernelGGL(k_mg_tohalf, dim3(std::min<unsigned>(grid_for(n0, 256), 4096u)), dim3(256), 0, c->stream, n0,
                       c->cd32, c->cx32, c->cy32, c->cz32, c->invd32, h[0], h[1], h[2], h[3], h[4], c->mg_mx,
                       c->mg_unscale);
    LAUNCH_CHECK(c);
  }
  MgLevel F = mg_level0(c);
  for (auto& b : c->mg) {
    MgLevel Cl = mg_level(b);
    const unsigned g = std::min<unsigned>(grid_for(b.d.nblocks * kBlockVol, 256), 4096u);
    hipLaunchKernelGGL(k_mg_coarsen, dim3(g), dim3(256), 0, c->stream, F, Cl, b.cd, b.cx, b.cy, b.cz);
    hipLaunchKernelGGL(k_mg_inverse, dim3(g), dim3(256), 0, c->stream, Cl, b.inv);
    LAUNCH_CHECK(c);
    F = Cl;
  }
  return PF_OK;
}

// workgroups of the level-0 kernels: a multiple of 8 (the XCD-aware list walk), at
// most 1024 (the partials of r.z)
unsigned mg_grid0(const pf_ctx* c) { return std::min(list_grid(c->npblk), 1024u); }

// z = V-cycle(r) on the assembled system; partial r.z (mg_grid0 workgroups) in
// part[2 * kMaxPartials ...]; every kernel is skipped while *done
int mg_apply(pf_ctx* c, const float* r, float* z, const int* done, double* part) {
  struct {
    int nu, coarsest;
    float omega, correction;
  } const P{c->mg_cfg.sweeps, c->mg_cfg.coarsest_sweeps, c->mg_cfg.omega, c->mg_cfg.correction};
  const int nl = (int)c->mg.size();
  MgWalk w0{c->pblk, c->npblk, done};
  MgWalk wd{nullptr, 0, done};
  const unsigned g0 = mg_grid0(c);
  // kernels that write no partial may take more workgroups (up to 16 list blocks each)
  const unsigned gw = list_grid(c->npblk);
  const MgLevel L0 = mg_level0(c);
  double* const rz = part + 2 * kMaxPartials;
  auto gdense = [](const Ctx::MgBuf& b) {
    return std::min<unsigned>(grid_for(b.d.nblocks * kBlockVol, 256), 2048u);
  };
  // level 0 by cell rows when the slabs are below 2 GiB and there is a coarse level
  const bool rows = nl > 0 && c->dc.nblocks * kBlockVol * 4 < (1LL << 31);
  MgRowArgs ra{};
  ra.list = c->pblk;
  ra.nlist = c->npblk;
  ra.done = done;
  ra.d = c->dc;
  ra.cd = reinterpret_cast<const _Float16*>(c->mg_h[0]);
  ra.cx = reinterpret_cast<const _Float16*>(c->mg_h[1]);
  ra.cy = reinterpret_cast<const _Float16*>(c->mg_h[2]);
  ra.cz = reinterpret_cast<const _Float16*>(c->mg_h[3]);
  ra.inv = reinterpret_cast<const _Float16*>(c->mg_h[4]);
  ra.unscale = c->mg_unscale;
  ra.r = r;
  ra.omega = P.omega;
  // the weights of a smoothing pass's sweeps: all omega, or (a Chebyshev ratio r > 1)
  // the reciprocals of the Chebyshev roots of degree nu on [lmax / r, lmax], lmax = 2
  // (Gershgorin bound of D^-1 A), which damp that whole interval; the post-pass takes
  // them in reverse order (they commute, so the cycle stays symmetric)
  const float cheb = c->mg_cfg.chebyshev_ratio;
  std::vector<float> wpre(P.nu, P.omega);
  if (cheb > 1.0f) {
    const double b = 2.0, lo = b / cheb;
    for (int q = 0; q < P.nu; ++q) {
      const double lam = 0.5 * (b + lo) + 0.5 * (b - lo) * std::cos((2.0 * q + 1.0) * M_PI / (2.0 * P.nu));
      wpre[q] = (float)(1.0 / lam);
    }
  }
  auto wpost = [&](int q) { return wpre[(size_t)(P.nu - 1 - q)]; };
  ra.omega_c = P.correction;
  if (nl > 0) ra.dc = c->mg[0].d;
  // the level-0 kernels that write an iterate alternate between the work array and
  // z; the first one is chosen so that the last lands in z
  const int pre = rows ? (P.nu >= 2 ? P.nu - 1 : 1) : P.nu;
  const int writes = pre + P.nu;
  float* cur = nullptr;
  float* nxt = (writes & 1) ? z : c->mg_t0;
  auto flip = [&]() {
    cur = nxt;
    nxt = (cur == z) ? c->mg_t0 : z;
  };
  // pre-smoothing
  if (rows && P.nu >= 2) {
    ra.x = nullptr; ra.out = nxt; ra.part = nullptr;
    ra.omega_u = wpre[0];
    ra.omega = wpre[1];
    hipLaunchKernelGGL(k_mg_row<MG_FIRST2>, dim3(gw), dim3(256), 0, c->stream, ra);
    flip();
    for (int s2 = 2; s2 < P.nu; ++s2) {
      ra.x = cur; ra.out = nxt;
      ra.omega = wpre[(size_t)s2];
      hipLaunchKernelGGL(k_mg_row<MG_SWEEP>, dim3(gw), dim3(256), 0, c->stream, ra);
      flip();
    }
  } else {
    hipLaunchKernelGGL(k_mg_first, dim3(g0), dim3(256), 0, c->stream, w0, L0, r, nxt, wpre[0]);
    flip();
    const int more = nl > 0 ? P.nu - 1 : 2 * P.nu - 1;
    for (int s2 = 0; s2 < more; ++s2) {
      const bool last = nl == 0 && s2 == more - 1;
      const int q = s2 + 1;  // sweep q of pre (q < nu) and post (q - nu) smoothing
      const float om = q < P.nu ? wpre[(size_t)q] : wpost(q - P.nu);
      hipLaunchKernelGGL(k_mg_sweep, dim3(g0), dim3(256), 0, c->stream, w0, L0, r, cur, nxt, om,
                         last ? rz : nullptr);
      flip();
    }
  }
  if (nl == 0) {
    LAUNCH_CHECK(c);
    return PF_OK;
  }
  // restriction to level 1
  if (rows) {
    ra.x = cur; ra.out = c->mg[0].r; ra.part = nullptr;
    hipLaunchKernelGGL(k_mg_row<MG_RESTRICT>, dim3(gw), dim3(256), 0, c->stream, ra);
  } else {
    hipLaunchKernelGGL(k_mg_restrict, dim3(g0), dim3(256), 0, c->stream, w0, L0, r, cur, mg_level(c->mg[0]),
                       c->mg[0].r);
  }
  // per-aggregate correction scaling: S on the restricted residual ...
  const long long n1 = c->mg[0].d.nblocks * kBlockVol;
  if (c->mg_cs)
    hipLaunchKernelGGL(k_mg_scale, dim3(std::min<unsigned>(grid_for(n1, 256), 2048u)), dim3(256), 0, c->stream, n1,
                       c->mg[0].r, c->mg_cs, done);
  // coarse levels: level l (1..nl) solves A_l x = r_l from zero, result in res[l].
  // (Two corrections per level -- a W-cycle -- cost more launches than they save
  // iterations at 256^3, and twice the 1.8 over-correction stops converging.)
  std::vector<float*> res(nl + 1), alt(nl + 1);
  std::function<void(int)> coarse = [&](int l) {
    Ctx::MgBuf& b = c->mg[(size_t)l - 1];
    const MgLevel L = mg_level(b);
    float* x = b.x;
    float* t = b.t;
    if (l == nl) {
      hipLaunchKernelGGL(k_mg_coarsest, dim3(1), dim3(64), 0, c->stream, done, L, b.r, x, P.omega, P.coarsest);
      res[l] = x;
      alt[l] = t;
      return;
    }
    const unsigned g = gdense(b);
    hipLaunchKernelGGL(k_mg_first, dim3(g), dim3(256), 0, c->stream, wd, L, b.r, x, wpre[0]);
    for (int s2 = 1; s2 < P.nu; ++s2) {
      hipLaunchKernelGGL(k_mg_sweep, dim3(g), dim3(256), 0, c->stream, wd, L, b.r, x, t, wpre[(size_t)s2],
                         nullptr);
      std::swap(x, t);
    }
    Ctx::MgBuf& nb = c->mg[(size_t)l];
    hipLaunchKernelGGL(k_mg_restrict, dim3(g), dim3(256), 0, c->stream, wd, L, b.r, x, mg_level(nb), nb.r);
    coarse(l + 1);
    hipLaunchKernelGGL(k_mg_prolong, dim3(g), dim3(256), 0, c->stream, wd, L, mg_level(nb), res[l + 1], x,
                       P.correction);
    for (int s2 = 0; s2 < P.nu; ++s2) {
      hipLaunchKernelGGL(k_mg_sweep, dim3(g), dim3(256), 0, c->stream, wd, L, b.r, x, t, wpost(s2), nullptr);
      std::swap(x, t);
    }
    res[l] = x;
    alt[l] = t;
  };
  coarse(1);
  // ... and on the correction (P S B_1 S P^T stays symmetric)
  if (c->mg_cs)
    hipLaunchKernelGGL(k_mg_scale, dim3(std::min<unsigned>(grid_for(n1, 256), 2048u)), dim3(256), 0, c->stream, n1,
                       res[1], c->mg_cs, done);
  // level 0: prolong and post-smooth, the last sweep into z with the partial r.z
  if (rows) {
    ra.x = cur; ra.out = nxt; ra.xc = res[1]; ra.part = P.nu == 1 ? rz : nullptr;
    ra.omega = wpost(0);
    hipLaunchKernelGGL(k_mg_row<MG_PSWEEP>, dim3(ra.part ? g0 : gw), dim3(256), 0, c->stream, ra);
    flip();
    for (int s2 = 1; s2 < P.nu; ++s2) {
      ra.x = cur; ra.out = nxt; ra.part = s2 == P.nu - 1 ? rz : nullptr;
      ra.omega = wpost(s2);
      hipLaunchKernelGGL(k_mg_row<MG_SWEEP>, dim3(ra.part ? g0 : gw), dim3(256), 0, c->stream, ra);
      flip();
    }
  } else {
    hipLaunchKernelGGL(k_mg_prolong, dim3(g0), dim3(256), 0, c->stream, w0, L0, mg_level(c->mg[0]), res[1], cur,
                       P.correction);
    for (int s2 = 0; s2 < P.nu; ++s2) {
      hipLaunchKernelGGL(k_mg_sweep, dim3(g0), dim3(256), 0, c->stream, w0, L0, r, cur, nxt, wpost(s2),
                         s2 == P.nu - 1 ? rz : nullptr);
      flip();
    }
  }
  LAUNCH_CHECK(c);
  if (cur != z) return fail(c, PF_ERR_STATE, "multigrid: cycle did not end in z");
  return PF_OK;
}

int run_project(pf_ctx* c, const pf_phase* ph, double dt, const pf_solve_cfg* cfg, double cdiv,
                pf_solve_stats* out, bool timing, int ev_assemble, int ev_solve, int ev_vel,
                bool single = false, double mr_rel_tol = 0.0) {
  const bool mr = c->mr;
  if (cfg->precond < PF_PRECOND_JACOBI || cfg->precond > PF_PRECOND_MULTIGRID)
    return fail(c, PF_ERR_ARG, "unknown preconditioner");
  if (cfg->precond == PF_PRECOND_MULTIGRID && mr)
    return fail(c, PF_ERR_ARG, "the multigrid preconditioner runs on the sparse path only");
  pf_solve_stats S;
  std::memset(&S, 0, sizeof(S));
  S.pinned_cell = -1;
  const Geom& g = c->geom;
  const long long nb = c->dc.nblocks;
  // 1. blocks of the previous solve return to zero; the pressure field is cleared
  if (c->npblk_prev > 0) {
    const bool warm = cfg->warm_start != 0;
    if (warm) {
      hipLaunchKernelGGL(k_warm_save, dim3(grid_for(c->npblk_prev, 4)), dim3(256), 0, c->stream, c->pblk_prev,
                         c->npblk_prev, c->f[F_P], c->x64);
      LAUNCH_CHECK(c);
    }
    ZeroArgs z{};
    z.list = c->pblk_prev;
    z.nlist = c->npblk_prev;
    // with the multigrid work array: its values at cells outside the next system are
    // read (times a zero coupling) by the level-0 sweeps, so none may stay non-finite
    float* f32[13] = {c->f[F_P], c->x32, c->r32, c->z32, c->pA, c->pB, c->Ap, c->cd32, c->cx32, c->cy32, c->cz32,
                      c->invd32, c->mg_t0};
    for (int q = 0; q < 13; ++q) z.f32[q] = f32[q];
    z.nf32 = c->mg_t0 ? 13 : 12;
    double* f64[7] = {c->b64, c->r64, c->cd64, c->cx64, c->cy64, c->cz64, c->x64};
    for (int q = 0; q < 7; ++q) z.f64[q] = f64[q];
    z.nf64 = warm ? 6 : 7;  // a warm start keeps x64 (the saved pressure)
    hipLaunchKernelGGL(k_zero_blocks, dim3(grid_for(c->npblk_prev, 4)), dim3(256), 0, c->stream, z);
    LAUNCH_CHECK(c);
  }
  HIP_TRY(c, fill_u8(c->stream, c->act[F_P], 0, nb));
  // 2. sorted list of blocks holding FLUID cells
  uint8_t* flag8 = reinterpret_cast<uint8_t*>(c->blk_flag32);
  hipLaunchKernelGGL(k_fluid_flags, dim3(grid_for(nb, 4)), dim3(256), 0, c->stream, c->dc, c->marker, flag8);
  LAUNCH_CHECK(c);
  HIP_TRY(c, compact_flags_morton(*c, c->dc, flag8, c->pblk, c->u32_dev + 2));
  HIP_TRY(c, fill_u32(c->stream, c->u32_dev, 0u, 1));
  HIP_TRY(c, fill_u32(c->stream, c->u32_dev + 1, 0xffffffffu, 1));
  HIP_TRY(c, fill_u32(c->stream, c->u32_dev + 4, 0u, 1));
  HIP_TRY(c, hipMemcpyAsync(c->u32_host, c->u32_dev, sizeof(uint32_t) * 4, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  const long long npb = c->u32_host[2];
  c->npblk = npb;
  if (npb == 0) {
    std::swap(c->pblk, c->pblk_prev);
    c->npblk_prev = 0;
    c->npblk = 0;
    S.converged = 1;
    record(c, timing, ev_assemble);
    record(c, timing, ev_solve);
    record(c, timing, ev_vel);
    if (out) *out = S;
    return PF_OK;
  }
  // learned preconditioner: tiles and features of this system (for the solve or for
  // recording training snapshots)
  const bool want_learned = cfg->precond == PF_PRECOND_LEARNED;
  if (want_learned && !c->lpc) return fail(c, PF_ERR_STATE, "precond = learned but no network attached (pf_lpc_attach)");
  if (want_learned && mr) return fail(c, PF_ERR_ARG, "the learned preconditioner runs on the sparse path only");
  const bool want_mg = cfg->precond == PF_PRECOND_MULTIGRID;
  const bool lpc_system = c->lpc && (want_learned || c->snap_cap > 0);
  c->tiles_valid = false;
  c->nsnap = 0;
  c->snap_last_iter = -1;
  if (lpc_system) {
    int s2 = lpc_features_alloc(c);
    if (!s2) s2 = lpc_build_tiles(c);
    if (s2) return s2;
  }
  ProjArgs pa;
  pa.g = g; pa.dc = c->dc; pa.du = c->du; pa.dv = c->dv; pa.dw = c->dw;
  pa.marker = c->marker; pa.list = c->pblk; pa.nlist = npb;
  const pf_phase unit_phase = {1.0, 1.0, 1.0, 1.0};
  pa.ph = phase_k(single ? &unit_phase : ph);
  pa.scale = dt / (g.dx * g.dx);
  pa.cdiv = single ? 0.0 : cdiv;
  pa.pin = c->u32_dev;
  pa.single = single ? 1 : 0;
  const unsigned G = list_grid(npb);
  const bool refmode = mr && c->mr_cfg && c->mr_cfg->mode == PF_MR_SOLVE_REFERENCE;
  if (mr) {
    c->mr_detail = pf_mr_solve_detail{};
    c->mr_detail.coarse_preconditioner_auto_disable_iteration = -1;
  }
  if (refmode) {
    if (!c->mrf) HIP_TRY(c, dalloc(&c->mrf, (long long)kMrfVectors * c->dc.nblocks * kBlockVol));
    if (!c->mr_pin_dev) HIP_TRY(c, dalloc(&c->mr_pin_dev, 1));
    if (!c->mr_pin_host) HIP_TRY(c, hipHostMalloc((void**)&c->mr_pin_host, sizeof(long long)));
    HIP_TRY(c, fill_u32(c->stream, reinterpret_cast<uint32_t*>(c->mr_pin_dev), 0xffffffffu, 2));
  }
  if (mr) {
    // multires: couplings are written per fine face, so the arrays start at zero
    const Dims* fd[3] = {&c->du, &c->dv, &c->dw};
    for (int a = 0; a < 3; ++a) {
      HIP_TRY(c, fill_f32(c->stream, c->mr_cf32[a], 0.f, fd[a]->nblocks * kBlockVol));
      HIP_TRY(c, fill_f64(c->stream, c->mr_cf64[a], 0.0, fd[a]->nblocks * kBlockVol));
    }
    MrAsm ma{};
    ma.m = mr_grid(c);
    ma.list = c->pblk; ma.nlist = npb;
    ma.ph = phase_k(ph);
    ma.dt = dt; ma.cdiv = cdiv; ma.pin = c->u32_dev;
    ma.u = c->f[F_U]; ma.v = c->f[F_V]; ma.w = c->f[F_W];
    ma.mu = c->f[F_MU]; ma.mv = c->f[F_MV]; ma.mw = c->f[F_MW];
    ma.b64 = c->b64; ma.cd64 = c->cd64; ma.vol64 = c->mr_vol64;
    ma.cd32 = c->cd32; ma.invd32 = c->invd32;
    for (int a = 0; a < 3; ++a) { ma.cf64[a] = c->mr_cf64[a]; ma.cf32[a] = c->mr_cf32[a]; }
    ma.part = c->partial;
    if (refmode) {
      const long long ncell = c->dc.nblocks * kBlockVol;
      ma.rref = c->mrf + 12 * ncell;
      ma.dref = c->mrf + 13 * ncell;
      ma.vtrue = c->mrf + 14 * ncell;
      ma.pin_cell = c->mr_pin_dev;
    }
    hipLaunchKernelGGL(k_mr_pin, dim3(G), dim3(256), 0, c->stream, ma);
    LAUNCH_CHECK(c);
    TIMED(c, PF_K_ASSEMBLE, hipLaunchKernelGGL(k_mr_assemble, dim3(G), dim3(256), 0, c->stream, ma));
    LAUNCH_CHECK(c);
  } else {
  hipLaunchKernelGGL(k_pin_detect, dim3(list_grid(npb)), dim3(256), 0, c->stream, pa);
  LAUNCH_CHECK(c);
  AsmOut ao;
  ao.b64 = c->b64; ao.cd64 = c->cd64; ao.cx64 = c->cx64; ao.cy64 = c->cy64; ao.cz64 = c->cz64;
  ao.cd32 = c->cd32; ao.cx32 = c->cx32; ao.cy32 = c->cy32; ao.cz32 = c->cz32; ao.invd32 = c->invd32;
  ao.u = c->f[F_U]; ao.v = c->f[F_V]; ao.w = c->f[F_W];
  ao.mu = c->f[F_MU]; ao.mv = c->f[F_MV]; ao.mw = c->f[F_MW];
  ao.part = c->partial;
  ao.fbeta = lpc_system ? c->fbeta : nullptr;
  ao.fphi = lpc_system ? c->fphi : nullptr;
  TIMED(c, PF_K_ASSEMBLE, hipLaunchKernelGGL(k_assemble, dim3(G), dim3(256), 0, c->stream, pa, ao));
  LAUNCH_CHECK(c);
  }
  hipLaunchKernelGGL(k_reduce_slots, dim3(1), dim3(256), 0, c->stream, c->partial, (int)G, 1, 0u, c->red + 2);
  LAUNCH_CHECK(c);
  record(c, timing, ev_assemble);

  // 3. FP64 iterative refinement around the FP32 Jacobi PCG
  VecArgs va;
  va.dc = c->dc; va.list = c->pblk; va.nlist = npb; va.sc = c->sc; va.part = c->partial;
  hipLaunchKernelGGL(k_ir_start, dim3(G), dim3(256), 0, c->stream, va, c->x64, c->invd32, cfg->warm_start);
  LAUNCH_CHECK(c);
  MrOp mop{};
  MrOp64 mop64{};
  if (mr) {
    mop.m = mop64.m = mr_grid(c);
    mop.cd = c->cd32; mop64.cd = c->cd64; mop64.vol = c->mr_vol64;
    for (int a = 0; a < 3; ++a) { mop.cf[a] = c->mr_cf32[a]; mop64.cf[a] = c->mr_cf64[a]; }
  }
  auto residual = [&]() -> int {
    if (mr)
      TIMED(c, PF_K_RESIDUAL64,
            hipLaunchKernelGGL(k_mr_residual64, dim3(G), dim3(256), 0, c->stream, va, mop64, c->b64, c->x64, c->r64));
    else
    TIMED(c, PF_K_RESIDUAL64,
          hipLaunchKernelGGL(k_residual64, dim3(G), dim3(256), 0, c->stream, va, c->b64, c->x64, c->r64, c->cd64,
                             c->cx64, c->cy64, c->cz64));
    LAUNCH_CHECK(c);
    hipLaunchKernelGGL(k_reduce_slots, dim3(1), dim3(256), 0, c->stream, c->partial, (int)G, 2, 2u, c->red);
    LAUNCH_CHECK(c);
    return read_small(c);
  };
  int st = residual();
  if (st) return st;
  const double bnorm2 = c->red_host[2];
  double rr = c->red_host[0], rmax = c->red_host[1];
  S.rhs_norm2 = bnorm2;
  S.initial_max_residual = rmax;
  S.pinned = c->u32_host[0] == 0;
  S.pinned_cell = S.pinned ? (int64_t)c->u32_host[1] : -1;
  const int max_iters = std::max(cfg->max_iters, 0);
  const int max_updates = cfg->max_outer > 0 ? cfg->max_outer : 200;
  // (a multigrid iteration costs several products: check at least every 4)
  const int check_base = cfg->check_every > 0 ? cfg->check_every : 8;
  const int check = want_mg ? std::min(check_base, 4) : check_base;
  const double delta = (cfg->inner_rel_tol > 0 && cfg->inner_rel_tol < 1) ? cfg->inner_rel_tol : 0.1;
  // the multires rule: max(absolute, relative * initial max residual), with x0 = 0
  // (multires_pressure3d.cpp:858-863)
  const double mr_tol = std::max(std::max(0.0, cfg->tol), std::max(0.0, mr_rel_tol) * rmax);
  auto met = [&]() {
    if (mr) return rmax < mr_tol;
    if (cfg->stop_rule == PF_STOP_REL2) return bnorm2 == 0.0 || std::sqrt(rr) <= cfg->tol * std::sqrt(bnorm2);
    return rmax < cfg->tol;
  };
  // FP32 residual level that triggers a (checking) replacement: the target itself
  // (none for the multires rows, whose FP32 residual is scaled by the cell volume)
  const double tol2_abs = mr ? 0.0
                             : (cfg->stop_rule == PF_STOP_REL2 ? cfg->tol * cfg->tol * bnorm2 : cfg->tol * cfg->tol);
  static const bool trace = std::getenv("PF_SOLVE_TRACE") != nullptr;
  // most FP32 iterations between two residual replacements
  const char* env_between = std::getenv("PF_SOLVE_REPLACE_EVERY");
  const int between = env_between ? std::max(8, std::atoi(env_between)) : 200;
  int total = 0, outer = 0;
  bool converged = met();
  if (lpc_system) c->tiles_valid = true;
  const long long ncell = c->dc.nblocks * kBlockVol;
  // residual snapshot (training data), at most one per `snap_every` iterations
  auto snapshot = [&](int iter) -> int {
    if (c->snap_cap <= 0 || c->nsnap >= c->snap_cap) return PF_OK;
    if (c->snap_last_iter >= 0 && iter - c->snap_last_iter < c->snap_every) return PF_OK;
    HIP_TRY(c, hipMemcpyAsync(c->snaps + (long long)c->nsnap * ncell, c->r32, sizeof(float) * (size_t)ncell,
                              hipMemcpyDeviceToDevice, c->stream));
    c->nsnap++;
    c->snap_last_iter = iter;
    return PF_OK;
  };
  // flexible PCG around a preconditioner that is not a diagonal: the learned network
  // or the multigrid V-cycle (z and partial r.z in slot 2, GT partials)
  bool learned = want_learned || want_mg;
  if (want_mg) {
    st = mg_build(c);
    if (st) return st;
  }
  const unsigned GT = want_mg ? mg_grid0(c) : want_learned ? lpc_scatter_grid(c->ntiles) : 1u;
  auto papply = [&](const int* done) -> int {
    if (want_mg) {
      long long ki = kbegin(c, PF_K_MG_APPLY);
      int s2 = mg_apply(c, c->r32, c->z32, done, c->partial);
      kend(c, ki);
      return s2;
    }
    long long ki = kbegin(c, PF_K_LPC_APPLY);
    int s2 = lpc_apply(c, c->r32, c->z32, done, c->partial, 2);
    kend(c, ki);
    return s2;
  };
  int* done_dev = &c->sc->done;
  if (refmode) {
    // the reference's FP64 iteration with its options; x in x64
    HIP_TRY(c, hipMemcpyAsync(c->mr_pin_host, c->mr_pin_dev, sizeof(long long), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->mr_detail.reference_order = 1;
    int ri = 0;
    bool rc = false, rb = false;
    st = mr_reference_solve(c, *c->mr_cfg, G, va, mop64, (long long)c->u32_host[4], ri, rc, rb);
    if (st) return st;
    total = ri;
    converged = rc;
    if (rb) S.breakdown = 1;
    st = residual();
    if (st) return st;
    rr = c->red_host[0];
    rmax = c->red_host[1];
  }
  if (!refmode && !converged && max_iters > 0 && std::isfinite(rmax)) {
    // (re)start: r32 from r64, x32 = 0, p = 0, z = M(r)
    auto start = [&](int budget) -> int {
      hipLaunchKernelGGL(k_pcg_load, dim3(G), dim3(256), 0, c->stream, va, c->r64, c->x32, c->r32, c->z32, c->pA,
                         c->invd32, learned ? 1 : 0);
      LAUNCH_CHECK(c);
      if (learned) {
        int s2 = papply(nullptr);
        if (s2) return s2;
        hipLaunchKernelGGL(k_pcg_start_flex, dim3(1), dim3(256), 0, c->stream, c->partial, (int)G, (int)GT, c->sc,
                           delta * delta, tol2_abs, budget, between);
      } else {
        hipLaunchKernelGGL(k_pcg_start, dim3(1), dim3(256), 0, c->stream, c->partial, (int)G, c->sc, delta * delta,
                           tol2_abs, budget, between);
      }
      LAUNCH_CHECK(c);
      return PF_OK;
    };
    st = start(max_iters);
    if (st) return st;
    st = snapshot(0);
    if (st) return st;
    // the two iteration kernels use at most 1024 workgroups, so the per-iteration
    // scalar kernels sum few partials
    const unsigned GS = std::min(G, 1024u);
    // the buffer-addressed SpMV takes 32-bit byte offsets: arrays below 2 GiB
    const bool spmv_buffers = c->dc.nblocks * kBlockVol * 4 < (1LL << 31);
    VecArgs vs = va;
    int iter_before = 0, iter_base = 0;
    double best_rr = rr, prev_rr = rr;
    int no_gain = 0, learned_slow = 0;
    for (;;) {
      const size_t rec0 = c->krec.size();
      for (int q = 0; q < check; ++q) {
        if (mr)
          TIMED(c, PF_K_PCG_SPMV,
                hipLaunchKernelGGL(k_mr_spmv, dim3(GS), dim3(256), 0, c->stream, vs, mop, c->z32, c->pA, c->pB, c->Ap));
        else if (spmv_buffers)
        TIMED(c, PF_K_PCG_SPMV,
              hipLaunchKernelGGL(k_pcg_spmv, dim3(GS), dim3(256), 0, c->stream, vs, c->z32, c->pA, c->pB, c->Ap,
                                 c->cd32, c->cx32, c->cy32, c->cz32));
        else
        TIMED(c, PF_K_PCG_SPMV,
              hipLaunchKernelGGL(k_pcg_spmv_gather, dim3(GS), dim3(256), 0, c->stream, vs, c->z32, c->pA, c->pB,
                                 c->Ap, c->cd32, c->cx32, c->cy32, c->cz32));
        hipLaunchKernelGGL(k_pcg_alpha, dim3(1), dim3(256), 0, c->stream, c->partial, (int)GS, c->sc);
        if (learned) {
          TIMED(c, PF_K_PCG_UPDATE,
                hipLaunchKernelGGL(k_pcg_update_flex, dim3(GS), dim3(256), 0, c->stream, vs, c->x32, c->r32, c->z32,
                                   c->pA, c->pB, c->Ap));
          st = papply(done_dev);
          if (st) return st;
          hipLaunchKernelGGL(k_pcg_beta_flex, dim3(1), dim3(256), 0, c->stream, c->partial, (int)GS, (int)GT, c->sc);
        } else {
          TIMED(c, PF_K_PCG_UPDATE,
                hipLaunchKernelGGL(k_pcg_update, dim3(GS), dim3(256), 0, c->stream, vs, c->x32, c->r32, c->z32, c->pA,
                                   c->pB, c->Ap, c->invd32));
          hipLaunchKernelGGL(k_pcg_beta, dim3(1), dim3(256), 0, c->stream, c->partial, (int)GS, c->sc);
        }
      }
      LAUNCH_CHECK(c);
      HIP_TRY(c, hipMemcpyAsync(c->done_host, &c->sc->iter, sizeof(int) * 8, hipMemcpyDeviceToHost, c->stream));
      HIP_TRY(c, hipStreamSynchronize(c->stream));
      // {iter, done, breakdown, stalled, iter_best, max_iter, stall_window, upd}
      // launches after a pause return at once: keep the timings of those that ran
      const int ran = c->done_host[0] - iter_before;
      iter_before = c->done_host[0];
      int kept_s = 0, kept_u = 0, kept_l = 0;
      for (size_t r = rec0; r < c->krec.size(); ++r) {
        int& k = c->krec[r].first;
        if (k == PF_K_PCG_SPMV && kept_s++ >= ran) k = -1;
        if (k == PF_K_PCG_UPDATE && kept_u++ >= ran) k = -1;
        if ((k == PF_K_LPC_APPLY || k == PF_K_MG_APPLY) && kept_l++ >= ran) k = -1;
      }
      if (ran > 0) {
        st = snapshot(iter_base + c->done_host[0]);
        if (st) return st;
      }
      if (!c->done_host[1]) continue;
      total = iter_base + c->done_host[0];
      const bool breakdown = c->done_host[2] != 0;
      // residual replacement: x64 += x32, r64 = b - A x64 in FP64
      hipLaunchKernelGGL(k_ir_accum, dim3(G), dim3(256), 0, c->stream, va, c->x64, c->x32);
      LAUNCH_CHECK(c);
      st = residual();
      if (st) return st;
      rr = c->red_host[0];
      rmax = c->red_host[1];
      ++outer;
      converged = met();
      if (trace) {
        double s_rr32 = 0.0;
        HIP_TRY(c, hipMemcpy(&s_rr32, &c->sc->rr, sizeof(double), hipMemcpyDeviceToHost));
        std::fprintf(stderr, "[pf solve] update %d iter %d fp32 rel %.3e true rel %.3e max %.3e%s%s\n", outer, total,
                     std::sqrt(s_rr32 / std::max(bnorm2, 1e-300)), std::sqrt(rr / std::max(bnorm2, 1e-300)), rmax,
                     breakdown ? " breakdown" : "", learned ? " (learned)" : "");
      }
      if (rr < best_rr * 0.999) { best_rr = rr; no_gain = 0; } else { ++no_gain; }
      // (no gain over 40 replacements: the solve has stagnated)
      if (converged || total >= max_iters || outer >= max_updates || !std::isfinite(rr) || no_gain >= 40) {
        if (breakdown) S.breakdown = 1;
        break;
      }
      // the learned preconditioner falls back to Jacobi when two replacements in a
      // row gained less than 10% (or the recursion broke down)
      if (learned) {
        learned_slow = (rr > 0.9 * prev_rr || breakdown) ? learned_slow + 1 : 0;
        if (learned_slow >= 2 || breakdown) {
          learned = false;
          S.fallback_used = 1;
          iter_base = total;
          iter_before = 0;
          prev_rr = rr;
          st = start(max_iters - total);
          if (st) return st;
          continue;
        }
      }
      prev_rr = rr;
      // resume from the true residual; after a breakdown restart the direction too
      hipLaunchKernelGGL(k_pcg_load, dim3(G), dim3(256), 0, c->stream, va, c->r64, c->x32, c->r32, c->z32,
                         breakdown ? c->pA : nullptr, c->invd32, learned ? 1 : 0);
      LAUNCH_CHECK(c);
      if (learned) {
        st = papply(nullptr);
        if (st) return st;
      }
      if (breakdown) {
        if (learned)
          hipLaunchKernelGGL(k_pcg_start_flex, dim3(1), dim3(256), 0, c->stream, c->partial, (int)G, (int)GT, c->sc,
                             delta * delta, tol2_abs, max_iters - total, between);
        else
          hipLaunchKernelGGL(k_pcg_start, dim3(1), dim3(256), 0, c->stream, c->partial, (int)G, c->sc, delta * delta,
                             tol2_abs, max_iters - total, between);
        iter_base = total;
        iter_before = 0;
      } else if (learned) {
        hipLaunchKernelGGL(k_pcg_resume_flex, dim3(1), dim3(256), 0, c->stream, c->partial, (int)G, (int)GT, c->sc);
      } else {
        hipLaunchKernelGGL(k_pcg_resume, dim3(1), dim3(256), 0, c->stream, c->partial, (int)G, c->sc);
      }
      LAUNCH_CHECK(c);
    }
  }
  S.iterations = total;
  S.outer_iterations = outer;
  S.converged = converged ? 1 : 0;
  S.final_max_residual = rmax;
  S.true_max_residual = rmax;
  S.true_rel_residual = bnorm2 > 0 ? std::sqrt(rr / bnorm2) : 0.0;
  S.fluid_cells = c->u32_host[4];
  if (mr) S.pinned_cell = S.pinned ? 0 : -1;  // index in the reference's FLUID cell order
  record(c, timing, ev_solve);

  // 4. pressure store and face velocity update
  if (mr) {
    const long long ncell = c->dc.nblocks * kBlockVol;
    hipLaunchKernelGGL(k_mr_store_pressure, dim3(grid_for(ncell, 256)), dim3(256), 0, c->stream, mr_grid(c), c->x64,
                       c->f[F_P], c->mr_pfull, c->act[F_P]);
    const Dims* fd[3] = {&c->du, &c->dv, &c->dw};
    const PhaseK pk = phase_k(ph);
    for (int a = 0; a < 3; ++a)
      TIMED(c, PF_K_VELOCITY,
            hipLaunchKernelGGL(k_mr_velocity, dim3(grid_for(fd[a]->nblocks * kBlockVol, 256)), dim3(256), 0, c->stream,
                               mr_grid(c), a, c->mr_pfull, c->f[F_U + a], c->f[F_MU + a], pk, dt));
    LAUNCH_CHECK(c);
    record(c, timing, ev_vel);
    std::swap(c->pblk, c->pblk_prev);
    c->npblk_prev = npb;
    if (out) *out = S;
    return PF_OK;
  }
  hipLaunchKernelGGL(k_store_pressure, dim3(G), dim3(256), 0, c->stream, va, c->marker, c->x64, c->f[F_P],
                     c->act[F_P]);
  LAUNCH_CHECK(c);
  VelArgs vl;
  vl.g = g; vl.dc = c->dc; vl.df[0] = c->du; vl.df[1] = c->dv; vl.df[2] = c->dw;
  vl.marker = c->marker; vl.p = c->f[F_P];
  for (int ax = 0; ax < 3; ++ax) {
    vl.vel[ax] = c->f[F_U + ax];
    vl.mass[ax] = c->f[F_MU + ax];
    vl.act_vel[ax] = c->act[F_U + ax];
  }
  vl.ph = pa.ph;
  vl.s = dt / g.dx;
  vl.list = c->pblk;
  vl.nlist = npb;
  vl.unit_beta = single ? 1 : 0;
  TIMED(c, PF_K_VELOCITY,
        hipLaunchKernelGGL(k_velocity_update, dim3(grid_for(npb, 4)), dim3(256), 0, c->stream, vl));
  LAUNCH_CHECK(c);
  record(c, timing, ev_vel);
  std::swap(c->pblk, c->pblk_prev);
  c->npblk_prev = npb;
  if (out) *out = S;
  return PF_OK;
}

int run_g2p_advect(pf_ctx* c, bool g2p, bool adv, double aL, double aG, double dt, int order) {
  if (c->np == 0) return PF_OK;
  G2PArgs a;
  a.g = c->geom;
  a.df[0] = c->du; a.df[1] = c->dv; a.df[2] = c->dw;
  for (int ax = 0; ax < 3; ++ax) {
    a.vel[ax] = c->f[F_U + ax];
    a.saved[ax] = c->f[F_SU + ax];
  }
  a.P = c->pa;
  a.n = c->np;
  a.inv_dx = (float)(1.0 / c->geom.dx);
  a.ox = (float)c->geom.ox; a.oy = (float)c->geom.oy; a.oz = (float)c->geom.oz;
  const Geom& g = c->geom;
  a.lo[0] = (float)(g.ox + 0.5 * g.dx); a.hi[0] = (float)(g.ox + (g.nx - 0.5) * g.dx);
  a.lo[1] = (float)(g.oy + 0.5 * g.dx); a.hi[1] = (float)(g.oy + (g.ny - 0.5) * g.dx);
  a.lo[2] = (float)(g.oz + 0.5 * g.dx); a.hi[2] = (float)(g.oz + (g.nz - 0.5) * g.dx);
  a.aL = (float)aL; a.aG = (float)aG; a.dt = (float)dt;
  a.order = order;
  a.do_g2p = g2p; a.do_advect = adv;
  a.esc = c->esc_dev;
  a.vmax_bits = reinterpret_cast<uint32_t*>(c->vmax_dev);
  TIMED(c, PF_K_G2P_ADVECT,
        hipLaunchKernelGGL(k_g2p_advect, dim3(xcd_grid(c->np, 256, 8192)), dim3(256), 0, c->stream, a));
  LAUNCH_CHECK(c);
  if (adv) c->binned = false;
  return PF_OK;
}

// calibrateRhoTilde0, src/transfer/transfer3d_tp.cpp:29-36: raw u-face mass at
// (4,4,4) of an 8^3 lattice with 8 liquid particles per cell, over rho_l.
double calibrate_host(const pf_phase* pp, double Vp) {
  const int n = 8;
  std::vector<double> M((size_t)(n + 1) * n * n, 0.0);
  const double mp = pp->rho_l * Vp;
  auto kern = [](double d2) {
    double q = d2 / (1.5 * 1.5), t = 1.0 - q;
    return t > 0.0 ? t * t * t : 0.0;
  };
  for (int k = 0; k < n; ++k)
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i)
        for (int s = 0; s < 8; ++s) {
          double gx = i + 0.25 + 0.5 * (s & 1), gy = j + 0.25 + 0.5 * ((s >> 1) & 1) - 0.5,
                 gz = k + 0.25 + 0.5 * ((s >> 2) & 1) - 0.5;
          int i0 = (int)std::floor(gx), j0 = (int)std::floor(gy), k0 = (int)std::floor(gz);
          double wsum = 0.0;
          for (int dk = -2; dk <= 3; ++dk)
            for (int dj = -2; dj <= 3; ++dj)
              for (int di = -2; di <= 3; ++di) {
                int ii = i0 + di, jj = j0 + dj, kk = k0 + dk;
                if (ii < 0 || ii >= n + 1 || jj < 0 || jj >= n || kk < 0 || kk >= n) continue;
                double dx = gx - ii, dy = gy - jj, dz = gz - kk;
                wsum += kern(dx * dx + dy * dy + dz * dz);
              }
          if (wsum <= 0.0) continue;
          for (int dk = -2; dk <= 3; ++dk)
            for (int dj = -2; dj <= 3; ++dj)
              for (int di = -2; di <= 3; ++di) {
                int ii = i0 + di, jj = j0 + dj, kk = k0 + dk;
                if (ii < 0 || ii >= n + 1 || jj < 0 || jj >= n || kk < 0 || kk >= n) continue;
                double dx = gx - ii, dy = gy - jj, dz = gz - kk;
                double w = kern(dx * dx + dy * dy + dz * dz) / wsum;
                if (w <= 0.0) continue;
                M[ii + (size_t)(n + 1) * (jj + (size_t)n * kk)] += w * mp;
              }
        }
  double raw = M[4 + (size_t)(n + 1) * (4 + (size_t)n * 4)];
  return raw > 0 ? raw / pp->rho_l : 1.0;
}

}  // namespace

extern "C" {

int pf_abi_version(void) { return PF_ABI_VERSION; }

void pf_mg_defaults(pf_mg_cfg* cfg) {
  if (cfg) *cfg = mg_default_cfg();
}

int pf_set_multigrid(pf_ctx* c, const pf_mg_cfg* cfg) {
  if (!c || !cfg) return PF_ERR_ARG;
  if (!mg_cfg_valid(*cfg)) return fail(c, PF_ERR_ARG, "invalid multigrid cycle");
  c->mg_cfg = *cfg;
  return PF_OK;
}

int pf_mg_set_correction_weights(pf_ctx* c, const float* w, int64_t n) {
  if (!c) return PF_ERR_ARG;
  if (!w) {
    dfree(c->mg_cs);
    return PF_OK;
  }
  const int ax = (c->geom.nx + 1) / 2, ay = (c->geom.ny + 1) / 2, az = (c->geom.nz + 1) / 2;
  if (std::max({c->geom.nx, c->geom.ny, c->geom.nz}) <= 4)
    return fail(c, PF_ERR_ARG, "pf_mg_set_correction_weights: the grid has no coarse level");
  if (n != (int64_t)ax * ay * az)
    return fail(c, PF_ERR_ARG, "pf_mg_set_correction_weights: n must be the aggregate count ceil(nx/2)*ceil(ny/2)*ceil(nz/2)");
  const Dims d1 = make_dims(ax, ay, az);
  std::vector<float> s((size_t)(d1.nblocks * kBlockVol), 1.0f);
  for (int k = 0; k < az; ++k)
    for (int j = 0; j < ay; ++j)
      for (int i = 0; i < ax; ++i) {
        const float v = w[i + (int64_t)ax * (j + (int64_t)ay * k)];
        if (!(v > 0.f) || !std::isfinite(v))
          return fail(c, PF_ERR_ARG, "pf_mg_set_correction_weights: weights must be positive and finite");
        s[(size_t)bm_index(d1, i, j, k)] = std::sqrt(v);
      }
  if (!c->mg_cs) HIP_TRY(c, dalloc(&c->mg_cs, d1.nblocks * kBlockVol));
  HIP_TRY(c, hipMemcpyAsync(c->mg_cs, s.data(), sizeof(float) * s.size(), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return PF_OK;
}

int pf_get_multigrid(const pf_ctx* c, pf_mg_cfg* cfg) {
  if (!c || !cfg) return PF_ERR_ARG;
  *cfg = c->mg_cfg;
  return PF_OK;
}

const char* pf_last_error(const pf_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

void* pf_stream(pf_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int pf_kernel_timing_mask(pf_ctx* c, uint32_t mask) {
  if (!c) return PF_ERR_ARG;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  c->ktime = mask & ((1u << PF_NKERNELS) - 1u);
  c->krec.clear();
  c->kev_used = 0;
  return PF_OK;
}

int pf_kernel_timing(pf_ctx* c, int32_t enable) {
  return pf_kernel_timing_mask(c, enable ? 0xffffffffu : 0u);
}

int pf_kernel_time(pf_ctx* c, int32_t kernel, int64_t* launches, double* total_ms) {
  if (!c || !launches || !total_ms || kernel < 0 || kernel >= PF_NKERNELS) return PF_ERR_ARG;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  int64_t n = 0;
  double t = 0.0;
  for (const auto& r : c->krec) {
    if (r.first != kernel) continue;
    float ms = 0.f;
    HIP_TRY(c, hipEventElapsedTime(&ms, c->kev[r.second], c->kev[r.second + 1]));
    t += ms;
    ++n;
  }
  *launches = n;
  *total_ms = t;
  return PF_OK;
}

int pf_copy_probe(pf_ctx* c, int64_t n, double* ms) {
  if (!c || n <= 0 || (n & 3) != 0 || n > (1LL << 31)) return PF_ERR_ARG;
  // its own buffers (allocated here: a measurement hook, not a step stage)
  float *src = nullptr, *dst = nullptr;
  HIP_TRY(c, dalloc(&src, n));
  HIP_TRY(c, dalloc(&dst, n));
  HIP_TRY(c, hipDeviceSynchronize());
  hipEvent_t a, b;
  HIP_TRY(c, hipEventCreate(&a));
  HIP_TRY(c, hipEventCreate(&b));
  const long long n4 = n / 4;
  long long blocks = std::min<long long>((n4 + 255) / 256, 1 << 16);
  float t = 1e30f;
  for (int rep = 0; rep < 4; ++rep) {  // first pass warms the TLB; keep the best
    HIP_TRY(c, hipEventRecord(a, c->stream));
    hipLaunchKernelGGL(k_copy_f32x4, dim3((unsigned)blocks), dim3(256), 0, c->stream,
                       reinterpret_cast<const float4*>(src), reinterpret_cast<float4*>(dst), n4);
    LAUNCH_CHECK(c);
    HIP_TRY(c, hipEventRecord(b, c->stream));
    HIP_TRY(c, hipEventSynchronize(b));
    float tr = 0.f;
    HIP_TRY(c, hipEventElapsedTime(&tr, a, b));
    t = std::min(t, tr);
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  HIP_TRY(c, hipFree(src));
  HIP_TRY(c, hipFree(dst));
  if (ms) *ms = t;
  return PF_OK;
}

int pf_synchronize(pf_ctx* c) {
  if (!c) return PF_ERR_ARG;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return PF_OK;
}

int pf_create(pf_ctx** out, const pf_grid_desc* d) {
  if (!out || !d) return PF_ERR_ARG;
  *out = nullptr;
  if (d->nx < 3 || d->ny < 3 || d->nz < 3 || !(d->dx > 0.0)) return PF_ERR_ARG;
  long long ncell_est = (long long)d->nx * d->ny * d->nz;
  if (ncell_est > (1LL << 31)) return PF_ERR_ARG;  // cell keys and ranks are 32-bit
  pf_ctx* c = new pf_ctx();
  c->mg_cfg = mg_default_cfg();
  HIP_TRY(c, hipGetDevice(&c->device));
  HIP_TRY(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  c->geom = Geom{d->nx, d->ny, d->nz, d->dx, d->ox, d->oy, d->oz};
  c->du = make_dims(d->nx + 1, d->ny, d->nz);
  c->dv = make_dims(d->nx, d->ny + 1, d->nz);
  c->dw = make_dims(d->nx, d->ny, d->nz + 1);
  c->dc = make_dims(d->nx, d->ny, d->nz);
  const Dims* fd[F_N] = {&c->du, &c->dv, &c->dw, &c->du, &c->dv, &c->dw, &c->dc, &c->dc, &c->du, &c->dv, &c->dw};
  for (int f = 0; f < F_N; ++f) {
    if (f != F_MK) HIP_TRY(c, dalloc(&c->f[f], fd[f]->nblocks * kBlockVol));
    HIP_TRY(c, dalloc(&c->act[f], fd[f]->nblocks));
  }
  const long long ncell = c->dc.nblocks * kBlockVol;
  HIP_TRY(c, dalloc(&c->marker, ncell));
  HIP_TRY(c, dalloc(&c->cell_count, ncell));
  HIP_TRY(c, dalloc(&c->cell_start, ncell + 1));
  c->color_cap = (long long)((c->dc.nbx + 1) / 2) * ((c->dc.nby + 1) / 2) * ((c->dc.nbz + 1) / 2);
  HIP_TRY(c, dalloc(&c->color_list, 8 * c->color_cap));
  HIP_TRY(c, dalloc(&c->color_cnt, 8));
  long long maxblk = std::max(std::max(c->du.nblocks, c->dv.nblocks), std::max(c->dw.nblocks, c->dc.nblocks));
  HIP_TRY(c, dalloc(&c->pblk, maxblk));
  HIP_TRY(c, dalloc(&c->pblk_prev, maxblk));
  HIP_TRY(c, dalloc(&c->blk_flag32, maxblk));
  c->morton_cap = morton_codes(c->dc) + 1;
  HIP_TRY(c, dalloc(&c->blk_scan, std::max(maxblk + 1, c->morton_cap)));
  HIP_TRY(c, dalloc(&c->morton_flags, c->morton_cap));
  float** f32[] = {&c->x32, &c->r32, &c->z32, &c->pA, &c->pB, &c->Ap, &c->cd32, &c->cx32, &c->cy32, &c->cz32, &c->invd32};
  for (float** p : f32) HIP_TRY(c, dalloc(p, ncell));
  double** f64[] = {&c->b64, &c->x64, &c->r64, &c->cd64, &c->cx64, &c->cy64, &c->cz64};
  for (double** p : f64) HIP_TRY(c, dalloc(p, ncell));
  HIP_TRY(c, dalloc(&c->sc, 1));
  HIP_TRY(c, dalloc(&c->partial, (long long)kMaxPartials * 4));
  HIP_TRY(c, dalloc(&c->red, 16));
  HIP_TRY(c, hipHostMalloc((void**)&c->red_host, sizeof(double) * 16));
  HIP_TRY(c, hipHostMalloc((void**)&c->done_host, sizeof(int) * 16));
  HIP_TRY(c, dalloc(&c->u32_dev, 64));
  HIP_TRY(c, hipHostMalloc((void**)&c->u32_host, sizeof(uint32_t) * 64));
  HIP_TRY(c, dalloc(&c->esc_dev, 8));
  HIP_TRY(c, dalloc(&c->vmax_dev, 1));
  c->scan_tile_cap = (std::max(ncell, maxblk) + 4095) / 4096 + 1;
  HIP_TRY(c, dalloc(&c->scan_tiles, c->scan_tile_cap));
  long long maxnodes = maxblk * kBlockVol;
  HIP_TRY(c, dalloc(&c->stage_f, maxnodes));
  c->stage_cap = maxnodes;
  for (int e = 0; e <= PF_NSTAGES; ++e) HIP_TRY(c, hipEventCreate(&c->ev[e]));
  // cell starts must be valid before the first binning (empty particle set)
  c->binned = false;
  *out = c;
  return PF_OK;
}

int pf_destroy(pf_ctx* c) {
  if (!c) return PF_OK;
  hipStreamSynchronize(c->stream);
  for (int f = 0; f < F_N; ++f) { hipFree(c->f[f]); hipFree(c->act[f]); }
  hipFree(c->marker); hipFree(c->cell_count); hipFree(c->cell_start);
  hipFree(c->color_list); hipFree(c->color_cnt); hipFree(c->pblk); hipFree(c->pblk_prev);
  hipFree(c->blk_flag32); hipFree(c->blk_scan); hipFree(c->morton_flags);
  PBuf* bufs[2] = {&c->pa, &c->pb};
  for (PBuf* b : bufs) {
    hipFree(b->x); hipFree(b->y); hipFree(b->z); hipFree(b->u); hipFree(b->v); hipFree(b->w);
    hipFree(b->type); hipFree(b->id); hipFree(b->vol);
  }
  hipFree(c->key); hipFree(c->rank);
  float* f32[] = {c->x32, c->r32, c->z32, c->pA, c->pB, c->Ap, c->cd32, c->cx32, c->cy32, c->cz32, c->invd32};
  for (float* p : f32) hipFree(p);
  double* f64[] = {c->b64, c->x64, c->r64, c->cd64, c->cx64, c->cy64, c->cz64};
  for (double* p : f64) hipFree(p);
  dfree(c->mr_lvl); dfree(c->mr_vol64); dfree(c->mr_pfull); dfree(c->mr_bbox); dfree(c->mr_leaves_dev);
  dfree(c->mrf); dfree(c->mrc); dfree(c->mr_pin_dev);
  dfree(c->iface); dfree(c->iface_bad);
  dfree(c->pa_alive); dfree(c->pa_key); dfree(c->pa_by_id); dfree(c->pa_new_id);
  dfree(c->pa_cflag); dfree(c->pa_d1); dfree(c->pa_d2); dfree(c->pa_wanted); dfree(c->pa_wscan); dfree(c->pa_cnt);
  for (auto& m : c->mg) {
    dfree(m.cd); dfree(m.cx); dfree(m.cy); dfree(m.cz); dfree(m.inv); dfree(m.x); dfree(m.r); dfree(m.t);
  }
  c->mg.clear();
  dfree(c->mg_t0);
  dfree(c->mg_cs);
  for (auto& h : c->mg_h) dfree(h);
  dfree(c->mg_mx);
  dfree(c->mg_unscale);
  if (c->mr_pin_host) hipHostFree(c->mr_pin_host);
  for (int a = 0; a < 3; ++a) { dfree(c->mr_valid[a]); dfree(c->mr_cf32[a]); dfree(c->mr_cf64[a]); }
  if (c->mr_bbox_host) hipHostFree(c->mr_bbox_host);
  dfree(c->fbeta); dfree(c->fphi); dfree(c->tiles); dfree(c->tile_flags); dfree(c->snaps);
  hipFree(c->sc); hipFree(c->partial); hipFree(c->red);
  hipHostFree(c->red_host); hipHostFree(c->done_host); hipHostFree(c->u32_host);
  hipFree(c->u32_dev); hipFree(c->esc_dev); hipFree(c->vmax_dev);
  hipFree(c->scan_tiles); hipFree(c->stage_f);
  for (int e = 0; e <= PF_NSTAGES; ++e)
    if (c->ev[e]) hipEventDestroy(c->ev[e]);
  for (hipEvent_t e : c->kev) hipEventDestroy(e);
  hipStreamDestroy(c->stream);
  delete c;
  return PF_OK;
}

int pf_particles_upload(pf_ctx* c, int64_t n, const float* x, const float* y, const float* z, const float* u,
                        const float* v, const float* w, const uint8_t* type, const float* volume) {
  if (!c || n < 0) return PF_ERR_ARG;
  if (n > 0 && (!x || !y || !z || !u || !v || !w || !type)) return fail(c, PF_ERR_ARG, "null particle array");
  if (n >= (1LL << 32)) return fail(c, PF_ERR_ARG, "particle count exceeds 32-bit ids");
  int st = grow_particles(c, n, volume != nullptr);
  if (st) return st;
  c->np = n;
  c->has_vol = volume != nullptr;
  if (n > 0) {
    size_t fb = sizeof(float) * (size_t)n;
    HIP_TRY(c, hipMemcpyAsync(c->pa.x, x, fb, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->pa.y, y, fb, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->pa.z, z, fb, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->pa.u, u, fb, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->pa.v, v, fb, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->pa.w, w, fb, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->pa.type, type, (size_t)n, hipMemcpyHostToDevice, c->stream));
    if (volume) HIP_TRY(c, hipMemcpyAsync(c->pa.vol, volume, fb, hipMemcpyHostToDevice, c->stream));
    std::vector<uint32_t> ids((size_t)n);
    for (int64_t p = 0; p < n; ++p) ids[(size_t)p] = (uint32_t)p;
    HIP_TRY(c, hipMemcpyAsync(c->pa.id, ids.data(), sizeof(uint32_t) * (size_t)n, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, fill_u32(c->stream, c->u32_dev, 0u, 1));
    hipLaunchKernelGGL(k_validate, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, (long long)n, c->pa, c->geom,
                       c->u32_dev);
    LAUNCH_CHECK(c);
    HIP_TRY(c, hipMemcpyAsync(c->u32_host, c->u32_dev, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (c->u32_host[0]) {
      c->np = 0;
      return fail(c, PF_ERR_ARG, "particle position outside the grid's cell range");
    }
  }
  double lv = 0.0;
  for (int64_t p = 0; p < n; ++p)
    if (type[p] == 0) lv += volume ? (double)volume[p] : 1.0;
  c->liquid_volume = lv;
  c->binned = false;
  return PF_OK;
}

int pf_particles_download(pf_ctx* c, float* x, float* y, float* z, float* u, float* v, float* w) {
  if (!c || !x || !y || !z || !u || !v || !w) return PF_ERR_ARG;
  const long long n = c->np;
  if (n == 0) return PF_OK;
  PBuf t = c->pb;
  hipLaunchKernelGGL(k_unsort, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, n, c->pa, t.x, t.y, t.z, t.u, t.v,
                     t.w);
  LAUNCH_CHECK(c);
  size_t fb = sizeof(float) * (size_t)n;
  HIP_TRY(c, hipMemcpyAsync(x, t.x, fb, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipMemcpyAsync(y, t.y, fb, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipMemcpyAsync(z, t.z, fb, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipMemcpyAsync(u, t.u, fb, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipMemcpyAsync(v, t.v, fb, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipMemcpyAsync(w, t.w, fb, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return PF_OK;
}

int64_t pf_particle_count(const pf_ctx* c) { return c ? c->np : -1; }

int pf_particles_download_ext(pf_ctx* c, uint8_t* type, float* volume) {
  if (!c) return PF_ERR_ARG;
  const long long n = c->np;
  if (n == 0 || (!type && !volume)) return PF_OK;
  PBuf t = c->pb;
  hipLaunchKernelGGL(k_unsort_ext, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, n, c->pa, c->has_vol, t.type,
                     t.x);
  LAUNCH_CHECK(c);
  if (type) HIP_TRY(c, hipMemcpyAsync(type, t.type, (size_t)n, hipMemcpyDeviceToHost, c->stream));
  if (volume) HIP_TRY(c, hipMemcpyAsync(volume, t.x, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return PF_OK;
}

double pf_liquid_volume(const pf_ctx* c) { return c ? c->liquid_volume : 0.0; }

int pf_scene_generate(pf_ctx* c, int32_t kind, uint64_t seed) {
  if (!c || kind < 0 || kind > 2) return PF_ERR_ARG;
  const Geom& g = c->geom;
  // filled cells form the box [1, ie) x [1, je) x [1, nz-1): cell-centre tests
  // y < 0.6 (bubble, dam-break) and x < 0.4 (dam-break)
  int ie = g.nx - 1, je = g.ny - 1;
  if (kind == 1 || kind == 2) {
    je = 1;
    while (je < g.ny - 1 && (je + 0.5) / g.ny < 0.6) ++je;
  }
  if (kind == 2) {
    ie = 1;
    while (ie < g.nx - 1 && (ie + 0.5) / g.nx < 0.4) ++ie;
  }
  long long ncells = (long long)(ie - 1) * (je - 1) * (g.nz - 2);
  long long n = 8 * ncells;
  int st = grow_particles(c, n, false);
  if (st) return st;
  c->np = n;
  c->has_vol = false;
  if (n > 0) {
    hipLaunchKernelGGL(k_scene, dim3(grid_for(ncells, 256)), dim3(256), 0, c->stream, (int)kind, g, ie, je, seed, c->pa);
    LAUNCH_CHECK(c);
  }
  // liquid volume (type 0 count): read back the types once
  std::vector<uint8_t> t((size_t)n);
  if (n > 0) {
    HIP_TRY(c, hipMemcpyAsync(t.data(), c->pa.type, (size_t)n, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
  }
  long long liq = 0;
  for (uint8_t q : t) liq += q == 0;
  c->liquid_volume = (double)liq;
  c->binned = false;
  return PF_OK;
}

int pf_mark_cells(pf_ctx* c) {
  if (!c) return PF_ERR_ARG;
  return run_mark(c);
}

int pf_p2g(pf_ctx* c, const pf_phase* ph, double Vp) {
  if (!c || !ph) return PF_ERR_ARG;
  int st = run_p2g_splat(c, ph, Vp);
  if (st) return st;
  return run_face_ops(c, FO_NORMALIZE, 0.f);
}

int pf_p2g_single(pf_ctx* c) {
  if (!c) return PF_ERR_ARG;
  int st = run_p2g_splat(c, nullptr, 1.0, true);
  if (st) return st;
  return run_face_ops(c, FO_NORMALIZE, 0.f);
}

int pf_save_velocity(pf_ctx* c) {
  if (!c) return PF_ERR_ARG;
  return run_face_ops(c, FO_SAVE, 0.f);
}

int pf_apply_gravity(pf_ctx* c, double dt, double gravity) {
  if (!c) return PF_ERR_ARG;
  return run_face_ops(c, FO_GRAVITY, (float)(dt * gravity));
}

int pf_apply_walls(pf_ctx* c) {
  if (!c) return PF_ERR_ARG;
  return run_face_ops(c, FO_WALLS, 0.f);
}

int pf_project(pf_ctx* c, const pf_phase* ph, double dt, const pf_solve_cfg* cfg, double div_correction,
               pf_solve_stats* stats) {
  if (!c || !ph || !cfg) return PF_ERR_ARG;
  return run_project(c, ph, dt, cfg, div_correction, stats, false, 0, 0, 0);
}

int pf_project_single(pf_ctx* c, double dt, const pf_solve_cfg* cfg, pf_solve_stats* stats) {
  if (!c || !cfg) return PF_ERR_ARG;
  return run_project(c, nullptr, dt, cfg, 0.0, stats, false, 0, 0, 0, true);
}

int pf_g2p(pf_ctx* c, double aL, double aG) {
  if (!c) return PF_ERR_ARG;
  return run_g2p_advect(c, true, false, aL, aG, 0.0, 2);
}

int pf_advect(pf_ctx* c, double dt, int32_t order, pf_escape_stats* stats) {
  if (!c) return PF_ERR_ARG;
  HIP_TRY(c, fill_u32(c->stream, reinterpret_cast<uint32_t*>(c->esc_dev), 0u, 8));
  int st = run_g2p_advect(c, false, true, 0.0, 0.0, dt, order);
  if (st) return st;
  if (stats) {
    HIP_TRY(c, hipMemcpyAsync(stats, c->esc_dev, sizeof(int32_t) * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
  }
  return PF_OK;
}

// ---------------------------------------------------------------- particle adaptivity
namespace {

// a particle buffer of cap entries (with volumes)
int alloc_pbuf(pf_ctx* c, PBuf& b, long long cap) {
  b = PBuf{};
  HIP_TRY(c, dalloc(&b.x, cap)); HIP_TRY(c, dalloc(&b.y, cap)); HIP_TRY(c, dalloc(&b.z, cap));
  HIP_TRY(c, dalloc(&b.u, cap)); HIP_TRY(c, dalloc(&b.v, cap)); HIP_TRY(c, dalloc(&b.w, cap));
  HIP_TRY(c, dalloc(&b.type, cap)); HIP_TRY(c, dalloc(&b.id, cap)); HIP_TRY(c, dalloc(&b.vol, cap));
  return PF_OK;
}

void free_pbuf(PBuf& b) {
  hipFree(b.x); hipFree(b.y); hipFree(b.z); hipFree(b.u); hipFree(b.v); hipFree(b.w);
  hipFree(b.type); hipFree(b.id); hipFree(b.vol);
  b = PBuf{};
}

// room in the scan's tile buffer for n entries
int ensure_scan_capacity(pf_ctx* c, long long n) {
  const long long need = (n + kScanTile - 1) / kScanTile + 1;
  if (need <= c->scan_tile_cap) return PF_OK;
  hipFree(c->scan_tiles);
  c->scan_tiles = nullptr;
  c->scan_tile_cap = 0;
  HIP_TRY(c, dalloc(&c->scan_tiles, need));
  c->scan_tile_cap = need;
  return PF_OK;
}

// the liquid particle volume after a change of the particle set (FP64 partial sums
// per workgroup, added in workgroup order)
int update_liquid_volume(pf_ctx* c) {
  if (c->np == 0) {
    c->liquid_volume = 0.0;
    return PF_OK;
  }
  const unsigned nwg = std::min<unsigned>(grid_for(c->np, 256), 1024u);
  hipLaunchKernelGGL(k_pa_liquid_volume, dim3(nwg), dim3(256), 0, c->stream, c->pa, c->np, c->has_vol, c->partial);
  const CombineOps ops{1, {0}};
  hipLaunchKernelGGL(k_iface_combine, dim3(1), dim3(64), 0, c->stream, c->partial, (int)nwg, ops, c->red);
  LAUNCH_CHECK(c);
  HIP_TRY(c, hipMemcpyAsync(c->red_host, c->red, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  c->liquid_volume = c->red_host[0];
  return PF_OK;
}

// applyParticleAdaptivity on the context's particles (see pf_adapt.hip)
int run_adaptivity(pf_ctx* c, const pf_adapt_cfg* cfg, pf_adapt_stats* out) {
  pf_adapt_stats S;
  std::memset(&S, 0, sizeof(S));
  const bool any = cfg->narrow_band_air || cfg->gas_particle_coarsening || cfg->liquid_particle_coarsening ||
                   cfg->liquid_particle_refill;
  if (!any || c->np == 0) {
    if (out) *out = S;
    return PF_OK;
  }
  int st = ensure_binned(c);
  if (st) return st;
  const long long np = c->np, ncell = c->dc.nblocks * kBlockVol;
  const long long nlin = (long long)c->geom.nx * c->geom.ny * c->geom.nz;
  if (np > c->pa_cap) {
    dfree(c->pa_alive); dfree(c->pa_key); dfree(c->pa_by_id); dfree(c->pa_new_id);
    const long long cap = np + np / 4;
    HIP_TRY(c, dalloc(&c->pa_alive, cap));
    HIP_TRY(c, dalloc(&c->pa_key, cap));
    HIP_TRY(c, dalloc(&c->pa_by_id, cap + 1));
    HIP_TRY(c, dalloc(&c->pa_new_id, cap + 1));
    c->pa_cap = cap;
  }
  if (!c->pa_cflag) {
    HIP_TRY(c, dalloc(&c->pa_cflag, ncell));
    HIP_TRY(c, dalloc(&c->pa_d1, ncell));
    HIP_TRY(c, dalloc(&c->pa_d2, ncell));
    HIP_TRY(c, dalloc(&c->pa_wanted, nlin + 1));
    HIP_TRY(c, dalloc(&c->pa_wscan, nlin + 1));
    HIP_TRY(c, dalloc(&c->pa_cnt, 32));
  }
  // the coarsening and refill stages write volumes: start from 1.0 where none were given
  if (!c->has_vol && (cfg->gas_particle_coarsening || cfg->liquid_particle_coarsening || cfg->liquid_particle_refill)) {
    st = grow_particles(c, np, true);
    if (st) return st;
    HIP_TRY(c, fill_f32(c->stream, c->pa.vol, 1.0f, np));
    c->has_vol = true;
  }
  st = ensure_scan_capacity(c, std::max(np, nlin));
  if (st) return st;
  HIP_TRY(c, fill_u8(c->stream, c->pa_alive, 1, np));
  HIP_TRY(c, fill_u32(c->stream, c->pa_cnt, 0u, 32));
  AdaptArgs a;
  a.g = c->geom;
  a.dc = c->dc;
  a.P = c->pa;
  a.np = np;
  a.start = c->cell_start;
  a.alive = c->pa_alive;
  a.key = c->pa_key;
  a.cflag = c->pa_cflag;
  a.cnt = c->pa_cnt;
  const dim3 gc(grid_for(ncell, 256)), gp(grid_for(np, 256)), b(256);
  // box dilation of the bits `mask` of cflag with radius r, into pa_d1
  auto dilate = [&](uint8_t mask, int r) -> int {
    hipLaunchKernelGGL(k_pa_dilate, gc, b, 0, c->stream, c->dc, c->geom, 0, r, mask, c->pa_cflag, c->pa_d1);
    hipLaunchKernelGGL(k_pa_dilate, gc, b, 0, c->stream, c->dc, c->geom, 1, r, (uint8_t)1, c->pa_d1, c->pa_d2);
    hipLaunchKernelGGL(k_pa_dilate, gc, b, 0, c->stream, c->dc, c->geom, 2, r, (uint8_t)1, c->pa_d2, c->pa_d1);
    LAUNCH_CHECK(c);
    return PF_OK;
  };
  uint32_t cnt[32];
  auto counters = [&]() -> int {
    HIP_TRY(c, hipMemcpyAsync(c->u32_host, c->pa_cnt, sizeof(uint32_t) * PA_NCOUNTERS, hipMemcpyDeviceToHost,
                              c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    std::memcpy(cnt, c->u32_host, sizeof(uint32_t) * PA_NCOUNTERS);
    return PF_OK;
  };
  // applyNarrowBandAir (:112-167): nothing at all when no cell holds liquid
  if (cfg->narrow_band_air) {
    hipLaunchKernelGGL(k_pa_flags, gc, b, 0, c->stream, a, (int)PA_NB_LIQUID_CELLS);
    LAUNCH_CHECK(c);
    st = counters();
    if (st) return st;
    if (cnt[PA_NB_LIQUID_CELLS] > 0) {
      st = dilate(1, std::max(0, cfg->narrow_band_air_radius));
      if (st) return st;
      hipLaunchKernelGGL(k_pa_narrow_band, gp, b, 0, c->stream, a, (const uint8_t*)c->pa_d1);
      LAUNCH_CHECK(c);
    }
  }
  // applyGasParticleCoarsening, then applyTypedParticleCoarsening for liquid (:169-241)
  const int types[2] = {1, 0};
  const int enabled[2] = {cfg->gas_particle_coarsening, cfg->liquid_particle_coarsening};
  const int targets[2] = {cfg->gas_particles_per_cell_target, cfg->liquid_particles_per_cell_target};
  const uint32_t seeds[2] = {cfg->gas_particle_coarsening_seed, cfg->liquid_particle_coarsening_seed};
  const int slot0[2] = {PA_GAS_REMOVED, PA_LIQ_REMOVED};
  for (int t = 0; t < 2; ++t) {
    if (!enabled[t]) continue;
    hipLaunchKernelGGL(k_pa_keys, gp, b, 0, c->stream, a, types[t], seeds[t], slot0[t]);
    hipLaunchKernelGGL(k_pa_coarsen, gc, b, 0, c->stream, a, types[t], std::max(1, targets[t]), slot0[t]);
    LAUNCH_CHECK(c);
  }
  // the survivors' new ids (their rank in id order)
  hipLaunchKernelGGL(k_pa_alive_by_id, gp, b, 0, c->stream, a, c->pa_by_id);
  LAUNCH_CHECK(c);
  HIP_TRY(c, scan_u32(*c, c->pa_by_id, c->pa_new_id, np));
  HIP_TRY(c, hipMemcpyAsync(c->u32_host + 8, c->pa_new_id + np, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
  // applyLiquidParticleRefill (:243-374)
  RefillArgs r;
  r.target = std::max(1, cfg->liquid_refill_particles_per_cell_target);
  r.interface_only = cfg->liquid_particle_refill_interface_only;
  r.seed = cfg->liquid_particle_refill_seed;
  r.budget = cfg->liquid_particle_refill_max_added_per_step > 0 ? cfg->liquid_particle_refill_max_added_per_step
                                                                : 2147483647LL;
  r.near_gas = c->pa_d1;
  r.wanted = c->pa_wanted;
  r.wscan = c->pa_wscan;
  long long added = 0;
  if (cfg->liquid_particle_refill) {
    hipLaunchKernelGGL(k_pa_flags, gc, b, 0, c->stream, a, -1);
    LAUNCH_CHECK(c);
    st = dilate(2, std::max(0, cfg->liquid_particle_refill_interface_radius));
    if (st) return st;
    HIP_TRY(c, fill_u32(c->stream, c->pa_wanted, 0u, nlin));
    hipLaunchKernelGGL(k_pa_refill_count, gc, b, 0, c->stream, a, r);
    LAUNCH_CHECK(c);
    HIP_TRY(c, scan_u32(*c, c->pa_wanted, c->pa_wscan, nlin));
    HIP_TRY(c, hipMemcpyAsync(c->u32_host + 9, c->pa_wscan + nlin, sizeof(uint32_t), hipMemcpyDeviceToHost,
                              c->stream));
  }
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  const long long n_kept = c->u32_host[8];
  if (cfg->liquid_particle_refill) added = std::min<long long>(r.budget, (long long)c->u32_host[9]);
  const long long n_new = n_kept + added;
  if (n_new >= (1LL << 32)) return fail(c, PF_ERR_ARG, "particle count exceeds 32-bit ids");
  // the output buffer: pb, or a larger new one
  PBuf outb = c->pb;
  bool fresh = false;
  if (n_new > c->pcap) {
    st = alloc_pbuf(c, outb, n_new + n_new / 4);
    if (st) return st;
    fresh = true;
  }
  r.n_kept = n_kept;
  r.out = outb;
  if (cfg->liquid_particle_refill) {
    hipLaunchKernelGGL(k_pa_refill_apply, gc, b, 0, c->stream, a, r);
    LAUNCH_CHECK(c);
  }
  hipLaunchKernelGGL(k_pa_compact, gp, b, 0, c->stream, a, (const uint32_t*)c->pa_new_id, outb, c->has_vol);
  LAUNCH_CHECK(c);
  st = counters();
  if (st) return st;
  if (fresh) {
    const long long cap = n_new + n_new / 4;
    free_pbuf(c->pa);
    free_pbuf(c->pb);
    c->pa = outb;
    st = alloc_pbuf(c, c->pb, cap);
    if (st) return st;
    hipFree(c->key); hipFree(c->rank);
    c->key = nullptr;
    c->rank = nullptr;
    HIP_TRY(c, dalloc(&c->key, cap));
    HIP_TRY(c, dalloc(&c->rank, cap));
    c->pcap = cap;
  } else {
    std::swap(c->pa, c->pb);
  }
  c->np = n_new;
  c->binned = false;
  st = update_liquid_volume(c);
  if (st) return st;
  S.narrow_band_air_liquid_cells = (int32_t)cnt[PA_NB_LIQUID_CELLS];
  S.narrow_band_air_gas_before = (int32_t)cnt[PA_NB_GAS_BEFORE];
  S.narrow_band_air_gas_after = (int32_t)cnt[PA_NB_GAS_AFTER];
  S.narrow_band_air_removed = S.narrow_band_air_gas_before - S.narrow_band_air_gas_after;
  S.gas_removed = (int32_t)cnt[PA_GAS_REMOVED];
  S.gas_cells = (int32_t)cnt[PA_GAS_CELLS];
  S.gas_overfull_cells = (int32_t)cnt[PA_GAS_OVERFULL];
  S.gas_before = (int32_t)cnt[PA_GAS_BEFORE];
  S.gas_after = (int32_t)cnt[PA_GAS_AFTER];
  S.liquid_removed = (int32_t)cnt[PA_LIQ_REMOVED];
  S.liquid_cells = (int32_t)cnt[PA_LIQ_CELLS];
  S.liquid_overfull_cells = (int32_t)cnt[PA_LIQ_OVERFULL];
  S.liquid_before = (int32_t)cnt[PA_LIQ_BEFORE];
  S.liquid_after = (int32_t)cnt[PA_LIQ_AFTER];
  if (cfg->liquid_particle_refill) {
    S.refill_added = (int32_t)added;
    S.refill_cells = (int32_t)cnt[PA_RF_CELLS];
    S.refill_interface_cells = (int32_t)cnt[PA_RF_INTERFACE];
    S.refill_underfull_cells = (int32_t)cnt[PA_RF_UNDERFULL];
    S.refill_budget_limited = cnt[PA_RF_BUDGET] > 0 ? 1 : 0;
    S.refill_before = (int32_t)cnt[PA_RF_BEFORE];
    S.refill_after = (int32_t)(cnt[PA_RF_BEFORE] + added);
  }
  if (out) *out = S;
  return PF_OK;
}

// updateVolumeCorrectionStats (sparse_sim3d_tp.cpp:368-379): c_div from the liquid
// volume; the given c_div when the correction is off
double volume_correction(const pf_ctx* c, int on, double strength, double target, double Vp, double dt,
                         double given, double* liquid_volume) {
  const double lv = c->liquid_volume * Vp;
  *liquid_volume = lv;
  if (!on) return given;
  if (strength > 0.0 && target > 0.0 && dt > 0.0) return strength * (target - lv) / (dt * target);
  return 0.0;
}

}  // namespace

int pf_particle_adaptivity(pf_ctx* c, const pf_adapt_cfg* cfg, pf_adapt_stats* stats) {
  if (!c || !cfg) return PF_ERR_ARG;
  return run_adaptivity(c, cfg, stats);
}

void pf_adapt_defaults(pf_adapt_cfg* cfg) {
  if (!cfg) return;
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->narrow_band_air_radius = 2;
  cfg->gas_particles_per_cell_target = 4;
  cfg->gas_particle_coarsening_seed = 0x9e3779b9u;
  cfg->liquid_particles_per_cell_target = 4;
  cfg->liquid_particle_coarsening_seed = 0x7f4a7c15u;
  cfg->liquid_particle_refill_interface_radius = 1;
  cfg->liquid_refill_particles_per_cell_target = 4;
  cfg->liquid_particle_refill_seed = 0x6a09e667u;
}

// ---------------------------------------------------------------- interface diagnostics, CSF
namespace {

constexpr int kIfaceFields = 9;

// the cell phase field, normals and curvature from the current face masses and
// markers (interface_diagnostics3d.cpp:137-208); mr: sample the multires leaf cells
int iface_fields(pf_ctx* c, const pf_phase* ph, IfaceArgs& a, bool mr) {
  const long long n = c->dc.nblocks * kBlockVol;
  if (!c->iface) {
    HIP_TRY(c, dalloc(&c->iface, n * kIfaceFields));
    HIP_TRY(c, dalloc(&c->iface_bad, 1));
  }
  a.dc = c->dc; a.du = c->du; a.dv = c->dv; a.dw = c->dw;
  a.g = c->geom;
  a.mu = c->f[F_MU]; a.mv = c->f[F_MV]; a.mw = c->f[F_MW];
  a.marker = c->marker;
  a.ph = phase_k(ph);
  a.phi = c->iface;
  a.gmag = c->iface + n;
  for (int d = 0; d < 3; ++d) a.nrm[d] = c->iface + (2 + d) * n;
  a.kap = c->iface + 5 * n;
  a.lvl = mr ? c->mr_lvl : nullptr;
  const dim3 g(grid_for(n, 256));
  const long long ki = kbegin(c, PF_K_INTERFACE);
  hipLaunchKernelGGL(k_iface_phi, g, dim3(256), 0, c->stream, a);
  hipLaunchKernelGGL(k_iface_normal, g, dim3(256), 0, c->stream, a);
  hipLaunchKernelGGL(k_iface_curvature, g, dim3(256), 0, c->stream, a);
  kend(c, ki);
  LAUNCH_CHECK(c);
  return PF_OK;
}

unsigned iface_groups(const pf_ctx* c) {
  return std::min<unsigned>(grid_for(c->dc.nblocks * kBlockVol, 256), 1024u);
}

// accumulate and finishStats (interface_diagnostics3d.cpp:26-66) over the FLUID cells
int iface_reduce(pf_ctx* c, const IfaceArgs& a, pf_iface_stats* out) {
  const unsigned nwg = iface_groups(c);
  const CombineOps ops{kIfaceSlots, {0, 1, 2, 0, 0, 0, 2, 0, 2, 0}};
  const long long ki = kbegin(c, PF_K_INTERFACE);
  hipLaunchKernelGGL(k_iface_diag, dim3(nwg), dim3(256), 0, c->stream, a, c->partial);
  hipLaunchKernelGGL(k_iface_combine, dim3(1), dim3(64), 0, c->stream, c->partial, (int)nwg, ops, c->red);
  kend(c, ki);
  LAUNCH_CHECK(c);
  HIP_TRY(c, hipMemcpyAsync(c->red_host, c->red, sizeof(double) * kIfaceSlots, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  const double* v = c->red_host;
  pf_iface_stats S;
  std::memset(&S, 0, sizeof(S));
  S.finite = v[9] == 0.0 ? 1 : 0;
  S.sample_cells = (int32_t)v[0];
  S.interface_cells = (int32_t)v[4];
  if (S.sample_cells > 0) {
    S.phi_min = v[1];
    S.phi_max = v[2];
    S.phi_mean = v[3] / v[0];
  }
  S.grad_max = v[6];
  S.curvature_abs_max = v[8];
  if (S.interface_cells > 0) {
    S.grad_mean = v[5] / v[4];
    S.curvature_abs_mean = v[7] / v[4];
  }
  S.surface_tension_candidate = S.finite && S.interface_cells > 0 && S.grad_max > kIfaceGradEps;
  *out = S;
  return PF_OK;
}

// initSurfaceStats (:87-102), the CSF increments (:404-440), finishSurfaceStats (:124-131)
int surface_apply(pf_ctx* c, const pf_phase* ph, const IfaceArgs& a, double dt, double strength, double maxd,
                  int radius, pf_surface_stats* out) {
  pf_surface_stats S;
  std::memset(&S, 0, sizeof(S));
  S.finite = 1;
  S.enabled = strength > 0.0 && dt > 0.0;
  S.strength = strength;
  S.max_delta_speed_limit = maxd;
  S.curvature_smoothing_radius = std::min(std::max(radius, 0), 3);
  const double rmin = std::min(ph->rho_l, ph->rho_g), dx = c->geom.dx;
  const double pi = 3.14159265358979323846;
  S.capillary_dt_limit = (dx <= 0.0 || strength <= 0.0 || rmin <= 0.0 || !std::isfinite(rmin))
                             ? 0.0 : std::sqrt((rmin * dx * dx * dx) / (2.0 * pi * strength));
  S.capillary_stable = !S.enabled || S.capillary_dt_limit <= 0.0 || dt <= S.capillary_dt_limit + 1e-12;
  if (S.enabled) {
    const long long n = c->dc.nblocks * kBlockVol;
    CsfArgs cs;
    cs.f = a;
    cs.dt = dt;
    cs.strength = strength;
    cs.maxd = maxd;
    cs.radius = S.curvature_smoothing_radius;
    for (int d = 0; d < 3; ++d) cs.delta[d] = c->iface + (6 + d) * n;
    cs.first_bad = c->iface_bad;
    const unsigned nwg = iface_groups(c);
    const CombineOps ops{kCsfSlots, {0, 0, 2, 0, 2, 0, 2, 0}};
    HIP_TRY(c, fill_u32(c->stream, reinterpret_cast<uint32_t*>(c->iface_bad), 0xffffffffu, 2));
    const long long ki = kbegin(c, PF_K_SURFACE);
    hipLaunchKernelGGL(k_csf_delta, dim3(nwg), dim3(256), 0, c->stream, cs, c->partial);
    hipLaunchKernelGGL(k_iface_combine, dim3(1), dim3(64), 0, c->stream, c->partial, (int)nwg, ops, c->red);
    const Dims* fd[3] = {&c->du, &c->dv, &c->dw};
    for (int ax = 0; ax < 3; ++ax)
      hipLaunchKernelGGL(k_csf_faces, dim3(grid_for(fd[ax]->nblocks * kBlockVol, 256)), dim3(256), 0, c->stream,
                         cs, ax, c->f[F_U + ax], static_cast<const float*>(c->f[F_MU + ax]));
    kend(c, ki);
    LAUNCH_CHECK(c);
    HIP_TRY(c, hipMemcpyAsync(c->red_host, c->red, sizeof(double) * kCsfSlots, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    const double* v = c->red_host;
    S.applied_cells = (int32_t)v[0];
    S.finite = v[7] == 0.0 ? 1 : 0;
    S.max_delta_speed = v[2];
    S.raw_curvature_abs_max = v[4];
    S.smoothed_curvature_abs_max = v[6];
    S.mean_delta_speed = v[1];
    S.raw_curvature_abs_mean = v[3];
    S.smoothed_curvature_abs_mean = v[5];
    if (S.applied_cells > 0) {
      S.mean_delta_speed /= v[0];
      S.raw_curvature_abs_mean /= v[0];
      S.smoothed_curvature_abs_mean /= v[0];
    }
  }
  if (out) *out = S;
  return PF_OK;
}

// normalisation, the saved copy and gravity, then interface_diagnostics_last,
// surface tension and the walls (sparse_sim3d_tp.cpp:865-874,
// multires_sim3d_tp.cpp:683-694).  The face operations leave the masses the
// diagnostics read unchanged, so they fuse into one pass unless the CSF increments
// have to land between gravity and the walls.
int step_face_stage(pf_ctx* c, const pf_phase* ph, double dt, double gravity, bool mr, bool diag, bool csf,
                    double strength, double maxd, int radius, pf_step_stats* S) {
  int st = run_face_ops(c, FO_NORMALIZE | FO_SAVE | FO_GRAVITY | (csf ? 0 : FO_WALLS), (float)(dt * gravity));
  if (st || !(diag || csf)) return st;
  IfaceArgs ia;
  st = iface_fields(c, ph, ia, mr);
  if (st) return st;
  if (diag) {
    st = iface_reduce(c, ia, &S->iface);
    if (st) return st;
  }
  if (!csf) return PF_OK;
  st = surface_apply(c, ph, ia, dt, strength, maxd, radius, &S->surface);
  if (st) return st;
  return run_face_ops(c, FO_WALLS, 0.f);
}

// the defaults of InterfaceDiagnostics3D and SurfaceTensionStats3D
// (interface_diagnostics3d.h:7-36), kept where a part is skipped
void default_interface_stats(pf_step_stats* S) {
  S->iface.finite = 1;
  S->surface.finite = 1;
  S->surface.capillary_stable = 1;
}

int interface_entry(pf_ctx* c, const pf_phase* ph, bool mr, IfaceArgs& a) {
  if (mr && !c->mr) return fail(c, PF_ERR_STATE, "no multires layout (pf_mr_set_layout)");
  if (!mr && c->mr) return fail(c, PF_ERR_STATE, "a multires context: use the pf_mr_ interface entry points");
  return iface_fields(c, ph, a, mr);
}

}  // namespace

int pf_interface_diagnostics(pf_ctx* c, const pf_phase* ph, pf_iface_stats* out) {
  if (!c || !ph || !out) return PF_ERR_ARG;
  IfaceArgs a;
  int st = interface_entry(c, ph, false, a);
  if (st) return st;
  return iface_reduce(c, a, out);
}

int pf_apply_surface_tension(pf_ctx* c, const pf_phase* ph, double dt, double strength, double max_delta_speed,
                             int32_t smoothing_radius, pf_surface_stats* out) {
  if (!c || !ph) return PF_ERR_ARG;
  IfaceArgs a;
  int st = interface_entry(c, ph, false, a);
  if (st) return st;
  return surface_apply(c, ph, a, dt, strength, max_delta_speed, smoothing_radius, out);
}

int pf_mr_interface_diagnostics(pf_ctx* c, const pf_phase* ph, pf_iface_stats* out) {
  if (!c || !ph || !out) return PF_ERR_ARG;
  IfaceArgs a;
  int st = interface_entry(c, ph, true, a);
  if (st) return st;
  return iface_reduce(c, a, out);
}

int pf_mr_apply_surface_tension(pf_ctx* c, const pf_phase* ph, double dt, double strength, double max_delta_speed,
                                int32_t smoothing_radius, pf_surface_stats* out) {
  if (!c || !ph) return PF_ERR_ARG;
  IfaceArgs a;
  int st = interface_entry(c, ph, true, a);
  if (st) return st;
  return surface_apply(c, ph, a, dt, strength, max_delta_speed, smoothing_radius, out);
}

// ---------------------------------------------------------------- render-cache export
void pf_render_default_camera(int32_t nx, int32_t ny, int32_t nz, double dx, pf_render_camera* out) {
  if (!out) return;
  // defaultRenderCacheCamera3D, render_cache3d.h:48-62
  const double cx = 0.5 * nx * dx, cy = 0.5 * ny * dx, cz = 0.5 * nz * dx;
  const double span = std::max({nx * dx, ny * dx, nz * dx, dx});
  *out = pf_render_camera{{cx, cy + 0.2 * span, cz + 2.0 * span}, {cx, cy, cz}, {0.0, 1.0, 0.0}, 45.0, 0.05 * dx,
                          5.0 * span, 50.0};
}

int pf_write_render_cache(pf_ctx* c, const char* path, const pf_render_frame* m, int64_t* bytes) {
  if (!c || !m) return PF_ERR_ARG;
  if (!path || !path[0]) return fail(c, PF_ERR_ARG, "render cache path is empty");
  const Geom& g = c->geom;
  const pf_render_camera& cam = m->camera;
  // writeHeader's checks, render_cache3d.h:373-381
  const bool ok = std::isfinite(m->dt) && std::isfinite(m->time) && finite3(cam.position) && finite3(cam.target) &&
                  finite3(cam.up) && cam.fov_degrees > 0.0 && std::isfinite(cam.fov_degrees) && cam.near_clip > 0.0 &&
                  cam.far_clip > cam.near_clip && std::isfinite(cam.near_clip) && std::isfinite(cam.far_clip) &&
                  cam.focal_length_mm > 0.0 && std::isfinite(cam.focal_length_mm);
  if (!ok) return fail(c, PF_ERR_ARG, "invalid render cache metadata");
  const long long ncell = c->dc.nblocks * kBlockVol;
  // the cell phase field, then the field and the markers to the host
  if (!c->iface) {
    HIP_TRY(c, dalloc(&c->iface, ncell * kIfaceFields));
    HIP_TRY(c, dalloc(&c->iface_bad, 1));
  }
  IfaceArgs a;
  a.dc = c->dc; a.du = c->du; a.dv = c->dv; a.dw = c->dw;
  a.g = g;
  a.mu = c->f[F_MU]; a.mv = c->f[F_MV]; a.mw = c->f[F_MW];
  a.marker = c->marker;
  a.ph = phase_k(&m->phase);
  a.phi = c->iface;
  hipLaunchKernelGGL(k_iface_phi, dim3(grid_for(ncell, 256)), dim3(256), 0, c->stream, a);
  LAUNCH_CHECK(c);
  std::vector<double> phi((size_t)ncell);
  std::vector<uint8_t> mk((size_t)ncell);
  HIP_TRY(c, hipMemcpyAsync(phi.data(), c->iface, sizeof(double) * (size_t)ncell, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipMemcpyAsync(mk.data(), c->marker, (size_t)ncell, hipMemcpyDeviceToHost, c->stream));
  std::vector<RenderCell> cells;
  double pv = 0.0;
  const Dims& dc = c->dc;
  if (!c->mr) {
    // sparsePhaseCells (render_cache3d.h:557-588): the marker blocks in allocation
    // order (the pressure blocks in id order when no block was marked)
    std::vector<std::pair<unsigned long long, long long>> order;
    std::vector<uint8_t> act((size_t)dc.nblocks);
    HIP_TRY(c, hipMemcpyAsync(act.data(), c->act[F_MK], (size_t)dc.nblocks, hipMemcpyDeviceToHost, c->stream));
    std::vector<unsigned long long> key;
    if (c->mark_order_valid && c->np > 0) {
      unsigned long long* kd = reinterpret_cast<unsigned long long*>(c->stage_f);
      if (c->stage_cap * (long long)sizeof(float) < dc.nblocks * (long long)sizeof(unsigned long long))
        return fail(c, PF_ERR_STATE, "render cache: staging buffer too small");
      hipLaunchKernelGGL(k_render_block_keys, dim3(grid_for(dc.nblocks, 256)), dim3(256), 0, c->stream, dc, g,
                         c->act[F_MK], c->cell_start, c->pa.id, kd);
      LAUNCH_CHECK(c);
      key.resize((size_t)dc.nblocks);
      HIP_TRY(c, hipMemcpyAsync(key.data(), kd, sizeof(unsigned long long) * (size_t)dc.nblocks,
                                hipMemcpyDeviceToHost, c->stream));
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    bool any = false;
    for (long long b = 0; b < dc.nblocks; ++b)
      if (act[(size_t)b]) {
        any = true;
        order.push_back({key.empty() ? (unsigned long long)b : key[(size_t)b], b});
      }
    if (!any) {
      HIP_TRY(c, hipMemcpy(act.data(), c->act[F_P], (size_t)dc.nblocks, hipMemcpyDeviceToHost));
      for (long long b = 0; b < dc.nblocks; ++b)
        if (act[(size_t)b]) order.push_back({(unsigned long long)b, b});
    }
    std::sort(order.begin(), order.end());
    const double cv = g.dx * g.dx * g.dx;
    for (const auto& o : order) {
      int bx, by, bz;
      block_coords(dc, o.second, bx, by, bz);
      for (int l = 0; l < 64; ++l) {
        const int i = bx * 4 + (l & 3), j = by * 4 + ((l >> 2) & 3), k = bz * 4 + (l >> 4);
        if (i >= g.nx || j >= g.ny || k >= g.nz) continue;
        const size_t q = (size_t)(o.second * kBlockVol + l);
        const double ph = phi[q];
        if (mk[q] == CELL_AIR && ph <= 1e-8) continue;
        const double wv = ph * cv;
        pv += wv;
        cells.push_back(RenderCell{i, j, k, 0, mk[q], ph, wv});
      }
    }
  } else {
    // mrPhaseCells (render_cache3d.h:590-613): leaf cells in leafCells order, the
    // phase field at the centre clamped to the grid
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    const std::vector<int32_t>& lv = c->mr_leaves;
    for (size_t a2 = 0; a2 + 3 < lv.size(); a2 += 4) {
      const int L = lv[a2], step = 1 << L;
      for (int l = 0; l < 64; ++l) {
        const int i = lv[a2 + 1] * 4 * step + (l & 3) * step, j = lv[a2 + 2] * 4 * step + ((l >> 2) & 3) * step;
        const int k = lv[a2 + 3] * 4 * step + (l >> 4) * step;
        if (i >= g.nx || j >= g.ny || k >= g.nz) continue;
        const int ci = std::min(g.nx - 1, i + step / 2), cj = std::min(g.ny - 1, j + step / 2);
        const int ck = std::min(g.nz - 1, k + step / 2);
        const int marker = mk[(size_t)bm_index(dc, i, j, k)];
        const double ph = phi[(size_t)bm_index(dc, ci, cj, ck)];
        if (marker == CELL_AIR && ph <= 1e-8) continue;
        const double cd = g.dx * (double)step;
        const double wv = ph * cd * cd * cd;
        pv += wv;
        cells.push_back(RenderCell{i, j, k, L, marker, ph, wv});
      }
    }
  }
  // the particles in the caller's order
  RenderParticles P;
  const size_t n = (size_t)c->np;
  P.x.resize(n); P.y.resize(n); P.z.resize(n); P.u.resize(n); P.v.resize(n); P.w.resize(n);
  P.vol.resize(n); P.type.resize(n);
  if (n > 0) {
    int st = pf_particles_download(c, P.x.data(), P.y.data(), P.z.data(), P.u.data(), P.v.data(), P.w.data());
    if (st) return st;
    st = pf_particles_download_ext(c, P.type.data(), P.vol.data());
    if (st) return st;
  }
  std::string err;
  if (!write_frame(g, path, c->mr ? "multires3d_tp" : "sparse3d_tp", m, cells, pv, P, bytes, err))
    return fail(c, PF_ERR_ARG, err);
  return PF_OK;
}

int pf_write_render_manifest(const char* path, const char* kind, int32_t nx, int32_t ny, int32_t nz, double dx,
                             const pf_render_manifest_frame* frames, int64_t nframes) {
  // requireValidPath and requireManifestArgs, render_cache3d.h:118-166
  if (!path || !path[0] || !kind || !kind[0] || nx <= 0 || ny <= 0 || nz <= 0 || !(dx > 0.0) || !std::isfinite(dx) ||
      nframes < 0 || (nframes > 0 && !frames))
    return PF_ERR_ARG;
  for (int64_t a2 = 0; a2 < nframes; ++a2) {
    const pf_render_manifest_frame& f = frames[a2];
    if (f.frame < 0 || f.step < 0 || !f.path || !f.path[0] || f.bytes < 0 || !std::isfinite(f.time)) return PF_ERR_ARG;
  }
  Text t;
  t.raw("{\n  \"lsfs_cache3d_manifest_version\":1,\n  \"cache_schema_version\":2,\n  \"world_units\":\"cell\",\n"
        "  \"sim_kind\":");
  json_string(t, kind);
  const double ext[3] = {nx * dx, ny * dx, nz * dx};
  auto dims = [&]() {
    t.raw("\"frame_bounds_max\":[");
    t.num(ext[0]);
    t.raw(",");
    t.num(ext[1]);
    t.raw(",");
    t.num(ext[2]);
    t.raw("]");
  };
  t.raw(",\n  \"dims\":[");
  t.num((long long)nx);
  t.raw(",");
  t.num((long long)ny);
  t.raw(",");
  t.num((long long)nz);
  t.raw("],\n  \"dx\":");
  t.num(dx);
  t.raw(",\n  \"frame_bounds_min\":[0,0,0],\n  ");
  dims();
  t.raw(",\n  \"frame_count\":");
  t.num((long long)nframes);
  t.raw(",\n  \"frames\":[\n");
  for (int64_t a2 = 0; a2 < nframes; ++a2) {
    const pf_render_manifest_frame& f = frames[a2];
    t.raw("    {\"frame\":");
    t.num((long long)f.frame);
    t.raw(",\"step\":");
    t.num((long long)f.step);
    t.raw(",\"time\":");
    t.num(f.time);
    t.raw(",\"shutter_open\":");
    t.num(f.time);
    t.raw(",\"shutter_close\":");
    t.num(f.time);
    t.raw(",\"frame_bounds_min\":[0,0,0],");
    dims();
    t.raw(",\"path\":");
    json_string(t, f.path);
    t.raw(",\"bytes\":");
    t.num((long long)f.bytes);
    t.raw(a2 + 1 < nframes ? "},\n" : "}\n");
  }
  t.raw("  ]\n}\n");
  FILE* fp = std::fopen(path, "wb");
  if (!fp) return PF_ERR_ARG;
  const bool w = std::fwrite(t.s.data(), 1, t.s.size(), fp) == t.s.size();
  return (std::fclose(fp) == 0 && w) ? PF_OK : PF_ERR_ARG;
}

int pf_timestep(pf_ctx* c, double requested_dt, int32_t enabled, double cfl, double min_dt, double* eff,
                double* vmax_out, double* cfl_dt, int32_t* limited) {
  if (!c) return PF_ERR_ARG;
  HIP_TRY(c, fill_u32(c->stream, reinterpret_cast<uint32_t*>(c->vmax_dev), 0u, 1));
  if (c->np > 0) {
    hipLaunchKernelGGL(k_max_speed, dim3(std::min(grid_for(c->np, 256), 8192u)), dim3(256), 0, c->stream, c->np, c->pa,
                       reinterpret_cast<uint32_t*>(c->vmax_dev));
    LAUNCH_CHECK(c);
  }
  float vm = 0.f;
  HIP_TRY(c, hipMemcpyAsync(&vm, c->vmax_dev, sizeof(float), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  // computeAdaptiveParticleTimestep3D, timestep3d.h:22-49
  double v = vm, e = requested_dt, cd = requested_dt;
  int lim = 0;
  if (enabled && requested_dt > 0.0 && c->geom.dx > 0.0 && cfl > 0.0 && v > 0.0) {
    cd = cfl * c->geom.dx / v;
    double l = std::min(requested_dt, cd);
    if (min_dt > 0.0) l = std::max(min_dt, l);
    e = std::min(requested_dt, l);
    lim = e < requested_dt ? 1 : 0;
  }
  if (eff) *eff = e;
  if (vmax_out) *vmax_out = v;
  if (cfl_dt) *cfl_dt = cd;
  if (limited) *limited = lim;
  return PF_OK;
}

int pf_step(pf_ctx* c, const pf_step_cfg* cfg, pf_step_stats* stats) {
  if (!c || !cfg) return PF_ERR_ARG;
  if (cfg->model != PF_MODEL_TWO_PHASE && cfg->model != PF_MODEL_SINGLE_PHASE)
    return fail(c, PF_ERR_ARG, "pf_step_cfg.model must be PF_MODEL_TWO_PHASE or PF_MODEL_SINGLE_PHASE");
  const bool single = cfg->model == PF_MODEL_SINGLE_PHASE;
  pf_step_stats S;
  std::memset(&S, 0, sizeof(S));
  default_interface_stats(&S);
  const bool timing = cfg->timing != 0;
  record(c, timing, 0);
  // applyParticleAdaptivity (sparse_sim3d_tp.cpp:855; the single-phase step has none)
  int st = single ? PF_OK : run_adaptivity(c, &cfg->adapt, &S.adapt);
  if (st) return st;
  st = ensure_binned(c);
  if (st) return st;
  record(c, timing, 1);
  // timestep3d.h: the speed scan is always done (stats); it gates dt only if enabled
  st = pf_timestep(c, cfg->dt, cfg->adaptive_timestep, cfg->adaptive_cfl, cfg->adaptive_min_dt, &S.effective_dt,
                   &S.max_particle_speed, &S.cfl_limit_dt, &S.limited);
  if (st) return st;
  const double dt = S.effective_dt;
  S.c_div = single ? cfg->c_div
                   : volume_correction(c, cfg->c_div_volume_correction, cfg->c_div_strength,
                                       cfg->liquid_volume_target, cfg->Vp, dt, cfg->c_div, &S.liquid_volume);
  st = run_mark(c);
  if (st) return st;
  record(c, timing, 2);
  st = run_p2g_splat(c, &cfg->phase, cfg->Vp, single);
  if (st) return st;
  record(c, timing, 3);
  // the single-phase step has no interface work (sparse_sim3d.cpp:106-115)
  st = step_face_stage(c, &cfg->phase, dt, cfg->gravity, false, !single && !cfg->skip_interface_diagnostics,
                       !single && cfg->surface_tension, cfg->surface_tension_strength,
                       cfg->surface_tension_max_delta_speed, cfg->surface_tension_smoothing_radius, &S);
  if (st) return st;
  record(c, timing, 4);
  st = run_project(c, &cfg->phase, dt, &cfg->solve, S.c_div, &S.solve, timing, 5, 6, 7, single);
  if (st) return st;
  HIP_TRY(c, fill_u32(c->stream, reinterpret_cast<uint32_t*>(c->esc_dev), 0u, 8));
  // the single-phase blend uses one alpha (SparseSim3D::alpha, sparse_sim3d.h:8)
  st = run_g2p_advect(c, true, true, cfg->alpha_liquid, single ? cfg->alpha_liquid : cfg->alpha_gas, dt,
                      cfg->advection_order);
  if (st) return st;
  record(c, timing, 8);
  HIP_TRY(c, hipMemcpyAsync(&S.escape, c->esc_dev, sizeof(int32_t) * 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (timing) {
    float ms;
    for (int s = 0; s < PF_NSTAGES; ++s) {
      hipEventElapsedTime(&ms, c->ev[s], c->ev[s + 1]);
      S.stage_ms[s] = ms;
    }
    hipEventElapsedTime(&ms, c->ev[0], c->ev[PF_NSTAGES]);
    S.total_ms = ms;
  }
  if (stats) *stats = S;
  return PF_OK;
}

int pf_field_dims(const pf_ctx* c, int32_t field, int32_t dims[3]) {
  if (!c || !dims || field_slot(field) < 0) return PF_ERR_ARG;
  const Dims& d = dims_of(c, field);
  dims[0] = d.nx; dims[1] = d.ny; dims[2] = d.nz;
  return PF_OK;
}

int pf_field_download(pf_ctx* c, int32_t field, float* dense) {
  int slot = field_slot(field);
  if (!c || !dense || slot < 0) return PF_ERR_ARG;
  const Dims& d = dims_of(c, field);
  long long n = (long long)d.nx * d.ny * d.nz;
  hipLaunchKernelGGL(k_to_dense, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, d,
                     slot == F_MK ? nullptr : c->f[slot], slot == F_MK ? c->marker : nullptr, c->stage_f);
  LAUNCH_CHECK(c);
  HIP_TRY(c, hipMemcpyAsync(dense, c->stage_f, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return PF_OK;
}

int pf_field_upload(pf_ctx* c, int32_t field, const float* dense, const uint8_t* block_active) {
  int slot = field_slot(field);
  if (!c || !dense || slot < 0) return PF_ERR_ARG;
  const Dims& d = dims_of(c, field);
  long long n = (long long)d.nx * d.ny * d.nz;
  if (slot == F_MK) {
    // markers are AIR/FLUID/SOLID and FLUID cells are interior: the pressure
    // stencil of a FLUID cell then never leaves the grid
    for (int k = 0; k < d.nz; ++k)
      for (int j = 0; j < d.ny; ++j)
        for (int i = 0; i < d.nx; ++i) {
          float m = dense[i + (long long)d.nx * (j + (long long)d.ny * k)];
          if (!(m == 0.f || m == 1.f || m == 2.f)) return fail(c, PF_ERR_ARG, "marker values must be 0, 1 or 2");
          bool edge = i == 0 || j == 0 || k == 0 || i == d.nx - 1 || j == d.ny - 1 || k == d.nz - 1;
          if (edge && m == 1.f) return fail(c, PF_ERR_ARG, "FLUID marker on a boundary cell");
        }
  }
  HIP_TRY(c, hipMemcpyAsync(c->stage_f, dense, sizeof(float) * (size_t)n, hipMemcpyHostToDevice, c->stream));
  uint8_t* act_in = nullptr;
  if (block_active) {
    act_in = reinterpret_cast<uint8_t*>(c->blk_flag32);
    HIP_TRY(c, hipMemcpyAsync(act_in, block_active, (size_t)d.nblocks, hipMemcpyHostToDevice, c->stream));
  }
  hipLaunchKernelGGL(k_from_dense, dim3(grid_for(d.nblocks, 4)), dim3(256), 0, c->stream, d, c->stage_f,
                     slot == F_MK ? nullptr : c->f[slot], slot == F_MK ? c->marker : nullptr, c->act[slot], act_in);
  if (slot == F_MK) c->mark_order_valid = false;
  LAUNCH_CHECK(c);
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return PF_OK;
}

int pf_active_blocks(pf_ctx* c, int32_t field, int32_t* bids, int64_t capacity, int64_t* n) {
  int slot = field_slot(field);
  if (!c || !n || slot < 0) return PF_ERR_ARG;
  const Dims& d = dims_of(c, field);
  uint32_t* list = reinterpret_cast<uint32_t*>(c->stage_f);
  HIP_TRY(c, compact_flags(*c, c->act[slot], d.nblocks, list, c->u32_dev + 3));
  HIP_TRY(c, hipMemcpyAsync(c->u32_host, c->u32_dev, sizeof(uint32_t) * 4, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  long long cnt = c->u32_host[3];
  *n = cnt;
  if (bids && capacity > 0) {
    long long m = std::min<long long>(cnt, capacity);
    HIP_TRY(c, hipMemcpyAsync(bids, list, sizeof(int32_t) * (size_t)m, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
  }
  return PF_OK;
}

int pf_bin_counts(pf_ctx* c, uint32_t* dense) {
  if (!c || !dense) return PF_ERR_ARG;
  int st = ensure_binned(c);
  if (st) return st;
  const long long n = (long long)c->dc.nx * c->dc.ny * c->dc.nz;
  uint32_t* out = reinterpret_cast<uint32_t*>(c->stage_f);  // stage_cap >= cells
  hipLaunchKernelGGL(k_counts_dense, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, c->dc, c->cell_count, out);
  LAUNCH_CHECK(c);
  HIP_TRY(c, hipMemcpyAsync(dense, out, sizeof(uint32_t) * (size_t)n, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return PF_OK;
}

int pf_color_lists(pf_ctx* c, int32_t* lists, int64_t capacity, int64_t counts[8]) {
  if (!c || !counts) return PF_ERR_ARG;
  int st = ensure_binned(c);
  if (st) return st;
  HIP_TRY(c, hipMemcpyAsync(c->u32_host, c->color_cnt, sizeof(uint32_t) * 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  long long total = 0;
  for (int k = 0; k < 8; ++k) {
    counts[k] = c->u32_host[k];
    total += counts[k];
  }
  if (!lists) return PF_OK;
  if (capacity < total) return fail(c, PF_ERR_ARG, "pf_color_lists: capacity below the summed list lengths");
  long long off = 0;
  for (int k = 0; k < 8; ++k) {
    if (counts[k] == 0) continue;
    HIP_TRY(c, hipMemcpyAsync(lists + off, c->color_list + (long long)k * c->color_cap,
                              sizeof(int32_t) * (size_t)counts[k], hipMemcpyDeviceToHost, c->stream));
    off += counts[k];
  }
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  off = 0;
  for (int k = 0; k < 8; ++k) {  // append order within a colour is free: return it sorted
    std::sort(lists + off, lists + off + counts[k]);
    off += counts[k];
  }
  return PF_OK;
}

int pf_face_phi(pf_ctx* c, const pf_phase* ph, int32_t axis, double* dense) {
  if (!c || !ph || !dense || axis < 0 || axis > 2) return PF_ERR_ARG;
  const Dims& d = axis == 0 ? c->du : (axis == 1 ? c->dv : c->dw);
  const long long n = (long long)d.nx * d.ny * d.nz;
  double* out = nullptr;
  HIP_TRY(c, dalloc(&out, n));
  hipLaunchKernelGGL(k_face_phi, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, d, c->f[F_MU + axis],
                     phase_k(ph), out);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(dense, out, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  hipFree(out);
  if (e != hipSuccess) return fail(c, PF_ERR_HIP, std::string("pf_face_phi: ") + hipGetErrorString(e));
  return PF_OK;
}

double pf_calibrate_rho_tilde0(const pf_phase* ph, double Vp) { return ph ? calibrate_host(ph, Vp) : 0.0; }

double pf_phi_from_raw(double rt, const pf_phase* pp) {
  double rmin = std::log(pp->rho_l / pp->rho_g) * pp->rho_g * pp->rho_tilde_0;
  if (rt < rmin) return 0.0;
  double v = std::sqrt(std::max(rt - rmin, 0.0) / (pp->alpha_phi * pp->rho_tilde_0 * pp->rho_l));
  return std::min(v, 1.0);
}

double pf_beta_from_phi(double phi, const pf_phase* pp) { return 1.0 / (phi * pp->rho_l + (1.0 - phi) * pp->rho_g); }

// ---------------------------------------------------------------- learned preconditioner

// splitmix64 stream and Box-Muller normals for the He initialisation
static double lpc_uniform(uint64_t& st) {
  uint64_t z = (st += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return ((z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

int pf_lpc_create(pf_lpc** out, uint64_t seed) {
  if (!out) return PF_ERR_ARG;
  *out = nullptr;
  pf_lpc* L = new pf_lpc();
  HIP_TRY(L, hipGetDevice(&L->device));
  HIP_TRY(L, hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking));
  HIP_TRY(L, dalloc(&L->w, kNW));
  HIP_TRY(L, dalloc(&L->m, kNW));
  HIP_TRY(L, dalloc(&L->v, kNW));
  HIP_TRY(L, dalloc(&L->grad, kNW));
  HIP_TRY(L, dalloc(&L->bf.w1f, 16 * kpad(4)));
  HIP_TRY(L, dalloc(&L->bf.w1b, 4 * kpad(16)));
  HIP_TRY(L, dalloc(&L->bf.w2f, 32 * kpad(16)));
  HIP_TRY(L, dalloc(&L->bf.w2b, 16 * kpad(32)));
  HIP_TRY(L, dalloc(&L->bf.w3f, 16 * kpad(32)));
  HIP_TRY(L, dalloc(&L->bf.w3b, 32 * kpad(16)));
  HIP_TRY(L, dalloc(&L->bf.wh, 16));
  // He initialisation: N(0, 2 / fan_in) for the ReLU layers, N(0, 1/16) for the head
  std::vector<float> w(kNW);
  uint64_t st = seed;
  const int off[4] = {kOffW1, kOffW2, kOffW3, kOffWH};
  const int cnt[4] = {kNW1, kNW2, kNW3, kNWH};
  const double sd[4] = {std::sqrt(2.0 / (27 * 4)), std::sqrt(2.0 / (27 * 16)), std::sqrt(2.0 / (27 * 32)),
                        std::sqrt(1.0 / 16)};
  for (int l = 0; l < 4; ++l)
    for (int i = 0; i < cnt[l]; ++i) {
      double u1 = lpc_uniform(st), u2 = lpc_uniform(st);
      w[off[l] + i] = (float)(sd[l] * std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
    }
  HIP_TRY(L, hipMemcpy(L->w, w.data(), sizeof(float) * kNW, hipMemcpyHostToDevice));
  *out = L;
  return PF_OK;
}

int pf_lpc_destroy(pf_lpc* L) {
  if (!L) return PF_OK;
  hipStreamSynchronize(L->stream);
  dfree(L->w); dfree(L->m); dfree(L->v); dfree(L->grad);
  dfree(L->bf.w1f); dfree(L->bf.w1b); dfree(L->bf.w2f); dfree(L->bf.w2b); dfree(L->bf.w3f); dfree(L->bf.w3b);
  dfree(L->bf.wh);
  dfree(L->X); dfree(L->h1); dfree(L->h2); dfree(L->h3); dfree(L->d3); dfree(L->d2); dfree(L->d1);
  dfree(L->o); dfree(L->go); dfree(L->scale); dfree(L->part);
  dfree(L->zbuf); dfree(L->qbuf); dfree(L->gz); dfree(L->stage);
  hipStreamDestroy(L->stream);
  delete L;
  return PF_OK;
}

const char* pf_lpc_last_error(const pf_lpc* L) { return L ? L->err.c_str() : "null network"; }

int64_t pf_lpc_num_weights(void) { return kNW; }

int pf_lpc_get_weights(pf_lpc* L, float* w, int64_t n) {
  if (!L || !w || n != kNW) return PF_ERR_ARG;
  HIP_TRY(L, hipStreamSynchronize(L->stream));
  HIP_TRY(L, hipMemcpy(w, L->w, sizeof(float) * kNW, hipMemcpyDeviceToHost));
  return PF_OK;
}

int pf_lpc_set_weights(pf_lpc* L, const float* w, int64_t n) {
  if (!L || !w || n != kNW) return PF_ERR_ARG;
  for (int64_t i = 0; i < n; ++i)
    if (!std::isfinite(w[i])) { L->err = "non-finite weight"; return PF_ERR_ARG; }
  HIP_TRY(L, hipStreamSynchronize(L->stream));
  HIP_TRY(L, hipMemcpy(L->w, w, sizeof(float) * kNW, hipMemcpyHostToDevice));
  HIP_TRY(L, hipMemset(L->m, 0, sizeof(float) * kNW));
  HIP_TRY(L, hipMemset(L->v, 0, sizeof(float) * kNW));
  L->adam_t = 0;
  L->prepped = false;
  return PF_OK;
}

static int lpc_stage_input(pf_lpc* L, int64_t nt, const float* X) {
  const long long n = nt * (long long)kHaloVol * 4;
  if (n > L->stage_cap) {
    dfree(L->stage);
    HIP_TRY(L, dalloc(&L->stage, n));
    L->stage_cap = n;
  }
  HIP_TRY(L, hipMemcpyAsync(L->stage, X, sizeof(float) * (size_t)n, hipMemcpyHostToDevice, L->stream));
  hipLaunchKernelGGL(k_f32_to_bf16, dim3((unsigned)std::min<long long>((n + 255) / 256, 65536)), dim3(256), 0,
                     L->stream, L->stage, L->X, n);
  HIP_TRY(L, hipGetLastError());
  return PF_OK;
}

int pf_lpc_forward_tiles(pf_lpc* L, int64_t nt, const float* X, float* O) {
  if (!L || !X || !O || nt <= 0) return PF_ERR_ARG;
  int st = lpc_grow(L, nt, false);
  if (!st) st = lpc_stage_input(L, nt, X);
  if (!st) st = lpc_forward(L, L->stream, nt, false, nullptr);
  if (st) return st;
  HIP_TRY(L, hipMemcpyAsync(O, L->o, sizeof(float) * (size_t)(nt * kTileVol), hipMemcpyDeviceToHost, L->stream));
  HIP_TRY(L, hipStreamSynchronize(L->stream));
  return PF_OK;
}

int pf_lpc_grad_tiles(pf_lpc* L, int64_t nt, const float* X, const float* gO, float* gW) {
  if (!L || !X || !gO || !gW || nt <= 0) return PF_ERR_ARG;
  int st = lpc_grow(L, nt, true);
  if (!st) st = lpc_stage_input(L, nt, X);
  if (!st) st = lpc_forward(L, L->stream, nt, true, nullptr);
  if (st) return st;
  HIP_TRY(L, hipMemcpyAsync(L->go, gO, sizeof(float) * (size_t)(nt * kTileVol), hipMemcpyHostToDevice, L->stream));
  st = lpc_backward(L, L->stream, nt);
  if (st) return st;
  HIP_TRY(L, hipMemcpyAsync(gW, L->grad, sizeof(float) * kNW, hipMemcpyDeviceToHost, L->stream));
  HIP_TRY(L, hipStreamSynchronize(L->stream));
  return PF_OK;
}

int pf_lpc_attach(pf_ctx* c, pf_lpc* L) {
  if (!c) return PF_ERR_ARG;
  if (L && L->device != c->device) return fail(c, PF_ERR_ARG, "network and context are on different devices");
  if (L) HIP_TRY(c, hipStreamSynchronize(L->stream));
  c->lpc = L;
  c->tiles_valid = false;
  return PF_OK;
}

int pf_lpc_record(pf_ctx* c, int32_t max_snapshots, int32_t every) {
  if (!c || max_snapshots < 0 || every < 0) return PF_ERR_ARG;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  dfree(c->snaps);
  c->snap_cap = 0;
  c->nsnap = 0;
  if (max_snapshots == 0) return PF_OK;
  HIP_TRY(c, dalloc(&c->snaps, (long long)max_snapshots * c->dc.nblocks * kBlockVol));
  c->snap_cap = max_snapshots;
  c->snap_every = std::max(every, 1);
  return PF_OK;
}

int32_t pf_lpc_snapshot_count(const pf_ctx* c) { return c ? c->nsnap : 0; }

// loss ||r - A z||^2 / ||r||^2 of z = M(r) (learned) or z = D^-1 r for snapshot k;
// leaves z in L->zbuf, q = r - A z in L->qbuf, ||r||^2 in *rr
static int lpc_loss_eval(pf_lpc* L, pf_ctx* c, int k, bool learned, bool store, double* loss, double* rr) {
  const long long ncell = c->dc.nblocks * kBlockVol;
  if (ncell > L->cell_cap) {
    dfree(L->zbuf); dfree(L->qbuf); dfree(L->gz);
    HIP_TRY(L, dalloc(&L->zbuf, ncell));
    HIP_TRY(L, dalloc(&L->qbuf, ncell));
    HIP_TRY(L, dalloc(&L->gz, ncell));
    L->cell_cap = ncell;
  }
  int st = lpc_grow(L, c->ntiles, store);
  if (st) return fail(c, st, "learned preconditioner: " + L->err);
  const float* r = c->snaps + (long long)k * ncell;
  LpcTiles T = lpc_tiles_args(c, r);
  hipLaunchKernelGGL(k_lpc_gather, dim3((unsigned)c->ntiles), dim3(256), 0, c->stream, T);
  if (learned) {
    st = lpc_forward(L, c->stream, c->ntiles, store, nullptr);
    if (st) return fail(c, st, "learned preconditioner: " + L->err);
  }
  hipLaunchKernelGGL(k_lpc_scatter, dim3(lpc_scatter_grid(c->ntiles)), dim3(256), 0, c->stream, T,
                     learned ? L->o : nullptr, L->zbuf, nullptr, 0);
  VecArgs va{};
  va.dc = c->dc; va.list = c->pblk_prev; va.nlist = c->npblk_prev; va.sc = c->sc; va.part = c->partial;
  const unsigned G = list_grid(c->npblk_prev);
  hipLaunchKernelGGL(k_lpc_apply_a, dim3(G), dim3(256), 0, c->stream, va, L->zbuf, r, L->qbuf, c->cd32, c->cx32,
                     c->cy32, c->cz32, 1.f);
  hipLaunchKernelGGL(k_reduce_slots, dim3(1), dim3(256), 0, c->stream, c->partial, (int)G, 2, 0u, c->red);
  LAUNCH_CHECK(c);
  st = read_small(c);
  if (st) return st;
  *rr = c->red_host[1];
  *loss = *rr > 0.0 ? c->red_host[0] / *rr : 0.0;
  return PF_OK;
}

static int lpc_check_system(pf_lpc* L, pf_ctx* c) {
  if (!L || !c) return PF_ERR_ARG;
  if (c->lpc != L) return fail(c, PF_ERR_STATE, "network is not attached to this context");
  if (!c->tiles_valid || c->nsnap <= 0)
    return fail(c, PF_ERR_STATE, "no recorded system: pf_lpc_record, then solve with the network attached");
  HIP_TRY(c, hipStreamSynchronize(L->stream));
  return PF_OK;
}

int pf_lpc_loss(pf_lpc* L, pf_ctx* c, int32_t k, double* loss_learned, double* loss_jacobi) {
  int st = lpc_check_system(L, c);
  if (st) return st;
  if (k < 0 || k >= c->nsnap) return fail(c, PF_ERR_ARG, "snapshot index out of range");
  double rr = 0.0, a = 0.0, b = 0.0;
  st = lpc_loss_eval(L, c, k, true, false, &a, &rr);
  if (!st) st = lpc_loss_eval(L, c, k, false, false, &b, &rr);
  if (st) return st;
  if (loss_learned) *loss_learned = a;
  if (loss_jacobi) *loss_jacobi = b;
  return PF_OK;
}

int pf_lpc_train(pf_lpc* L, pf_ctx* c, int32_t steps, double lr, uint64_t seed, double* mean_loss) {
  int st = lpc_check_system(L, c);
  if (st) return st;
  if (steps < 0 || !(lr > 0.0)) return fail(c, PF_ERR_ARG, "steps >= 0 and lr > 0 required");
  uint64_t rs = seed;
  double lsum = 0.0;
  for (int it = 0; it < steps; ++it) {
    const int k = (int)(lpc_uniform(rs) * c->nsnap) % c->nsnap;
    double loss = 0.0, rr = 0.0;
    st = lpc_loss_eval(L, c, k, true, true, &loss, &rr);
    if (st) return st;
    lsum += loss;
    if (!(rr > 0.0)) continue;
    // dL/dz = -2 A q / ||r||^2 (A symmetric), then through z = r/diag + s * o
    VecArgs va{};
    va.dc = c->dc; va.list = c->pblk_prev; va.nlist = c->npblk_prev; va.sc = c->sc; va.part = c->partial;
    const unsigned G = list_grid(c->npblk_prev);
    hipLaunchKernelGGL(k_lpc_apply_a, dim3(G), dim3(256), 0, c->stream, va, L->qbuf, nullptr, L->gz, c->cd32,
                       c->cx32, c->cy32, c->cz32, (float)(-2.0 / rr));
    LpcTiles T = lpc_tiles_args(c, c->snaps + (long long)k * c->dc.nblocks * kBlockVol);
    hipLaunchKernelGGL(k_lpc_gz_to_tiles, dim3(lpc_scatter_grid(c->ntiles)), dim3(256), 0, c->stream, T, L->gz,
                       L->go);
    LAUNCH_CHECK(c);
    st = lpc_backward(L, c->stream, c->ntiles);
    if (st) return fail(c, st, "learned preconditioner: " + L->err);
    L->adam_t++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float bc1 = 1.f - std::pow(b1, (float)L->adam_t), bc2 = 1.f - std::pow(b2, (float)L->adam_t);
    hipLaunchKernelGGL(k_lpc_adam, dim3((kNW + 255) / 256), dim3(256), 0, c->stream, L->w, L->m, L->v, L->grad, kNW,
                       (float)lr, b1, b2, 1e-8f, bc1, bc2);
    LAUNCH_CHECK(c);
    L->prepped = false;
    st = lpc_prep(L, c->stream);
    if (st) return fail(c, st, "learned preconditioner: " + L->err);
  }
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (mean_loss) *mean_loss = steps > 0 ? lsum / steps : 0.0;
  return PF_OK;
}

// ---------------------------------------------------------------- multires

int pf_mr_set_layout(pf_ctx* c, int64_t n, const int32_t* leaves) {
  if (!c || n <= 0 || !leaves) return PF_ERR_ARG;
  if (c->geom.ox != 0.0 || c->geom.oy != 0.0 || c->geom.oz != 0.0)
    return fail(c, PF_ERR_ARG, "the multires path takes positions relative to the origin 0");
  std::vector<Leaf> lv((size_t)n);
  for (int64_t a = 0; a < n; ++a)
    for (int q = 0; q < 4; ++q) lv[(size_t)a][q] = leaves[4 * a + q];
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return mr_apply_layout(c, lv);
}

int pf_mr_get_layout(pf_ctx* c, int32_t* leaves, int64_t cap, int64_t* n) {
  if (!c || !n) return PF_ERR_ARG;
  *n = (int64_t)c->mr_leaves.size() / 4;
  if (leaves) std::copy(c->mr_leaves.begin(), c->mr_leaves.begin() + 4 * std::min<int64_t>(cap, *n), leaves);
  return PF_OK;
}

int64_t pf_mr_layout_fine_box(int32_t nx, int32_t ny, int32_t nz, int32_t level, int32_t box_valid, const int32_t* box,
                              int32_t* leaves, int64_t cap) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || level < 0 || level > 6 || (box_valid && !box)) return -1;
  std::vector<Leaf> lv = mr_layout_box(nx, ny, nz, level, box_valid != 0, box);
  if (leaves)
    for (int64_t a = 0; a < std::min<int64_t>(cap, (int64_t)lv.size()); ++a)
      for (int q = 0; q < 4; ++q) leaves[4 * a + q] = lv[(size_t)a][q];
  return (int64_t)lv.size();
}

int64_t pf_mr_layout_refine(int32_t nx, int32_t ny, int32_t nz, const int32_t* in, int64_t n, const int32_t* box,
                            int32_t* leaves, int64_t cap) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || n < 0 || (n > 0 && !in) || !box) return -1;
  std::vector<Leaf> lv((size_t)n);
  for (int64_t a = 0; a < n; ++a)
    for (int q = 0; q < 4; ++q) lv[(size_t)a][q] = in[4 * a + q];
  std::sort(lv.begin(), lv.end());
  lv.erase(std::unique(lv.begin(), lv.end()), lv.end());
  const int x0 = std::clamp(box[0], 0, nx), y0 = std::clamp(box[1], 0, ny), z0 = std::clamp(box[2], 0, nz);
  const int x1 = std::clamp(box[3], 0, nx), y1 = std::clamp(box[4], 0, ny), z1 = std::clamp(box[5], 0, nz);
  if (x0 < x1 && y0 < y1 && z0 < z1) mr_refine(lv, nx, ny, nz, x0, y0, z0, x1, y1, z1);
  if (leaves)
    for (int64_t a = 0; a < std::min<int64_t>(cap, (int64_t)lv.size()); ++a)
      for (int q = 0; q < 4; ++q) leaves[4 * a + q] = lv[(size_t)a][q];
  return (int64_t)lv.size();
}

int pf_mr_face_valid(pf_ctx* c, int32_t axis, uint8_t* dense) {
  if (!c || !dense || axis < 0 || axis > 2) return PF_ERR_ARG;
  if (!c->mr) return fail(c, PF_ERR_STATE, "no multires layout (pf_mr_set_layout)");
  const Dims& d = axis == 0 ? c->du : (axis == 1 ? c->dv : c->dw);
  std::vector<uint8_t> bm((size_t)(d.nblocks * kBlockVol));
  HIP_TRY(c, hipMemcpyAsync(bm.data(), c->mr_valid[axis], bm.size(), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  for (int k = 0; k < d.nz; ++k)
    for (int j = 0; j < d.ny; ++j)
      for (int i = 0; i < d.nx; ++i) dense[i + (long long)d.nx * (j + (long long)d.ny * k)] = bm[(size_t)bm_index(d, i, j, k)];
  return PF_OK;
}

int pf_mr_mark_cells(pf_ctx* c) {
  if (!c) return PF_ERR_ARG;
  if (!c->mr) return fail(c, PF_ERR_STATE, "no multires layout (pf_mr_set_layout)");
  int st = ensure_binned(c);
  if (st) return st;
  const long long ncell = c->dc.nblocks * kBlockVol;
  HIP_TRY(c, fill_u8(c->stream, c->act[F_MK], 0, c->dc.nblocks));
  hipLaunchKernelGGL(k_mr_mark, dim3(grid_for(ncell, 256)), dim3(256), 0, c->stream, mr_grid(c), c->cell_start,
                     c->marker, c->act[F_MK]);
  LAUNCH_CHECK(c);
  return PF_OK;
}

int pf_mr_p2g(pf_ctx* c, const pf_phase* ph, double Vp) {
  if (!c || !ph) return PF_ERR_ARG;
  if (!c->mr) return fail(c, PF_ERR_STATE, "no multires layout (pf_mr_set_layout)");
  int st = run_p2g_splat(c, ph, Vp, false, true);
  if (st) return st;
  return run_face_ops(c, FO_NORMALIZE, 0.f);
}

static pf_solve_cfg mr_to_solve(const pf_mr_solve_cfg* m) {
  pf_solve_cfg s{};
  s.max_iters = m->max_iterations;
  s.stop_rule = PF_STOP_REF_MAXABS;
  s.tol = m->absolute_tolerance;
  s.inner_rel_tol = m->inner_rel_tol > 0 ? m->inner_rel_tol : 0.1;
  s.max_outer = m->max_outer > 0 ? m->max_outer : 10000;
  s.check_every = m->check_every > 0 ? m->check_every : 8;
  s.precond = PF_PRECOND_JACOBI;
  return s;
}

// the multires projection with `cfg` in force for this solve only
static int mr_project_with(pf_ctx* c, const pf_phase* ph, double dt, const pf_mr_solve_cfg* cfg, double c_div,
                           pf_solve_stats* stats, bool timing, int e0, int e1, int e2) {
  if (cfg->max_iterations < 0 || cfg->absolute_tolerance < 0 || cfg->relative_tolerance < 0)
    return fail(c, PF_ERR_ARG, "invalid multires solve configuration");
  if (cfg->mode != PF_MR_SOLVE_FAST && cfg->mode != PF_MR_SOLVE_REFERENCE)
    return fail(c, PF_ERR_ARG, "unknown multires solve mode");
  if (cfg->mode == PF_MR_SOLVE_REFERENCE && (!mr_cfg_valid(*cfg) || !std::isfinite(c_div)))
    return fail(c, PF_ERR_ARG, "projectMR3D invalid solve config");
  const pf_solve_cfg s = mr_to_solve(cfg);
  c->mr_cfg = cfg;
  const int st = run_project(c, ph, dt, &s, c_div, stats, timing, e0, e1, e2, false, cfg->relative_tolerance);
  c->mr_cfg = nullptr;
  return st;
}

int pf_mr_project(pf_ctx* c, const pf_phase* ph, double dt, const pf_mr_solve_cfg* cfg, double c_div,
                  pf_solve_stats* stats) {
  if (!c || !ph || !cfg) return PF_ERR_ARG;
  if (!c->mr) return fail(c, PF_ERR_STATE, "no multires layout (pf_mr_set_layout)");
  return mr_project_with(c, ph, dt, cfg, c_div, stats, false, 0, 0, 0);
}

void pf_mr_solve_defaults(pf_mr_solve_cfg* m) {
  if (!m) return;
  *m = pf_mr_solve_cfg{};
  m->max_iterations = 160;
  m->check_every = 8;
  m->absolute_tolerance = 1e-7;
  m->relative_tolerance = 1e-6;
  m->inner_rel_tol = 0.1;
  m->max_outer = 10000;
  m->mode = PF_MR_SOLVE_FAST;
  m->use_jacobi_preconditioner = 1;
  m->use_flexible_cg_beta = 0;
  m->adaptive_restart = 1;
  m->relaxation_sweeps = 0;
  m->restart_growth_threshold = 10.0;
  m->relaxation_omega = 0.67;
  m->relaxation_min_omega = 0.05;
  m->coarse_correction_iterations = 32;
  m->coarse_correction_sweeps = 1;
  m->coarse_correction_absolute_tolerance = 0.0;
  m->coarse_correction_relative_tolerance = 1e-3;
  m->coarse_correction_min_scale = 1.0 / 64.0;
  m->coarse_preconditioner_iterations = 8;
  m->coarse_preconditioner_absolute_tolerance = 0.0;
  m->coarse_preconditioner_relative_tolerance = 1e-2;
  m->coarse_preconditioner_scale = 1.0;
  m->coarse_preconditioner_min_rz_gain = 0.0;
  m->coarse_preconditioner_max_work_ratio = 0.0;
  m->coarse_preconditioner_auto_disable = 0;
  m->coarse_preconditioner_auto_disable_after = 4;
}

int pf_mr_last_solve_detail(pf_ctx* c, pf_mr_solve_detail* d) {
  if (!c || !d) return PF_ERR_ARG;
  *d = c->mr_detail;
  return PF_OK;
}

int pf_mr_update_refinement(pf_ctx* c, const pf_mr_refine_cfg* rc, int32_t* fine_leaves, int32_t* budget_limited) {
  if (!c || !rc) return PF_ERR_ARG;
  if (c->geom.ox != 0.0 || c->geom.oy != 0.0 || c->geom.oz != 0.0)
    return fail(c, PF_ERR_ARG, "the multires path takes positions relative to the origin 0");
  int st = mr_dynamic(c, rc);
  if (st) return st;
  if (fine_leaves) *fine_leaves = c->mr_last_fine_leaves;
  if (budget_limited) *budget_limited = c->mr_budget_limited;
  return PF_OK;
}

int pf_mr_retained_box(pf_ctx* c, int32_t* box, int32_t* valid) {
  if (!c || !box || !valid) return PF_ERR_ARG;
  *valid = c->mr_retained_valid ? 1 : 0;
  for (int q = 0; q < 6; ++q) box[q] = c->mr_retained_valid ? c->mr_retained[q] : 0;
  return PF_OK;
}

int pf_mr_step(pf_ctx* c, const pf_mr_step_cfg* cfg, pf_step_stats* stats) {
  if (!c || !cfg) return PF_ERR_ARG;
  if (c->geom.ox != 0.0 || c->geom.oy != 0.0 || c->geom.oz != 0.0)
    return fail(c, PF_ERR_ARG, "the multires path takes positions relative to the origin 0");
  pf_step_stats S;
  std::memset(&S, 0, sizeof(S));
  default_interface_stats(&S);
  const bool timing = cfg->timing != 0;
  record(c, timing, 0);
  // applyParticleAdaptivity, then updateDynamicRefinement (multires_sim3d_tp.cpp:672-673)
  int st = run_adaptivity(c, &cfg->adapt, &S.adapt);
  if (st) return st;
  st = mr_dynamic(c, &cfg->refine);
  if (st) return st;
  if (!c->mr) return fail(c, PF_ERR_STATE, "no multires layout (pf_mr_set_layout or dynamic refinement)");
  st = ensure_binned(c);
  if (st) return st;
  record(c, timing, 1);
  st = pf_timestep(c, cfg->dt, cfg->adaptive_timestep, cfg->adaptive_cfl, cfg->adaptive_min_dt, &S.effective_dt,
                   &S.max_particle_speed, &S.cfl_limit_dt, &S.limited);
  if (st) return st;
  const double dt = S.effective_dt;
  st = pf_mr_mark_cells(c);
  if (st) return st;
  record(c, timing, 2);
  st = run_p2g_splat(c, &cfg->phase, cfg->Vp, false, true);
  if (st) return st;
  record(c, timing, 3);
  st = step_face_stage(c, &cfg->phase, dt, cfg->gravity, true, !cfg->skip_interface_diagnostics,
                       cfg->surface_tension != 0, cfg->surface_tension_strength, cfg->surface_tension_max_delta_speed,
                       cfg->surface_tension_smoothing_radius, &S);
  if (st) return st;
  record(c, timing, 4);
  S.c_div = volume_correction(c, cfg->c_div_volume_correction, cfg->c_div_strength, cfg->liquid_volume_target,
                              cfg->Vp, dt, cfg->c_div, &S.liquid_volume);
  st = mr_project_with(c, &cfg->phase, dt, &cfg->solve, S.c_div, &S.solve, timing, 5, 6, 7);
  if (st) return st;
  HIP_TRY(c, fill_u32(c->stream, reinterpret_cast<uint32_t*>(c->esc_dev), 0u, 8));
  st = run_g2p_advect(c, true, true, cfg->alpha_liquid, cfg->alpha_gas, dt, cfg->advection_order);
  if (st) return st;
  record(c, timing, 8);
  HIP_TRY(c, hipMemcpyAsync(&S.escape, c->esc_dev, sizeof(int32_t) * 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (timing) {
    float ms;
    for (int q = 0; q < PF_NSTAGES; ++q) {
      hipEventElapsedTime(&ms, c->ev[q], c->ev[q + 1]);
      S.stage_ms[q] = ms;
    }
    hipEventElapsedTime(&ms, c->ev[0], c->ev[PF_NSTAGES]);
    S.total_ms = ms;
  }
  if (stats) *stats = S;
  return PF_OK;
}

}  // extern "C"
