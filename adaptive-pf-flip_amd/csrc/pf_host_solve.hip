// pf_host_solve.hip -- the aggregation multigrid preconditioner (levels, cycle), the
// pressure projection with its PCG and FP64 refinement loop, and their entry points.
//
// One of the library's host files: pf_api.hip includes them in order into a single
// translation unit together with the kernel files, so every launch is a direct call.

namespace {

// ---- aggregation multigrid preconditioner (pf_mg.hip)

bool mg_cfg_valid(const pf_mg_cfg& q) {
  return q.sweeps >= 1 && q.sweeps <= 16 && q.coarsest_sweeps >= 1 && q.coarsest_sweeps <= 1024 &&
         q.omega > 0.0f && q.omega < 1.0f && q.correction > 0.0f && q.correction < 2.0f &&
         (q.chebyshev_ratio == 0.0f || (q.chebyshev_ratio > 1.0f && q.chebyshev_ratio <= 1000.0f));
}

// the built-in cycle (pf_set_multigrid changes a context's)
const pf_mg_cfg& mg_default_cfg() {
  static const pf_mg_cfg p{2, 32, 0.9f, 1.8f, 6.0f};
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
  for (auto& p : b.h) dfree(p);
  dfree(b.mx);
  dfree(b.unscale);
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
    for (auto& p : b.h)
      if (e == hipSuccess) e = dalloc(&p, n);
    if (e == hipSuccess) e = dalloc(&b.mx, 2);
    if (e == hipSuccess) e = dalloc(&b.unscale, 2);
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

int mg_coarsen_from(pf_ctx* c, MgLevel F, size_t from);

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
    if (!c->mg_mx_valid) {
      HIP_TRY(c, fill_u32(c->stream, c->mg_mx, 0u, 2));
      hipLaunchKernelGGL(k_mg_absmax, dim3(std::min<unsigned>(grid_for(n0, 256), 2048u)), dim3(256), 0, c->stream,
                         n0, c->cd32, c->cx32, c->cy32, c->cz32, c->invd32, c->mg_mx);
      c->mg_mx_valid = true;
    }
    _Float16* h[5];
    for (int q = 0; q < 5; ++q) h[q] = reinterpret_cast<_Float16*>(c->mg_h[q]);
    hipLaunchKernelGGL(k_mg_tohalf, dim3(std::min<unsigned>(grid_for(n0, 256), 4096u)), dim3(256), 0, c->stream, n0,
                       c->cd32, c->cx32, c->cy32, c->cz32, c->invd32, h[0], h[1], h[2], h[3], h[4], c->mg_mx,
                       c->mg_unscale);
    LAUNCH_CHECK(c);
  }
  return mg_coarsen_from(c, mg_level0(c), 0);
}

// a level's inverse diagonal and the FP16 copy its row kernels read (as level 0's)
int mg_level_finish(pf_ctx* c, Ctx::MgBuf& b) {
  const long long nl = b.d.nblocks * kBlockVol;
  const unsigned g = std::min<unsigned>(grid_for(nl, 256), 4096u);
  hipLaunchKernelGGL(k_mg_inverse, dim3(g), dim3(256), 0, c->stream, mg_level(b), b.inv);
  HIP_TRY(c, fill_u32(c->stream, b.mx, 0u, 2));
  hipLaunchKernelGGL(k_mg_absmax, dim3(std::min<unsigned>(grid_for(nl, 256), 2048u)), dim3(256), 0, c->stream, nl,
                     b.cd, b.cx, b.cy, b.cz, b.inv, b.mx);
  _Float16* h[5];
  for (int q = 0; q < 5; ++q) h[q] = reinterpret_cast<_Float16*>(b.h[q]);
  hipLaunchKernelGGL(k_mg_tohalf, dim3(std::min<unsigned>(grid_for(nl, 256), 4096u)), dim3(256), 0, c->stream, nl,
                     b.cd, b.cx, b.cy, b.cz, b.inv, h[0], h[1], h[2], h[3], h[4], b.mx, b.unscale);
  LAUNCH_CHECK(c);
  return PF_OK;
}

// levels c->mg[from..] by aggregation from F (the level above c->mg[from])
int mg_coarsen_from(pf_ctx* c, MgLevel F, size_t from) {
  for (size_t l = from; l < c->mg.size(); ++l) {
    Ctx::MgBuf& b = c->mg[l];
    const MgLevel Cl = mg_level(b);
    const unsigned g = std::min<unsigned>(grid_for(b.d.nblocks * kBlockVol, 256), 4096u);
    hipLaunchKernelGGL(k_mg_coarsen, dim3(g), dim3(256), 0, c->stream, F, Cl, b.cd, b.cx, b.cy, b.cz);
    int st = mg_level_finish(c, b);
    if (st) return st;
    F = Cl;
  }
  return PF_OK;
}

// The multires fluid list split by level for the fast solves: blocks of level 0-1 (a
// wave each) and level-2 blocks (one cell each, four to a wave).  The level-2 kernels
// write a block's cell entry only, so the PCG's z, p and A p start as zeros.
int mr_split_lists(pf_ctx* c, const MrGrid& m) {
  const long long nb = c->dc.nblocks;
  if (!c->mr_list2) {
    HIP_TRY(c, dalloc(&c->mr_f01, nb));
    HIP_TRY(c, dalloc(&c->mr_f2, nb));
    HIP_TRY(c, dalloc(&c->mr_list01, nb));
    HIP_TRY(c, dalloc(&c->mr_list2, nb));
  }
  HIP_TRY(c, fill_u8(c->stream, c->mr_f01, 0, nb));
  HIP_TRY(c, fill_u8(c->stream, c->mr_f2, 0, nb));
  hipLaunchKernelGGL(k_mr_mg_split, dim3(grid_for(c->npblk, 256)), dim3(256), 0, c->stream, c->pblk, c->npblk, m,
                     c->mr_f01, c->mr_f2);
  LAUNCH_CHECK(c);
  HIP_TRY(c, compact_flags_morton(*c, c->dc, c->mr_f01, c->mr_list01, c->u32_dev + 12));
  HIP_TRY(c, compact_flags_morton(*c, c->dc, c->mr_f2, c->mr_list2, c->u32_dev + 13));
  HIP_TRY(c, hipMemcpyAsync(c->u32_host + 12, c->u32_dev + 12, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost,
                            c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  c->mr_n01 = c->u32_host[12];
  c->mr_n2 = c->u32_host[13];
  if (c->mr_n01 + c->mr_n2 != c->npblk) return fail(c, PF_ERR_STATE, "multires block split lost blocks");
  float* arrs[4] = {c->z32, c->pA, c->pB, c->Ap};
  for (float* q : arrs) HIP_TRY(c, fill_f32(c->stream, q, 0.f, nb * kBlockVol));
  return PF_OK;
}

// The multires cycle's levels: the block grid (the uniform hierarchy's second level,
// c->mg[1]) from R S P, then the uniform levels below it.  Leaves of level 0..2 only.
int mr_mg_build(pf_ctx* c, const MrOp& op) {
  for (size_t a = 0; a < c->mr_leaves.size(); a += 4)
    if (c->mr_leaves[a] > 2)
      return fail(c, PF_ERR_ARG, "the multires multigrid takes leaves of level 0 to 2");
  if (!c->mg_t0) {
    int st = mg_alloc(c);
    if (st) return st;
  }
  if (c->mg.size() < 2) return fail(c, PF_ERR_ARG, "the multires multigrid needs a grid above 8 cells per axis");
  Ctx::MgBuf& b = c->mg[1];
  const long long nl = b.d.nblocks * kBlockVol;
  // block-grid cells of no listed block keep a zero residual through the solve
  HIP_TRY(c, fill_f32(c->stream, b.r, 0.f, nl));
  float** arrs[4] = {&b.cd, &b.cx, &b.cy, &b.cz};
  for (float** q : arrs) HIP_TRY(c, fill_f32(c->stream, *q, 0.f, nl));
  hipLaunchKernelGGL(k_mr_mg_coarse, dim3(grid_for(c->dc.nblocks, 4)), dim3(256), 0, c->stream, op, b.d, b.cd, b.cx,
                     b.cy, b.cz);
  LAUNCH_CHECK(c);
  // the level-2 kernels write the cell entry of a block only: the other entries of the
  // cycle's work array stay zero
  HIP_TRY(c, fill_f32(c->stream, c->mg_t0, 0.f, c->dc.nblocks * kBlockVol));
  int st = mg_level_finish(c, b);
  if (st) return st;
  return mg_coarsen_from(c, mg_level(b), 2);
}

// workgroups of the level-0 kernels: a multiple of 8 (the XCD-aware list walk), at
// most 1024 (the partials of r.z)
unsigned mg_grid0n(long long nlist) { return std::min(list_grid(nlist), 1024u); }
unsigned mg_grid0(const pf_ctx* c) { return mg_grid0n(c->npblk); }

// the weights of a smoothing pass's sweeps: all omega, or (a Chebyshev ratio r > 1)
// the reciprocals of the Chebyshev roots of degree nu on [lmax / r, lmax], lmax = 2
// (Gershgorin bound of D^-1 A), which damp that whole interval; the post-pass takes
// them in reverse order (they commute, so the cycle stays symmetric)
std::vector<float> mg_sweep_weights(const pf_ctx* c) {
  const int nu = c->mg_cfg.sweeps;
  const float cheb = c->mg_cfg.chebyshev_ratio;
  std::vector<float> w(nu, c->mg_cfg.omega);
  if (cheb > 1.0f) {
    const double b = 2.0, lo = b / cheb;
    for (int q = 0; q < nu; ++q) {
      const double lam = 0.5 * (b + lo) + 0.5 * (b - lo) * std::cos((2.0 * q + 1.0) * M_PI / (2.0 * nu));
      w[q] = (float)(1.0 / lam);
    }
  }
  return w;
}

// coarse levels: level l (1..nl) solves A_l x = r_l from zero (its buffer r holds r_l);
// returns the buffer that holds x.  (Two corrections per level -- a W-cycle -- cost
// more launches than they save iterations at 256^3, and twice the 1.8 over-correction
// stops converging.)
float* mg_coarse(pf_ctx* c, int l, const int* done, const std::vector<float>& wpre) {
  const int nl = (int)c->mg.size();
  const int nu = c->mg_cfg.sweeps;
  MgWalk wd{nullptr, 0, done};
  Ctx::MgBuf& b = c->mg[(size_t)l - 1];
  const MgLevel L = mg_level(b);
  float* x = b.x;
  float* t = b.t;
  if (l == nl) {
    hipLaunchKernelGGL(k_mg_coarsest, dim3(1), dim3(64), 0, c->stream, done, L, b.r, x, c->mg_cfg.omega,
                       c->mg_cfg.coarsest_sweeps);
    return x;
  }
  if (nu >= 2) {
    // by cell rows, as level 0 (k_mg_row over every block of the slab, the level's
    // FP16 operator): the two first sweeps fused, the restriction, then the
    // prolongation fused with the first post-sweep -- four passes instead of six
    Ctx::MgBuf& nb = c->mg[(size_t)l];
    MgRowArgs ra{};
    ra.list = nullptr;
    ra.nlist = b.d.nblocks;
    ra.done = done;
    ra.d = b.d;
    ra.dc = nb.d;
    ra.cd = reinterpret_cast<const _Float16*>(b.h[0]);
    ra.cx = reinterpret_cast<const _Float16*>(b.h[1]);
    ra.cy = reinterpret_cast<const _Float16*>(b.h[2]);
    ra.cz = reinterpret_cast<const _Float16*>(b.h[3]);
    ra.inv = reinterpret_cast<const _Float16*>(b.h[4]);
    ra.unscale = b.unscale;
    ra.r = b.r;
    ra.omega_c = c->mg_cfg.correction;
    ra.part = nullptr;
    const unsigned gw = list_grid(b.d.nblocks);
    float* cur = nullptr;
    float* nxt = x;
    auto flip = [&]() {
      cur = nxt;
      nxt = (cur == x) ? t : x;
    };
    ra.x = nullptr; ra.out = nxt;
    ra.omega_u = wpre[0];
    ra.omega = wpre[1];
    hipLaunchKernelGGL(k_mg_row<MG_FIRST2>, dim3(gw), dim3(256), 0, c->stream, ra);
    flip();
    for (int s2 = 2; s2 < nu; ++s2) {
      ra.x = cur; ra.out = nxt;
      ra.omega = wpre[(size_t)s2];
      hipLaunchKernelGGL(k_mg_row<MG_SWEEP>, dim3(gw), dim3(256), 0, c->stream, ra);
      flip();
    }
    ra.x = cur; ra.out = nb.r;
    hipLaunchKernelGGL(k_mg_row<MG_RESTRICT>, dim3(gw), dim3(256), 0, c->stream, ra);
    float* below = mg_coarse(c, l + 1, done, wpre);
    ra.x = cur; ra.out = nxt; ra.xc = below;
    ra.omega = wpre[(size_t)(nu - 1)];
    hipLaunchKernelGGL(k_mg_row<MG_PSWEEP>, dim3(gw), dim3(256), 0, c->stream, ra);
    flip();
    for (int s2 = 1; s2 < nu; ++s2) {
      ra.x = cur; ra.out = nxt;
      ra.omega = wpre[(size_t)(nu - 1 - s2)];
      hipLaunchKernelGGL(k_mg_row<MG_SWEEP>, dim3(gw), dim3(256), 0, c->stream, ra);
      flip();
    }
    return cur;
  }
  const unsigned g = std::min<unsigned>(grid_for(b.d.nblocks * kBlockVol, 256), 2048u);
  hipLaunchKernelGGL(k_mg_first, dim3(g), dim3(256), 0, c->stream, wd, L, b.r, x, wpre[0]);
  for (int s2 = 1; s2 < nu; ++s2) {
    hipLaunchKernelGGL(k_mg_sweep, dim3(g), dim3(256), 0, c->stream, wd, L, b.r, x, t, wpre[(size_t)s2], nullptr);
    std::swap(x, t);
  }
  Ctx::MgBuf& nb = c->mg[(size_t)l];
  hipLaunchKernelGGL(k_mg_restrict, dim3(g), dim3(256), 0, c->stream, wd, L, b.r, x, mg_level(nb), nb.r);
  float* below = mg_coarse(c, l + 1, done, wpre);
  hipLaunchKernelGGL(k_mg_prolong, dim3(g), dim3(256), 0, c->stream, wd, L, mg_level(nb), below, x,
                     c->mg_cfg.correction);
  for (int s2 = 0; s2 < nu; ++s2) {
    hipLaunchKernelGGL(k_mg_sweep, dim3(g), dim3(256), 0, c->stream, wd, L, b.r, x, t,
                       wpre[(size_t)(nu - 1 - s2)], nullptr);
    std::swap(x, t);
  }
  return x;
}

// level-1 quantities of one cycle, for training the correction weights: the restricted
// residual y before the scaling and the coarse solution B (S y) before the scaling
struct MgTrace {
  float* y = nullptr;
  float* h = nullptr;
};

// z = V-cycle(r) on the assembled system whose FLUID blocks are list[0..nlist);
// partial r.z (mg_grid0n(nlist) workgroups) in part[2 * kMaxPartials ...]; every
// kernel is skipped while *done
int mg_apply_on(pf_ctx* c, const uint32_t* list, long long nlist, const float* r, float* z, const int* done,
                double* part, const MgTrace* tr) {
  struct {
    int nu, coarsest;
    float omega, correction;
  } const P{c->mg_cfg.sweeps, c->mg_cfg.coarsest_sweeps, c->mg_cfg.omega, c->mg_cfg.correction};
  const int nl = (int)c->mg.size();
  MgWalk w0{list, nlist, done};
  const unsigned g0 = mg_grid0n(nlist);
  // kernels that write no partial may take more workgroups (up to 16 list blocks each)
  const unsigned gw = list_grid(nlist);
  const MgLevel L0 = mg_level0(c);
  double* const rz = part + 2 * kMaxPartials;
  // level 0 by cell rows when the slabs are below 2 GiB and there is a coarse level
  const bool rows = nl > 0 && c->dc.nblocks * kBlockVol * 4 < (1LL << 31);
  MgRowArgs ra{};
  ra.list = list;
  ra.nlist = nlist;
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
  const std::vector<float> wpre = mg_sweep_weights(c);
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
  if (tr && tr->y)
    HIP_TRY(c, hipMemcpyAsync(tr->y, c->mg[0].r, sizeof(float) * (size_t)n1, hipMemcpyDeviceToDevice, c->stream));
  if (c->mg_cs_active)
    hipLaunchKernelGGL(k_mg_scale, dim3(std::min<unsigned>(grid_for(n1, 256), 2048u)), dim3(256), 0, c->stream, n1,
                       c->mg[0].r, c->mg_cs_active, done);
  float* const res1 = mg_coarse(c, 1, done, wpre);
  if (tr && tr->h)
    HIP_TRY(c, hipMemcpyAsync(tr->h, res1, sizeof(float) * (size_t)n1, hipMemcpyDeviceToDevice, c->stream));
  // ... and on the correction (P S B_1 S P^T stays symmetric)
  if (c->mg_cs_active)
    hipLaunchKernelGGL(k_mg_scale, dim3(std::min<unsigned>(grid_for(n1, 256), 2048u)), dim3(256), 0, c->stream, n1,
                       res1, c->mg_cs_active, done);
  // level 0: prolong and post-smooth, the last sweep into z with the partial r.z
  if (rows) {
    ra.x = cur; ra.out = nxt; ra.xc = res1; ra.part = P.nu == 1 ? rz : nullptr;
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
    hipLaunchKernelGGL(k_mg_prolong, dim3(g0), dim3(256), 0, c->stream, w0, L0, mg_level(c->mg[0]), res1, cur,
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

int mg_apply(pf_ctx* c, const float* r, float* z, const int* done, double* part) {
  return mg_apply_on(c, c->pblk, c->npblk, r, z, done, part, nullptr);
}

// z = the multires cycle on r over the listed blocks: nu sweeps on S, the block sums of
// the residual onto the block grid, the uniform cycle from there, the prolongation and
// nu sweeps back, the last into z with the partial r.z (mg_grid0n workgroups) in slot 2
// of part; skipped while *done
int mr_mg_apply(pf_ctx* c, const MrOp& op, const VecArgs& va, const float* r, float* z, const int* done,
                double* part) {
  const int nu = std::max(1, c->mg_cfg.sweeps);
  std::vector<float> wpre = mg_sweep_weights(c);
  wpre.resize((size_t)nu, c->mg_cfg.omega);
  const unsigned g0 = mg_grid0n(va.nlist);
  double* const rz = part + 2 * kMaxPartials;
  const float* inv = c->invd32;
  // the blocks of level 0-1 (a wave each) and of level 2 (four to a wave)
  VecArgs v01 = va;
  v01.list = c->mr_list01;
  v01.nlist = c->mr_n01;
  const uint32_t* l2 = c->mr_list2;
  const long long n2 = c->mr_n2;
  const unsigned gw = list_grid(v01.nlist);
  const unsigned gw2 = (unsigned)std::clamp<long long>((n2 + 15) / 16, 8, 4096);
  // 2 nu writes (the prolongation adds in place): the first lands in the work array
  float* cur = nullptr;
  float* nxt = c->mg_t0;
  auto flip = [&]() {
    cur = nxt;
    nxt = (cur == z) ? c->mg_t0 : z;
  };
  // one sweep over both lists; the last one leaves the partial r.z in slot 2 (the
  // level-2 kernel adds its share to the partials the other one wrote)
  auto sweep = [&](float w, bool last) {
    hipLaunchKernelGGL(k_mr_mg_sweep, dim3(last ? g0 : gw), dim3(256), 0, c->stream, v01, op, done, inv, r, cur, nxt, w,
                       last ? rz : nullptr);
    hipLaunchKernelGGL(k_mr_mg_sweep2, dim3(last ? g0 : gw2), dim3(256), 0, c->stream, l2, n2, op, done, inv, r, cur,
                       nxt, w, last ? rz : nullptr);
    flip();
  };
  for (int q = 0; q < nu; ++q) sweep(wpre[(size_t)q], false);
  Ctx::MgBuf& b = c->mg[1];
  hipLaunchKernelGGL(k_mr_mg_restrict, dim3(gw), dim3(256), 0, c->stream, v01, op, done, r, cur, b.d, b.r);
  hipLaunchKernelGGL(k_mr_mg_restrict2, dim3(gw2), dim3(256), 0, c->stream, l2, n2, op, done, r, cur, b.d, b.r);
  float* const xc = mg_coarse(c, 2, done, wpre);
  hipLaunchKernelGGL(k_mr_mg_prolong, dim3(gw), dim3(256), 0, c->stream, v01, op, done, b.d, xc, cur,
                     c->mg_cfg.correction);
  hipLaunchKernelGGL(k_mr_mg_prolong2, dim3(gw2), dim3(256), 0, c->stream, l2, n2, op, done, b.d, xc, cur,
                     c->mg_cfg.correction);
  for (int q = 0; q < nu; ++q) sweep(wpre[(size_t)(nu - 1 - q)], q == nu - 1);
  LAUNCH_CHECK(c);
  if (cur != z) return fail(c, PF_ERR_STATE, "multires multigrid: cycle did not end in z");
  return PF_OK;
}

// single: the single-phase unit-density projection (ph unused, cdiv ignored);
// c->mr: the multires finite-volume projection, stopping when
// |rhs - A x|_inf < max(cfg->tol, mr_rel_tol * |rhs|_inf)
int run_project(pf_ctx* c, const pf_phase* ph, double dt, const pf_solve_cfg* cfg, double cdiv,
                pf_solve_stats* out, bool timing, int ev_assemble, int ev_solve, int ev_vel,
                bool single = false, double mr_rel_tol = 0.0) {
  const bool mr = c->mr;
  if (cfg->precond < PF_PRECOND_JACOBI || cfg->precond > PF_PRECOND_MULTIGRID)
    return fail(c, PF_ERR_ARG, "unknown preconditioner");
  pf_solve_stats S;
  std::memset(&S, 0, sizeof(S));
  S.pinned_cell = -1;
  const Geom& g = c->geom;
  const long long nb = c->dc.nblocks;
  // 1. the blocks holding FLUID cells (flags; sorted list below)
  c->mg_mx_valid = false;
  uint8_t* flag8 = reinterpret_cast<uint8_t*>(c->blk_flag32);
  hipLaunchKernelGGL(k_fluid_flags, dim3(grid_for(nb, 4)), dim3(256), 0, c->stream, c->dc, c->marker, flag8);
  LAUNCH_CHECK(c);
  // 2. blocks of the previous solve return to zero; the pressure field is cleared
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
    // Cleared on every block of the previous solve: the pressure and the iteration
    // vectors (with the multigrid work array: its values at cells outside the next
    // system are read, times a zero coupling, by the level-0 sweeps, so none may stay
    // non-finite).  Only on the blocks that leave the system (sparse path): the
    // assembly's outputs, which k_assemble writes on all 64 cells of every listed
    // block, x64 (k_ir_start writes it there) and r64 (written before it is read).
    int nf = 0;
    for (float* p : {c->f[F_P], c->x32, c->r32, c->z32, c->pA, c->pB, c->Ap, c->mg_t0})
      if (p) z.f32[nf++] = p;
    z.nf32_all = nf;
    for (float* p : {c->cd32, c->cx32, c->cy32, c->cz32, c->invd32}) z.f32[nf++] = p;
    z.nf32 = nf;
    double* f64[7] = {c->b64, c->r64, c->cd64, c->cx64, c->cy64, c->cz64, c->x64};
    for (int q = 0; q < 7; ++q) z.f64[q] = f64[q];
    z.nf64 = warm ? 6 : 7;  // a warm start keeps x64 (the saved pressure)
    z.keep = mr ? nullptr : flag8;
    hipLaunchKernelGGL(k_zero_blocks, dim3(grid_for(c->npblk_prev, 4)), dim3(256), 0, c->stream, z);
    LAUNCH_CHECK(c);
  }
  HIP_TRY(c, fill_u8(c->stream, c->act[F_P], 0, nb));
  // 3. the sorted list
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
  // learned preconditioner: the multigrid cycle with per-aggregate correction weights
  // from the attached network; its tiles and features of this system are built for
  // the solve or for recording training snapshots
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
    if (!s2) s2 = lpc_cw_tiles(c, c->pblk, npb);
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
  // the FP16 scaling of a multigrid build that may follow needs the operator's maxima:
  // the assembly reduces them from the values it writes (cells outside the list are zero)
  const bool want_mx = want_mg || want_learned || lpc_system;
  ao.mx = nullptr;
  if (want_mx) {
    if (!c->mg_t0) {
      int s2 = mg_alloc(c);
      if (s2) return s2;
    }
    HIP_TRY(c, fill_u32(c->stream, c->mg_mx, 0u, 2));
    ao.mx = c->mg_mx;
  }
  TIMED(c, PF_K_ASSEMBLE, hipLaunchKernelGGL(k_assemble, dim3(G), dim3(256), 0, c->stream, pa, ao));
  LAUNCH_CHECK(c);
  c->mg_mx_valid = want_mx;
  }
  // red[2] = sum b^2; red[3] = max |b| (sparse assembly only)
  hipLaunchKernelGGL(k_reduce_slots, dim3(1), dim3(256), 0, c->stream, c->partial, (int)G, mr ? 1 : 2, 2u,
                     c->red + 2);
  LAUNCH_CHECK(c);
  record(c, timing, ev_assemble);

  // 4. FP64 iterative refinement around the FP32 Jacobi PCG
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
  const double* r_src = c->r64;  // the FP64 residual the next (re)start loads
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
    r_src = c->r64;
    return read_small(c);
  };
  // From x = 0 (no warm start) the FP64 residual is b itself: its sum of squares and max
  // norm come from the assembly's reduction (the same partition and order, so the same
  // bits), and the first load reads b64 -- no residual pass.
  const bool zero_start = !mr && !cfg->warm_start;
  int st = zero_start ? read_small(c) : residual();
  if (st) return st;
  if (zero_start) r_src = c->b64;
  const double bnorm2 = c->red_host[2];
  double rr = zero_start ? c->red_host[2] : c->red_host[0], rmax = zero_start ? c->red_host[3] : c->red_host[1];
  S.rhs_norm2 = bnorm2;
  S.initial_max_residual = rmax;
  S.pinned = c->u32_host[0] == 0;
  S.pinned_cell = S.pinned ? (int64_t)c->u32_host[1] : -1;
  const int max_iters = std::max(cfg->max_iters, 0);
  const int max_updates = cfg->max_outer > 0 ? cfg->max_outer : 200;
  // (a multigrid iteration costs several products: check at least every 4)
  const int check_base = cfg->check_every > 0 ? cfg->check_every : 8;
  const int check = (want_mg || want_learned) ? std::min(check_base, 4) : check_base;
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
  constexpr int between = 200;
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
  // flexible PCG around the multigrid V-cycle, plain or with the network's correction
  // weights (z and partial r.z in slot 2, GT partials)
  bool learned = want_learned || want_mg;
  c->mg_cs_active = c->mg_cs;  // caller-set weights (pf_mg_set_correction_weights) or none
  if (mr && !refmode) {
    st = mr_split_lists(c, mop.m);
    if (st) return st;
  }
  if (learned) {
    st = mr ? mr_mg_build(c, mop) : mg_build(c);
    if (st) return st;
  }
  if (want_learned && !c->mg.empty()) {
    // the network runs once per solve; the cycle then costs what the plain one does
    long long ki = kbegin(c, PF_K_LPC_APPLY);
    st = lpc_cw_weights(c, false);
    kend(c, ki);
    if (st) return st;
    c->mg_cs_active = c->mg_cs_net;
  }
  const unsigned GT = learned ? mg_grid0(c) : 1u;
  auto papply = [&](const int* done) -> int {
    long long ki = kbegin(c, PF_K_MG_APPLY);
    int s2 = mr ? mr_mg_apply(c, mop, va, c->r32, c->z32, done, c->partial)
                : mg_apply(c, c->r32, c->z32, done, c->partial);
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
      hipLaunchKernelGGL(k_pcg_load, dim3(G), dim3(256), 0, c->stream, va, r_src, c->x32, c->r32, c->z32, c->pA,
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
    // (the multires SpMV waits on dependent loads per block: it takes more workgroups)
    const unsigned GS = std::min(G, mr ? 4096u : 1024u);
    // the buffer-addressed SpMV takes 32-bit byte offsets: arrays below 2 GiB
    const bool spmv_buffers = c->dc.nblocks * kBlockVol * 4 < (1LL << 31);
    VecArgs vs = va;
    // the multires SpMV: blocks of level 0-1 here, level-2 blocks in a second launch
    VecArgs vs01 = va;
    vs01.list = c->mr_list01;
    vs01.nlist = c->mr_n01;
    int iter_before = 0, iter_base = 0;
    double best_rr = rr, prev_rr = rr;
    int no_gain = 0, learned_slow = 0;
    // With a cycle preconditioner an iteration costs ~0.8 ms at 256^3, so iterations
    // queued past the replacement trigger are worth avoiding: the next look at the done
    // flag comes after the predicted number of iterations, from the FP32 residual's
    // observed decay per iteration toward max(tol2 rr_ref, tol2_abs) (1..check)
    const bool predict = want_mg || want_learned;
    int q_next = check;
    double rate = -1.0, rr_last = -1.0;
    int iter_last = 0;
    for (;;) {
      const size_t rec0 = c->krec.size();
      for (int q = 0; q < q_next; ++q) {
        if (mr)
          TIMED(c, PF_K_PCG_SPMV,
                hipLaunchKernelGGL(k_mr_spmv, dim3(GS), dim3(256), 0, c->stream, vs01, mop, c->z32, c->pA, c->pB,
                                   c->Ap);
                hipLaunchKernelGGL(k_mr_spmv2, dim3(GS), dim3(256), 0, c->stream, c->mr_list2, c->mr_n2, vs, mop,
                                   c->z32, c->pA, c->pB, c->Ap));
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
                hipLaunchKernelGGL(k_pcg_update_flex, dim3(GS), dim3(256), 0, c->stream, mr ? vs01 : vs, c->x32,
                                   c->r32, c->z32, c->pA, c->pB, c->Ap);
                if (mr) hipLaunchKernelGGL(k_mr_update2<true>, dim3(GS), dim3(256), 0, c->stream, c->mr_list2,
                                           c->mr_n2, vs, c->x32, c->r32, c->z32, c->pA, c->pB, c->Ap, c->invd32));
          st = papply(done_dev);
          if (st) return st;
          hipLaunchKernelGGL(k_pcg_beta_flex, dim3(1), dim3(256), 0, c->stream, c->partial, (int)GS, (int)GT, c->sc);
        } else {
          TIMED(c, PF_K_PCG_UPDATE,
                hipLaunchKernelGGL(k_pcg_update, dim3(GS), dim3(256), 0, c->stream, mr ? vs01 : vs, c->x32, c->r32,
                                   c->z32, c->pA, c->pB, c->Ap, c->invd32);
                if (mr) hipLaunchKernelGGL(k_mr_update2<false>, dim3(GS), dim3(256), 0, c->stream, c->mr_list2,
                                           c->mr_n2, vs, c->x32, c->r32, c->z32, c->pA, c->pB, c->Ap, c->invd32));
          hipLaunchKernelGGL(k_pcg_beta, dim3(1), dim3(256), 0, c->stream, c->partial, (int)GS, c->sc);
        }
      }
      LAUNCH_CHECK(c);
      HIP_TRY(c, hipMemcpyAsync(c->sc_host, c->sc, sizeof(SolveScalars), hipMemcpyDeviceToHost, c->stream));
      HIP_TRY(c, hipStreamSynchronize(c->stream));
      std::memcpy(c->done_host, &c->sc_host->iter, sizeof(int) * 8);
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
      if (!c->done_host[1]) {
        q_next = check;
        if (predict) {
          const SolveScalars& H = *c->sc_host;
          const int it = c->done_host[0];
          if (rr_last > 0.0 && it > iter_last && H.rr > 0.0 && H.rr < rr_last)
            rate = std::pow(H.rr / rr_last, 1.0 / (it - iter_last));
          rr_last = H.rr;
          iter_last = it;
          const double target = std::max(H.tol2 * H.rr_ref, H.tol2_abs);
          if (rate > 0.0 && rate < 1.0 && H.rr > target)
            q_next = std::max(1, std::min(check, (int)std::ceil(std::log(target / H.rr) / std::log(rate) - 1e-9)));
        }
        continue;
      }
      // a new segment follows (residual replacement): it ends after about
      // log(tol2) / log(rate) iterations; its first look measures the decay afresh
      rr_last = -1.0;
      q_next = check;
      if (predict && rate > 0.0 && rate < 1.0 && c->sc_host->tol2 > 0.0 && c->sc_host->tol2 < 1.0)
        q_next = std::max(1, std::min(check, (int)std::ceil(std::log(c->sc_host->tol2) / std::log(rate) - 1e-9)));
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

  // 5. pressure store and face velocity update
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

}  // namespace

extern "C" {

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

int pf_project(pf_ctx* c, const pf_phase* ph, double dt, const pf_solve_cfg* cfg, double div_correction,
               pf_solve_stats* stats) {
  if (!c || !ph || !cfg) return PF_ERR_ARG;
  return run_project(c, ph, dt, cfg, div_correction, stats, false, 0, 0, 0);
}

int pf_project_single(pf_ctx* c, double dt, const pf_solve_cfg* cfg, pf_solve_stats* stats) {
  if (!c || !cfg) return PF_ERR_ARG;
  return run_project(c, nullptr, dt, cfg, 0.0, stats, false, 0, 0, 0, true);
}

}  // extern "C"
