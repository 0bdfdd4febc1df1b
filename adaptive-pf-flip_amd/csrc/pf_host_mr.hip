// pf_host_mr.hip -- the multires layout (leaves, refinement, dynamic refinement), the
// reference-order FP64 multires solve and the pf_mr_ entry points.
//
// One of the library's host files: pf_api.hip includes them in order into a single
// translation unit together with the kernel files, so every launch is a direct call.

namespace {

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

// The multires multi-bubble tank (a three-level generalisation of
// initBubbleTankInterfaceBand, multires_sim3d_tp.cpp:457-500): water in the fine cells
// j < ny/2 and nb bubbles drawn from the seed, radius U[0.02, 0.06] nx dx, centres
// uniform over the positions that keep each bubble two cells inside the water.
MrScene mr_draw_scene(const Geom& g, uint64_t seed, int nb) {
  MrScene sc{};
  sc.nb = nb;
  sc.water = g.ny / 2;
  uint64_t s = seed ^ 0x5DEECE66DULL;
  auto uni = [&]() { return (double)(splitmix64(s) >> 11) * 0x1.0p-53; };
  const double L = g.nx * g.dx, m = 2.0 * g.dx;
  const double lo[3] = {0.0, 0.0, 0.0}, hi[3] = {g.nx * g.dx, sc.water * g.dx, g.nz * g.dx};
  for (int b = 0; b < nb; ++b) {
    const double r = (0.02 + 0.04 * uni()) * L;
    double c[3];
    for (int a = 0; a < 3; ++a) {
      const double span = hi[a] - lo[a] - 2.0 * (r + m);
      c[a] = span > 0.0 ? lo[a] + r + m + uni() * span : 0.5 * (lo[a] + hi[a]);
    }
    sc.b[b] = make_float4((float)c[0], (float)c[1], (float)c[2], (float)r);
  }
  return sc;
}

// Distance in fine cells from the cell centres of the fine box [x0, x1) x [y0, y1) x
// [z0, z1) to the nearest interface of the scene: the free surface y = water and the
// bubble spheres (0 where the box straddles one).
double mr_scene_distance(const Geom& g, const MrScene& sc, int x0, int y0, int z0, int x1, int y1, int z1) {
  const double lo[3] = {x0 + 0.5, y0 + 0.5, z0 + 0.5}, hi[3] = {x1 - 0.5, y1 - 0.5, z1 - 0.5};
  double d = std::max(0.0, std::max(lo[1] - sc.water, sc.water - hi[1]));
  for (int b = 0; b < sc.nb; ++b) {
    const double c[3] = {sc.b[b].x / g.dx, sc.b[b].y / g.dx, sc.b[b].z / g.dx}, r = sc.b[b].w / g.dx;
    double near = 0.0, far = 0.0;
    for (int a = 0; a < 3; ++a) {
      const double e = std::max(0.0, std::max(lo[a] - c[a], c[a] - hi[a]));
      const double f = std::max(std::abs(lo[a] - c[a]), std::abs(hi[a] - c[a]));
      near += e * e;
      far += f * f;
    }
    d = std::min(d, std::max(0.0, std::max(std::sqrt(near) - r, r - std::sqrt(far))));
  }
  return d;
}

// Three levels by distance to the nearest interface: a level-2 leaf (16 fine cells)
// within band + 16 cells splits into level-1 leaves, and those within band cells into
// level-0 leaves.  A level-2 leaf never touches a level-0 one (2:1 balance): the
// distance over cell centres is 1-Lipschitz, a level-0 leaf's level-1 parent holds a
// centre within band, and every centre next to that parent lies within 8 sqrt(3) < 16
// of it.
std::vector<Leaf> mr_layout_scene(const Geom& g, const MrScene& sc, int band) {
  std::vector<Leaf> out;
  auto dist = [&](const Leaf& k) {
    const int s = 4 << k[0];
    return mr_scene_distance(g, sc, k[1] * s, k[2] * s, k[3] * s, std::min(g.nx, k[1] * s + s),
                             std::min(g.ny, k[2] * s + s), std::min(g.nz, k[3] * s + s));
  };
  std::function<void(const Leaf&)> place = [&](const Leaf& k) {
    const double lim = k[0] == 2 ? band + 16.0 : (double)band;
    if (k[0] == 0 || dist(k) > lim) {
      out.push_back(k);
      return;
    }
    for (int c = 0; c < 8; ++c) {
      const Leaf ch = {k[0] - 1, 2 * k[1] + (c & 1), 2 * k[2] + ((c >> 1) & 1), 2 * k[3] + (c >> 2)};
      if (mr_block_in_domain(g.nx, g.ny, g.nz, ch)) place(ch);
    }
  };
  for (int bz = 0; bz < (g.nz + 15) / 16; ++bz)
    for (int by = 0; by < (g.ny + 15) / 16; ++by)
      for (int bx = 0; bx < (g.nx + 15) / 16; ++bx) place({2, bx, by, bz});
  std::sort(out.begin(), out.end());
  return out;
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
    // level1ParentTuple (:39-43) throws above level 1: the coarse space has no
    // aggregates for leaves of level 2 and up
    for (size_t a = 0; a < c->mr_leaves.size(); a += 4)
      if (c->mr_leaves[a] > 1)
        return fail(c, PF_ERR_ARG, "the level-1 coarse space takes leaves of level 0 and 1 only");
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

}  // namespace

