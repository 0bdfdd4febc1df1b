// pf_host_mr_api.hip -- the pf_mr_ entry points (multires layout, marking, P2G, projection,
// dynamic refinement and the multires step).
//
// One of the library's host files: pf_api.hip includes them in order into a single
// translation unit together with the kernel files, so every launch is a direct call.
extern "C" {

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

int pf_mr_scene_bubbles(pf_ctx* c, uint64_t seed, int32_t nb, int32_t band, float* bubbles) {
  if (!c || nb < 0 || nb > kMaxSceneBubbles || band < 1) return PF_ERR_ARG;
  const Geom& g = c->geom;
  if (g.ox != 0.0 || g.oy != 0.0 || g.oz != 0.0)
    return fail(c, PF_ERR_ARG, "the multires path takes positions relative to the origin 0");
  const MrScene sc = mr_draw_scene(g, seed, nb);
  if (bubbles)
    for (int b = 0; b < nb; ++b) {
      bubbles[4 * b] = sc.b[b].x; bubbles[4 * b + 1] = sc.b[b].y;
      bubbles[4 * b + 2] = sc.b[b].z; bubbles[4 * b + 3] = sc.b[b].w;
    }
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  int st = mr_apply_layout(c, mr_layout_scene(g, sc, band));
  if (st) return st;
  // particles: count per leaf, prefix on the host, then write
  const int nl = (int)(c->mr_leaves.size() / 4);
  const unsigned grid = (unsigned)((nl + 3) / 4);
  uint32_t* cnt = nullptr;
  long long* off = nullptr;
  HIP_TRY(c, dalloc(&cnt, nl));
  if (hipMalloc((void**)&off, sizeof(long long) * (size_t)nl) != hipSuccess) {
    dfree(cnt);
    return fail(c, PF_ERR_HIP, "pf_mr_scene_bubbles: allocation failed");
  }
  std::vector<uint32_t> hc((size_t)nl);
  std::vector<long long> ho((size_t)nl);
  long long n = 0;
  auto run = [&]() -> int {
    hipLaunchKernelGGL(k_mr_scene<true>, dim3(grid), dim3(256), 0, c->stream, c->mr_leaves_dev, nl, g, sc, seed, cnt,
                       nullptr, c->pa);
    LAUNCH_CHECK(c);
    HIP_TRY(c, hipMemcpyAsync(hc.data(), cnt, sizeof(uint32_t) * (size_t)nl, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    for (int a = 0; a < nl; ++a) {
      ho[(size_t)a] = n;
      n += hc[(size_t)a];
    }
    HIP_TRY(c, hipMemcpyAsync(off, ho.data(), sizeof(long long) * (size_t)nl, hipMemcpyHostToDevice, c->stream));
    int s2 = grow_particles(c, n, false);
    if (s2) return s2;
    if (n > 0) {
      hipLaunchKernelGGL(k_mr_scene<false>, dim3(grid), dim3(256), 0, c->stream, c->mr_leaves_dev, nl, g, sc, seed,
                         cnt, off, c->pa);
      LAUNCH_CHECK(c);
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return PF_OK;
  };
  st = run();
  dfree(cnt);
  dfree(off);
  if (st) return st;
  c->np = n;
  c->has_vol = false;
  c->vmax_valid = false;
  c->binned = false;
  std::vector<uint8_t> t((size_t)n);
  if (n > 0) {
    HIP_TRY(c, hipMemcpyAsync(t.data(), c->pa.type, (size_t)n, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
  }
  long long liq = 0;
  for (uint8_t q : t) liq += q == 0;
  c->liquid_volume = (double)liq;
  return PF_OK;
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
  s.precond = m->mode == PF_MR_SOLVE_MULTIGRID ? PF_PRECOND_MULTIGRID : PF_PRECOND_JACOBI;
  return s;
}

// the multires projection with `cfg` in force for this solve only
static int mr_project_with(pf_ctx* c, const pf_phase* ph, double dt, const pf_mr_solve_cfg* cfg, double c_div,
                           pf_solve_stats* stats, bool timing, int e0, int e1, int e2) {
  if (cfg->max_iterations < 0 || cfg->absolute_tolerance < 0 || cfg->relative_tolerance < 0)
    return fail(c, PF_ERR_ARG, "invalid multires solve configuration");
  if (cfg->mode != PF_MR_SOLVE_FAST && cfg->mode != PF_MR_SOLVE_REFERENCE && cfg->mode != PF_MR_SOLVE_MULTIGRID)
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
