// pf_host_particles.hip -- binning, marking, the P2G splat launches, the face-field pass,
// G2P and advection, and the rho_tilde_0 calibration.
//
// One of the library's host files: pf_api.hip includes them in order into a single
// translation unit together with the kernel files, so every launch is a direct call.

namespace {

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
    hipFree(c->rank);
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
  // the cell index in exact FP32 where the origin is 0 and dx a power of two
  const Geom& g = c->geom;
  const float fscale =
      (g.ox == 0.0 && g.oy == 0.0 && g.oz == 0.0 && is_pow2(g.dx)) ? (float)(1.0 / g.dx) : 0.f;
  if (c->np > 0) {
    TIMED(c, PF_K_BIN_COUNT,
          hipLaunchKernelGGL(k_bin_count, dim3(grid_for(c->np, 256)), dim3(256), 0, c->stream, c->np, c->pa,
                             c->geom, c->dc, fscale, c->rank, c->cell_count));
    LAUNCH_CHECK(c);
  }
  HIP_TRY(c, scan_u32(*c, c->cell_count, c->cell_start, ncell));
  if (c->np > 0) {
    TIMED(c, PF_K_BIN_SCATTER,
          hipLaunchKernelGGL(k_bin_scatter, dim3(grid_for(c->np, 256)), dim3(256), 0, c->stream, c->np, c->pa,
                             c->pb, c->has_vol, c->geom, c->dc, fscale, c->rank, c->cell_start));
    LAUNCH_CHECK(c);
    std::swap(c->pa, c->pb);
    // cells of more than kRankMax particles (piles against walls) sorted by caller id,
    // so that P2G walks them in order without a quadratic selection
    const long long cap = c->np / (kRankMax + 1) + 1;
    if (cap > c->crowd_cap) {
      dfree(c->crowd_list);
      HIP_TRY(c, dalloc(&c->crowd_list, cap));
      c->crowd_cap = cap;
    }
    HIP_TRY(c, fill_u32(c->stream, c->u32_dev + 16, 0u, 1));
    hipLaunchKernelGGL(k_crowded_cells, dim3(grid_for(ncell, 256)), dim3(256), 0, c->stream, ncell, c->cell_start,
                       (uint32_t)kRankMax, (uint32_t)kSortedMax, c->crowd_list, c->u32_dev + 16, c->crowd_cap);
    hipLaunchKernelGGL(k_sort_crowded, dim3(512), dim3(256), 0, c->stream, c->crowd_list, c->u32_dev + 16,
                       c->cell_start, c->pa, c->has_vol);
    LAUNCH_CHECK(c);
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
    const P2GShared sh = p2g_shared_args(a);
    const P2GField f0 = p2g_field_args(a, 0), f1 = p2g_field_args(a, 1), f2 = p2g_field_args(a, 2);
    if (mr)
      TIMED(c, PF_K_P2G, hipLaunchKernelGGL((k_p2g_color<2, false>), grid, dim3(64), 0, c->stream, sh, f0, f1, f2));
    else if (single)
      TIMED(c, PF_K_P2G, hipLaunchKernelGGL((k_p2g_color<1, false>), grid, dim3(64), 0, c->stream, sh, f0, f1, f2));
    else if (a.fast32)
      TIMED(c, PF_K_P2G, hipLaunchKernelGGL((k_p2g_color<0, true>), grid, dim3(64), 0, c->stream, sh, f0, f1, f2));
    else
      TIMED(c, PF_K_P2G, hipLaunchKernelGGL((k_p2g_color<0, false>), grid, dim3(64), 0, c->stream, sh, f0, f1, f2));
    LAUNCH_CHECK(c);
  }
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
  const Dims* fd[3] = {&c->du, &c->dv, &c->dw};
  if (g2p)
    for (int ax = 0; ax < 3; ++ax) {
      if (!c->g2p_quad[ax]) HIP_TRY(c, dalloc(&c->g2p_quad[ax], fd[ax]->nblocks * kBlockVol));
      a.quad[ax] = c->g2p_quad[ax];
    }
  // the G2P reduces the max speed of the velocities it writes (the next CFL scan's value)
  if (g2p) HIP_TRY(c, fill_u32(c->stream, reinterpret_cast<uint32_t*>(c->vmax_dev), 0u, 1));
  {
    // (the kernel timer covers the packing and the kernel)
    const long long ki = kbegin(c, PF_K_G2P_ADVECT);
    const unsigned grid = xcd_grid(c->np, 256, 8192);
    if (g2p) {
      for (int ax = 0; ax < 3; ++ax)
        hipLaunchKernelGGL(k_pack_quad, dim3(std::min<unsigned>(grid_for(fd[ax]->nblocks * kBlockVol, 256), 8192u)),
                           dim3(256), 0, c->stream, *fd[ax], c->f[F_U + ax], c->f[F_SU + ax], c->g2p_quad[ax]);
      hipLaunchKernelGGL(k_g2p_advect<true>, dim3(grid), dim3(256), 0, c->stream, a);
    } else {
      hipLaunchKernelGGL(k_g2p_advect<false>, dim3(grid), dim3(256), 0, c->stream, a);
    }
    kend(c, ki);
  }
  LAUNCH_CHECK(c);
  if (g2p) c->vmax_valid = true;
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
