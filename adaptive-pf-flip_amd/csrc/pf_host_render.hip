// pf_host_render.hip -- the render-cache JSONL export and its manifest.
//
// One of the library's host files: pf_api.hip includes them in order into a single
// translation unit together with the kernel files, so every launch is a direct call.

extern "C" {

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

}  // extern "C"
