// pf_host_adapt.hip -- particle adaptivity (narrow-band air, coarsening, refill) and its entry point.
//
// One of the library's host files: pf_api.hip includes them in order into a single
// translation unit together with the kernel files, so every launch is a direct call.

extern "C" {

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
// per workgroup, combined in a fixed order)
int update_liquid_volume(pf_ctx* c) {
  if (c->np == 0) {
    c->liquid_volume = 0.0;
    return PF_OK;
  }
  const unsigned nwg = std::min<unsigned>(grid_for(c->np, 256), 1024u);
  hipLaunchKernelGGL(k_pa_liquid_volume, dim3(nwg), dim3(256), 0, c->stream, c->pa, c->np, c->has_vol, c->partial);
  const CombineOps ops{1, {0}};
  hipLaunchKernelGGL(k_iface_combine, dim3(1), dim3(kCombineThreads), 0, c->stream, c->partial, (int)nwg, ops, c->red);
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
    hipFree(c->rank);
    c->rank = nullptr;
    HIP_TRY(c, dalloc(&c->rank, cap));
    c->pcap = cap;
  } else {
    std::swap(c->pa, c->pb);
  }
  c->np = n_new;
  c->binned = false;
  c->vmax_valid = false;
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

}  // extern "C"
