// pf_host_lpc.hip -- the learned preconditioner: network buffers, forward and backward
// launches, tiles and features of a system, training, and the pf_lpc_ entry points.
//
// One of the library's host files: pf_api.hip includes them in order into a single
// translation unit together with the kernel files, so every launch is a direct call.

namespace {

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

// level-1 extents (the 2x2x2 aggregates of the multigrid cycle)
void lpc_level1_extents(const pf_ctx* c, int& nx1, int& ny1, int& nz1) {
  nx1 = (c->geom.nx + 1) / 2;
  ny1 = (c->geom.ny + 1) / 2;
  nz1 = (c->geom.nz + 1) / 2;
}

// 16^3 tiles of aggregates holding the FLUID blocks list[0..nlist)
int lpc_cw_tiles(pf_ctx* c, const uint32_t* list, long long nlist) {
  int nx1, ny1, nz1;
  lpc_level1_extents(c, nx1, ny1, nz1);
  c->ntx = (nx1 + 15) / 16; c->nty = (ny1 + 15) / 16; c->ntz = (nz1 + 15) / 16;
  const long long nall = (long long)c->ntx * c->nty * c->ntz;
  if (nall > c->ntiles_all) {
    dfree(c->tiles); dfree(c->tile_flags);
    HIP_TRY(c, dalloc(&c->tiles, nall));
    HIP_TRY(c, dalloc(&c->tile_flags, nall));
    c->ntiles_all = nall;
  }
  HIP_TRY(c, fill_u8(c->stream, c->tile_flags, 0, nall));
  if (nlist > 0) {
    hipLaunchKernelGGL(k_lpc_cw_tile_flags, dim3(grid_for(nlist, 256)), dim3(256), 0, c->stream, list, nlist, c->dc,
                       c->ntx, c->nty, c->tile_flags);
    LAUNCH_CHECK(c);
  }
  HIP_TRY(c, compact_flags(*c, c->tile_flags, nall, c->tiles, c->u32_dev + 8));
  HIP_TRY(c, hipMemcpyAsync(c->u32_host + 8, c->u32_dev + 8, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  c->ntiles = c->u32_host[8];
  return PF_OK;
}

CwArgs lpc_cw_args(pf_ctx* c) {
  CwArgs a{};
  a.dc = c->dc;
  a.g = c->geom;
  a.d1 = c->mg[0].d;
  lpc_level1_extents(c, a.nx1, a.ny1, a.nz1);
  a.ntx = c->ntx; a.nty = c->nty;
  a.tiles = c->tiles; a.ntiles = c->ntiles;
  a.marker = c->marker; a.fbeta = c->fbeta; a.fphi = c->fphi;
  a.cd = c->cd32; a.invd = c->invd32;
  a.X = c->lpc->X;
  return a;
}

unsigned lpc_tile_grid(long long nt) { return (unsigned)std::max<long long>(1, std::min<long long>(nt, 1024)); }

// the network's per-aggregate weights of the last assembled system into c->mg_cs_net
// (s = sqrt(w), 1 outside the tiles); store: keep the activations for a backward pass.
// Needs the multigrid levels (mg_build) and the tiles and features of the system.
int lpc_cw_weights(pf_ctx* c, bool store) {
  pf_lpc* L = c->lpc;
  if (!L) return fail(c, PF_ERR_STATE, "no network attached (pf_lpc_attach)");
  if (c->mg.empty()) return fail(c, PF_ERR_STATE, "learned preconditioner: the grid has no coarse level");
  const long long n1 = c->mg[0].d.nblocks * kBlockVol;
  if (!c->mg_cs_net) HIP_TRY(c, dalloc(&c->mg_cs_net, n1));
  HIP_TRY(c, fill_f32(c->stream, c->mg_cs_net, 1.f, n1));
  if (c->ntiles == 0) return PF_OK;
  int st = lpc_grow(L, c->ntiles, store);
  if (st) return fail(c, st, "learned preconditioner: " + L->err);
  CwArgs a = lpc_cw_args(c);
  hipLaunchKernelGGL(k_lpc_cw_gather, dim3((unsigned)c->ntiles), dim3(256), 0, c->stream, a);
  LAUNCH_CHECK(c);
  st = lpc_forward(L, c->stream, c->ntiles, store, nullptr);
  if (st) return fail(c, st, "learned preconditioner: " + L->err);
  hipLaunchKernelGGL(k_lpc_cw_scatter, dim3(lpc_tile_grid(c->ntiles)), dim3(256), 0, c->stream, a, L->o,
                     c->mg_cs_net);
  LAUNCH_CHECK(c);
  return PF_OK;
}

int lpc_features_alloc(pf_ctx* c) {
  if (c->fbeta) return PF_OK;
  const long long ncell = c->dc.nblocks * kBlockVol;
  HIP_TRY(c, dalloc(&c->fbeta, ncell));
  HIP_TRY(c, dalloc(&c->fphi, ncell));
  return PF_OK;
}

}  // namespace

