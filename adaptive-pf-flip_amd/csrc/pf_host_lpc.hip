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

}  // namespace

