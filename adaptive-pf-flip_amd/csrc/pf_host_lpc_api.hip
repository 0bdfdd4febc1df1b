// pf_host_lpc_api.hip -- the pf_lpc_ entry points of the learned preconditioner (weights,
// tiles forward and backward, attach, recording, loss and training).
//
// One of the library's host files: pf_api.hip includes them in order into a single
// translation unit together with the kernel files, so every launch is a direct call.
extern "C" {

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

}  // extern "C"
