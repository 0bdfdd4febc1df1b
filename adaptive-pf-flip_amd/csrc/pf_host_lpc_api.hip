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
  // He initialisation: N(0, 2 / fan_in) for the ReLU layers; the head starts at zero,
  // so a new network outputs 0 everywhere (weights w = 1: the plain cycle)
  std::vector<float> w(kNW);
  uint64_t st = seed;
  const int off[4] = {kOffW1, kOffW2, kOffW3, kOffWH};
  const int cnt[4] = {kNW1, kNW2, kNW3, kNWH};
  const double sd[4] = {std::sqrt(2.0 / (27 * 4)), std::sqrt(2.0 / (27 * 16)), std::sqrt(2.0 / (27 * 32)),
                        0.0};
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
  dfree(L->zbuf); dfree(L->qbuf); dfree(L->gz); dfree(L->ebuf); dfree(L->rbuf); dfree(L->stage);
  for (auto& p : L->l1) dfree(p);
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
  // (on the network's stream, so the next training step sees the cleared moments)
  HIP_TRY(L, hipMemsetAsync(L->m, 0, sizeof(float) * kNW, L->stream));
  HIP_TRY(L, hipMemsetAsync(L->v, 0, sizeof(float) * kNW, L->stream));
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

// the loss of the network's objective for snapshot k of the last assembled system, with
// M the multigrid cycle with the network's correction weights (learned) or the plain
// cycle; store keeps what a backward pass needs (activations, level-1 trace).
//   residual: ||r - A z||^2 / ||r||^2, z = M r, r the snapshot
//   energy:   (e - z)^T A (e - z) / e^T A e, z = M A e, e the snapshot after
//             power_steps steps e <- e - M A e
// Leaves z in L->zbuf, q = r - A z in L->qbuf (r = A e for the energy form) and the
// loss's denominator in *den.
static int lpc_loss_eval(pf_lpc* L, pf_ctx* c, int k, bool learned, bool store, double* loss, double* den) {
  const long long ncell = c->dc.nblocks * kBlockVol;
  const long long n1 = c->mg[0].d.nblocks * kBlockVol;
  const bool energy = L->objective == PF_LPC_OBJECTIVE_ENERGY;
  if (ncell > L->cell_cap) {
    dfree(L->zbuf); dfree(L->qbuf); dfree(L->gz); dfree(L->ebuf); dfree(L->rbuf);
    HIP_TRY(L, dalloc(&L->zbuf, ncell));
    HIP_TRY(L, dalloc(&L->qbuf, ncell));
    HIP_TRY(L, dalloc(&L->gz, ncell));
    HIP_TRY(L, dalloc(&L->ebuf, ncell));
    HIP_TRY(L, dalloc(&L->rbuf, ncell));
    L->cell_cap = ncell;
  }
  if (n1 > L->l1_cap) {
    for (auto& p : L->l1) dfree(p);
    for (auto& p : L->l1) HIP_TRY(L, dalloc(&p, n1));
    L->l1_cap = n1;
  }
  // stale values outside the system must stay finite: clear the work vectors
  HIP_TRY(c, fill_f32(c->stream, L->zbuf, 0.f, ncell));
  HIP_TRY(c, fill_f32(c->stream, L->qbuf, 0.f, ncell));
  const float* r = c->snaps + (long long)k * ncell;
  int st = PF_OK;
  if (learned) {
    st = lpc_cw_weights(c, store);
    if (st) return st;
  }
  c->mg_cs_active = learned ? c->mg_cs_net : nullptr;
  VecArgs va{};
  va.dc = c->dc; va.list = c->pblk_prev; va.nlist = c->npblk_prev; va.sc = c->sc; va.part = c->partial;
  const unsigned G = list_grid(c->npblk_prev);
  const unsigned gdn = std::min<unsigned>(grid_for(ncell, 256), 4096u);
  if (energy) {
    HIP_TRY(c, hipMemcpyAsync(L->ebuf, r, sizeof(float) * (size_t)ncell, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, fill_f32(c->stream, L->rbuf, 0.f, ncell));
    for (int p = 0; p <= L->power_steps && !st; ++p) {
      hipLaunchKernelGGL(k_lpc_apply_a, dim3(G), dim3(256), 0, c->stream, va, L->ebuf, nullptr, L->rbuf, c->cd32,
                         c->cx32, c->cy32, c->cz32, 1.f);
      if (p == L->power_steps) break;
      st = mg_apply_on(c, c->pblk_prev, c->npblk_prev, L->rbuf, L->zbuf, nullptr, c->partial, nullptr);
      hipLaunchKernelGGL(k_lpc_axpy, dim3(gdn), dim3(256), 0, c->stream, ncell, L->ebuf, L->zbuf, -1.f);
    }
    r = L->rbuf;
  }
  MgTrace tr;
  tr.y = L->l1[0];
  tr.h = L->l1[1];
  if (!st) st = mg_apply_on(c, c->pblk_prev, c->npblk_prev, r, L->zbuf, nullptr, c->partial, store ? &tr : nullptr);
  c->mg_cs_active = nullptr;
  if (st) return st;
  hipLaunchKernelGGL(k_lpc_apply_a, dim3(G), dim3(256), 0, c->stream, va, L->zbuf, r, L->qbuf, c->cd32, c->cx32,
                     c->cy32, c->cz32, 1.f);
  if (energy) {
    hipLaunchKernelGGL(k_lpc_dots3, dim3(G), dim3(256), 0, c->stream, va, L->ebuf, L->qbuf, L->zbuf, r);
    hipLaunchKernelGGL(k_reduce_slots, dim3(1), dim3(256), 0, c->stream, c->partial, (int)G, 3, 0u, c->red);
  } else {
    hipLaunchKernelGGL(k_reduce_slots, dim3(1), dim3(256), 0, c->stream, c->partial, (int)G, 2, 0u, c->red);
  }
  LAUNCH_CHECK(c);
  st = read_small(c);
  if (st) return st;
  if (energy) {
    *den = c->red_host[2];
    *loss = *den > 0.0 ? (c->red_host[0] - c->red_host[1]) / *den : 0.0;
  } else {
    *den = c->red_host[1];
    *loss = *den > 0.0 ? c->red_host[0] / *den : 0.0;
  }
  return PF_OK;
}

static int lpc_check_system(pf_lpc* L, pf_ctx* c) {
  if (!L || !c) return PF_ERR_ARG;
  if (c->lpc != L) return fail(c, PF_ERR_STATE, "network is not attached to this context");
  if (!c->tiles_valid || c->nsnap <= 0)
    return fail(c, PF_ERR_STATE, "no recorded system: pf_lpc_record, then solve with the network attached");
  HIP_TRY(c, hipStreamSynchronize(L->stream));
  // the cycle of the recorded system (its operator stays until the next projection)
  int st = mg_build(c);
  if (st) return st;
  if (c->mg.empty()) return fail(c, PF_ERR_STATE, "learned preconditioner: the grid has no coarse level");
  return PF_OK;
}

int pf_lpc_loss(pf_lpc* L, pf_ctx* c, int32_t k, double* loss_learned, double* loss_plain) {
  int st = lpc_check_system(L, c);
  if (st) return st;
  if (k < 0 || k >= c->nsnap) return fail(c, PF_ERR_ARG, "snapshot index out of range");
  double rr = 0.0, a = 0.0, b = 0.0;
  st = lpc_loss_eval(L, c, k, true, false, &a, &rr);
  if (!st) st = lpc_loss_eval(L, c, k, false, false, &b, &rr);
  if (st) return st;
  if (loss_learned) *loss_learned = a;
  if (loss_plain) *loss_plain = b;
  return PF_OK;
}

// L->grad = d(loss)/dw of the network's objective for snapshot k, with the loss (see
// pf_lpc_cw.hip for the chain rule through the cycle)
static int lpc_grad_eval(pf_lpc* L, pf_ctx* c, int k, double* loss_out) {
  const long long ncell = c->dc.nblocks * kBlockVol;
  const long long n1 = c->mg[0].d.nblocks * kBlockVol;
  const unsigned gdn = std::min<unsigned>(grid_for(ncell, 256), 4096u);
  const unsigned gd1 = std::min<unsigned>(grid_for(n1, 256), 2048u);
  const std::vector<float> wpre = mg_sweep_weights(c);
  VecArgs va{};
  va.dc = c->dc; va.list = c->pblk_prev; va.nlist = c->npblk_prev; va.sc = c->sc; va.part = c->partial;
  const unsigned G = list_grid(c->npblk_prev);
  double loss = 0.0, rr = 0.0;
  int st = lpc_loss_eval(L, c, k, true, true, &loss, &rr);
  if (st) return st;
  if (loss_out) *loss_out = loss;
  HIP_TRY(c, fill_f32(c->stream, L->grad, 0.f, kNW));
  if (!(rr > 0.0) || c->ntiles == 0) return PF_OK;
  // dL/dz = -2 A q / ||r||^2 (residual) or -2 q / e^T A e (energy), then through the
  // post-smoothing (transposed sweeps I - theta A D^-1, first sweep's weight last)
  HIP_TRY(c, fill_f32(c->stream, L->gz, 0.f, ncell));
  if (L->objective == PF_LPC_OBJECTIVE_ENERGY)
    hipLaunchKernelGGL(k_lpc_axpy, dim3(gdn), dim3(256), 0, c->stream, ncell, L->gz, L->qbuf, (float)(-2.0 / rr));
  else
    hipLaunchKernelGGL(k_lpc_apply_a, dim3(G), dim3(256), 0, c->stream, va, L->qbuf, nullptr, L->gz, c->cd32,
                       c->cx32, c->cy32, c->cz32, (float)(-2.0 / rr));
  for (size_t j = 0; j < wpre.size(); ++j) {
    hipLaunchKernelGGL(k_lpc_dscale, dim3(gdn), dim3(256), 0, c->stream, ncell, L->gz, c->invd32, L->zbuf);
    hipLaunchKernelGGL(k_lpc_apply_a, dim3(G), dim3(256), 0, c->stream, va, L->zbuf, nullptr, L->qbuf, c->cd32,
                       c->cx32, c->cy32, c->cz32, 1.f);
    hipLaunchKernelGGL(k_lpc_axpy, dim3(gdn), dim3(256), 0, c->stream, ncell, L->gz, L->qbuf, -wpre[j]);
  }
  // dL/dg = w_c P^T (...) on level 1
  int nx1, ny1, nz1;
  lpc_level1_extents(c, nx1, ny1, nz1);
  HIP_TRY(c, fill_f32(c->stream, L->l1[2], 0.f, n1));
  hipLaunchKernelGGL(k_lpc_restrict_sum, dim3(gd1), dim3(256), 0, c->stream, c->dc, c->geom, L->gz, c->invd32,
                     c->mg[0].d, nx1, ny1, nz1, c->mg_cfg.correction, L->l1[2]);
  // B (S dL/dg) by the cycle's coarse levels
  hipLaunchKernelGGL(k_lpc_mul, dim3(gd1), dim3(256), 0, c->stream, n1, c->mg_cs_net, L->l1[2], c->mg[0].r);
  float* bw = mg_coarse(c, 1, nullptr, wpre);
  // dL/ds = dL/dg (.) h + y (.) B S dL/dg, then dL/do in tile layout
  hipLaunchKernelGGL(k_lpc_cw_grad, dim3(gd1), dim3(256), 0, c->stream, n1, L->l1[2], L->l1[1], L->l1[0], bw,
                     L->l1[4]);
  CwArgs a = lpc_cw_args(c);
  hipLaunchKernelGGL(k_lpc_cw_gout, dim3(lpc_tile_grid(c->ntiles)), dim3(256), 0, c->stream, a, L->o, L->l1[4],
                     c->mg_cs_net, L->go);
  LAUNCH_CHECK(c);
  st = lpc_backward(L, c->stream, c->ntiles);
  if (st) return fail(c, st, "learned preconditioner: " + L->err);
  return PF_OK;
}

int pf_lpc_grad(pf_lpc* L, pf_ctx* c, int32_t k, float* grad, double* loss) {
  int st = lpc_check_system(L, c);
  if (st) return st;
  if (k < 0 || k >= c->nsnap) return fail(c, PF_ERR_ARG, "snapshot index out of range");
  st = lpc_grad_eval(L, c, k, loss);
  if (st) return st;
  if (grad) HIP_TRY(c, hipMemcpyAsync(grad, L->grad, sizeof(float) * kNW, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return PF_OK;
}

// one Adam step per sampled snapshot on the gradient of the network's objective
int pf_lpc_set_objective(pf_lpc* L, int32_t objective, int32_t power_steps) {
  if (!L) return PF_ERR_ARG;
  if ((objective != PF_LPC_OBJECTIVE_RESIDUAL && objective != PF_LPC_OBJECTIVE_ENERGY) || power_steps < 0 ||
      power_steps > 64) {
    L->err = "objective must be PF_LPC_OBJECTIVE_RESIDUAL or PF_LPC_OBJECTIVE_ENERGY with 0..64 power steps";
    return PF_ERR_ARG;
  }
  L->objective = objective;
  L->power_steps = power_steps;
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
    double loss = 0.0;
    st = lpc_grad_eval(L, c, k, &loss);
    if (st) return st;
    lsum += loss;
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
