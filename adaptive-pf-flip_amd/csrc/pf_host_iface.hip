// pf_host_iface.hip -- interface diagnostics and CSF surface tension, sparse and multires.
//
// One of the library's host files: pf_api.hip includes them in order into a single
// translation unit together with the kernel files, so every launch is a direct call.

extern "C" {

// ---------------------------------------------------------------- interface diagnostics, CSF
namespace {

constexpr int kIfaceFields = 9;

// the cell phase field, normals and curvature from the current face masses and
// markers (interface_diagnostics3d.cpp:137-208); mr: sample the multires leaf cells
int iface_fields(pf_ctx* c, const pf_phase* ph, IfaceArgs& a, bool mr) {
  const long long n = c->dc.nblocks * kBlockVol;
  if (!c->iface) {
    HIP_TRY(c, dalloc(&c->iface, n * kIfaceFields));
    HIP_TRY(c, dalloc(&c->iface_bad, 1));
  }
  a.dc = c->dc; a.du = c->du; a.dv = c->dv; a.dw = c->dw;
  a.g = c->geom;
  a.mu = c->f[F_MU]; a.mv = c->f[F_MV]; a.mw = c->f[F_MW];
  a.marker = c->marker;
  a.ph = phase_k(ph);
  a.phi = c->iface;
  a.gmag = c->iface + n;
  for (int d = 0; d < 3; ++d) a.nrm[d] = c->iface + (2 + d) * n;
  a.kap = c->iface + 5 * n;
  a.lvl = mr ? c->mr_lvl : nullptr;
  const dim3 g(grid_for(n, 256));
  const long long ki = kbegin(c, PF_K_INTERFACE);
  hipLaunchKernelGGL(k_iface_phi, g, dim3(256), 0, c->stream, a);
  hipLaunchKernelGGL(k_iface_normal, g, dim3(256), 0, c->stream, a);
  hipLaunchKernelGGL(k_iface_curvature, g, dim3(256), 0, c->stream, a);
  kend(c, ki);
  LAUNCH_CHECK(c);
  return PF_OK;
}

unsigned iface_groups(const pf_ctx* c) {
  return std::min<unsigned>(grid_for(c->dc.nblocks * kBlockVol, 256), 1024u);
}

// accumulate and finishStats (interface_diagnostics3d.cpp:26-66) over the FLUID cells
int iface_reduce(pf_ctx* c, const IfaceArgs& a, pf_iface_stats* out) {
  const unsigned nwg = iface_groups(c);
  const CombineOps ops{kIfaceSlots, {0, 1, 2, 0, 0, 0, 2, 0, 2, 0}};
  const long long ki = kbegin(c, PF_K_INTERFACE);
  hipLaunchKernelGGL(k_iface_diag, dim3(nwg), dim3(256), 0, c->stream, a, c->partial);
  hipLaunchKernelGGL(k_iface_combine, dim3(1), dim3(kCombineThreads), 0, c->stream, c->partial, (int)nwg, ops, c->red);
  kend(c, ki);
  LAUNCH_CHECK(c);
  HIP_TRY(c, hipMemcpyAsync(c->red_host, c->red, sizeof(double) * kIfaceSlots, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  const double* v = c->red_host;
  pf_iface_stats S;
  std::memset(&S, 0, sizeof(S));
  S.finite = v[9] == 0.0 ? 1 : 0;
  S.sample_cells = (int32_t)v[0];
  S.interface_cells = (int32_t)v[4];
  if (S.sample_cells > 0) {
    S.phi_min = v[1];
    S.phi_max = v[2];
    S.phi_mean = v[3] / v[0];
  }
  S.grad_max = v[6];
  S.curvature_abs_max = v[8];
  if (S.interface_cells > 0) {
    S.grad_mean = v[5] / v[4];
    S.curvature_abs_mean = v[7] / v[4];
  }
  S.surface_tension_candidate = S.finite && S.interface_cells > 0 && S.grad_max > kIfaceGradEps;
  *out = S;
  return PF_OK;
}

// initSurfaceStats (:87-102), the CSF increments (:404-440), finishSurfaceStats (:124-131)
int surface_apply(pf_ctx* c, const pf_phase* ph, const IfaceArgs& a, double dt, double strength, double maxd,
                  int radius, pf_surface_stats* out) {
  pf_surface_stats S;
  std::memset(&S, 0, sizeof(S));
  S.finite = 1;
  S.enabled = strength > 0.0 && dt > 0.0;
  S.strength = strength;
  S.max_delta_speed_limit = maxd;
  S.curvature_smoothing_radius = std::min(std::max(radius, 0), 3);
  const double rmin = std::min(ph->rho_l, ph->rho_g), dx = c->geom.dx;
  const double pi = 3.14159265358979323846;
  S.capillary_dt_limit = (dx <= 0.0 || strength <= 0.0 || rmin <= 0.0 || !std::isfinite(rmin))
                             ? 0.0 : std::sqrt((rmin * dx * dx * dx) / (2.0 * pi * strength));
  S.capillary_stable = !S.enabled || S.capillary_dt_limit <= 0.0 || dt <= S.capillary_dt_limit + 1e-12;
  if (S.enabled) {
    const long long n = c->dc.nblocks * kBlockVol;
    CsfArgs cs;
    cs.f = a;
    cs.dt = dt;
    cs.strength = strength;
    cs.maxd = maxd;
    cs.radius = S.curvature_smoothing_radius;
    for (int d = 0; d < 3; ++d) cs.delta[d] = c->iface + (6 + d) * n;
    cs.first_bad = c->iface_bad;
    const unsigned nwg = iface_groups(c);
    const CombineOps ops{kCsfSlots, {0, 0, 2, 0, 2, 0, 2, 0}};
    HIP_TRY(c, fill_u32(c->stream, reinterpret_cast<uint32_t*>(c->iface_bad), 0xffffffffu, 2));
    const long long ki = kbegin(c, PF_K_SURFACE);
    hipLaunchKernelGGL(k_csf_delta, dim3(nwg), dim3(256), 0, c->stream, cs, c->partial);
    hipLaunchKernelGGL(k_iface_combine, dim3(1), dim3(kCombineThreads), 0, c->stream, c->partial, (int)nwg, ops, c->red);
    const Dims* fd[3] = {&c->du, &c->dv, &c->dw};
    for (int ax = 0; ax < 3; ++ax)
      hipLaunchKernelGGL(k_csf_faces, dim3(grid_for(fd[ax]->nblocks * kBlockVol, 256)), dim3(256), 0, c->stream,
                         cs, ax, c->f[F_U + ax], static_cast<const float*>(c->f[F_MU + ax]));
    kend(c, ki);
    LAUNCH_CHECK(c);
    HIP_TRY(c, hipMemcpyAsync(c->red_host, c->red, sizeof(double) * kCsfSlots, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    const double* v = c->red_host;
    S.applied_cells = (int32_t)v[0];
    S.finite = v[7] == 0.0 ? 1 : 0;
    S.max_delta_speed = v[2];
    S.raw_curvature_abs_max = v[4];
    S.smoothed_curvature_abs_max = v[6];
    S.mean_delta_speed = v[1];
    S.raw_curvature_abs_mean = v[3];
    S.smoothed_curvature_abs_mean = v[5];
    if (S.applied_cells > 0) {
      S.mean_delta_speed /= v[0];
      S.raw_curvature_abs_mean /= v[0];
      S.smoothed_curvature_abs_mean /= v[0];
    }
  }
  if (out) *out = S;
  return PF_OK;
}

// normalisation, the saved copy and gravity, then interface_diagnostics_last,
// surface tension and the walls (sparse_sim3d_tp.cpp:865-874,
// multires_sim3d_tp.cpp:683-694).  The face operations leave the masses the
// diagnostics read unchanged, so they fuse into one pass unless the CSF increments
// have to land between gravity and the walls.
int step_face_stage(pf_ctx* c, const pf_phase* ph, double dt, double gravity, bool mr, bool diag, bool csf,
                    double strength, double maxd, int radius, pf_step_stats* S) {
  int st = run_face_ops(c, FO_NORMALIZE | FO_SAVE | FO_GRAVITY | (csf ? 0 : FO_WALLS), (float)(dt * gravity));
  if (st || !(diag || csf)) return st;
  IfaceArgs ia;
  st = iface_fields(c, ph, ia, mr);
  if (st) return st;
  if (diag) {
    st = iface_reduce(c, ia, &S->iface);
    if (st) return st;
  }
  if (!csf) return PF_OK;
  st = surface_apply(c, ph, ia, dt, strength, maxd, radius, &S->surface);
  if (st) return st;
  return run_face_ops(c, FO_WALLS, 0.f);
}

// the defaults of InterfaceDiagnostics3D and SurfaceTensionStats3D
// (interface_diagnostics3d.h:7-36), kept where a part is skipped
void default_interface_stats(pf_step_stats* S) {
  S->iface.finite = 1;
  S->surface.finite = 1;
  S->surface.capillary_stable = 1;
}

int interface_entry(pf_ctx* c, const pf_phase* ph, bool mr, IfaceArgs& a) {
  if (mr && !c->mr) return fail(c, PF_ERR_STATE, "no multires layout (pf_mr_set_layout)");
  if (!mr && c->mr) return fail(c, PF_ERR_STATE, "a multires context: use the pf_mr_ interface entry points");
  return iface_fields(c, ph, a, mr);
}

}  // namespace

int pf_interface_diagnostics(pf_ctx* c, const pf_phase* ph, pf_iface_stats* out) {
  if (!c || !ph || !out) return PF_ERR_ARG;
  IfaceArgs a;
  int st = interface_entry(c, ph, false, a);
  if (st) return st;
  return iface_reduce(c, a, out);
}

int pf_apply_surface_tension(pf_ctx* c, const pf_phase* ph, double dt, double strength, double max_delta_speed,
                             int32_t smoothing_radius, pf_surface_stats* out) {
  if (!c || !ph) return PF_ERR_ARG;
  IfaceArgs a;
  int st = interface_entry(c, ph, false, a);
  if (st) return st;
  return surface_apply(c, ph, a, dt, strength, max_delta_speed, smoothing_radius, out);
}

int pf_mr_interface_diagnostics(pf_ctx* c, const pf_phase* ph, pf_iface_stats* out) {
  if (!c || !ph || !out) return PF_ERR_ARG;
  IfaceArgs a;
  int st = interface_entry(c, ph, true, a);
  if (st) return st;
  return iface_reduce(c, a, out);
}

int pf_mr_apply_surface_tension(pf_ctx* c, const pf_phase* ph, double dt, double strength, double max_delta_speed,
                                int32_t smoothing_radius, pf_surface_stats* out) {
  if (!c || !ph) return PF_ERR_ARG;
  IfaceArgs a;
  int st = interface_entry(c, ph, true, a);
  if (st) return st;
  return surface_apply(c, ph, a, dt, strength, max_delta_speed, smoothing_radius, out);
}

}  // extern "C"
