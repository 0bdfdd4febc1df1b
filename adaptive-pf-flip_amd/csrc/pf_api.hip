// pf_api.hip -- the C ABI of include/pf_gpu.h: the translation unit, the context,
// particle transfers, the stage entry points, the step and field access.
// Built as one translation unit with the kernel files so every launch is a direct
// call (no relocatable device code).  All work is stream-ordered on the context's
// stream; the only host synchronisations are the ones a result needs: the
// fluid-block count before the solve, the refinement loop's residual check, the
// convergence flag every `check_every` inner iterations, and downloads.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "pf_internal.h"
#include "pf_util.hip"
#include "pf_bin.hip"
#include "pf_grid.hip"
#include "pf_p2g.hip"
#include "pf_project.hip"
#include "pf_g2p.hip"
#include "pf_lpc.hip"
#include "pf_mr.hip"
#include "pf_mr_ref.hip"
#include "pf_iface.hip"
#include "pf_adapt.hip"
#include "pf_render.hip"
#include "pf_mg.hip"
#include "pf_mr_mg.hip"
#include "pf_lpc_cw.hip"

using namespace pf;

#include "pf_host.h"
#include "pf_host_particles.hip"
#include "pf_host_mr.hip"
#include "pf_host_lpc.hip"
#include "pf_host_solve.hip"
#include "pf_host_adapt.hip"
#include "pf_host_iface.hip"
#include "pf_host_render.hip"
#include "pf_host_lpc_api.hip"
#include "pf_host_mr_api.hip"

extern "C" {

int pf_abi_version(void) { return PF_ABI_VERSION; }

const char* pf_last_error(const pf_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

void* pf_stream(pf_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int pf_kernel_timing_mask(pf_ctx* c, uint32_t mask) {
  if (!c) return PF_ERR_ARG;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  c->ktime = mask & ((1u << PF_NKERNELS) - 1u);
  c->krec.clear();
  c->kev_used = 0;
  return PF_OK;
}

int pf_kernel_timing(pf_ctx* c, int32_t enable) {
  return pf_kernel_timing_mask(c, enable ? 0xffffffffu : 0u);
}

int pf_kernel_time(pf_ctx* c, int32_t kernel, int64_t* launches, double* total_ms) {
  if (!c || !launches || !total_ms || kernel < 0 || kernel >= PF_NKERNELS) return PF_ERR_ARG;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  int64_t n = 0;
  double t = 0.0;
  for (const auto& r : c->krec) {
    if (r.first != kernel) continue;
    float ms = 0.f;
    HIP_TRY(c, hipEventElapsedTime(&ms, c->kev[r.second], c->kev[r.second + 1]));
    t += ms;
    ++n;
  }
  *launches = n;
  *total_ms = t;
  return PF_OK;
}

int pf_copy_probe(pf_ctx* c, int64_t n, double* ms) {
  if (!c || n <= 0 || (n & 3) != 0 || n > (1LL << 31)) return PF_ERR_ARG;
  // its own buffers (allocated here: a measurement hook, not a step stage)
  float *src = nullptr, *dst = nullptr;
  HIP_TRY(c, dalloc(&src, n));
  HIP_TRY(c, dalloc(&dst, n));
  HIP_TRY(c, hipDeviceSynchronize());
  hipEvent_t a, b;
  HIP_TRY(c, hipEventCreate(&a));
  HIP_TRY(c, hipEventCreate(&b));
  const long long n4 = n / 4;
  long long blocks = std::min<long long>((n4 + 255) / 256, 1 << 16);
  float t = 1e30f;
  for (int rep = 0; rep < 4; ++rep) {  // first pass warms the TLB; keep the best
    HIP_TRY(c, hipEventRecord(a, c->stream));
    hipLaunchKernelGGL(k_copy_f32x4, dim3((unsigned)blocks), dim3(256), 0, c->stream,
                       reinterpret_cast<const float4*>(src), reinterpret_cast<float4*>(dst), n4);
    LAUNCH_CHECK(c);
    HIP_TRY(c, hipEventRecord(b, c->stream));
    HIP_TRY(c, hipEventSynchronize(b));
    float tr = 0.f;
    HIP_TRY(c, hipEventElapsedTime(&tr, a, b));
    t = std::min(t, tr);
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  HIP_TRY(c, hipFree(src));
  HIP_TRY(c, hipFree(dst));
  if (ms) *ms = t;
  return PF_OK;
}

int pf_synchronize(pf_ctx* c) {
  if (!c) return PF_ERR_ARG;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return PF_OK;
}

int pf_create(pf_ctx** out, const pf_grid_desc* d) {
  if (!out || !d) return PF_ERR_ARG;
  *out = nullptr;
  if (d->nx < 3 || d->ny < 3 || d->nz < 3 || !(d->dx > 0.0)) return PF_ERR_ARG;
  long long ncell_est = (long long)d->nx * d->ny * d->nz;
  if (ncell_est > (1LL << 31)) return PF_ERR_ARG;  // cell keys and ranks are 32-bit
  pf_ctx* c = new pf_ctx();
  c->mg_cfg = mg_default_cfg();
  HIP_TRY(c, hipGetDevice(&c->device));
  HIP_TRY(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  c->geom = Geom{d->nx, d->ny, d->nz, d->dx, d->ox, d->oy, d->oz};
  c->du = make_dims(d->nx + 1, d->ny, d->nz);
  c->dv = make_dims(d->nx, d->ny + 1, d->nz);
  c->dw = make_dims(d->nx, d->ny, d->nz + 1);
  c->dc = make_dims(d->nx, d->ny, d->nz);
  const Dims* fd[F_N] = {&c->du, &c->dv, &c->dw, &c->du, &c->dv, &c->dw, &c->dc, &c->dc, &c->du, &c->dv, &c->dw};
  for (int f = 0; f < F_N; ++f) {
    if (f != F_MK) HIP_TRY(c, dalloc(&c->f[f], fd[f]->nblocks * kBlockVol));
    HIP_TRY(c, dalloc(&c->act[f], fd[f]->nblocks));
  }
  const long long ncell = c->dc.nblocks * kBlockVol;
  HIP_TRY(c, dalloc(&c->marker, ncell));
  HIP_TRY(c, dalloc(&c->cell_count, ncell));
  HIP_TRY(c, dalloc(&c->cell_start, ncell + 1));
  c->color_cap = (long long)((c->dc.nbx + 1) / 2) * ((c->dc.nby + 1) / 2) * ((c->dc.nbz + 1) / 2);
  HIP_TRY(c, dalloc(&c->color_list, 8 * c->color_cap));
  HIP_TRY(c, dalloc(&c->color_cnt, 8));
  long long maxblk = std::max(std::max(c->du.nblocks, c->dv.nblocks), std::max(c->dw.nblocks, c->dc.nblocks));
  HIP_TRY(c, dalloc(&c->pblk, maxblk));
  HIP_TRY(c, dalloc(&c->pblk_prev, maxblk));
  HIP_TRY(c, dalloc(&c->blk_flag32, maxblk));
  c->morton_cap = morton_codes(c->dc) + 1;
  HIP_TRY(c, dalloc(&c->blk_scan, std::max(maxblk + 1, c->morton_cap)));
  HIP_TRY(c, dalloc(&c->morton_flags, c->morton_cap));
  float** f32[] = {&c->x32, &c->r32, &c->z32, &c->pA, &c->pB, &c->Ap, &c->cd32, &c->cx32, &c->cy32, &c->cz32, &c->invd32};
  for (float** p : f32) HIP_TRY(c, dalloc(p, ncell));
  double** f64[] = {&c->b64, &c->x64, &c->r64, &c->cd64, &c->cx64, &c->cy64, &c->cz64};
  for (double** p : f64) HIP_TRY(c, dalloc(p, ncell));
  HIP_TRY(c, dalloc(&c->sc, 1));
  HIP_TRY(c, dalloc(&c->partial, (long long)kMaxPartials * 4));
  HIP_TRY(c, dalloc(&c->red, 16));
  HIP_TRY(c, hipHostMalloc((void**)&c->red_host, sizeof(double) * 16));
  HIP_TRY(c, hipHostMalloc((void**)&c->done_host, sizeof(int) * 16));
  HIP_TRY(c, hipHostMalloc((void**)&c->sc_host, sizeof(SolveScalars)));
  HIP_TRY(c, dalloc(&c->u32_dev, 64));
  HIP_TRY(c, hipHostMalloc((void**)&c->u32_host, sizeof(uint32_t) * 64));
  HIP_TRY(c, dalloc(&c->esc_dev, 8));
  HIP_TRY(c, dalloc(&c->vmax_dev, 1));
  c->scan_tile_cap = (std::max(ncell, maxblk) + 4095) / 4096 + 1;
  HIP_TRY(c, dalloc(&c->scan_tiles, c->scan_tile_cap));
  long long maxnodes = maxblk * kBlockVol;
  HIP_TRY(c, dalloc(&c->stage_f, maxnodes));
  c->stage_cap = maxnodes;
  for (int e = 0; e <= PF_NSTAGES; ++e) HIP_TRY(c, hipEventCreate(&c->ev[e]));
  // cell starts must be valid before the first binning (empty particle set)
  c->binned = false;
  *out = c;
  return PF_OK;
}

int pf_destroy(pf_ctx* c) {
  if (!c) return PF_OK;
  hipStreamSynchronize(c->stream);
  for (int f = 0; f < F_N; ++f) { hipFree(c->f[f]); hipFree(c->act[f]); }
  hipFree(c->marker); hipFree(c->cell_count); hipFree(c->cell_start);
  hipFree(c->color_list); hipFree(c->color_cnt); dfree(c->crowd_list); hipFree(c->pblk); hipFree(c->pblk_prev);
  hipFree(c->blk_flag32); hipFree(c->blk_scan); hipFree(c->morton_flags);
  PBuf* bufs[2] = {&c->pa, &c->pb};
  for (PBuf* b : bufs) {
    hipFree(b->x); hipFree(b->y); hipFree(b->z); hipFree(b->u); hipFree(b->v); hipFree(b->w);
    hipFree(b->type); hipFree(b->id); hipFree(b->vol);
  }
  hipFree(c->rank);
  float* f32[] = {c->x32, c->r32, c->z32, c->pA, c->pB, c->Ap, c->cd32, c->cx32, c->cy32, c->cz32, c->invd32};
  for (float* p : f32) hipFree(p);
  double* f64[] = {c->b64, c->x64, c->r64, c->cd64, c->cx64, c->cy64, c->cz64};
  for (double* p : f64) hipFree(p);
  dfree(c->mr_lvl); dfree(c->mr_vol64); dfree(c->mr_pfull); dfree(c->mr_bbox); dfree(c->mr_leaves_dev);
  dfree(c->mrf); dfree(c->mrc); dfree(c->mr_pin_dev);
  dfree(c->mr_f01); dfree(c->mr_f2); dfree(c->mr_list01); dfree(c->mr_list2);
  dfree(c->iface); dfree(c->iface_bad);
  dfree(c->pa_alive); dfree(c->pa_key); dfree(c->pa_by_id); dfree(c->pa_new_id);
  dfree(c->pa_cflag); dfree(c->pa_d1); dfree(c->pa_d2); dfree(c->pa_wanted); dfree(c->pa_wscan); dfree(c->pa_cnt);
  for (auto& m : c->mg) {
    dfree(m.cd); dfree(m.cx); dfree(m.cy); dfree(m.cz); dfree(m.inv); dfree(m.x); dfree(m.r); dfree(m.t);
  }
  c->mg.clear();
  dfree(c->mg_t0);
  dfree(c->mg_cs);
  dfree(c->mg_cs_net);
  for (auto& h : c->mg_h) dfree(h);
  dfree(c->mg_mx);
  dfree(c->mg_unscale);
  if (c->mr_pin_host) hipHostFree(c->mr_pin_host);
  for (int a = 0; a < 3; ++a) { dfree(c->mr_valid[a]); dfree(c->mr_cf32[a]); dfree(c->mr_cf64[a]); }
  if (c->mr_bbox_host) hipHostFree(c->mr_bbox_host);
  dfree(c->fbeta); dfree(c->fphi); dfree(c->tiles); dfree(c->tile_flags); dfree(c->snaps);
  hipFree(c->sc); hipFree(c->partial); hipFree(c->red);
  hipHostFree(c->red_host); hipHostFree(c->done_host); hipHostFree(c->sc_host); hipHostFree(c->u32_host);
  hipFree(c->u32_dev); hipFree(c->esc_dev); hipFree(c->vmax_dev);
  for (auto* q : c->g2p_quad) hipFree(q);
  hipFree(c->scan_tiles); hipFree(c->stage_f);
  for (int e = 0; e <= PF_NSTAGES; ++e)
    if (c->ev[e]) hipEventDestroy(c->ev[e]);
  for (hipEvent_t e : c->kev) hipEventDestroy(e);
  hipStreamDestroy(c->stream);
  delete c;
  return PF_OK;
}

int pf_particles_upload(pf_ctx* c, int64_t n, const float* x, const float* y, const float* z, const float* u,
                        const float* v, const float* w, const uint8_t* type, const float* volume) {
  if (!c || n < 0) return PF_ERR_ARG;
  if (n > 0 && (!x || !y || !z || !u || !v || !w || !type)) return fail(c, PF_ERR_ARG, "null particle array");
  if (n >= (1LL << 32)) return fail(c, PF_ERR_ARG, "particle count exceeds 32-bit ids");
  int st = grow_particles(c, n, volume != nullptr);
  if (st) return st;
  c->np = n;
  c->has_vol = volume != nullptr;
  c->vmax_valid = false;
  if (n > 0) {
    size_t fb = sizeof(float) * (size_t)n;
    HIP_TRY(c, hipMemcpyAsync(c->pa.x, x, fb, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->pa.y, y, fb, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->pa.z, z, fb, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->pa.u, u, fb, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->pa.v, v, fb, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->pa.w, w, fb, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->pa.type, type, (size_t)n, hipMemcpyHostToDevice, c->stream));
    if (volume) HIP_TRY(c, hipMemcpyAsync(c->pa.vol, volume, fb, hipMemcpyHostToDevice, c->stream));
    std::vector<uint32_t> ids((size_t)n);
    for (int64_t p = 0; p < n; ++p) ids[(size_t)p] = (uint32_t)p;
    HIP_TRY(c, hipMemcpyAsync(c->pa.id, ids.data(), sizeof(uint32_t) * (size_t)n, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, fill_u32(c->stream, c->u32_dev, 0u, 1));
    hipLaunchKernelGGL(k_validate, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, (long long)n, c->pa, c->geom,
                       c->u32_dev);
    LAUNCH_CHECK(c);
    HIP_TRY(c, hipMemcpyAsync(c->u32_host, c->u32_dev, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (c->u32_host[0]) {
      c->np = 0;
      return fail(c, PF_ERR_ARG, "particle position outside the grid's cell range");
    }
  }
  double lv = 0.0;
  for (int64_t p = 0; p < n; ++p)
    if (type[p] == 0) lv += volume ? (double)volume[p] : 1.0;
  c->liquid_volume = lv;
  c->binned = false;
  return PF_OK;
}

int pf_particles_download(pf_ctx* c, float* x, float* y, float* z, float* u, float* v, float* w) {
  if (!c || !x || !y || !z || !u || !v || !w) return PF_ERR_ARG;
  const long long n = c->np;
  if (n == 0) return PF_OK;
  PBuf t = c->pb;
  hipLaunchKernelGGL(k_unsort, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, n, c->pa, t.x, t.y, t.z, t.u, t.v,
                     t.w);
  LAUNCH_CHECK(c);
  size_t fb = sizeof(float) * (size_t)n;
  HIP_TRY(c, hipMemcpyAsync(x, t.x, fb, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipMemcpyAsync(y, t.y, fb, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipMemcpyAsync(z, t.z, fb, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipMemcpyAsync(u, t.u, fb, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipMemcpyAsync(v, t.v, fb, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipMemcpyAsync(w, t.w, fb, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return PF_OK;
}

int64_t pf_particle_count(const pf_ctx* c) { return c ? c->np : -1; }

int pf_particles_download_ext(pf_ctx* c, uint8_t* type, float* volume) {
  if (!c) return PF_ERR_ARG;
  const long long n = c->np;
  if (n == 0 || (!type && !volume)) return PF_OK;
  PBuf t = c->pb;
  hipLaunchKernelGGL(k_unsort_ext, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, n, c->pa, c->has_vol, t.type,
                     t.x);
  LAUNCH_CHECK(c);
  if (type) HIP_TRY(c, hipMemcpyAsync(type, t.type, (size_t)n, hipMemcpyDeviceToHost, c->stream));
  if (volume) HIP_TRY(c, hipMemcpyAsync(volume, t.x, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return PF_OK;
}

double pf_liquid_volume(const pf_ctx* c) { return c ? c->liquid_volume : 0.0; }

int pf_scene_generate(pf_ctx* c, int32_t kind, uint64_t seed) {
  if (!c || kind < 0 || kind > 2) return PF_ERR_ARG;
  const Geom& g = c->geom;
  // filled cells form the box [1, ie) x [1, je) x [1, nz-1): cell-centre tests
  // y < 0.6 (bubble, dam-break) and x < 0.4 (dam-break)
  int ie = g.nx - 1, je = g.ny - 1;
  if (kind == 1 || kind == 2) {
    je = 1;
    while (je < g.ny - 1 && (je + 0.5) / g.ny < 0.6) ++je;
  }
  if (kind == 2) {
    ie = 1;
    while (ie < g.nx - 1 && (ie + 0.5) / g.nx < 0.4) ++ie;
  }
  long long ncells = (long long)(ie - 1) * (je - 1) * (g.nz - 2);
  long long n = 8 * ncells;
  int st = grow_particles(c, n, false);
  if (st) return st;
  c->np = n;
  c->has_vol = false;
  c->vmax_valid = false;
  if (n > 0) {
    hipLaunchKernelGGL(k_scene, dim3(grid_for(ncells, 256)), dim3(256), 0, c->stream, (int)kind, g, ie, je, seed, c->pa);
    LAUNCH_CHECK(c);
  }
  // liquid volume (type 0 count): read back the types once
  std::vector<uint8_t> t((size_t)n);
  if (n > 0) {
    HIP_TRY(c, hipMemcpyAsync(t.data(), c->pa.type, (size_t)n, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
  }
  long long liq = 0;
  for (uint8_t q : t) liq += q == 0;
  c->liquid_volume = (double)liq;
  c->binned = false;
  return PF_OK;
}

int pf_mark_cells(pf_ctx* c) {
  if (!c) return PF_ERR_ARG;
  return run_mark(c);
}

int pf_p2g(pf_ctx* c, const pf_phase* ph, double Vp) {
  if (!c || !ph) return PF_ERR_ARG;
  int st = run_p2g_splat(c, ph, Vp);
  if (st) return st;
  return run_face_ops(c, FO_NORMALIZE, 0.f);
}

int pf_p2g_single(pf_ctx* c) {
  if (!c) return PF_ERR_ARG;
  int st = run_p2g_splat(c, nullptr, 1.0, true);
  if (st) return st;
  return run_face_ops(c, FO_NORMALIZE, 0.f);
}

int pf_save_velocity(pf_ctx* c) {
  if (!c) return PF_ERR_ARG;
  return run_face_ops(c, FO_SAVE, 0.f);
}

int pf_apply_gravity(pf_ctx* c, double dt, double gravity) {
  if (!c) return PF_ERR_ARG;
  return run_face_ops(c, FO_GRAVITY, (float)(dt * gravity));
}

int pf_apply_walls(pf_ctx* c) {
  if (!c) return PF_ERR_ARG;
  return run_face_ops(c, FO_WALLS, 0.f);
}

int pf_g2p(pf_ctx* c, double aL, double aG) {
  if (!c) return PF_ERR_ARG;
  return run_g2p_advect(c, true, false, aL, aG, 0.0, 2);
}

int pf_advect(pf_ctx* c, double dt, int32_t order, pf_escape_stats* stats) {
  if (!c) return PF_ERR_ARG;
  HIP_TRY(c, fill_u32(c->stream, reinterpret_cast<uint32_t*>(c->esc_dev), 0u, 8));
  int st = run_g2p_advect(c, false, true, 0.0, 0.0, dt, order);
  if (st) return st;
  if (stats) {
    HIP_TRY(c, hipMemcpyAsync(stats, c->esc_dev, sizeof(int32_t) * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
  }
  return PF_OK;
}

int pf_timestep(pf_ctx* c, double requested_dt, int32_t enabled, double cfl, double min_dt, double* eff,
                double* vmax_out, double* cfl_dt, int32_t* limited) {
  if (!c) return PF_ERR_ARG;
  // the scan is skipped while the last G2P's own maximum (or the last scan) still
  // describes the particles' velocities
  if (!c->vmax_valid) {
    HIP_TRY(c, fill_u32(c->stream, reinterpret_cast<uint32_t*>(c->vmax_dev), 0u, 1));
    if (c->np > 0) {
      hipLaunchKernelGGL(k_max_speed, dim3(std::min(grid_for(c->np, 256), 8192u)), dim3(256), 0, c->stream, c->np,
                         c->pa, reinterpret_cast<uint32_t*>(c->vmax_dev));
      LAUNCH_CHECK(c);
    }
    c->vmax_valid = true;
  }
  float vm = 0.f;
  HIP_TRY(c, hipMemcpyAsync(&vm, c->vmax_dev, sizeof(float), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  // computeAdaptiveParticleTimestep3D, timestep3d.h:22-49
  double v = vm, e = requested_dt, cd = requested_dt;
  int lim = 0;
  if (enabled && requested_dt > 0.0 && c->geom.dx > 0.0 && cfl > 0.0 && v > 0.0) {
    cd = cfl * c->geom.dx / v;
    double l = std::min(requested_dt, cd);
    if (min_dt > 0.0) l = std::max(min_dt, l);
    e = std::min(requested_dt, l);
    lim = e < requested_dt ? 1 : 0;
  }
  if (eff) *eff = e;
  if (vmax_out) *vmax_out = v;
  if (cfl_dt) *cfl_dt = cd;
  if (limited) *limited = lim;
  return PF_OK;
}

int pf_step(pf_ctx* c, const pf_step_cfg* cfg, pf_step_stats* stats) {
  if (!c || !cfg) return PF_ERR_ARG;
  if (cfg->model != PF_MODEL_TWO_PHASE && cfg->model != PF_MODEL_SINGLE_PHASE)
    return fail(c, PF_ERR_ARG, "pf_step_cfg.model must be PF_MODEL_TWO_PHASE or PF_MODEL_SINGLE_PHASE");
  const bool single = cfg->model == PF_MODEL_SINGLE_PHASE;
  pf_step_stats S;
  std::memset(&S, 0, sizeof(S));
  default_interface_stats(&S);
  const bool timing = cfg->timing != 0;
  record(c, timing, 0);
  // applyParticleAdaptivity (sparse_sim3d_tp.cpp:855; the single-phase step has none)
  int st = single ? PF_OK : run_adaptivity(c, &cfg->adapt, &S.adapt);
  if (st) return st;
  st = ensure_binned(c);
  if (st) return st;
  record(c, timing, 1);
  // timestep3d.h: the speed scan is always done (stats); it gates dt only if enabled
  st = pf_timestep(c, cfg->dt, cfg->adaptive_timestep, cfg->adaptive_cfl, cfg->adaptive_min_dt, &S.effective_dt,
                   &S.max_particle_speed, &S.cfl_limit_dt, &S.limited);
  if (st) return st;
  const double dt = S.effective_dt;
  S.c_div = single ? cfg->c_div
                   : volume_correction(c, cfg->c_div_volume_correction, cfg->c_div_strength,
                                       cfg->liquid_volume_target, cfg->Vp, dt, cfg->c_div, &S.liquid_volume);
  st = run_mark(c);
  if (st) return st;
  record(c, timing, 2);
  st = run_p2g_splat(c, &cfg->phase, cfg->Vp, single);
  if (st) return st;
  record(c, timing, 3);
  // the single-phase step has no interface work (sparse_sim3d.cpp:106-115)
  st = step_face_stage(c, &cfg->phase, dt, cfg->gravity, false, !single && !cfg->skip_interface_diagnostics,
                       !single && cfg->surface_tension, cfg->surface_tension_strength,
                       cfg->surface_tension_max_delta_speed, cfg->surface_tension_smoothing_radius, &S);
  if (st) return st;
  record(c, timing, 4);
  st = run_project(c, &cfg->phase, dt, &cfg->solve, S.c_div, &S.solve, timing, 5, 6, 7, single);
  if (st) return st;
  HIP_TRY(c, fill_u32(c->stream, reinterpret_cast<uint32_t*>(c->esc_dev), 0u, 8));
  // the single-phase blend uses one alpha (SparseSim3D::alpha, sparse_sim3d.h:8)
  st = run_g2p_advect(c, true, true, cfg->alpha_liquid, single ? cfg->alpha_liquid : cfg->alpha_gas, dt,
                      cfg->advection_order);
  if (st) return st;
  record(c, timing, 8);
  HIP_TRY(c, hipMemcpyAsync(&S.escape, c->esc_dev, sizeof(int32_t) * 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (timing) {
    float ms;
    for (int s = 0; s < PF_NSTAGES; ++s) {
      hipEventElapsedTime(&ms, c->ev[s], c->ev[s + 1]);
      S.stage_ms[s] = ms;
    }
    hipEventElapsedTime(&ms, c->ev[0], c->ev[PF_NSTAGES]);
    S.total_ms = ms;
  }
  if (stats) *stats = S;
  return PF_OK;
}

int pf_field_dims(const pf_ctx* c, int32_t field, int32_t dims[3]) {
  if (!c || !dims || field_slot(field) < 0) return PF_ERR_ARG;
  const Dims& d = dims_of(c, field);
  dims[0] = d.nx; dims[1] = d.ny; dims[2] = d.nz;
  return PF_OK;
}

int pf_field_download(pf_ctx* c, int32_t field, float* dense) {
  int slot = field_slot(field);
  if (!c || !dense || slot < 0) return PF_ERR_ARG;
  const Dims& d = dims_of(c, field);
  long long n = (long long)d.nx * d.ny * d.nz;
  hipLaunchKernelGGL(k_to_dense, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, d,
                     slot == F_MK ? nullptr : c->f[slot], slot == F_MK ? c->marker : nullptr, c->stage_f);
  LAUNCH_CHECK(c);
  HIP_TRY(c, hipMemcpyAsync(dense, c->stage_f, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return PF_OK;
}

int pf_field_upload(pf_ctx* c, int32_t field, const float* dense, const uint8_t* block_active) {
  int slot = field_slot(field);
  if (!c || !dense || slot < 0) return PF_ERR_ARG;
  const Dims& d = dims_of(c, field);
  long long n = (long long)d.nx * d.ny * d.nz;
  if (slot == F_MK) {
    // markers are AIR/FLUID/SOLID and FLUID cells are interior: the pressure
    // stencil of a FLUID cell then never leaves the grid
    for (int k = 0; k < d.nz; ++k)
      for (int j = 0; j < d.ny; ++j)
        for (int i = 0; i < d.nx; ++i) {
          float m = dense[i + (long long)d.nx * (j + (long long)d.ny * k)];
          if (!(m == 0.f || m == 1.f || m == 2.f)) return fail(c, PF_ERR_ARG, "marker values must be 0, 1 or 2");
          bool edge = i == 0 || j == 0 || k == 0 || i == d.nx - 1 || j == d.ny - 1 || k == d.nz - 1;
          if (edge && m == 1.f) return fail(c, PF_ERR_ARG, "FLUID marker on a boundary cell");
        }
  }
  HIP_TRY(c, hipMemcpyAsync(c->stage_f, dense, sizeof(float) * (size_t)n, hipMemcpyHostToDevice, c->stream));
  uint8_t* act_in = nullptr;
  if (block_active) {
    act_in = reinterpret_cast<uint8_t*>(c->blk_flag32);
    HIP_TRY(c, hipMemcpyAsync(act_in, block_active, (size_t)d.nblocks, hipMemcpyHostToDevice, c->stream));
  }
  hipLaunchKernelGGL(k_from_dense, dim3(grid_for(d.nblocks, 4)), dim3(256), 0, c->stream, d, c->stage_f,
                     slot == F_MK ? nullptr : c->f[slot], slot == F_MK ? c->marker : nullptr, c->act[slot], act_in);
  if (slot == F_MK) c->mark_order_valid = false;
  LAUNCH_CHECK(c);
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return PF_OK;
}

int pf_active_blocks(pf_ctx* c, int32_t field, int32_t* bids, int64_t capacity, int64_t* n) {
  int slot = field_slot(field);
  if (!c || !n || slot < 0) return PF_ERR_ARG;
  const Dims& d = dims_of(c, field);
  uint32_t* list = reinterpret_cast<uint32_t*>(c->stage_f);
  HIP_TRY(c, compact_flags(*c, c->act[slot], d.nblocks, list, c->u32_dev + 3));
  HIP_TRY(c, hipMemcpyAsync(c->u32_host, c->u32_dev, sizeof(uint32_t) * 4, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  long long cnt = c->u32_host[3];
  *n = cnt;
  if (bids && capacity > 0) {
    long long m = std::min<long long>(cnt, capacity);
    HIP_TRY(c, hipMemcpyAsync(bids, list, sizeof(int32_t) * (size_t)m, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
  }
  return PF_OK;
}

int pf_bin_counts(pf_ctx* c, uint32_t* dense) {
  if (!c || !dense) return PF_ERR_ARG;
  int st = ensure_binned(c);
  if (st) return st;
  const long long n = (long long)c->dc.nx * c->dc.ny * c->dc.nz;
  uint32_t* out = reinterpret_cast<uint32_t*>(c->stage_f);  // stage_cap >= cells
  hipLaunchKernelGGL(k_counts_dense, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, c->dc, c->cell_count, out);
  LAUNCH_CHECK(c);
  HIP_TRY(c, hipMemcpyAsync(dense, out, sizeof(uint32_t) * (size_t)n, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return PF_OK;
}

int pf_color_lists(pf_ctx* c, int32_t* lists, int64_t capacity, int64_t counts[8]) {
  if (!c || !counts) return PF_ERR_ARG;
  int st = ensure_binned(c);
  if (st) return st;
  HIP_TRY(c, hipMemcpyAsync(c->u32_host, c->color_cnt, sizeof(uint32_t) * 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  long long total = 0;
  for (int k = 0; k < 8; ++k) {
    counts[k] = c->u32_host[k];
    total += counts[k];
  }
  if (!lists) return PF_OK;
  if (capacity < total) return fail(c, PF_ERR_ARG, "pf_color_lists: capacity below the summed list lengths");
  long long off = 0;
  for (int k = 0; k < 8; ++k) {
    if (counts[k] == 0) continue;
    HIP_TRY(c, hipMemcpyAsync(lists + off, c->color_list + (long long)k * c->color_cap,
                              sizeof(int32_t) * (size_t)counts[k], hipMemcpyDeviceToHost, c->stream));
    off += counts[k];
  }
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  off = 0;
  for (int k = 0; k < 8; ++k) {  // append order within a colour is free: return it sorted
    std::sort(lists + off, lists + off + counts[k]);
    off += counts[k];
  }
  return PF_OK;
}

int pf_face_phi(pf_ctx* c, const pf_phase* ph, int32_t axis, double* dense) {
  if (!c || !ph || !dense || axis < 0 || axis > 2) return PF_ERR_ARG;
  const Dims& d = axis == 0 ? c->du : (axis == 1 ? c->dv : c->dw);
  const long long n = (long long)d.nx * d.ny * d.nz;
  double* out = nullptr;
  HIP_TRY(c, dalloc(&out, n));
  hipLaunchKernelGGL(k_face_phi, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, d, c->f[F_MU + axis],
                     phase_k(ph), out);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(dense, out, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  hipFree(out);
  if (e != hipSuccess) return fail(c, PF_ERR_HIP, std::string("pf_face_phi: ") + hipGetErrorString(e));
  return PF_OK;
}

double pf_calibrate_rho_tilde0(const pf_phase* ph, double Vp) { return ph ? calibrate_host(ph, Vp) : 0.0; }

double pf_phi_from_raw(double rt, const pf_phase* pp) {
  double rmin = std::log(pp->rho_l / pp->rho_g) * pp->rho_g * pp->rho_tilde_0;
  if (rt < rmin) return 0.0;
  double v = std::sqrt(std::max(rt - rmin, 0.0) / (pp->alpha_phi * pp->rho_tilde_0 * pp->rho_l));
  return std::min(v, 1.0);
}

double pf_beta_from_phi(double phi, const pf_phase* pp) { return 1.0 / (phi * pp->rho_l + (1.0 - phi) * pp->rho_g); }

}  // extern "C"
