/* pf_oracle_render.c -- oracle (test infrastructure, not the product) for the
 * render-cache JSONL export (src/driver/render_cache3d.h): writeSparseRenderCache3D,
 * writeMRRenderCache3D and writeRenderCacheManifest3D, restated with stdio.  Numbers
 * print as %.17g, which is what the reference's std::setprecision(17) stream gives;
 * this library has no secondary particles, so their sections carry empty sets. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "pf_oracle.h"

#define NUM "%.17g"

/* defaultRenderCacheCamera3D, render_cache3d.h:48-62 */
void orc_render_default_camera(int nx, int ny, int nz, double dx, double cam[13]) {
  const double cx = 0.5 * nx * dx, cy = 0.5 * ny * dx, cz = 0.5 * nz * dx;
  double span = nx * dx;
  if (ny * dx > span) span = ny * dx;
  if (nz * dx > span) span = nz * dx;
  if (dx > span) span = dx;
  const double c[13] = {cx, cy + 0.2 * span, cz + 2.0 * span, cx, cy, cz, 0.0, 1.0, 0.0, 45.0, 0.05 * dx, 5.0 * span, 50.0};
  memcpy(cam, c, sizeof(c));
}

typedef struct {
  double min[3], max[3];
  int valid;
} bounds;

static void include_point(bounds* b, double x, double y, double z) {
  if (!isfinite(x) || !isfinite(y) || !isfinite(z)) return;   /* includePoint, :96-109 */
  const double p[3] = {x, y, z};
  if (!b->valid) {
    for (int a = 0; a < 3; ++a) b->min[a] = b->max[a] = p[a];
    b->valid = 1;
    return;
  }
  for (int a = 0; a < 3; ++a) {
    if (p[a] < b->min[a]) b->min[a] = p[a];   /* std::min(min, p): keeps min unless p < min */
    if (b->max[a] < p[a]) b->max[a] = p[a];
  }
}

static void put_vec(FILE* f, const char* key, const double v[3]) {
  fprintf(f, "\"%s\":[" NUM "," NUM "," NUM "]", key, v[0], v[1], v[2]);
}

typedef struct {
  int i, j, k, level, marker;
  double phi, vol;
} rcell;

/* the frame file shared by the sparse and multires writers (:592-711) */
static int write_frame(const char* path, const char* kind, int nx, int ny, int nz, double dx,
                       const orc_render_meta* m, const rcell* cells, int64_t ncell, double phase_vol, int64_t n,
                       const double* x, const double* y, const double* z, const double* u, const double* v,
                       const double* w, const uint8_t* type, const double* vol) {
  const double* cam = m->camera;
  int ok = nx > 0 && ny > 0 && nz > 0 && dx > 0.0 && isfinite(dx) && isfinite(m->dt) && isfinite(m->time);
  for (int a = 0; a < 9; ++a) ok = ok && isfinite(cam[a]);
  ok = ok && cam[9] > 0.0 && isfinite(cam[9]) && cam[10] > 0.0 && cam[11] > cam[10] && isfinite(cam[10]) &&
       isfinite(cam[11]) && cam[12] > 0.0 && isfinite(cam[12]);
  if (!ok || !path || !path[0]) return -1;   /* writeHeader's invalid_argument, :372-381 */
  FILE* f = fopen(path, "w");
  if (!f) return -2;
  const double dmin[3] = {0.0, 0.0, 0.0}, dmax[3] = {nx * dx, ny * dx, nz * dx};
  const double close = m->time + (m->dt > 0.0 ? m->dt : 0.0);
  /* writeHeader, :362-418 */
  fprintf(f, "{\"section\":\"header\",\"lsfs_cache3d_version\":1,\"cache_schema_version\":2,\"sim_kind\":\"%s\","
             "\"frame\":%d,\"time\":" NUM ",\"dt\":" NUM ",\"dims\":[%d,%d,%d],\"dx\":" NUM
             ",\"world_units\":\"cell\",\"shutter_open\":" NUM ",\"shutter_close\":" NUM ",",
          kind, m->frame, m->time, m->dt, nx, ny, nz, dx, m->time, close);
  put_vec(f, "frame_bounds_min", dmin);
  fputc(',', f);
  put_vec(f, "frame_bounds_max", dmax);
  fprintf(f, ",\"phase\":{\"rho_l\":" NUM ",\"rho_g\":" NUM ",\"alpha_phi\":" NUM ",\"rho_tilde_0\":" NUM "}}\n",
          m->phase.rho_l, m->phase.rho_g, m->phase.alpha_phi, m->phase.rho_tilde_0);
  fputs("{\"section\":\"camera\",", f);
  put_vec(f, "position", cam);
  fputc(',', f);
  put_vec(f, "target", cam + 3);
  fputc(',', f);
  put_vec(f, "up", cam + 6);
  fprintf(f, ",\"fov_degrees\":" NUM ",\"vertical_fov_degrees\":" NUM ",\"near_clip\":" NUM ",\"far_clip\":" NUM
             ",\"focal_length_mm\":" NUM "}\n", cam[9], cam[9], cam[10], cam[11], cam[12]);
  /* writeWaterSummary, :420-455 */
  double lvol = 0.0, gvol = 0.0;
  int64_t nl = 0, ngas = 0;
  for (int64_t p = 0; p < n; ++p) {
    if (type[p] == 0) { lvol += vol[p] * m->Vp; ++nl; }
    if (type[p] == 1) { gvol += vol[p] * m->Vp; ++ngas; }
  }
  fprintf(f, "{\"section\":\"water_volume\",\"primary_liquid_volume\":" NUM ",\"primary_gas_volume\":" NUM
             ",\"secondary_droplet_volume\":0,\"secondary_bubble_volume\":0,\"phase_field_liquid_volume\":" NUM
             ",\"phase_field_cells\":%lld,\"primary_liquid_count\":%lld,\"primary_gas_count\":%lld"
             ",\"secondary_droplet_count\":0,\"secondary_bubble_count\":0,\"secondary_particle_count\":0"
             ",\"secondary_droplet_age_min\":0,\"secondary_droplet_age_max\":0,\"secondary_bubble_age_min\":0"
             ",\"secondary_bubble_age_max\":0}\n",
          lvol, gvol, phase_vol, (long long)ncell, (long long)nl, (long long)ngas);
  /* writeSecondaryChannels, :457-474 */
  fputs("{\"section\":\"secondary_channels\",\"encoding\":\"render_channel_counts\",\"channel_names\":[\"droplet\","
        "\"spray\",\"foam\",\"bubble\"],\"droplet_count\":0,\"spray_count\":0,\"foam_count\":0,\"bubble_count\":0,"
        "\"total_count\":0}\n", f);
  /* writeCinematicMetadata with waterBounds, :476-516, :265-277 */
  bounds wb = {{0, 0, 0}, {0, 0, 0}, 0};
  for (int64_t p = 0; p < n; ++p)
    if (type[p] == 0) include_point(&wb, x[p], y[p], z[p]);
  for (int64_t c = 0; c < ncell; ++c) {
    if (cells[c].phi <= 0.0 && cells[c].vol <= 0.0) continue;
    const int s = 1 << (cells[c].level > 0 ? cells[c].level : 0);
    include_point(&wb, cells[c].i * dx, cells[c].j * dx, cells[c].k * dx);
    include_point(&wb, (cells[c].i + s) * dx, (cells[c].j + s) * dx, (cells[c].k + s) * dx);
  }
  fprintf(f, "{\"section\":\"cinematic_metadata\",\"cache_schema_version\":2,\"world_units\":\"cell\","
             "\"shutter_open\":" NUM ",\"shutter_close\":" NUM ",\"frame_time\":" NUM
             ",\"phase_field_sampling_stride\":1,\"water_bounds_valid\":%s,\"secondary_bounds_valid\":false"
             ",\"primary_particle_count\":%lld,\"primary_liquid_count\":%lld,\"primary_gas_count\":%lld"
             ",\"secondary_droplet_count\":0,\"secondary_bubble_count\":0,\"secondary_spray_count\":0"
             ",\"secondary_foam_count\":0,",
          m->time, close, m->time, wb.valid ? "true" : "false", (long long)n, (long long)nl, (long long)ngas);
  put_vec(f, "frame_bounds_min", dmin);
  fputc(',', f);
  put_vec(f, "frame_bounds_max", dmax);
  fputc(',', f);
  put_vec(f, "water_bounds_min", wb.valid ? wb.min : dmin);
  fputc(',', f);
  put_vec(f, "water_bounds_max", wb.valid ? wb.max : dmax);
  fputc(',', f);
  put_vec(f, "secondary_bounds_min", dmin);
  fputc(',', f);
  put_vec(f, "secondary_bounds_max", dmax);
  fputs("}\n", f);
  /* writePhaseField, :518-532 */
  fprintf(f, "{\"section\":\"phase_field\",\"encoding\":\"jsonl_cells\",\"count\":%lld}\n", (long long)ncell);
  for (int64_t c = 0; c < ncell; ++c)
    fprintf(f, "{\"section\":\"phase_cell\",\"i\":%d,\"j\":%d,\"k\":%d,\"level\":%d,\"marker\":%d,\"phi\":" NUM
               ",\"liquid_volume\":" NUM "}\n",
            cells[c].i, cells[c].j, cells[c].k, cells[c].level, cells[c].marker, cells[c].phi, cells[c].vol);
  /* writeParticleSection, :534-555: the primary particles, then the empty secondary sets */
  fprintf(f, "{\"section\":\"particles\",\"kind\":\"primary\",\"count\":%lld}\n", (long long)n);
  for (int64_t p = 0; p < n; ++p)
    fprintf(f, "{\"section\":\"particle\",\"kind\":\"primary\",\"render_channel\":\"%s\",\"index\":%lld,"
               "\"phase\":\"%s\",\"position\":[" NUM "," NUM "," NUM "],\"velocity\":[" NUM "," NUM "," NUM
               "],\"volume\":" NUM "}\n",
            type[p] == 0 ? "water" : "air", (long long)p, type[p] == 0 ? "liquid" : "gas", x[p], y[p], z[p], u[p],
            v[p], w[p], vol[p] * m->Vp);
  fputs("{\"section\":\"particles\",\"kind\":\"secondary_droplet\",\"count\":0}\n", f);
  fputs("{\"section\":\"particles\",\"kind\":\"secondary_bubble\",\"count\":0}\n", f);
  const int err = ferror(f);
  fclose(f);
  return err ? -2 : 0;
}

/* writeRenderCacheManifest3D, :713-755; the JSON string escapes of writeJsonString, :124-147 */
static void put_json_string(FILE* f, const char* s) {
  static const char* hex = "0123456789abcdef";
  fputc('"', f);
  for (const unsigned char* c = (const unsigned char*)s; *c; ++c) {
    switch (*c) {
      case '"': fputs("\\\"", f); break;
      case '\\': fputs("\\\\", f); break;
      case '\b': fputs("\\b", f); break;
      case '\f': fputs("\\f", f); break;
      case '\n': fputs("\\n", f); break;
      case '\r': fputs("\\r", f); break;
      case '\t': fputs("\\t", f); break;
      default:
        if (*c < 0x20) fprintf(f, "\\u00%c%c", hex[(*c >> 4) & 15], hex[*c & 15]);
        else fputc(*c, f);
    }
  }
  fputc('"', f);
}

int orc_render_manifest(const char* path, const char* kind, int nx, int ny, int nz, double dx, int64_t nframes,
                        const int32_t* frame, const int32_t* step, const double* time, const char* const* paths,
                        const int64_t* bytes) {
  if (!path || !path[0] || !kind || !kind[0] || nx <= 0 || ny <= 0 || nz <= 0 || dx <= 0.0 || !isfinite(dx)) return -1;
  for (int64_t a = 0; a < nframes; ++a)
    if (frame[a] < 0 || step[a] < 0 || !paths[a] || !paths[a][0] || bytes[a] < 0 || !isfinite(time[a])) return -1;
  FILE* f = fopen(path, "w");
  if (!f) return -2;
  fputs("{\n  \"lsfs_cache3d_manifest_version\":1,\n  \"cache_schema_version\":2,\n  \"world_units\":\"cell\",\n"
        "  \"sim_kind\":", f);
  put_json_string(f, kind);
  fprintf(f, ",\n  \"dims\":[%d,%d,%d],\n  \"dx\":" NUM ",\n  \"frame_bounds_min\":[0,0,0],\n"
             "  \"frame_bounds_max\":[" NUM "," NUM "," NUM "],\n  \"frame_count\":%lld,\n  \"frames\":[\n",
          nx, ny, nz, dx, nx * dx, ny * dx, nz * dx, (long long)nframes);
  for (int64_t a = 0; a < nframes; ++a) {
    fprintf(f, "    {\"frame\":%d,\"step\":%d,\"time\":" NUM ",\"shutter_open\":" NUM ",\"shutter_close\":" NUM
               ",\"frame_bounds_min\":[0,0,0],\"frame_bounds_max\":[" NUM "," NUM "," NUM "],\"path\":",
            frame[a], step[a], time[a], time[a], time[a], nx * dx, ny * dx, nz * dx);
    put_json_string(f, paths[a]);
    fprintf(f, ",\"bytes\":%lld}%s\n", (long long)bytes[a], a + 1 < nframes ? "," : "");
  }
  fputs("  ]\n}\n", f);
  const int err = ferror(f);
  fclose(f);
  return err ? -2 : 0;
}

/* the multires cells (mrPhaseCells, :590-613) are built in pf_oracle_mr.c, which owns
 * the leaf layout; it calls this writer */
int orc_render_frame_cells(const char* path, const char* kind, int nx, int ny, int nz, double dx,
                           const orc_render_meta* m, const int32_t* cell_ijk_level_marker, const double* cell_phi_vol,
                           int64_t ncell, double phase_vol, int64_t n, const double* x, const double* y,
                           const double* z, const double* u, const double* v, const double* w, const uint8_t* type,
                           const double* vol) {
  rcell* cells = (rcell*)malloc(sizeof(rcell) * (size_t)(ncell > 0 ? ncell : 1));
  for (int64_t c = 0; c < ncell; ++c) {
    const int32_t* q = cell_ijk_level_marker + 5 * c;
    cells[c].i = q[0]; cells[c].j = q[1]; cells[c].k = q[2]; cells[c].level = q[3]; cells[c].marker = q[4];
    cells[c].phi = cell_phi_vol[2 * c]; cells[c].vol = cell_phi_vol[2 * c + 1];
  }
  const int r = write_frame(path, kind, nx, ny, nz, dx, m, cells, ncell, phase_vol, n, x, y, z, u, v, w, type, vol);
  free(cells);
  return r;
}
