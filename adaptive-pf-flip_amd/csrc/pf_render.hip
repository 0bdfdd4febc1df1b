// pf_render.hip -- the render-cache JSONL export (src/driver/render_cache3d.h):
// one frame file per call (writeSparseRenderCache3D / writeMRRenderCache3D) and the
// frame manifest (writeRenderCacheManifest3D).
//
// The device computes what needs the grid: the cell phase field (the mean of the six
// face phi, as the interface kernels), and on the sparse grid the order in which
// the reference's markCells allocated the marker blocks, which is the order its
// phase cells are written in.  That order follows from the marking itself: the
// first event that touched a block, counting the three wall-plane loops in their
// loop order (sparse_sim3d_tp.cpp:23-42) and then the particles in their order
// (:43-50), so a block's key is its earliest wall event, or the wall-event count plus
// the smallest id of the particles binned into it.  The host sorts the active blocks
// by key, builds the cell records, and formats the file with worker threads; numbers
// use the shortest form of 17 significant digits (std::to_chars general, precision
// 17), which is what the reference's setprecision(17) stream prints.
#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "pf_internal.h"

namespace pf {

// per block: the allocation key of a marker-active block (~0 for the others)
__global__ void __launch_bounds__(256) k_render_block_keys(Dims dc, Geom g, const uint8_t* __restrict__ act_mk,
                                                           const uint32_t* __restrict__ start,
                                                           const uint32_t* __restrict__ ids,
                                                           unsigned long long* __restrict__ key) {
  const long long b = (long long)blockIdx.x * 256 + threadIdx.x;
  if (b >= dc.nblocks) return;
  if (!act_mk[b]) {
    key[b] = ~0ull;
    return;
  }
  int bx, by, bz;
  block_coords(dc, b, bx, by, bz);
  const unsigned long long nx = g.nx, ny = g.ny, nz = g.nz;
  const unsigned long long e1 = nz * ny * 2, e2 = nz * nx * 2, e3 = ny * nx * 2;
  unsigned long long best = ~0ull, pmin = ~0ull;
  for (int l = 0; l < 64; ++l) {
    const int i = bx * 4 + (l & 3), j = by * 4 + ((l >> 2) & 3), k = bz * 4 + (l >> 4);
    if (i >= g.nx || j >= g.ny || k >= g.nz) continue;
    // the first wall loop that sets this cell, and its event index in that loop
    unsigned long long e = ~0ull;
    if (i == 0 || i == g.nx - 1)
      e = ((unsigned long long)k * ny + j) * 2 + (i == 0 ? 0 : 1);
    else if (j == 0 || j == g.ny - 1)
      e = e1 + ((unsigned long long)k * nx + i) * 2 + (j == 0 ? 0 : 1);
    else if (k == 0 || k == g.nz - 1)
      e = e1 + e2 + ((unsigned long long)j * nx + i) * 2 + (k == 0 ? 0 : 1);
    best = e < best ? e : best;
    const long long c = b * kBlockVol + l;
    for (uint32_t p = start[c]; p < start[c + 1]; ++p) pmin = ids[p] < pmin ? ids[p] : pmin;
  }
  if (best == ~0ull) best = pmin != ~0ull ? e1 + e2 + e3 + pmin : e1 + e2 + e3 + (1ull << 33) + (unsigned long long)b;
  key[b] = best;
}

}  // namespace pf

namespace {

using namespace pf;

// text with the number formats of the reference's stream
struct Text {
  std::string s;
  void raw(const char* t) { s += t; }
  void num(double v) {
    char b[40];
    const auto r = std::to_chars(b, b + sizeof(b), v, std::chars_format::general, 17);
    s.append(b, r.ptr);
  }
  void num(long long v) {
    char b[24];
    const auto r = std::to_chars(b, b + sizeof(b), v);
    s.append(b, r.ptr);
  }
  void vec(const char* key, const double v[3]) {
    s += '"';
    s += key;
    s += "\":[";
    num(v[0]);
    s += ',';
    num(v[1]);
    s += ',';
    num(v[2]);
    s += ']';
  }
};

struct RenderCell {
  int i, j, k, level, marker;
  double phi, vol;
};

struct RenderParticles {
  std::vector<float> x, y, z, u, v, w, vol;
  std::vector<uint8_t> type;
};

unsigned worker_count() {
  unsigned t = std::thread::hardware_concurrency();
  return t < 1 ? 1 : (t > 16 ? 16 : t);
}

// format items [0, n) by `one(text, index)` with worker threads in chunks, writing
// each chunk to f in order; returns false on a write error
template <class F>
bool format_parallel(FILE* f, long long n, F one, long long& bytes) {
  const unsigned nt = worker_count();
  const long long chunk = 1 << 20;
  std::vector<Text> parts(nt);
  for (long long base = 0; base < n; base += chunk) {
    const long long m = std::min(chunk, n - base);
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t) {
      const long long lo = base + m * t / nt, hi = base + m * (t + 1) / nt;
      parts[t].s.clear();
      th.emplace_back([&parts, t, lo, hi, &one]() {
        for (long long q = lo; q < hi; ++q) one(parts[t], q);
      });
    }
    for (auto& x : th) x.join();
    for (unsigned t = 0; t < nt; ++t) {
      if (fwrite(parts[t].s.data(), 1, parts[t].s.size(), f) != parts[t].s.size()) return false;
      bytes += (long long)parts[t].s.size();
    }
  }
  return true;
}

bool finite3(const double* v) { return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]); }

struct Bounds {
  double mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0};
  bool valid = false;
  void add(double x, double y, double z) {   // includePoint, render_cache3d.h:96-109
    const double p[3] = {x, y, z};
    if (!finite3(p)) return;
    if (!valid) {
      for (int a = 0; a < 3; ++a) mn[a] = mx[a] = p[a];
      valid = true;
      return;
    }
    for (int a = 0; a < 3; ++a) {
      mn[a] = std::min(mn[a], p[a]);
      mx[a] = std::max(mx[a], p[a]);
    }
  }
};

// the frame file (render_cache3d.h:362-555, in the order of :592-711); false with
// err set on an I/O failure
bool write_frame(const Geom& g, const char* path, const char* kind, const pf_render_frame* m,
                 const std::vector<RenderCell>& cells, double phase_vol, const RenderParticles& P, int64_t* bytes_out,
                 std::string& err) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const double dx = g.dx;
  const long long n = (long long)P.type.size();
  FILE* f = std::fopen(path, "wb");
  if (!f) {
    err = std::string("render cache open failed: ") + path;
    return false;
  }
  const double dmin[3] = {0.0, 0.0, 0.0}, dmax[3] = {nx * dx, ny * dx, nz * dx};
  const double close = m->time + std::max(0.0, m->dt);
  const pf_render_camera& cam = m->camera;
  long long nl = 0, ng = 0;
  double lvol = 0.0, gvol = 0.0;
  Bounds wb;
  for (long long p = 0; p < n; ++p) {
    if (P.type[p] == 0) {
      lvol += (double)P.vol[p] * m->Vp;
      ++nl;
      wb.add(P.x[p], P.y[p], P.z[p]);
    } else if (P.type[p] == 1) {
      gvol += (double)P.vol[p] * m->Vp;
      ++ng;
    }
  }
  for (const RenderCell& q : cells) {   // waterBounds, :265-277
    if (q.phi <= 0.0 && q.vol <= 0.0) continue;
    const int s = 1 << std::max(0, q.level);
    wb.add(q.i * dx, q.j * dx, q.k * dx);
    wb.add((q.i + s) * dx, (q.j + s) * dx, (q.k + s) * dx);
  }
  Text t;
  t.raw("{\"section\":\"header\",\"lsfs_cache3d_version\":1,\"cache_schema_version\":2,\"sim_kind\":\"");
  t.raw(kind);
  t.raw("\",\"frame\":");
  t.num((long long)m->frame);
  t.raw(",\"time\":");
  t.num(m->time);
  t.raw(",\"dt\":");
  t.num(m->dt);
  t.raw(",\"dims\":[");
  t.num((long long)nx);
  t.raw(",");
  t.num((long long)ny);
  t.raw(",");
  t.num((long long)nz);
  t.raw("],\"dx\":");
  t.num(dx);
  t.raw(",\"world_units\":\"cell\",\"shutter_open\":");
  t.num(m->time);
  t.raw(",\"shutter_close\":");
  t.num(close);
  t.raw(",");
  t.vec("frame_bounds_min", dmin);
  t.raw(",");
  t.vec("frame_bounds_max", dmax);
  t.raw(",\"phase\":{\"rho_l\":");
  t.num(m->phase.rho_l);
  t.raw(",\"rho_g\":");
  t.num(m->phase.rho_g);
  t.raw(",\"alpha_phi\":");
  t.num(m->phase.alpha_phi);
  t.raw(",\"rho_tilde_0\":");
  t.num(m->phase.rho_tilde_0);
  t.raw("}}\n{\"section\":\"camera\",");
  t.vec("position", cam.position);
  t.raw(",");
  t.vec("target", cam.target);
  t.raw(",");
  t.vec("up", cam.up);
  t.raw(",\"fov_degrees\":");
  t.num(cam.fov_degrees);
  t.raw(",\"vertical_fov_degrees\":");
  t.num(cam.fov_degrees);
  t.raw(",\"near_clip\":");
  t.num(cam.near_clip);
  t.raw(",\"far_clip\":");
  t.num(cam.far_clip);
  t.raw(",\"focal_length_mm\":");
  t.num(cam.focal_length_mm);
  t.raw("}\n{\"section\":\"water_volume\",\"primary_liquid_volume\":");
  t.num(lvol);
  t.raw(",\"primary_gas_volume\":");
  t.num(gvol);
  t.raw(",\"secondary_droplet_volume\":0,\"secondary_bubble_volume\":0,\"phase_field_liquid_volume\":");
  t.num(phase_vol);
  t.raw(",\"phase_field_cells\":");
  t.num((long long)cells.size());
  t.raw(",\"primary_liquid_count\":");
  t.num(nl);
  t.raw(",\"primary_gas_count\":");
  t.num(ng);
  t.raw(",\"secondary_droplet_count\":0,\"secondary_bubble_count\":0,\"secondary_particle_count\":0"
        ",\"secondary_droplet_age_min\":0,\"secondary_droplet_age_max\":0,\"secondary_bubble_age_min\":0"
        ",\"secondary_bubble_age_max\":0}\n");
  t.raw("{\"section\":\"secondary_channels\",\"encoding\":\"render_channel_counts\",\"channel_names\":[\"droplet\","
        "\"spray\",\"foam\",\"bubble\"],\"droplet_count\":0,\"spray_count\":0,\"foam_count\":0,\"bubble_count\":0,"
        "\"total_count\":0}\n");
  t.raw("{\"section\":\"cinematic_metadata\",\"cache_schema_version\":2,\"world_units\":\"cell\",\"shutter_open\":");
  t.num(m->time);
  t.raw(",\"shutter_close\":");
  t.num(close);
  t.raw(",\"frame_time\":");
  t.num(m->time);
  t.raw(",\"phase_field_sampling_stride\":1,\"water_bounds_valid\":");
  t.raw(wb.valid ? "true" : "false");
  t.raw(",\"secondary_bounds_valid\":false,\"primary_particle_count\":");
  t.num(n);
  t.raw(",\"primary_liquid_count\":");
  t.num(nl);
  t.raw(",\"primary_gas_count\":");
  t.num(ng);
  t.raw(",\"secondary_droplet_count\":0,\"secondary_bubble_count\":0,\"secondary_spray_count\":0"
        ",\"secondary_foam_count\":0,");
  t.vec("frame_bounds_min", dmin);
  t.raw(",");
  t.vec("frame_bounds_max", dmax);
  t.raw(",");
  t.vec("water_bounds_min", wb.valid ? wb.mn : dmin);
  t.raw(",");
  t.vec("water_bounds_max", wb.valid ? wb.mx : dmax);
  t.raw(",");
  t.vec("secondary_bounds_min", dmin);
  t.raw(",");
  t.vec("secondary_bounds_max", dmax);
  t.raw("}\n{\"section\":\"phase_field\",\"encoding\":\"jsonl_cells\",\"count\":");
  t.num((long long)cells.size());
  t.raw("}\n");
  long long bytes = (long long)t.s.size();
  bool ok = std::fwrite(t.s.data(), 1, t.s.size(), f) == t.s.size();
  ok = ok && format_parallel(f, (long long)cells.size(), [&cells](Text& o, long long q) {
    const RenderCell& r = cells[(size_t)q];
    o.raw("{\"section\":\"phase_cell\",\"i\":");
    o.num((long long)r.i);
    o.raw(",\"j\":");
    o.num((long long)r.j);
    o.raw(",\"k\":");
    o.num((long long)r.k);
    o.raw(",\"level\":");
    o.num((long long)r.level);
    o.raw(",\"marker\":");
    o.num((long long)r.marker);
    o.raw(",\"phi\":");
    o.num(r.phi);
    o.raw(",\"liquid_volume\":");
    o.num(r.vol);
    o.raw("}\n");
  }, bytes);
  Text h;
  h.raw("{\"section\":\"particles\",\"kind\":\"primary\",\"count\":");
  h.num(n);
  h.raw("}\n");
  ok = ok && std::fwrite(h.s.data(), 1, h.s.size(), f) == h.s.size();
  bytes += (long long)h.s.size();
  const double Vp = m->Vp;
  ok = ok && format_parallel(f, n, [&P, Vp](Text& o, long long p) {
    const bool liquid = P.type[(size_t)p] == 0;
    o.raw(liquid ? "{\"section\":\"particle\",\"kind\":\"primary\",\"render_channel\":\"water\",\"index\":"
                 : "{\"section\":\"particle\",\"kind\":\"primary\",\"render_channel\":\"air\",\"index\":");
    o.num(p);
    o.raw(liquid ? ",\"phase\":\"liquid\",\"position\":[" : ",\"phase\":\"gas\",\"position\":[");
    o.num((double)P.x[(size_t)p]);
    o.raw(",");
    o.num((double)P.y[(size_t)p]);
    o.raw(",");
    o.num((double)P.z[(size_t)p]);
    o.raw("],\"velocity\":[");
    o.num((double)P.u[(size_t)p]);
    o.raw(",");
    o.num((double)P.v[(size_t)p]);
    o.raw(",");
    o.num((double)P.w[(size_t)p]);
    o.raw("],\"volume\":");
    o.num((double)P.vol[(size_t)p] * Vp);
    o.raw("}\n");
  }, bytes);
  const char* tail = "{\"section\":\"particles\",\"kind\":\"secondary_droplet\",\"count\":0}\n"
                     "{\"section\":\"particles\",\"kind\":\"secondary_bubble\",\"count\":0}\n";
  ok = ok && std::fwrite(tail, 1, std::strlen(tail), f) == std::strlen(tail);
  bytes += (long long)std::strlen(tail);
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) {
    err = std::string("render cache write failed: ") + path;
    return false;
  }
  if (bytes_out) *bytes_out = bytes;
  return true;
}

// writeJsonString, render_cache3d.h:124-147
void json_string(Text& t, const char* s) {
  static const char* hex = "0123456789abcdef";
  t.s += '"';
  for (const unsigned char* p = (const unsigned char*)s; *p; ++p) {
    switch (*p) {
      case '"': t.s += "\\\""; break;
      case '\\': t.s += "\\\\"; break;
      case '\b': t.s += "\\b"; break;
      case '\f': t.s += "\\f"; break;
      case '\n': t.s += "\\n"; break;
      case '\r': t.s += "\\r"; break;
      case '\t': t.s += "\\t"; break;
      default:
        if (*p < 0x20) {
          t.s += "\\u00";
          t.s += hex[(*p >> 4) & 15];
          t.s += hex[*p & 15];
        } else {
          t.s += (char)*p;
        }
    }
  }
  t.s += '"';
}

}  // namespace
