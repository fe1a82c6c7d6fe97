// Seeded procedural scenes for the GI update: the atrium (20 x 20 x 10 m, walls, 24 pillars, tessellated arches, 16
// rigidly moving boxes) and the city block (seeded buildings that translate and rotate each frame, street furniture).
//
// The output is the reference's scene data model: positions and normals per vertex, the triangle table
// (v0, v1, v2, material) of triangleTable (src/core/voxel.ts:730-739) and one material per triangle whose
// baseColorFactor is the triangle's albedo, uniform in [0.1, 0.9] per channel (the white default texture makes the
// voxel albedo equal to baseColorFactor, src/core/scene_gpu.ts:64-68). Quads follow the Builder idiom of
// src/core/scenes.ts:176-199 (outward faces, two triangles (0,1,2), (0,2,3) per quad), subdivided into a grid so that
// triangle counts reach the configured sizes. Moving parts are objects (DynamicObject, src/core/scenes.ts:41) built
// at rest, each with a seeded schedule (DynamicSpec.at); the scene's positions and normals are the rest mesh posed at
// the requested frame by posedVertices (src/core/scenes.ts:80-102): p = pivot + rotY(p - pivot) + offset, from the
// float rest data in double, rounded to float once.
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "rtgi.h"
#include "rtgi_internal.h"

namespace {

struct V3 {
  double x, y, z;
};
V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
double len(V3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
V3 unit(V3 a) {
  double l = len(a);
  return l > 0 ? a * (1.0 / l) : V3{0, 1, 0};
}

// splitmix64: a small, fully specified generator, so scenes are identical on every platform.
struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed ^ 0x9e3779b97f4a7c15ull) {}
  uint64_t next() {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  }
  double uni() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  double uni(double a, double b) { return a + (b - a) * uni(); }
};

struct Builder {
  std::vector<float> pos, nrm;
  std::vector<uint32_t> tri;
  std::vector<rtgi_object> objects;
  std::vector<rtgi_motion> motion;
  double step = 0.5;       // target edge of a sub-quad (m)

  // The vertices made between begin_object and end_object form one object turning about pivot on its schedule.
  void begin_object(V3 pivot, const rtgi_motion& m) {
    rtgi_object o{};
    o.first_vertex = (uint32_t)(pos.size() / 3);
    o.pivot[0] = pivot.x;
    o.pivot[1] = pivot.y;
    o.pivot[2] = pivot.z;
    objects.push_back(o);
    motion.push_back(m);
  }
  void end_object() { objects.back().vertex_count = (uint32_t)(pos.size() / 3) - objects.back().first_vertex; }

  uint32_t vert(V3 p, V3 n) {
    uint32_t i = (uint32_t)(pos.size() / 3);
    pos.push_back((float)p.x);
    pos.push_back((float)p.y);
    pos.push_back((float)p.z);
    nrm.push_back((float)n.x);
    nrm.push_back((float)n.y);
    nrm.push_back((float)n.z);
    return i;
  }
  void triangle(uint32_t a, uint32_t b, uint32_t c) {
    tri.push_back(a);
    tri.push_back(b);
    tri.push_back(c);
    tri.push_back(0);  // material: set to the triangle index at the end
  }
  // Planar quad a b c d (counter-clockwise seen from the outside), subdivided into a grid of about step-sized cells.
  void quad(V3 a, V3 b, V3 c, V3 d, double stp = -1) {
    if (stp <= 0) stp = step;
    V3 n = unit(cross(b - a, c - a));
    int nu = std::max(1, (int)std::ceil(len(b - a) / stp - 1e-9));
    int nv = std::max(1, (int)std::ceil(len(d - a) / stp - 1e-9));
    uint32_t base = (uint32_t)(pos.size() / 3);
    for (int j = 0; j <= nv; j++) {
      double v = (double)j / nv;
      for (int i = 0; i <= nu; i++) {
        double u = (double)i / nu;
        V3 p = a * ((1 - u) * (1 - v)) + b * (u * (1 - v)) + c * (u * v) + d * ((1 - u) * v);
        vert(p, n);
      }
    }
    for (int j = 0; j < nv; j++)
      for (int i = 0; i < nu; i++) {
        uint32_t k0 = base + j * (nu + 1) + i, k1 = k0 + 1, k3 = k0 + (nu + 1), k2 = k3 + 1;
        triangle(k0, k1, k2);
        triangle(k0, k2, k3);
      }
  }
  // Axis-aligned box with outward faces (Builder.box, src/core/scenes.ts:190-199).
  void box(V3 c, V3 h, double stp = -1) {
    double x0 = c.x - h.x, y0 = c.y - h.y, z0 = c.z - h.z, x1 = c.x + h.x, y1 = c.y + h.y, z1 = c.z + h.z;
    quad({x0, y0, z1}, {x1, y0, z1}, {x1, y1, z1}, {x0, y1, z1}, stp);
    quad({x1, y0, z0}, {x0, y0, z0}, {x0, y1, z0}, {x1, y1, z0}, stp);
    quad({x0, y0, z0}, {x0, y0, z1}, {x0, y1, z1}, {x0, y1, z0}, stp);
    quad({x1, y0, z1}, {x1, y0, z0}, {x1, y1, z0}, {x1, y1, z1}, stp);
    quad({x0, y1, z1}, {x1, y1, z1}, {x1, y1, z0}, {x0, y1, z0}, stp);
    quad({x0, y0, z0}, {x1, y0, z0}, {x1, y0, z1}, {x0, y0, z1}, stp);
  }
  // Hexahedron from 8 corners (bottom 0-3, top 4-7, both counter-clockwise seen from above).
  void hexa(const V3 k[8], double stp = -1) {
    quad(k[0], k[3], k[2], k[1], stp);  // bottom (outward -y)
    quad(k[4], k[5], k[6], k[7], stp);  // top
    quad(k[0], k[1], k[5], k[4], stp);
    quad(k[1], k[2], k[6], k[5], stp);
    quad(k[2], k[3], k[7], k[6], stp);
    quad(k[3], k[0], k[4], k[7], stp);
  }
  // Semicircular arch between two points at springing height y, radius = half their distance, swept as nseg
  // hexahedral voussoirs of the given thickness (radial) and depth (across the span).
  void arch(V3 a, V3 b, double y, double thick, double depth, int nseg) {
    V3 m = (a + b) * 0.5;
    m.y = y;
    V3 ax = unit(V3{b.x - a.x, 0, b.z - a.z});
    V3 ac = V3{-ax.z, 0, ax.x};  // across the span
    double r = 0.5 * len(V3{b.x - a.x, 0, b.z - a.z});
    for (int s = 0; s < nseg; s++) {
      double t0 = M_PI * s / nseg, t1 = M_PI * (s + 1) / nseg;
      V3 k[8];
      double rs[2] = {r - thick, r};
      // bottom face: inner ring; top: outer ring (radially).
      for (int q = 0; q < 2; q++) {
        double rr = rs[q];
        V3 p0 = m + ax * (-rr * std::cos(t0)) + V3{0, rr * std::sin(t0), 0};
        V3 p1 = m + ax * (-rr * std::cos(t1)) + V3{0, rr * std::sin(t1), 0};
        k[q * 4 + 0] = p0 + ac * (-depth * 0.5);
        k[q * 4 + 1] = p1 + ac * (-depth * 0.5);
        k[q * 4 + 2] = p1 + ac * (depth * 0.5);
        k[q * 4 + 3] = p0 + ac * (depth * 0.5);
      }
      hexa(k);
    }
  }
};

// posedVertices (src/core/scenes.ts:80-102) of objects [0, n) from rest data starting at vertex `first` into pos / nrm
// (whole-scene arrays), in double with the reference's operation order, each value rounded to float once.
void pose_objects(const rtgi_object* obj, uint32_t n, const rtgi_pose* ps, const float* rest_pos,
                  const float* rest_nrm, uint32_t first, float* pos, float* nrm) {
  for (uint32_t i = 0; i < n; i++) {
    const rtgi_object& o = obj[i];
    const double c = ps[i].c, s = ps[i].s;
    const double* off = ps[i].offset;
    for (uint32_t v = o.first_vertex; v < o.first_vertex + o.vertex_count; v++) {
      const float* rp = rest_pos + (size_t)(v - first) * 3;
      const float* rn = rest_nrm + (size_t)(v - first) * 3;
      double x = (double)rp[0] - o.pivot[0], y = (double)rp[1] - o.pivot[1], z = (double)rp[2] - o.pivot[2];
      double nx = rn[0], ny = rn[1], nz = rn[2];
      pos[(size_t)v * 3 + 0] = (float)(o.pivot[0] + c * x + s * z + off[0]);
      pos[(size_t)v * 3 + 1] = (float)(o.pivot[1] + y + off[1]);
      pos[(size_t)v * 3 + 2] = (float)(o.pivot[2] - s * x + c * z + off[2]);
      nrm[(size_t)v * 3 + 0] = (float)(c * nx + s * nz);
      nrm[(size_t)v * 3 + 1] = (float)ny;
      nrm[(size_t)v * 3 + 2] = (float)(-s * nx + c * nz);
    }
  }
}

rtgi_scene* finish(Builder& b, Rng& rng, const double eye[3], const float sun[3], double frame,
                   uint32_t first_dynamic = ~0u) {
  rtgi_scene* s = (rtgi_scene*)std::calloc(1, sizeof(rtgi_scene));
  if (!s) return nullptr;
  uint32_t nt = (uint32_t)(b.tri.size() / 4), nv = (uint32_t)(b.pos.size() / 3);
  const uint32_t no = (uint32_t)b.objects.size();
  const uint32_t rfirst = no ? b.objects[0].first_vertex : 0u;
  const uint32_t rcount = no ? b.objects[no - 1].first_vertex + b.objects[no - 1].vertex_count - rfirst : 0u;
  s->ntris = nt;
  s->first_dynamic_tri = std::min(first_dynamic, nt);
  s->nverts = nv;
  s->nmat = nt;
  s->nobj = no;
  s->pos = (float*)std::malloc(sizeof(float) * 3 * (size_t)std::max(nv, 1u));
  s->nrm = (float*)std::malloc(sizeof(float) * 3 * (size_t)std::max(nv, 1u));
  s->tri = (uint32_t*)std::malloc(sizeof(uint32_t) * 4 * (size_t)std::max(nt, 1u));
  s->materials = (float*)std::malloc(sizeof(float) * 4 * (size_t)std::max(nt, 1u));
  s->objects = (rtgi_object*)std::malloc(sizeof(rtgi_object) * std::max(no, 1u));
  s->motion = (rtgi_motion*)std::malloc(sizeof(rtgi_motion) * std::max(no, 1u));
  s->rest_pos = (float*)std::malloc(sizeof(float) * 3 * (size_t)std::max(rcount, 1u));
  s->rest_nrm = (float*)std::malloc(sizeof(float) * 3 * (size_t)std::max(rcount, 1u));
  std::vector<rtgi_pose> ps(no);
  if (!s->pos || !s->nrm || !s->tri || !s->materials || !s->objects || !s->motion || !s->rest_pos || !s->rest_nrm) {
    rtgi_scene_free(s);
    return nullptr;
  }
  std::memcpy(s->pos, b.pos.data(), sizeof(float) * b.pos.size());
  std::memcpy(s->nrm, b.nrm.data(), sizeof(float) * b.nrm.size());
  if (no) {
    std::memcpy(s->objects, b.objects.data(), sizeof(rtgi_object) * no);
    std::memcpy(s->motion, b.motion.data(), sizeof(rtgi_motion) * no);
    std::memcpy(s->rest_pos, b.pos.data() + (size_t)rfirst * 3, sizeof(float) * 3 * (size_t)rcount);
    std::memcpy(s->rest_nrm, b.nrm.data() + (size_t)rfirst * 3, sizeof(float) * 3 * (size_t)rcount);
    rtgi_motion_pose_at(s->motion, no, frame, ps.data());
    pose_objects(s->objects, no, ps.data(), s->rest_pos, s->rest_nrm, rfirst, s->pos, s->nrm);
  }
  for (uint32_t t = 0; t < nt; t++) {
    s->tri[4 * t + 0] = b.tri[4 * t + 0];
    s->tri[4 * t + 1] = b.tri[4 * t + 1];
    s->tri[4 * t + 2] = b.tri[4 * t + 2];
    s->tri[4 * t + 3] = t;
    for (int ch = 0; ch < 3; ch++) s->materials[4 * t + ch] = (float)rng.uni(0.1, 0.9);
    s->materials[4 * t + 3] = 1.0f;
  }
  for (int a = 0; a < 3; a++) {
    s->bounds_lo[a] = 1e30;
    s->bounds_hi[a] = -1e30;
  }
  for (uint32_t v = 0; v < nv; v++)
    for (int a = 0; a < 3; a++) {
      s->bounds_lo[a] = std::min(s->bounds_lo[a], (double)s->pos[3 * v + a]);
      s->bounds_hi[a] = std::max(s->bounds_hi[a], (double)s->pos[3 * v + a]);
    }
  for (int a = 0; a < 3; a++) {
    s->eye[a] = eye[a];
    s->sun_dir[a] = sun[a];
  }
  return s;
}

void normalizef(float v[3]) {
  double l = std::sqrt((double)v[0] * v[0] + (double)v[1] * v[1] + (double)v[2] * v[2]);
  for (int a = 0; a < 3; a++) v[a] = (float)(v[a] / l);
}

// Atrium: about 50k triangles at detail 1.
rtgi_scene* atrium(const rtgi_scene_params& p) {
  Rng rng(p.seed * 0x100000001b3ull + 1);
  Builder b;
  double det = p.detail > 0 ? p.detail : 1.0;
  b.step = 0.5 / det;
  // Floor, roof ring around a skylight, perimeter walls (0.3 m thick).
  b.box({0, -0.15, 0}, {10.3, 0.15, 10.3});
  b.box({0, 10.15, -7}, {10.3, 0.15, 3.3});
  b.box({0, 10.15, 7}, {10.3, 0.15, 3.3});
  b.box({-7, 10.15, 0}, {3.3, 0.15, 3.7});
  b.box({7, 10.15, 0}, {3.3, 0.15, 3.7});
  b.box({0, 5, -10.15}, {10.3, 5, 0.15});
  b.box({0, 5, 10.15}, {10.3, 5, 0.15});
  b.box({-10.15, 5, 0}, {0.15, 5, 10.0});
  b.box({10.15, 5, 0}, {0.15, 5, 10.0});
  // 24 pillars on the perimeter of a 7 x 7 ring (2 m apart, +-6 m), 7 m tall.
  std::vector<V3> ring;
  for (int i = 0; i < 7; i++) ring.push_back({-6.0 + 2.0 * i, 0, -6});
  for (int i = 1; i < 7; i++) ring.push_back({6, 0, -6.0 + 2.0 * i});
  for (int i = 1; i < 7; i++) ring.push_back({6.0 - 2.0 * i, 0, 6});
  for (int i = 1; i < 6; i++) ring.push_back({-6, 0, 6.0 - 2.0 * i});
  for (const V3& q : ring) b.box({q.x, 3.5, q.z}, {0.3, 3.5, 0.3}, b.step * 0.5);
  // Tessellated arches between neighbouring pillars, springing at 7 m.
  for (size_t i = 0; i < ring.size(); i++) {
    V3 a = ring[i], c = ring[(i + 1) % ring.size()];
    b.arch(a, c, 7.0, 0.3, 0.6, 16);
  }
  // Gallery slab above the arches.
  b.box({0, 8.15, -6}, {6.3, 0.15, 0.3});
  b.box({0, 8.15, 6}, {6.3, 0.15, 0.3});
  b.box({-6, 8.15, 0}, {0.3, 0.15, 5.7});
  b.box({6, 8.15, 0}, {0.3, 0.15, 5.7});
  // 16 boxes under seeded rigid motion (rotation about y and translation): the tail of the triangle table.
  const uint32_t first_dynamic = (uint32_t)(b.tri.size() / 4);
  for (int k = 0; k < 16; k++) {
    V3 c{rng.uni(-8, 8), rng.uni(0.4, 3.0), rng.uni(-8, 8)};
    V3 h{rng.uni(0.15, 0.4), rng.uni(0.15, 0.4), rng.uni(0.15, 0.4)};
    double w = rng.uni(-0.05, 0.05), ph = rng.uni(0, 2 * M_PI);
    V3 amp{rng.uni(-1, 1), rng.uni(-0.3, 0.3), rng.uni(-1, 1)};
    double wt = rng.uni(0.01, 0.04);
    b.begin_object(c, rtgi_motion{ph, w, wt, ph, {amp.x, amp.y, amp.z}});
    b.box(c, h, b.step * 0.5);
    b.end_object();
  }
  double eye[3] = {0.0, 5.0, 0.0};
  float sun[3] = {-0.3f, 1.0f, 0.25f};
  normalizef(sun);
  return finish(b, rng, eye, sun, p.frame, first_dynamic);
}

// City block: about 500k triangles at detail 1 with 64 buildings on a 200 m block.
rtgi_scene* city(const rtgi_scene_params& p) {
  Rng rng(p.seed * 0x100000001b3ull + 2);
  Builder b;
  double det = p.detail > 0 ? p.detail : 1.0;
  double ext = p.extent > 0 ? p.extent : 200.0;
  int nb = p.buildings > 0 ? p.buildings : 64;
  int side = (int)std::ceil(std::sqrt((double)nb));
  double lot = ext / side;
  b.step = 0.71 / det;
  // Ground, tessellated coarsely.
  b.quad({-ext / 2, 0, ext / 2}, {ext / 2, 0, ext / 2}, {ext / 2, 0, -ext / 2}, {-ext / 2, 0, -ext / 2}, 2.0 / det);
  // The buildings, one object each (contiguous vertices, as posedVertices expects of its objects).
  for (int k = 0; k < nb; k++) {
    int gx = k % side, gz = k / side;
    V3 c{-ext / 2 + (gx + 0.5) * lot, 0, -ext / 2 + (gz + 0.5) * lot};
    double wx = rng.uni(0.35, 0.7) * lot, wz = rng.uni(0.35, 0.7) * lot, hgt = rng.uni(8, 40);
    double rot0 = rng.uni(-0.3, 0.3), w = rng.uni(-0.01, 0.01), ph = rng.uni(0, 2 * M_PI);
    V3 amp{rng.uni(-0.5, 0.5), 0, rng.uni(-0.5, 0.5)};
    b.begin_object(c, rtgi_motion{rot0, w, 0.05, ph, {amp.x, amp.y, amp.z}});
    b.box({c.x, hgt / 2, c.z}, {wx / 2, hgt / 2, wz / 2});
    // A roof block and ledges.
    b.box({c.x, hgt + 1.0, c.z}, {wx / 4, 1.0, wz / 4}, b.step * 2);
    int nled = (int)(hgt / 4);
    for (int l = 1; l < nled; l++) b.box({c.x, 4.0 * l, c.z + wz / 2 + 0.15}, {wx / 2, 0.1, 0.15}, b.step * 2);
    b.end_object();
  }
  for (int k = 0; k < nb; k++) {
    int gx = k % side, gz = k / side;
    V3 c{-ext / 2 + (gx + 0.5) * lot, 0, -ext / 2 + (gz + 0.5) * lot};
    // Street furniture along the lot edge: lamp posts and benches.
    for (int f = 0; f < 6; f++) {
      double t = rng.uni(-0.45, 0.45) * lot;
      V3 q{c.x + t, 0, c.z - lot / 2 + 1.5};
      if (f % 2 == 0)
        b.box({q.x, 2.0, q.z}, {0.1, 2.0, 0.1}, 1.0);
      else
        b.box({q.x, 0.25, q.z}, {0.8, 0.25, 0.25}, 1.0);
    }
  }
  double eye[3] = {1.0, 32.0, 1.0};  // above the roofs: the probe grid spans y 0..63 m
  float sun[3] = {-0.4f, 1.0f, 0.3f};
  normalizef(sun);
  return finish(b, rng, eye, sun, p.frame);
}

}  // namespace

extern "C" int rtgi_scene_generate(const rtgi_scene_params* p, rtgi_scene** out) {
  if (!p || !out) return rtgi_fail(RTGI_E_INVALID, "rtgi_scene_generate: null argument");
  *out = nullptr;
  rtgi_scene* s = nullptr;
  try {
    if (p->kind == RTGI_SCENE_ATRIUM)
      s = atrium(*p);
    else if (p->kind == RTGI_SCENE_CITY)
      s = city(*p);
    else
      return rtgi_fail(RTGI_E_INVALID, "rtgi_scene_generate: unknown scene kind %d", p->kind);
  } catch (const std::bad_alloc&) {
    return rtgi_fail(RTGI_E_NOMEM, "rtgi_scene_generate: out of host memory");
  }
  if (!s) return rtgi_fail(RTGI_E_NOMEM, "rtgi_scene_generate: out of host memory");
  *out = s;
  return RTGI_OK;
}

extern "C" int rtgi_scene_free(rtgi_scene* s) {
  if (!s) return RTGI_OK;
  std::free(s->pos);
  std::free(s->nrm);
  std::free(s->tri);
  std::free(s->materials);
  std::free(s->objects);
  std::free(s->motion);
  std::free(s->rest_pos);
  std::free(s->rest_nrm);
  std::free(s);
  return RTGI_OK;
}

extern "C" int rtgi_motion_pose_at(const rtgi_motion* m, uint32_t n, double frame, rtgi_pose* out) {
  if (n && (!m || !out)) return rtgi_fail(RTGI_E_INVALID, "rtgi_motion_pose_at: null array");
  for (uint32_t i = 0; i < n; i++) {
    const double rot = m[i].rot0 + m[i].w * frame, a = std::sin(m[i].wt * frame + m[i].ph);
    out[i].c = std::cos(rot);
    out[i].s = std::sin(rot);
    for (int k = 0; k < 3; k++) out[i].offset[k] = m[i].amp[k] * a;
  }
  return RTGI_OK;
}
