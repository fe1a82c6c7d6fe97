// Host side of librtgi: the C ABI of include/rtgi.h. A world owns its HIP allocations (clipmap, cascades, scene copy)
// and a stream; each update is a fixed kernel sequence with no host round trip (launch_* in build.hip / trace.hip).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "rtgi.h"
#include "rtgi_internal.h"
#include "world.h"

namespace {
thread_local char g_err[512] = "";
}

int rtgi_fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

namespace {

template <class T>
int dalloc(rtgi_world* w, T** p, size_t count) {
  size_t b = std::max<size_t>(count, 1) * sizeof(T);
  hipError_t e = hipMalloc((void**)p, b);
  if (e != hipSuccess) {
    *p = nullptr;
    return rtgi_fail(RTGI_E_NOMEM, "hipMalloc(%zu bytes): %s", b, hipGetErrorString(e));
  }
  w->bytes += b;
  return RTGI_OK;
}

template <class T>
void dfree(rtgi_world* w, T*& p, size_t count) {
  if (p) {
    (void)hipFree(p);
    w->bytes -= std::max<size_t>(count, 1) * sizeof(T);
  }
  p = nullptr;
}

// octEncode (src/core/frame.ts:97-103) in double.
void oct_encode_d(double x, double y, double z, double& u, double& v) {
  double s = std::fabs(x) + std::fabs(y) + std::fabs(z);
  double px = x / s, py = y / s;
  if (z < 0) {
    double ox = px, oy = py;
    px = (1 - std::fabs(oy)) * (ox >= 0 ? 1 : -1);
    py = (1 - std::fabs(ox)) * (oy >= 0 ? 1 : -1);
  }
  u = px * 0.5 + 0.5;
  v = py * 0.5 + 0.5;
}

// Radiance and SH buffer sizes in floats: f32 rgba / 27 f32 per probe, or halves in the precision form.
size_t rad_floats(const rtgi_world* w) { return (size_t)w->ray_off[w->N] * ((w->rp.flags & RTGI_RC_FP16) ? 2 : 4); }
float half_to_float(uint16_t h) {
  uint32_t s = (uint32_t)(h >> 15) << 31, e = (h >> 10) & 31u, m = h & 1023u, bits;
  if (e == 0) {
    if (m == 0) {
      bits = s;
    } else {  // subnormal: renormalise
      e = 127 - 15 + 1;
      while (!(m & 1024u)) {
        m <<= 1;
        e--;
      }
      bits = s | (e << 23) | ((m & 1023u) << 13);
    }
  } else if (e == 31) {
    bits = s | 0x7f800000u | (m << 13);
  } else {
    bits = s | ((e + 127 - 15) << 23) | (m << 13);
  }
  float f;
  std::memcpy(&f, &bits, 4);
  return f;
}

bool is_pow2(double v) {
  int e;
  return std::frexp(v, &e) == 0.5;
}

}  // namespace

size_t sh_floats_of(const rtgi_world* w) {
  size_t np0 = (size_t)w->n[0] * w->n[0] * w->n[0];
  return (w->rp.flags & RTGI_RC_FP16) ? (np0 * 27 + 1) / 2 : np0 * 27;
}
int sh_buffers(const rtgi_rc_params* rp) {
  return (rp->flags & (RTGI_RC_NO_HISTORY | RTGI_RC_NO_MERGE)) ? 1 : 2;
}

VoxK vox_args(const rtgi_world* w) {
  VoxK k{};
  k.v0 = (float)w->vp.v0;
  k.pow2 = is_pow2(w->vp.v0) && (double)k.v0 == w->vp.v0;
  k.inv_v0 = k.pow2 ? 1.0f / k.v0 : 0.0f;
  k.NL = w->NL;
  k.G = w->G;
  k.E = w->E;
  k.per = w->per;
  k.max_bricks = w->vp.max_bricks;
  k.max_voxels = w->vp.max_voxels;
  std::memcpy(k.origin, w->origin, sizeof(k.origin));
  k.hdr = w->hdr;
  k.ptr = w->ptr;
  k.occ = w->occ;
  k.battr = w->battr;
  k.wslot = w->wslot;
  k.bcode = w->bcode;
  k.rep = w->rep;
  k.attr = w->attr;
  k.plane = w->plane;
  k.vslot = w->vslot;
  k.alb16 = w->alb16;
  k.big_vox = w->big_vox;
  k.big_mark = w->big_mark;
  k.big_cap = big_cap_for(w->ntris);
  k.pos = w->pos;
  k.nrm = w->nrm;
  k.tri = w->tri;
  k.mat = w->mat;
  k.ntris = std::min(w->ntris, w->dyn_first);  // static triangles only (voxel.ts:822)
  k.lds_marks = 0;
  k.lds_marks = vox_mark_lds_bytes(k) > 0;
  k.pose = w->d_pose;
  k.rest = w->rest;
  k.npose = w->pose_set && w->pose_span ? (uint32_t)w->objs.size() : 0u;
  k.pose_first = w->pose_first;
  k.pose_span = w->pose_span;
  k.stats = w->stats;
  if (w->surf_on) k.surf = SurfK{w->tex, w->uv, w->mext, w->lut, w->tex_log2, w->tex_mips, w->tex_layer_words,
                                 w->tex_layers};
  if (w->dyn_hdr) {
    k.dyn.hdr = w->dyn_hdr;
    k.dyn.ptr = w->dyn_ptr;
    k.dyn.bits = w->dyn_bits;
    k.dyn.meta = w->dyn_meta;
    k.dyn.mask = w->dyn_mask;
    k.dyn.data = w->dyn_data;
    k.dyn.cap = w->dyn_cap;
    k.dyn.nw = (uint32_t)(w->NL * w->per / 32);
    k.dyn.first_tri = std::min(w->dyn_first, w->ntris);
    k.dyn.end_tri = w->ntris;
  }
  return k;
}

TraceK trace_args(const rtgi_world* w) {
  TraceK t{};
  t.vx = vox_args(w);
  t.N = w->N;
  for (int c = 0; c < w->N; c++) {
    t.c[c].n = w->n[c];
    t.c[c].res = w->res[c];
    t.c[c].vox_level = w->vlev[c];
    t.c[c].b0 = w->b0[c];
    t.c[c].b1 = w->b1[c];
    t.c[c].ray_off = w->ray_off[c];
    t.c[c].np = (uint32_t)w->n[c] * w->n[c] * w->n[c];
    t.c[c].dir_off = w->dir_off[c];
    t.c[c].avg_off = w->avg_off[c];
    int nl = 0;
    while ((1 << nl) < w->n[c]) nl++;
    t.c[c].n_log2 = (1 << nl) == w->n[c] ? nl : -1;
    int rl = 0;
    while ((1 << rl) < w->res[c]) rl++;
    t.c[c].res_log2 = rl;
    t.c[c].tile_log2 = probe_tile_log2(w, c);
    t.c[c].probe_off = w->probe_off[c];
    t.c[c].rays = w->ray_off[c + 1] - w->ray_off[c];
  }
  for (int a = 0; a < 3; a++) t.G0[a] = w->G0[a];
  t.s0 = (float)w->rp.s0;
  t.vox_skip = (float)w->rp.vox_skip;
  t.vox_offset = (float)w->rp.vox_offset;
  for (int a = 0; a < 3; a++) {
    t.sun_dir[a] = w->frame.sun_dir[a];
    t.sun_e[a] = w->frame.sun_e[a];
    t.sky[a] = w->frame.sky[a];
  }
  t.sun_vis = w->frame.sun_vis;
  t.n_point = w->frame.n_point;
  std::memcpy(t.point_pos, w->frame.point_pos, sizeof(t.point_pos));
  std::memcpy(t.point_i, w->frame.point_i, sizeof(t.point_i));
  std::memcpy(t.omega0, w->omega0, sizeof(t.omega0));
  t.dirs = w->dirs;
  t.fp16 = (w->rp.flags & RTGI_RC_FP16) ? 1 : 0;
  t.bf16 = (w->rp.flags & RTGI_RC_BF16) ? 1 : 0;
  t.objects = ((w->rp.flags & RTGI_RC_OBJECTS) && (w->rp.flags & RTGI_RC_BOUNCE) && w->dyn_hdr &&
               w->dyn_first < w->ntris) ? 1 : 0;
  t.no_merge = (w->rp.flags & RTGI_RC_NO_MERGE) ? 1 : 0;
  t.reloc = (w->rp.flags & RTGI_RC_RELOCATE) ? 1 : 0;
  t.lift = (float)w->rp.lift;
  t.ppos = w->ppos;
  t.pnrm = w->pnrm;
  t.pvis = w->pvis;
  t.sunvis = w->sunvis;
  t.bounce = w->bounce;
  if (w->bounce) {
    t.bn_cur = w->bnvol[w->bn_par];
    t.bn_prev = w->bnvol[1 - w->bn_par];
    for (int a = 0; a < 3; a++) {
      t.bn_org[a] = w->bn_org[w->bn_par][a];
      t.bn_prev_org[a] = w->bn_org[1 - w->bn_par][a];
    }
    t.bn_frame = w->bn_frame;
    // levelOf (bounce.ts:110): round(log2(max(size, 1e-4) / v0)) clamped to the levels.
    double lv = std::nearbyint(std::log2((double)(BN_CELL / (float)w->vp.v0)));
    t.bn_level = (int32_t)std::min<double>(std::max<double>(lv, 0.0), (double)(w->NL - 1));
  }
  t.bn_dir_off = w->dir_off[w->N];
  t.total_probes = w->probe_off[w->N];
  t.rad = w->rad;
  t.avg = w->avg;
  t.hits = w->hits;
  t.sh = w->shb[w->sh_par];
  t.sh_prev = w->hist_prev ? w->shb[1 - w->sh_par] : nullptr;
  for (int a = 0; a < 3; a++) t.hist_d[a] = w->G0[a] - w->hist_G0[a];
  t.sh_fuse = 0;
  t.total_rays = w->ray_off[w->N];
  return t;
}

namespace {

}  // namespace

int probe_tile_log2(const rtgi_world* w, int c) {
  int T = 0;
  while ((1 << (T + 1)) <= w->data_tile) T++;
  int n = w->n[c], nl = 0;
  while ((1 << nl) < n) nl++;
  if ((1 << nl) != n || nl < T) return 0;  // rows when n is not a power of two or is smaller than the block
  return T;
}

uint32_t probe_index_host(const rtgi_world* w, int c, uint32_t x, uint32_t y, uint32_t z) {
  uint32_t n = (uint32_t)w->n[c];
  int T = probe_tile_log2(w, c);
  if (T == 0) return x + n * (y + n * z);
  int nl = 0;
  while ((1u << nl) < n) nl++;
  int nt = nl - T;
  uint32_t mT = (1u << T) - 1u;
  uint32_t t = (x >> T) | ((y >> T) << nt) | ((z >> T) << (2 * nt));
  return (t << (3 * T)) | (x & mT) | ((y & mT) << T) | ((z & mT) << (2 * T));
}

namespace {

void dyn_free(rtgi_world* w) {
  const size_t NLper = (size_t)w->NL * w->per, cap = (size_t)w->dyn_cap;
  dfree(w, w->dyn_hdr, DYN_HDR);
  dfree(w, w->dyn_ptr, NLper);
  dfree(w, w->dyn_bits, 2 * NLper / 32);
  dfree(w, w->dyn_meta, cap);
  dfree(w, w->dyn_mask, cap * 16);
  dfree(w, w->dyn_data, cap * 512 * DYN_WORDS);
}

// The world's fixed buffers live in one device allocation (the arena): one hipMalloc and one hipFree per world,
// each buffer 256-byte aligned. carve_fixed lays them out in a fixed order; with base == null it only adds up the
// bytes (the estimate and the arena's size).
constexpr size_t ARENA_ALIGN = 256;
struct Carver {
  char* base = nullptr;
  size_t off = 0;
};
template <class T>
void carve(Carver& cv, T** p, size_t count, bool on = true) {
  if (!on) return;
  size_t b = (std::max<size_t>(count, 1) * sizeof(T) + ARENA_ALIGN - 1) / ARENA_ALIGN * ARENA_ALIGN;
  if (cv.base) *p = reinterpret_cast<T*>(cv.base + cv.off);
  cv.off += b;
}
void carve_fixed(rtgi_world* w, Carver& cv) {
  const rtgi_rc_params* rp = &w->rp;
  const int Nc = w->N;
  size_t NLper = (size_t)w->NL * w->per, mb = (size_t)w->vp.max_bricks, mv = (size_t)w->vp.max_voxels;
  size_t np = (size_t)w->probe_off[Nc];
  carve(cv, &w->hdr, HDR_WORDS);
  carve(cv, &w->ptr, NLper);
  carve(cv, &w->occ, mb * 16);
  carve(cv, &w->battr, mb);
  carve(cv, &w->wslot, mb * 16);
  carve(cv, &w->bcode, mb);
  carve(cv, &w->rep, mb * 512);
  carve(cv, &w->attr, mv * 2);
  carve(cv, &w->plane, mv);
  carve(cv, &w->rad, rad_floats(w));
  carve(cv, &w->hits, (size_t)w->ray_off[Nc], (rp->flags & RTGI_RC_RECORD_HITS) != 0);
  for (int i = 0; i < sh_buffers(rp); i++) carve(cv, &w->shb[i], sh_floats_of(w));
  carve(cv, &w->dirs, (size_t)w->dir_off[Nc] + BN_RAYS);
  carve(cv, &w->avg, (size_t)w->avg_off[Nc] * 4);
  carve(cv, &w->sunvis, mv, (rp->flags & RTGI_RC_SUN_TRACE) != 0);
  carve(cv, &w->bounce, mv * 4, (rp->flags & RTGI_RC_BOUNCE) != 0);
  for (int i = 0; i < 2; i++) carve(cv, &w->bnvol[i], BN_WORDS, (rp->flags & RTGI_RC_BOUNCE) != 0);
  carve(cv, &w->vslot, mv, (rp->flags & (RTGI_RC_SUN_TRACE | RTGI_RC_BOUNCE)) != 0);
  carve(cv, &w->ppos, np * 4, (rp->flags & RTGI_RC_RELOCATE) != 0);
  carve(cv, &w->pnrm, np * 4, (rp->flags & RTGI_RC_RELOCATE) != 0);
  carve(cv, &w->pvis, np, (rp->flags & RTGI_RC_RELOCATE) != 0);
  carve(cv, &w->stats, 3 * RTGI_MAX_CASCADES, (rp->flags & RTGI_RC_STATS) != 0);
  carve(cv, &w->alb16, mv * 2, (rp->flags & RTGI_RC_BF16) != 0);
}

void objects_free(rtgi_world* w) {
  dfree(w, w->rest, (size_t)w->pose_span * 6);
  dfree(w, w->d_pose, w->objs.size());
  for (int i = 0; i < 2; i++) {
    if (w->h_pose[i]) (void)hipHostFree(w->h_pose[i]);
    w->h_pose[i] = nullptr;
  }
  w->objs.clear();
  w->pose_first = w->pose_span = 0;
  w->pose_set = false;
}

void surface_free(rtgi_world* w) {
  dfree(w, w->tex, w->tex_words);
  dfree(w, w->uv, (size_t)w->uv_verts * 2);
  dfree(w, w->mext, (size_t)w->mext_n * 4);
  dfree(w, w->lut, 511);
  w->tex_words = 0;
  w->uv_verts = w->mext_n = 0;
  w->tex_log2 = w->tex_mips = w->tex_layer_words = w->tex_layers = 0;
  w->surf_on = false;
}

// The sRGB tables of the texture sampler: the decoded value of each 8-bit code (the sRGB transfer function,
// IEC 61966-2-1, evaluated in double and rounded to f32), then the 255 encode thresholds (code k + 1 starts at the
// decoded value of (k + 0.5) / 255: storing the nearest code in the encoded scale, as an sRGB render target does).
double srgb_decode(double c) { return c <= 0.04045 ? c / 12.92 : std::pow((c + 0.055) / 1.055, 2.4); }
void srgb_tables(float* lut) {
  for (int c = 0; c < 256; c++) lut[c] = (float)srgb_decode(c / 255.0);
  for (int k = 0; k < 255; k++) lut[256 + k] = (float)srgb_decode((k + 0.5) / 255.0);
}

void free_all(rtgi_world* w) {
  objects_free(w);
  surface_free(w);
  dfree(w, w->big_vox, (size_t)w->NL * big_cap_for(w->cap_tris));
  dfree(w, w->big_mark, (size_t)w->NL * big_cap_for(w->cap_tris));
  dfree(w, w->tri, (size_t)w->cap_tris * 4);
  dfree(w, w->pos, (size_t)w->cap_verts * 3);
  dfree(w, w->nrm, (size_t)w->cap_verts * 3);
  dfree(w, w->mat, (size_t)w->cap_mat * 4);
  dyn_free(w);
  if (w->arena) {
    (void)hipFree(w->arena);
    w->bytes -= w->arena_bytes;
    w->arena = nullptr;
  }
}

}  // namespace

void sh_advance(rtgi_world* w) {
  if (w->rp.flags & RTGI_RC_NO_MERGE) return;
  const bool hist = sh_buffers(&w->rp) == 2;
  w->hist_prev = hist && w->sh_written;
  if (w->hist_prev) {
    w->sh_par = 1 - w->sh_par;
    for (int a = 0; a < 3; a++) w->hist_G0[a] = w->sh_G0[a];
  }
  for (int a = 0; a < 3; a++) w->sh_G0[a] = w->G0[a];
  w->sh_written = true;
}

void bounce_advance(rtgi_world* w) {
  if (!w->bounce) return;
  w->bn_par = 1 - w->bn_par;
  w->bn_frame++;
  // bnOriginNow (bounce.ts:103): floor(eye / BN_CELL) - dims / 2, eye as the f32 Frame uniform holds it.
  const int dims[3] = {BN_DX, BN_DY, BN_DZ};
  for (int a = 0; a < 3; a++) w->bn_org[w->bn_par][a] = (int)std::floor((float)w->frame.eye[a] / BN_CELL) - dims[a] / 2;
}

hipError_t world_wait(rtgi_world* w) {
  hipError_t e = hipStreamSynchronize(w->stream);
  if (e == hipSuccess && w->ext_done) e = hipEventSynchronize(w->ext_done);
  return e;
}

namespace {

int stage_launch(rtgi_world* w, uint32_t stage, hipStream_t s) {
  int e = 0;
  if (stage == RTGI_STAGE_VOX) {
    e = launch_pose(vox_args(w), s);
    if (!e) e = launch_vox_build(vox_args(w), s);
    if (!e && w->dyn_hdr) e = launch_dyn_build(vox_args(w), s);
    if (!e && w->dyn_hdr) e = launch_dyn_splat(trace_args(w), s);
  } else if (stage == RTGI_STAGE_TRACE && w->tuning.fused == 1) {
    // Fused launches also project the c0 SH in the trace stage (the SH stage is then empty).
    trace_begin(w);
    bounce_advance(w);
    sh_advance(w);
    TraceK t = trace_args(w);
    t.sh_fuse = 1;
    e = launch_sun(t, s);
    if (!e) e = launch_dyn_inject(t, s);
    if (!e) e = launch_bounce(t, s);
    if (!e) e = launch_relocate(t, s);
    for (int c = w->N - 1; c >= 0 && e == 0; c--) e = launch_trace_merge(t, c, s);
  } else if (stage == RTGI_STAGE_TRACE) {
    trace_begin(w);
    bounce_advance(w);
    TraceK t = trace_args(w);
    e = launch_sun(t, s);
    if (!e) e = launch_dyn_inject(t, s);
    if (!e) e = launch_bounce(t, s);
    if (!e) e = launch_relocate(t, s);
    if (!e) e = launch_trace(t, s);
  } else if (stage == RTGI_STAGE_MERGE) {
    if (!(w->rp.flags & RTGI_RC_NO_MERGE) && w->tuning.fused != 1) {
      TraceK t = trace_args(w);
      for (int c = w->N - 1; c >= 0 && e == 0; c--) e = launch_merge(t, c, s);
    }
  } else if (stage == RTGI_STAGE_SH) {
    if (!(w->rp.flags & RTGI_RC_NO_MERGE) && w->tuning.fused != 1) {
      sh_advance(w);
      e = launch_sh(trace_args(w), s);
    }
  }
  if (e != 0) return rtgi_fail(RTGI_E_HIP, "kernel launch failed: %s", hipGetErrorString((hipError_t)e));
  return RTGI_OK;
}

}  // namespace

// dirOf (src/core/rc.ts:291-298, 308) and the slab-safe reciprocal of traceVox (src/core/voxel.ts:128-130) in f32,
// the same operation sequence as the device evaluates (this file builds with -ffp-contract=off; SSE division and
// square root are IEEE-correct), so the table is bit-identical to per-ray evaluation.
void dir_entry(int di, int res, DirEntry* out) {
  float u = ((float)(di % res) + 0.5f) / (float)res, v = ((float)(di / res) + 0.5f) / (float)res;
  float fx = u * 2.0f - 1.0f, fy = v * 2.0f - 1.0f;
  float n[3] = {fx, fy, 1.0f - std::fabs(fx) - std::fabs(fy)};
  float t = std::max(-n[2], 0.0f);
  n[0] += n[0] >= 0.0f ? -t : t;
  n[1] += n[1] >= 0.0f ? -t : t;
  float l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  for (int a = 0; a < 3; a++) {
    out->d[a] = n[a] / l;
    float dd = std::fabs(out->d[a]) < 1e-9f ? 1e-9f : out->d[a];
    out->inv[a] = 1.0f / dd;
  }
  out->d[3] = out->inv[3] = 0.0f;
}

void bn_dir_entry(int i, int n, DirEntry* out) {
  double z = 1.0 - (2.0 * i + 1.0) / n;
  double r = std::sqrt(std::max(0.0, 1.0 - z * z));
  double f = i * 0.618033988749;
  double phi = 2.0 * 3.14159265 * (f - std::floor(f));
  float d[3] = {(float)(r * std::cos(phi)), (float)(r * std::sin(phi)), (float)z};
  for (int a = 0; a < 3; a++) {
    out->d[a] = d[a];
    float dd = std::fabs(d[a]) < 1e-9f ? 1e-9f : d[a];
    out->inv[a] = 1.0f / dd;
  }
  out->d[3] = out->inv[3] = 0.0f;
}

extern "C" const char* rtgi_last_error(void) { return g_err; }

extern "C" const char* rtgi_version(void) { return "rtgi-mi355x 0.1 (gfx950)"; }

extern "C" int rtgi_device_count(int32_t* n) {
  if (!n) return rtgi_fail(RTGI_E_INVALID, "rtgi_device_count: null argument");
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
  *n = c;
  return RTGI_OK;
}

extern "C" int rtgi_oct_solid_angles(int32_t res, int32_t samples, float* out) {
  if (res <= 0 || samples <= 0 || !out) return rtgi_fail(RTGI_E_INVALID, "rtgi_oct_solid_angles: bad argument");
  std::vector<double> counts((size_t)res * res, 0.0);
  const double golden = M_PI * (3 - std::sqrt(5.0));
  for (int i = 0; i < samples; i++) {
    double z = 1 - (2.0 * i + 1) / samples, r = std::sqrt(1 - z * z), phi = golden * i;
    double u, v;
    oct_encode_d(r * std::cos(phi), r * std::sin(phi), z, u, v);
    int iu = std::min(res - 1, (int)std::floor(u * res)), iv = std::min(res - 1, (int)std::floor(v * res));
    counts[(size_t)iv * res + iu] += 1;
  }
  for (size_t i = 0; i < counts.size(); i++) out[i] = (float)(counts[i] / samples * 4 * M_PI);
  return RTGI_OK;
}

namespace {

// Checks the parameters and fills the world's layout (levels, cascades, ray/direction/average offsets).
int world_layout(rtgi_world* w, const rtgi_vox_params* vp, const rtgi_rc_params* rp) {
  if (vp->levels < 1 || vp->levels > RTGI_MAX_LEVELS) return rtgi_fail(RTGI_E_INVALID, "at most 8 levels");
  if (vp->G < 1 || vp->G > 64 || ((int64_t)vp->G * vp->G * vp->G) % 64 != 0)
    return rtgi_fail(RTGI_E_INVALID, "G must be in 1..64 with G^3 a multiple of 64 (got %d)", vp->G);
  if (!(vp->v0 > 0)) return rtgi_fail(RTGI_E_INVALID, "v0 must be positive");
  if (vp->max_bricks < 1 || vp->max_voxels < 1) return rtgi_fail(RTGI_E_INVALID, "pool sizes must be positive");
  if ((int64_t)vp->max_bricks * 512 > 0xffffffffLL || (int64_t)vp->max_voxels * 2 > 0xffffffffLL)
    return rtgi_fail(RTGI_E_INVALID, "pool sizes overflow 32-bit indices");
  if (rp->cascades < 1 || rp->cascades > RTGI_MAX_CASCADES) return rtgi_fail(RTGI_E_INVALID, "at most 8 cascades");
  int Nc = rp->cascades;
  if (rp->probes0 < 1 || rp->probes0 % (1 << (Nc - 1)) != 0)
    return rtgi_fail(RTGI_E_INVALID, "probes0 (%d) must be a positive multiple of 2^(cascades-1)", rp->probes0);
  if (!(rp->s0 > 0)) return rtgi_fail(RTGI_E_INVALID, "s0 must be positive");
  if (rp->res_cap < 0) return rtgi_fail(RTGI_E_INVALID, "res_cap must be >= 0");
  if ((rp->flags & RTGI_RC_RELOCATE) && !(rp->lift >= 0)) return rtgi_fail(RTGI_E_INVALID, "lift must be >= 0");
  w->vp = *vp;
  w->rp = *rp;
  w->NL = vp->levels;
  w->G = vp->G;
  w->E = vp->G * 8;
  w->per = vp->G * vp->G * vp->G;
  w->N = Nc;
  uint64_t off = 0;
  for (int c = 0; c < Nc; c++) {
    w->n[c] = rp->probes0 >> c;
    w->res[c] = 4 << std::min(c, rp->res_cap);
    auto b = [&](int k) { return (float)(rp->k_interval * (std::pow(4.0, k) - 1) / 3); };
    w->b0[c] = b(c);
    w->b1[c] = b(c + 1);
    // voxLevel (src/core/rc.ts:210-213), evaluated in double once per cascade; round() is half-to-even as in WGSL.
    double v = rp->vox_scale * rp->s0 * std::exp2(rp->level_map * c);
    double L = std::nearbyint(std::log2(v / vp->v0));
    w->vlev[c] = (int)std::min(std::max(L, 0.0), (double)(vp->levels - 1));
    w->ray_off[c] = off;
    const uint64_t rays_c = (uint64_t)w->n[c] * w->n[c] * w->n[c] * w->res[c] * w->res[c];
    if (rays_c >= (1ull << 32))  // a cascade's rays are indexed in 32 bits
      return rtgi_fail(RTGI_E_INVALID, "cascade %d has %llu rays (n^3 res^2 must stay below 2^32)", c,
                       (unsigned long long)rays_c);
    off += rays_c;
    w->dir_off[c + 1] = w->dir_off[c] + (uint32_t)(w->res[c] * w->res[c]);
    uint64_t npc = (uint64_t)w->n[c] * w->n[c] * w->n[c];
    // The merge reads a cascade's averages through a buffer resource with 32-bit byte offsets.
    if (c > 0 && npc * w->res[c - 1] * w->res[c - 1] * 16u >= (1ull << 32))
      return rtgi_fail(RTGI_E_INVALID, "cascade %d's 2x2 averages exceed 4 GiB (n^3 res_{c-1}^2 x 16 B)", c);
    w->avg_off[c + 1] = w->avg_off[c] + (c > 0 ? npc * w->res[c - 1] * w->res[c - 1] : 0);
    w->probe_off[c + 1] = w->probe_off[c] + npc;
  }
  w->ray_off[Nc] = off;
  return RTGI_OK;
}

template <class T>
uint64_t nbytes(size_t count) {
  return (uint64_t)std::max<size_t>(count, 1) * sizeof(T);
}

}  // namespace

uint64_t world_bytes(const rtgi_vox_params* vp, const rtgi_rc_params* rp, uint32_t nverts, uint32_t ntris,
                     uint32_t nmat) {
  rtgi_world w;
  if (world_layout(&w, vp, rp)) return 0;
  Carver cv;
  carve_fixed(&w, cv);
  uint64_t b = cv.off;
  if (ntris) b += nbytes<uint32_t>((size_t)ntris * 4) + 2 * nbytes<uint32_t>((size_t)w.NL * big_cap_for(ntris));
  if (nverts) b += 2 * nbytes<float>((size_t)nverts * 3);
  if (nmat) b += nbytes<float>((size_t)nmat * 4);
  return b;
}

extern "C" int rtgi_world_estimate_bytes(const rtgi_vox_params* vp, const rtgi_rc_params* rp, uint32_t nverts,
                                         uint32_t ntris, uint32_t nmat, uint64_t* bytes) {
  if (!vp || !rp || !bytes) return rtgi_fail(RTGI_E_INVALID, "rtgi_world_estimate_bytes: null argument");
  rtgi_world w;
  int e = world_layout(&w, vp, rp);
  if (e) return e;
  *bytes = world_bytes(vp, rp, nverts, ntris, nmat);
  return RTGI_OK;
}

extern "C" int rtgi_world_create(const rtgi_vox_params* vp, const rtgi_rc_params* rp, rtgi_world** out) {
  if (!vp || !rp || !out) return rtgi_fail(RTGI_E_INVALID, "rtgi_world_create: null argument");
  *out = nullptr;
  {
    rtgi_world probe;
    int e = world_layout(&probe, vp, rp);
    if (e) return e;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return rtgi_fail(RTGI_E_NODEVICE, "no HIP device visible");

  rtgi_world* w = new (std::nothrow) rtgi_world();
  if (!w) return rtgi_fail(RTGI_E_NOMEM, "out of host memory");
  (void)hipGetDevice(&w->device);
  world_layout(w, vp, rp);
  int Nc = w->N;
  {
    // octSolidAngles(4) is a constant (a 2^20-point count, about 50 ms of host time): computed once per process.
    static const std::vector<float> omega0 = [] {
      std::vector<float> o(16);
      rtgi_oct_solid_angles(4, 1 << 20, o.data());
      return o;
    }();
    std::memcpy(w->omega0, omega0.data(), sizeof(w->omega0));
  }

  int e = RTGI_OK;
  {
    Carver cv;
    carve_fixed(w, cv);
    hipError_t he = hipMalloc((void**)&w->arena, cv.off);
    if (he != hipSuccess) {
      w->arena = nullptr;
      e = rtgi_fail(RTGI_E_NOMEM, "hipMalloc(%zu bytes): %s", cv.off, hipGetErrorString(he));
    } else {
      w->arena_bytes = cv.off;
      w->bytes += cv.off;
      cv = Carver{};
      cv.base = w->arena;
      carve_fixed(w, cv);
    }
  }
  if (!e && w->stats && hipMemset(w->stats, 0, 3 * RTGI_MAX_CASCADES * 8) != hipSuccess)
    e = rtgi_fail(RTGI_E_HIP, "hipMemset");
  if (!e) {
    std::vector<DirEntry> tab((size_t)w->dir_off[Nc] + BN_RAYS);
    for (int c = 0; c < Nc; c++)
      for (int di = 0; di < w->res[c] * w->res[c]; di++) dir_entry(di, w->res[c], &tab[w->dir_off[c] + di]);
    for (uint32_t i = 0; i < BN_RAYS; i++) bn_dir_entry((int)i, (int)BN_RAYS, &tab[w->dir_off[Nc] + i]);
    if (hipMemcpy(w->dirs, tab.data(), tab.size() * sizeof(DirEntry), hipMemcpyHostToDevice) != hipSuccess)
      e = rtgi_fail(RTGI_E_HIP, "direction table upload failed");
  }
  if (!e && hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking) != hipSuccess)
    e = rtgi_fail(RTGI_E_HIP, "hipStreamCreate failed");
  for (int i = 0; i < 5 && !e; i++)
    if (hipEventCreate(&w->ev[i]) != hipSuccess) e = rtgi_fail(RTGI_E_HIP, "hipEventCreate failed");
  if (!e && hipMemsetAsync(w->hdr, 0, HDR_WORDS * 4, w->stream) != hipSuccess)
    e = rtgi_fail(RTGI_E_HIP, "hipMemsetAsync failed");
  for (int i = 0; i < 2 && !e && w->bnvol[i]; i++)  // no volume before the first update: every probe inactive
    if (hipMemsetAsync(w->bnvol[i], 0, (size_t)BN_WORDS * 4, w->stream) != hipSuccess)
      e = rtgi_fail(RTGI_E_HIP, "hipMemsetAsync failed");
  if (!e && hipStreamSynchronize(w->stream) != hipSuccess) e = rtgi_fail(RTGI_E_HIP, "hipStreamSynchronize failed");
  if (e) {
    rtgi_world_destroy(w);
    return e;
  }
  *out = w;
  return RTGI_OK;
}

extern "C" int rtgi_world_upload_scene(rtgi_world* w, const float* pos, const float* nrm, uint32_t nverts,
                                       const uint32_t* tri, uint32_t ntris, const float* materials, uint32_t nmat) {
  if (!w) return rtgi_fail(RTGI_E_INVALID, "null world");
  if ((nverts && (!pos || !nrm)) || (ntris && !tri) || (nmat && !materials))
    return rtgi_fail(RTGI_E_INVALID, "rtgi_world_upload_scene: null array");
  for (uint32_t t = 0; t < ntris; t++) {
    if (tri[4 * t] >= nverts || tri[4 * t + 1] >= nverts || tri[4 * t + 2] >= nverts)
      return rtgi_fail(RTGI_E_INVALID, "triangle %u indexes a vertex out of range", t);
    if (tri[4 * t + 3] >= nmat) return rtgi_fail(RTGI_E_INVALID, "triangle %u indexes a material out of range", t);
  }
  if ((uint64_t)ntris >= 0xffffffffull) return rtgi_fail(RTGI_E_INVALID, "too many triangles");
  HIPCHK(hipSetDevice(w->device));
  HIPCHK(world_wait(w));
  objects_free(w);  // the objects and the surface belonged to the previous scene
  surface_free(w);
  int e = RTGI_OK;
  if (ntris > w->cap_tris) {
    dfree(w, w->tri, (size_t)w->cap_tris * 4);
    dfree(w, w->big_vox, (size_t)w->NL * big_cap_for(w->cap_tris));
    dfree(w, w->big_mark, (size_t)w->NL * big_cap_for(w->cap_tris));
    w->cap_tris = 0;
    if (!e) e = dalloc(w, &w->tri, (size_t)ntris * 4);
    if (!e) e = dalloc(w, &w->big_vox, (size_t)w->NL * big_cap_for(ntris));
    if (!e) e = dalloc(w, &w->big_mark, (size_t)w->NL * big_cap_for(ntris));
    if (e) return e;
    w->cap_tris = ntris;
  }
  if (nverts > w->cap_verts) {
    dfree(w, w->pos, (size_t)w->cap_verts * 3);
    dfree(w, w->nrm, (size_t)w->cap_verts * 3);
    w->cap_verts = 0;
    if (!e) e = dalloc(w, &w->pos, (size_t)nverts * 3);
    if (!e) e = dalloc(w, &w->nrm, (size_t)nverts * 3);
    if (e) return e;
    w->cap_verts = nverts;
  }
  if (nmat > w->cap_mat) {
    dfree(w, w->mat, (size_t)w->cap_mat * 4);
    w->cap_mat = 0;
    if ((e = dalloc(w, &w->mat, (size_t)nmat * 4))) return e;
    w->cap_mat = nmat;
  }
  if (ntris) HIPCHK(hipMemcpy(w->tri, tri, (size_t)ntris * 16, hipMemcpyHostToDevice));
  if (nverts) {
    HIPCHK(hipMemcpy(w->pos, pos, (size_t)nverts * 12, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(w->nrm, nrm, (size_t)nverts * 12, hipMemcpyHostToDevice));
  }
  if (nmat) HIPCHK(hipMemcpy(w->mat, materials, (size_t)nmat * 16, hipMemcpyHostToDevice));
  w->ntris = ntris;
  w->nverts = nverts;
  w->nmat = nmat;
  w->scene_set = true;
  return RTGI_OK;
}

extern "C" int rtgi_world_update_vertices(rtgi_world* w, uint32_t first, uint32_t count, const float* pos,
                                          const float* nrm) {
  if (!w) return rtgi_fail(RTGI_E_INVALID, "null world");
  if (!w->scene_set) return rtgi_fail(RTGI_E_STATE, "no scene uploaded");
  if ((uint64_t)first + count > w->nverts) return rtgi_fail(RTGI_E_INVALID, "vertex range out of bounds");
  if (count && (!pos || !nrm)) return rtgi_fail(RTGI_E_INVALID, "null array");
  HIPCHK(hipSetDevice(w->device));
  HIPCHK(world_wait(w));
  if (count) {
    HIPCHK(hipMemcpy(w->pos + (size_t)first * 3, pos, (size_t)count * 12, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(w->nrm + (size_t)first * 3, nrm, (size_t)count * 12, hipMemcpyHostToDevice));
  }
  return RTGI_OK;
}

extern "C" int rtgi_world_set_surface(rtgi_world* w, const float* uvs, uint32_t nverts,
                                      const rtgi_surface_material* mats, uint32_t nmat, const uint8_t* texels,
                                      uint32_t size, uint32_t layers) {
  if (!w) return rtgi_fail(RTGI_E_INVALID, "null world");
  if (!w->scene_set) return rtgi_fail(RTGI_E_STATE, "no scene uploaded");
  HIPCHK(hipSetDevice(w->device));
  HIPCHK(world_wait(w));
  if (!uvs) {  // untextured again
    surface_free(w);
    return RTGI_OK;
  }
  if (nverts != w->nverts) return rtgi_fail(RTGI_E_INVALID, "uvs: %u vertices, the scene has %u", nverts, w->nverts);
  if (nmat != w->nmat) return rtgi_fail(RTGI_E_INVALID, "%u surface materials, the scene has %u", nmat, w->nmat);
  if (nmat && !mats) return rtgi_fail(RTGI_E_INVALID, "null materials");
  if (size == 0 || (size & (size - 1)) != 0 || size > (1u << SURF_MAX_LOG2))
    return rtgi_fail(RTGI_E_INVALID, "texture size must be a power of two in 1..%u (got %u)", 1u << SURF_MAX_LOG2, size);
  if (layers && !texels) return rtgi_fail(RTGI_E_INVALID, "null texels");
  for (uint32_t m = 0; m < nmat; m++)
    if (mats[m].base_color_layer > layers)
      return rtgi_fail(RTGI_E_INVALID, "material %u samples layer %u of %u", m, mats[m].base_color_layer, layers);
  uint32_t lg = 0;
  while ((1u << lg) < size) lg++;
  uint64_t per = 0;
  for (uint32_t m = 0; m <= lg; m++) per += (uint64_t)1 << (2 * (lg - m));
  if (per * layers >= (1ull << 32)) return rtgi_fail(RTGI_E_INVALID, "texture array too large (%u layers)", layers);
  surface_free(w);
  int e = dalloc(w, &w->uv, (size_t)nverts * 2);
  if (!e) e = dalloc(w, &w->mext, (size_t)nmat * 4);
  if (!e) e = dalloc(w, &w->lut, 511);
  if (!e) e = dalloc(w, &w->tex, (size_t)(per * layers));
  if (e) {
    surface_free(w);
    return e;
  }
  w->uv_verts = nverts;
  w->mext_n = nmat;
  w->tex_words = (size_t)(per * layers);
  w->tex_log2 = lg;
  w->tex_mips = lg + 1;
  w->tex_layer_words = (uint32_t)per;
  w->tex_layers = layers;
  std::vector<uint32_t> me((size_t)nmat * 4, 0u);
  for (uint32_t m = 0; m < nmat; m++) {
    me[m * 4] = mats[m].base_color_layer;
    me[m * 4 + 1] = mats[m].flags;
    std::memcpy(&me[m * 4 + 2], &mats[m].alpha_cutoff, 4);
  }
  float lut[511];
  srgb_tables(lut);
  HIPCHK(hipMemcpy(w->uv, uvs, (size_t)nverts * 8, hipMemcpyHostToDevice));
  if (nmat) HIPCHK(hipMemcpy(w->mext, me.data(), me.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(w->lut, lut, sizeof(lut), hipMemcpyHostToDevice));
  for (uint32_t l = 0; l < layers; l++)
    HIPCHK(hipMemcpy(w->tex + (size_t)l * per, texels + (size_t)l * size * size * 4, (size_t)size * size * 4,
                     hipMemcpyHostToDevice));
  w->surf_on = true;
  VoxK k = vox_args(w);
  HIPCHK((hipError_t)launch_surf_mips(k.surf, w->tex, w->stream));
  HIPCHK(hipStreamSynchronize(w->stream));
  return RTGI_OK;
}

extern "C" int rtgi_world_set_objects(rtgi_world* w, const rtgi_object* objects, uint32_t n, const float* rest_pos,
                                      const float* rest_nrm) {
  if (!w) return rtgi_fail(RTGI_E_INVALID, "null world");
  if (n && (!objects || !rest_pos || !rest_nrm)) return rtgi_fail(RTGI_E_INVALID, "rtgi_world_set_objects: null array");
  if (n && !w->scene_set) return rtgi_fail(RTGI_E_STATE, "no scene uploaded");
  uint64_t end = 0;
  for (uint32_t i = 0; i < n; i++) {
    const rtgi_object& o = objects[i];
    if (i && o.first_vertex < end) return rtgi_fail(RTGI_E_INVALID, "object %u overlaps or precedes object %u", i, i - 1);
    end = (uint64_t)o.first_vertex + o.vertex_count;
    if (end > w->nverts) return rtgi_fail(RTGI_E_INVALID, "object %u: vertices out of range", i);
    for (int a = 0; a < 3; a++)
      if (!std::isfinite(o.pivot[a])) return rtgi_fail(RTGI_E_INVALID, "object %u: pivot not finite", i);
  }
  HIPCHK(hipSetDevice(w->device));
  HIPCHK(world_wait(w));
  objects_free(w);
  if (!n) return RTGI_OK;
  const uint32_t first = objects[0].first_vertex, span = (uint32_t)(end - first);
  int e = dalloc(w, &w->rest, (size_t)span * 6);
  if (!e) e = dalloc(w, &w->d_pose, n);
  for (int i = 0; i < 2 && !e; i++) {
    if (hipHostMalloc((void**)&w->h_pose[i], sizeof(PoseRec) * n, hipHostMallocDefault) != hipSuccess)
      e = rtgi_fail(RTGI_E_NOMEM, "hipHostMalloc");
    if (!e && !w->pose_copied[i] && hipEventCreateWithFlags(&w->pose_copied[i], hipEventDisableTiming) != hipSuccess)
      e = rtgi_fail(RTGI_E_HIP, "hipEventCreate");
  }
  w->objs.assign(objects, objects + n);
  w->pose_first = first;
  w->pose_span = span;
  if (!e && span) {
    if (hipMemcpy(w->rest, rest_pos, (size_t)span * 12, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(w->rest + (size_t)span * 3, rest_nrm, (size_t)span * 12, hipMemcpyHostToDevice) != hipSuccess)
      e = rtgi_fail(RTGI_E_HIP, "rest mesh upload failed");
  }
  if (e) objects_free(w);
  return e;
}

extern "C" int rtgi_world_set_pose(rtgi_world* w, const rtgi_pose* poses) {
  if (!w) return rtgi_fail(RTGI_E_INVALID, "null world");
  if (w->objs.empty()) return rtgi_fail(RTGI_E_STATE, "no objects registered");
  if (!poses) return rtgi_fail(RTGI_E_INVALID, "null poses");
  const size_t n = w->objs.size();
  for (size_t i = 0; i < n; i++) {
    const rtgi_pose& q = poses[i];
    if (!std::isfinite(q.c) || !std::isfinite(q.s) || !std::isfinite(q.offset[0]) || !std::isfinite(q.offset[1]) ||
        !std::isfinite(q.offset[2]))
      return rtgi_fail(RTGI_E_INVALID, "pose %zu not finite", i);
  }
  HIPCHK(hipSetDevice(w->device));
  HIPCHK(hipEventSynchronize(w->pose_copied[w->pose_cur]));
  PoseRec* h = w->h_pose[w->pose_cur];
  for (size_t i = 0; i < n; i++) {
    const rtgi_object& o = w->objs[i];
    for (int a = 0; a < 3; a++) {
      h[i].pivot[a] = o.pivot[a];
      h[i].off[a] = poses[i].offset[a];
    }
    h[i].c = poses[i].c;
    h[i].s = poses[i].s;
    h[i].first = o.first_vertex;
    h[i].count = o.vertex_count;
  }
  // A batched update of the world may still read the records being replaced.
  if (w->ext_done) HIPCHK(hipStreamWaitEvent(w->stream, w->ext_done, 0));
  HIPCHK(hipMemcpyAsync(w->d_pose, h, sizeof(PoseRec) * n, hipMemcpyHostToDevice, w->stream));
  HIPCHK(hipEventRecord(w->pose_copied[w->pose_cur], w->stream));
  w->pose_cur ^= 1;
  w->pose_set = true;
  return RTGI_OK;
}

int frame_origins(const rtgi_vox_params* vp, const rtgi_rc_params* rp, const rtgi_frame* f,
                  int origin[RTGI_MAX_LEVELS][3], int G0[3]) {
  if (f->n_point < 0 || f->n_point > RTGI_MAX_POINT_LIGHTS) return rtgi_fail(RTGI_E_INVALID, "0..8 point lights");
  for (int a = 0; a < 3; a++)
    if (!std::isfinite(f->eye[a])) return rtgi_fail(RTGI_E_INVALID, "eye must be finite");
  // Clipmap windows centred on the eye: floor(eye / bs) - G/2 (writeUniform, src/core/voxel.ts:872-875).
  for (int L = 0; L < vp->levels; L++) {
    double bs = vp->v0 * std::exp2((double)L) * 8;
    for (int a = 0; a < 3; a++) {
      double o = std::floor(f->eye[a] / bs) - vp->G / 2;
      if (std::fabs(o) > 1e6) return rtgi_fail(RTGI_E_INVALID, "eye too far from the origin");
      origin[L][a] = (int)o;
    }
  }
  // Dense probe grid, snapped to the coarsest cascade's spacing so that every cascade's grid nests in the next.
  int top = rp->cascades - 1;
  double st = rp->s0 * std::exp2((double)top);
  int ntop = rp->probes0 >> top;
  for (int a = 0; a < 3; a++) {
    double g = std::floor(f->eye[a] / st) - ntop / 2;
    if (std::fabs(g) > 1e6) return rtgi_fail(RTGI_E_INVALID, "eye too far from the origin");
    G0[a] = (int)g * (1 << top);
  }
  return RTGI_OK;
}

extern "C" int rtgi_world_set_frame(rtgi_world* w, const rtgi_frame* f) {
  if (!w || !f) return rtgi_fail(RTGI_E_INVALID, "null argument");
  int origin[RTGI_MAX_LEVELS][3] = {}, G0[3];
  int e = frame_origins(&w->vp, &w->rp, f, origin, G0);
  if (e) return e;  // the world keeps its previous frame
  w->frame = *f;
  std::memcpy(w->origin, origin, sizeof(origin));
  for (int a = 0; a < 3; a++) w->G0[a] = G0[a];
  w->frame_set = true;
  return RTGI_OK;
}

extern "C" int rtgi_world_enqueue_stages(rtgi_world* w, uint32_t stages, void* stream) {
  if (!w) return rtgi_fail(RTGI_E_INVALID, "null world");
  if (!w->scene_set) return rtgi_fail(RTGI_E_STATE, "no scene uploaded");
  if (!w->frame_set) return rtgi_fail(RTGI_E_STATE, "no frame set");
  if (stages & ~RTGI_STAGE_ALL) return rtgi_fail(RTGI_E_INVALID, "unknown stage bits");
  HIPCHK(hipSetDevice(w->device));
  hipStream_t s = stream ? (hipStream_t)stream : w->stream;
  // A batched update that included this world may still be running; the last pose copy runs on the world's stream.
  if (w->ext_done) HIPCHK(hipStreamWaitEvent(s, w->ext_done, 0));
  if (s != w->stream && w->pose_set) HIPCHK(hipStreamWaitEvent(s, w->pose_copied[w->pose_cur ^ 1], 0));
  const uint32_t order[4] = {RTGI_STAGE_VOX, RTGI_STAGE_TRACE, RTGI_STAGE_MERGE, RTGI_STAGE_SH};
  if (w->timing) HIPCHK(hipEventRecord(w->ev[0], s));
  for (int i = 0; i < 4; i++) {
    if (stages & order[i]) {
      int e = stage_launch(w, order[i], s);
      if (e) return e;
    }
    if (w->timing) HIPCHK(hipEventRecord(w->ev[i + 1], s));
  }
  w->last_stages = stages;
  if (s != w->stream) {
    // Later calls on the world's own stream (readback, sync) must see this work.
    HIPCHK(hipEventRecord(w->ev[4], s));
    HIPCHK(hipStreamWaitEvent(w->stream, w->ev[4], 0));
  }
  return RTGI_OK;
}

extern "C" int rtgi_world_enqueue(rtgi_world* w, void* stream) {
  return rtgi_world_enqueue_stages(w, RTGI_STAGE_ALL, stream);
}

extern "C" int rtgi_world_set_timing(rtgi_world* w, int32_t on) {
  if (!w) return rtgi_fail(RTGI_E_INVALID, "null world");
  w->timing = on != 0;
  return RTGI_OK;
}

extern "C" int rtgi_world_stage_ms(rtgi_world* w, float ms[4]) {
  if (!w || !ms) return rtgi_fail(RTGI_E_INVALID, "null argument");
  if (!w->timing) return rtgi_fail(RTGI_E_STATE, "timing is off");
  HIPCHK(hipSetDevice(w->device));
  HIPCHK(hipEventSynchronize(w->ev[4]));
  for (int i = 0; i < 4; i++) HIPCHK(hipEventElapsedTime(&ms[i], w->ev[i], w->ev[i + 1]));
  return RTGI_OK;
}

extern "C" int rtgi_world_sync(rtgi_world* w) {
  if (!w) return rtgi_fail(RTGI_E_INVALID, "null world");
  HIPCHK(hipSetDevice(w->device));
  HIPCHK(world_wait(w));
  return RTGI_OK;
}

extern "C" int rtgi_world_read_counters(rtgi_world* w, rtgi_counters* out) {
  if (!w || !out) return rtgi_fail(RTGI_E_INVALID, "null argument");
  HIPCHK(hipSetDevice(w->device));
  HIPCHK(world_wait(w));
  uint32_t h[HDR_WORDS];
  HIPCHK(hipMemcpyAsync(h, w->hdr, sizeof(h), hipMemcpyDeviceToHost, w->stream));
  HIPCHK(hipStreamSynchronize(w->stream));
  std::memset(out, 0, sizeof(*out));
  out->marked = h[H_MARKED];
  out->bricks = std::min(h[H_MARKED], (uint32_t)w->vp.max_bricks);
  out->voxels = h[H_VOXELS];
  out->brick_overflow = h[H_BRICK_OVF];
  out->big_overflow = h[H_BIG_OVF];
  out->voxel_overflow = h[H_VOXEL_OVF];
  for (int L = 0; L < w->NL; L++) {
    out->bricks_per_level[L] = h[H_BRICKS_L + L];
    out->big_per_level[L] = h[H_BIGVOX_L + L];
  }
  return RTGI_OK;
}

extern "C" int rtgi_world_get_layout(const rtgi_world* w, rtgi_layout* out) {
  if (!w || !out) return rtgi_fail(RTGI_E_INVALID, "null argument");
  std::memset(out, 0, sizeof(*out));
  out->levels = w->NL;
  out->G = w->G;
  out->E = w->E;
  std::memcpy(out->window_origin, w->origin, sizeof(out->window_origin));
  out->cascades = w->N;
  for (int c = 0; c < w->N; c++) {
    out->probes[c] = w->n[c];
    out->res[c] = w->res[c];
    out->vox_level[c] = w->vlev[c];
    out->rays[c] = w->ray_off[c + 1] - w->ray_off[c];
    out->b0[c] = w->b0[c];
    out->b1[c] = w->b1[c];
  }
  out->total_rays = w->ray_off[w->N];
  for (int a = 0; a < 3; a++) out->grid_origin[a] = w->G0[a];
  std::memcpy(out->omega0, w->omega0, sizeof(out->omega0));
  out->device_bytes = w->bytes;
  return RTGI_OK;
}

extern "C" int rtgi_world_read_clipmap(rtgi_world* w, int32_t level, uint32_t* occ_bits, uint32_t* rep,
                                       uint32_t* albedo, uint32_t* normal, float* plane, uint8_t* brick_state) {
  if (!w) return rtgi_fail(RTGI_E_INVALID, "null world");
  if (level < 0 || level >= w->NL) return rtgi_fail(RTGI_E_INVALID, "level out of range");
  HIPCHK(hipSetDevice(w->device));
  HIPCHK(world_wait(w));
  uint32_t h[HDR_WORDS];
  HIPCHK(hipMemcpy(h, w->hdr, sizeof(h), hipMemcpyDeviceToHost));
  uint32_t nb = std::min(h[H_MARKED], (uint32_t)w->vp.max_bricks);
  uint32_t nvox = std::min(h[H_VOXELS], (uint32_t)w->vp.max_voxels);
  const int G = w->G, E = w->E, per = w->per;
  std::vector<uint32_t> ptr(per), occ((size_t)nb * 16), battr(nb), rp, at;
  std::vector<float> pl;
  HIPCHK(hipMemcpy(ptr.data(), w->ptr + (size_t)level * per, (size_t)per * 4, hipMemcpyDeviceToHost));
  if (nb) {
    HIPCHK(hipMemcpy(occ.data(), w->occ, (size_t)nb * 64, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(battr.data(), w->battr, (size_t)nb * 4, hipMemcpyDeviceToHost));
  }
  if (rep && nb) {
    rp.resize((size_t)nb * 512);
    HIPCHK(hipMemcpy(rp.data(), w->rep, (size_t)nb * 2048, hipMemcpyDeviceToHost));
  }
  if ((albedo || normal) && nvox) {
    at.resize((size_t)nvox * 2);
    HIPCHK(hipMemcpy(at.data(), w->attr, (size_t)nvox * 8, hipMemcpyDeviceToHost));
  }
  if (plane && nvox) {
    pl.resize(nvox);
    HIPCHK(hipMemcpy(pl.data(), w->plane, (size_t)nvox * 4, hipMemcpyDeviceToHost));
  }
  size_t ncell = (size_t)E * E * E;
  if (occ_bits) std::memset(occ_bits, 0, ncell / 8);
  if (rep) std::memset(rep, 0, ncell * 4);
  if (albedo) std::memset(albedo, 0, ncell * 4);
  if (normal) std::memset(normal, 0, ncell * 4);
  if (plane) std::memset(plane, 0, ncell * 4);
  for (int bz = 0; bz < G; bz++)
    for (int by = 0; by < G; by++)
      for (int bx = 0; bx < G; bx++) {
        size_t bi = ((size_t)bz * G + by) * G + bx;
        uint32_t p = ptr[bi];
        uint8_t state = p == PTR_DROP ? 2 : (p >= PTR_MARK ? 0 : 1);
        if (state == 1 && p >= nb) return rtgi_fail(RTGI_E_STATE, "corrupt pointer table (brick %u >= %u)", p, nb);
        if (state == 1 && battr[p] == ATTR_NONE) state = 3;
        if (brick_state) brick_state[bi] = state;
        if (state != 1 && state != 3) continue;
        uint32_t rank = 0;
        for (int bit = 0; bit < 512; bit++) {
          uint32_t word = occ[(size_t)p * 16 + (bit >> 5)];
          if (!(word & (1u << (bit & 31)))) continue;
          int x = bx * 8 + (bit & 7), y = by * 8 + ((bit >> 3) & 7), z = bz * 8 + (bit >> 6);
          size_t ci = (size_t)x + (size_t)E * ((size_t)y + (size_t)E * z);
          if (occ_bits) occ_bits[ci >> 5] |= 1u << (ci & 31);
          if (rep) rep[ci] = rp[(size_t)p * 512 + bit];
          if (state == 1) {
            uint32_t ai = battr[p] + rank;
            if (ai < nvox) {
              if (albedo) albedo[ci] = at[(size_t)ai * 2];
              if (normal) normal[ci] = at[(size_t)ai * 2 + 1];
              if (plane) plane[ci] = pl[ai];
            }
          }
          rank++;
        }
      }
  return RTGI_OK;
}

extern "C" int rtgi_world_read_cascade(rtgi_world* w, int32_t c, float* rgba, uint32_t* hits) {
  if (!w) return rtgi_fail(RTGI_E_INVALID, "null world");
  if (c < 0 || c >= w->N) return rtgi_fail(RTGI_E_INVALID, "cascade out of range");
  if (hits && !w->hits) return rtgi_fail(RTGI_E_STATE, "hit ids were not recorded (RTGI_RC_RECORD_HITS)");
  HIPCHK(hipSetDevice(w->device));
  HIPCHK(world_wait(w));
  size_t n = w->ray_off[c + 1] - w->ray_off[c];
  size_t np = (size_t)w->n[c] * w->n[c] * w->n[c], dirs = (size_t)w->res[c] * w->res[c];
  // Device order is direction-major (ray = di * np + device probe index); the readback is probe-major with probes
  // in grid order x + n (y + n z).
  std::vector<uint32_t> dev(np);
  for (uint32_t z = 0, p = 0; z < (uint32_t)w->n[c]; z++)
    for (uint32_t y = 0; y < (uint32_t)w->n[c]; y++)
      for (uint32_t x = 0; x < (uint32_t)w->n[c]; x++, p++) dev[p] = probe_index_host(w, c, x, y, z);
  if (rgba && (w->rp.flags & RTGI_RC_FP16)) {
    std::vector<uint16_t> t(n * 4);
    HIPCHK(hipMemcpy(t.data(), (const uint16_t*)w->rad + w->ray_off[c] * 4, n * 8, hipMemcpyDeviceToHost));
    for (size_t d = 0; d < dirs; d++)
      for (size_t p = 0; p < np; p++)
        for (int q = 0; q < 4; q++) rgba[(p * dirs + d) * 4 + q] = half_to_float(t[(d * np + dev[p]) * 4 + q]);
  } else if (rgba) {
    std::vector<float> t(n * 4);
    HIPCHK(hipMemcpy(t.data(), w->rad + w->ray_off[c] * 4, n * 16, hipMemcpyDeviceToHost));
    for (size_t d = 0; d < dirs; d++)
      for (size_t p = 0; p < np; p++) std::memcpy(rgba + (p * dirs + d) * 4, &t[(d * np + dev[p]) * 4], 16);
  }
  if (hits) {
    std::vector<uint32_t> t(n);
    HIPCHK(hipMemcpy(t.data(), w->hits + w->ray_off[c], n * 4, hipMemcpyDeviceToHost));
    for (size_t d = 0; d < dirs; d++)
      for (size_t p = 0; p < np; p++) hits[p * dirs + d] = t[d * np + dev[p]];
  }
  return RTGI_OK;
}

extern "C" int rtgi_world_trace_rays(rtgi_world* w, int32_t c, const float* o, const float* d, uint32_t n, float* rgba,
                                     uint32_t* hits) {
  if (!w || (n && (!o || !d || !rgba))) return rtgi_fail(RTGI_E_INVALID, "rtgi_world_trace_rays: null argument");
  if (c < 0 || c >= w->N) return rtgi_fail(RTGI_E_INVALID, "cascade %d out of range", c);
  if (!w->scene_set || !w->frame_set) return rtgi_fail(RTGI_E_STATE, "no scene or no frame");
  if (n == 0) return RTGI_OK;
  HIPCHK(hipSetDevice(w->device));
  HIPCHK(world_wait(w));
  float *dod = nullptr, *dout = nullptr;
  uint32_t* dh = nullptr;
  int e = RTGI_OK;
  if (hipMalloc((void**)&dod, (size_t)n * 24) != hipSuccess || hipMalloc((void**)&dout, (size_t)n * 16) != hipSuccess ||
      hipMalloc((void**)&dh, (size_t)n * 4) != hipSuccess)
    e = rtgi_fail(RTGI_E_NOMEM, "hipMalloc for %u rays", n);
  if (!e && (hipMemcpy(dod, o, (size_t)n * 12, hipMemcpyHostToDevice) != hipSuccess ||
             hipMemcpy(dod + (size_t)n * 3, d, (size_t)n * 12, hipMemcpyHostToDevice) != hipSuccess))
    e = rtgi_fail(RTGI_E_HIP, "ray upload failed");
  if (!e) {
    int le = launch_trace_rays(trace_args(w), c, dod, dod + (size_t)n * 3, n, dout, dh, w->stream);
    if (le) e = rtgi_fail(RTGI_E_HIP, "ray query launch failed: %s", hipGetErrorString((hipError_t)le));
  }
  if (!e && hipStreamSynchronize(w->stream) != hipSuccess) e = rtgi_fail(RTGI_E_HIP, "ray query failed");
  if (!e && hipMemcpy(rgba, dout, (size_t)n * 16, hipMemcpyDeviceToHost) != hipSuccess)
    e = rtgi_fail(RTGI_E_HIP, "readback failed");
  if (!e && hits && hipMemcpy(hits, dh, (size_t)n * 4, hipMemcpyDeviceToHost) != hipSuccess)
    e = rtgi_fail(RTGI_E_HIP, "readback failed");
  if (dod) (void)hipFree(dod);
  if (dout) (void)hipFree(dout);
  if (dh) (void)hipFree(dh);
  return e;
}

extern "C" int rtgi_world_read_probes(rtgi_world* w, int32_t c, float* pos, float* nrm, uint32_t* parents) {
  if (!w) return rtgi_fail(RTGI_E_INVALID, "null world");
  if (c < 0 || c >= w->N) return rtgi_fail(RTGI_E_INVALID, "cascade out of range");
  if (!(w->rp.flags & RTGI_RC_RELOCATE)) return rtgi_fail(RTGI_E_STATE, "probes are not relocated (RTGI_RC_RELOCATE)");
  HIPCHK(hipSetDevice(w->device));
  HIPCHK(world_wait(w));
  size_t np = (size_t)w->n[c] * w->n[c] * w->n[c];
  std::vector<float> P(np * 4), N(np * 4);
  std::vector<uint32_t> V(np);
  HIPCHK(hipMemcpy(P.data(), w->ppos + w->probe_off[c] * 4, np * 16, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(N.data(), w->pnrm + w->probe_off[c] * 4, np * 16, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(V.data(), w->pvis + w->probe_off[c], np * 4, hipMemcpyDeviceToHost));
  for (uint32_t z = 0, p = 0; z < (uint32_t)w->n[c]; z++)
    for (uint32_t y = 0; y < (uint32_t)w->n[c]; y++)
      for (uint32_t x = 0; x < (uint32_t)w->n[c]; x++, p++) {
        uint32_t d = probe_index_host(w, c, x, y, z);
        for (int a = 0; a < 3; a++) {
          if (pos) pos[p * 3 + a] = P[d * 4 + a];
          if (nrm) nrm[p * 3 + a] = N[d * 4 + a];
        }
        if (parents) parents[p] = c + 1 < w->N ? V[d] : 0u;
      }
  return RTGI_OK;
}

extern "C" int rtgi_world_set_dynamic(rtgi_world* w, uint32_t first_dynamic_tri, int32_t dyn_bricks) {
  if (!w) return rtgi_fail(RTGI_E_INVALID, "null world");
  if (dyn_bricks < 1 || dyn_bricks > (1 << 20))
    return rtgi_fail(RTGI_E_INVALID, "dynamic layer: dyn_bricks must be in 1..2^20");
  if (w->per % 32) return rtgi_fail(RTGI_E_INVALID, "dynamic layer: G^3 must be a multiple of 32");
  HIPCHK(hipSetDevice(w->device));
  HIPCHK(world_wait(w));
  dyn_free(w);
  w->dyn_first = first_dynamic_tri;
  w->dyn_cap = dyn_bricks;
  const size_t NLper = (size_t)w->NL * w->per, cap = (size_t)dyn_bricks;
  int e = dalloc(w, &w->dyn_hdr, DYN_HDR);
  if (!e) e = dalloc(w, &w->dyn_ptr, NLper);
  if (!e) e = dalloc(w, &w->dyn_bits, 2 * NLper / 32);
  if (!e) e = dalloc(w, &w->dyn_meta, cap);
  if (!e) e = dalloc(w, &w->dyn_mask, cap * 16);
  if (!e) e = dalloc(w, &w->dyn_data, cap * 512 * DYN_WORDS);
  if (e) {
    dyn_free(w);
    w->dyn_first = 0xffffffffu;
    return e;
  }
  // Initial state (createDynLayer, dynlayer.ts:670-673): every pointer empty, no bits, bounds being marked at
  // their identities.
  std::vector<uint32_t> hdr(DYN_HDR, 0u);
  hdr[32] = hdr[33] = hdr[34] = 0xffffffffu;
  HIPCHK(hipMemcpyAsync(w->dyn_hdr, hdr.data(), DYN_HDR * 4, hipMemcpyHostToDevice, w->stream));
  HIPCHK(hipMemsetD32Async((hipDeviceptr_t)w->dyn_ptr, (int)DYN_EMPTY, NLper, w->stream));
  HIPCHK(hipMemsetAsync(w->dyn_bits, 0, 2 * NLper / 32 * 4, w->stream));
  HIPCHK(hipStreamSynchronize(w->stream));
  return RTGI_OK;
}

extern "C" int rtgi_world_read_dyn_counters(rtgi_world* w, uint32_t* bricks, uint32_t* brick_overflow,
                                            uint32_t* outside) {
  if (!w) return rtgi_fail(RTGI_E_INVALID, "null world");
  if (!w->dyn_hdr) return rtgi_fail(RTGI_E_STATE, "no dynamic layer (rtgi_world_set_dynamic)");
  HIPCHK(hipSetDevice(w->device));
  HIPCHK(world_wait(w));
  uint32_t h[DYN_HDR];
  HIPCHK(hipMemcpy(h, w->dyn_hdr, sizeof h, hipMemcpyDeviceToHost));
  if (bricks) *bricks = h[8];
  if (brick_overflow) *brick_overflow = h[1];
  if (outside) *outside = h[9];
  return RTGI_OK;
}

extern "C" int rtgi_world_read_dynamic(rtgi_world* w, uint32_t* hdr, uint32_t* ptr, uint32_t* meta, uint32_t* mask,
                                       uint32_t* data) {
  if (!w) return rtgi_fail(RTGI_E_INVALID, "null world");
  if (!w->dyn_hdr) return rtgi_fail(RTGI_E_STATE, "no dynamic layer (rtgi_world_set_dynamic)");
  HIPCHK(hipSetDevice(w->device));
  HIPCHK(world_wait(w));
  const size_t NLper = (size_t)w->NL * w->per, cap = (size_t)w->dyn_cap;
  if (hdr) HIPCHK(hipMemcpy(hdr, w->dyn_hdr, DYN_HDR * 4, hipMemcpyDeviceToHost));
  if (ptr) HIPCHK(hipMemcpy(ptr, w->dyn_ptr, NLper * 4, hipMemcpyDeviceToHost));
  if (meta) HIPCHK(hipMemcpy(meta, w->dyn_meta, cap * 4, hipMemcpyDeviceToHost));
  if (mask) HIPCHK(hipMemcpy(mask, w->dyn_mask, cap * 16 * 4, hipMemcpyDeviceToHost));
  if (data) HIPCHK(hipMemcpy(data, w->dyn_data, cap * 512 * DYN_WORDS * 4, hipMemcpyDeviceToHost));
  return RTGI_OK;
}

extern "C" int rtgi_world_read_bounce_volume(rtgi_world* w, uint16_t* texels, int32_t origin[3]) {
  if (!w || !texels || !origin) return rtgi_fail(RTGI_E_INVALID, "rtgi_world_read_bounce_volume: null argument");
  if (!w->bounce) return rtgi_fail(RTGI_E_STATE, "no bounce volume (RTGI_RC_BOUNCE)");
  HIPCHK(hipSetDevice(w->device));
  HIPCHK(world_wait(w));
  HIPCHK(hipMemcpy(texels, w->bnvol[w->bn_par], (size_t)BN_WORDS * 4, hipMemcpyDeviceToHost));
  for (int a = 0; a < 3; a++) origin[a] = w->bn_org[w->bn_par][a];
  return RTGI_OK;
}

extern "C" int rtgi_world_gather(rtgi_world* w, const float* pos, const float* nrm, uint32_t n, uint32_t flags,
                                 float* irr) {
  if (!w || (n && (!pos || !nrm || !irr))) return rtgi_fail(RTGI_E_INVALID, "rtgi_world_gather: null argument");
  if (flags & ~RTGI_GATHER_OCCLUSION) return rtgi_fail(RTGI_E_INVALID, "unknown gather flags");
  if (w->rp.flags & RTGI_RC_NO_MERGE) return rtgi_fail(RTGI_E_STATE, "no irradiance field (RTGI_RC_NO_MERGE)");
  if (!w->scene_set || !w->frame_set) return rtgi_fail(RTGI_E_STATE, "no scene or no frame");
  if (n == 0) return RTGI_OK;
  HIPCHK(hipSetDevice(w->device));
  HIPCHK(world_wait(w));
  float *din = nullptr, *dout = nullptr;
  int e = RTGI_OK;
  if (hipMalloc((void**)&din, (size_t)n * 24) != hipSuccess || hipMalloc((void**)&dout, (size_t)n * 12) != hipSuccess)
    e = rtgi_fail(RTGI_E_NOMEM, "hipMalloc for %u points", n);
  if (!e && (hipMemcpy(din, pos, (size_t)n * 12, hipMemcpyHostToDevice) != hipSuccess ||
             hipMemcpy(din + (size_t)n * 3, nrm, (size_t)n * 12, hipMemcpyHostToDevice) != hipSuccess))
    e = rtgi_fail(RTGI_E_HIP, "point upload failed");
  if (!e) {
    int le = launch_gather(trace_args(w), din, din + (size_t)n * 3, n, (flags & RTGI_GATHER_OCCLUSION) ? 1 : 0, dout,
                           w->stream);
    if (le) e = rtgi_fail(RTGI_E_HIP, "gather launch failed: %s", hipGetErrorString((hipError_t)le));
  }
  if (!e && hipStreamSynchronize(w->stream) != hipSuccess) e = rtgi_fail(RTGI_E_HIP, "gather failed");
  if (!e && hipMemcpy(irr, dout, (size_t)n * 12, hipMemcpyDeviceToHost) != hipSuccess)
    e = rtgi_fail(RTGI_E_HIP, "readback failed");
  if (din) (void)hipFree(din);
  if (dout) (void)hipFree(dout);
  return e;
}

extern "C" int rtgi_world_gather_async(rtgi_world* w, const float* d_pos, const float* d_nrm, uint32_t n,
                                       uint32_t flags, float* d_irr, void* stream) {
  if (!w || (n && (!d_pos || !d_nrm || !d_irr))) return rtgi_fail(RTGI_E_INVALID, "rtgi_world_gather_async: null argument");
  if (flags & ~RTGI_GATHER_OCCLUSION) return rtgi_fail(RTGI_E_INVALID, "unknown gather flags");
  if (w->rp.flags & RTGI_RC_NO_MERGE) return rtgi_fail(RTGI_E_STATE, "no irradiance field (RTGI_RC_NO_MERGE)");
  if (!w->scene_set || !w->frame_set) return rtgi_fail(RTGI_E_STATE, "no scene or no frame");
  if (n == 0) return RTGI_OK;
  HIPCHK(hipSetDevice(w->device));
  hipStream_t s = stream ? (hipStream_t)stream : w->stream;
  if (s != w->stream) {  // after the world's own work and its last batched update
    if (!w->order_ev) HIPCHK(hipEventCreateWithFlags(&w->order_ev, hipEventDisableTiming));
    HIPCHK(hipEventRecord(w->order_ev, w->stream));
    HIPCHK(hipStreamWaitEvent(s, w->order_ev, 0));
  }
  if (w->ext_done) HIPCHK(hipStreamWaitEvent(s, w->ext_done, 0));
  int le = launch_gather(trace_args(w), d_pos, d_nrm, n, (flags & RTGI_GATHER_OCCLUSION) ? 1 : 0, d_irr, s);
  if (le) return rtgi_fail(RTGI_E_HIP, "gather launch failed: %s", hipGetErrorString((hipError_t)le));
  if (s != w->stream) {  // the world's next update (which rewrites the field) waits for the gather
    HIPCHK(hipEventRecord(w->order_ev, s));
    HIPCHK(hipStreamWaitEvent(w->stream, w->order_ev, 0));
  }
  return RTGI_OK;
}

extern "C" int rtgi_world_gather_noalloc(rtgi_world* w, const float* pos, const float* nrm, uint32_t n, uint32_t flags,
                                         float* irr) {
  if (!w || (n && (!pos || !nrm || !irr))) return rtgi_fail(RTGI_E_INVALID, "rtgi_world_gather_noalloc: null argument");
  if (flags & ~RTGI_GATHER_OCCLUSION) return rtgi_fail(RTGI_E_INVALID, "unknown gather flags");
  if (n > RTGI_GATHER_SCRATCH_POINTS)
    return rtgi_fail(RTGI_E_INVALID, "rtgi_world_gather_noalloc: at most %u points per call (got %u)",
                     RTGI_GATHER_SCRATCH_POINTS, n);
  if (w->rp.flags & RTGI_RC_NO_MERGE) return rtgi_fail(RTGI_E_STATE, "no irradiance field (RTGI_RC_NO_MERGE)");
  if (!w->scene_set || !w->frame_set) return rtgi_fail(RTGI_E_STATE, "no scene or no frame");
  if (n == 0) return RTGI_OK;
  HIPCHK(hipSetDevice(w->device));
  HIPCHK(world_wait(w));
  float* sp = nullptr;  // this device's copy of the staging array
  if (int se = gather_scratch(w->device, &sp))
    return rtgi_fail(RTGI_E_HIP, "gather staging address: %s", hipGetErrorString((hipError_t)se));
  const size_t M = RTGI_GATHER_SCRATCH_POINTS;
  float *dpos = sp, *dnrm = sp + 3 * M, *dirr = sp + 6 * M;
  HIPCHK(hipMemcpyAsync(dpos, pos, (size_t)n * 12, hipMemcpyHostToDevice, w->stream));
  HIPCHK(hipMemcpyAsync(dnrm, nrm, (size_t)n * 12, hipMemcpyHostToDevice, w->stream));
  int le = launch_gather(trace_args(w), dpos, dnrm, n, (flags & RTGI_GATHER_OCCLUSION) ? 1 : 0, dirr, w->stream);
  if (le) return rtgi_fail(RTGI_E_HIP, "gather launch failed: %s", hipGetErrorString((hipError_t)le));
  HIPCHK(hipMemcpyAsync(irr, dirr, (size_t)n * 12, hipMemcpyDeviceToHost, w->stream));
  HIPCHK(hipStreamSynchronize(w->stream));
  return RTGI_OK;
}

extern "C" int rtgi_world_read_trace_stats(rtgi_world* w, rtgi_trace_stats* out) {
  if (!w || !out) return rtgi_fail(RTGI_E_INVALID, "null argument");
  if (!w->stats) return rtgi_fail(RTGI_E_STATE, "tracer statistics are off (RTGI_RC_STATS)");
  HIPCHK(hipSetDevice(w->device));
  HIPCHK(world_wait(w));
  unsigned long long v[3 * RTGI_MAX_CASCADES];
  HIPCHK(hipMemcpy(v, w->stats, sizeof(v), hipMemcpyDeviceToHost));
  for (int c = 0; c < RTGI_MAX_CASCADES; c++) {
    out->traces[c] = v[c];
    out->steps[c] = v[RTGI_MAX_CASCADES + c];
    out->capped[c] = v[2 * RTGI_MAX_CASCADES + c];
  }
  return RTGI_OK;
}

extern "C" int rtgi_world_read_sh(rtgi_world* w, float* sh) {
  if (!w || !sh) return rtgi_fail(RTGI_E_INVALID, "null argument");
  HIPCHK(hipSetDevice(w->device));
  HIPCHK(world_wait(w));
  size_t np = (size_t)w->n[0] * w->n[0] * w->n[0];
  const float* src = w->shb[w->sh_par];  // planes [27][np] in grid-row order
  if (w->rp.flags & RTGI_RC_FP16) {
    std::vector<uint16_t> planes(np * 27);
    HIPCHK(hipMemcpy(planes.data(), src, np * 27 * 2, hipMemcpyDeviceToHost));
    for (size_t p = 0; p < np; p++)
      for (int q = 0; q < 27; q++) sh[p * 27 + q] = half_to_float(planes[(size_t)q * np + p]);
  } else {
    std::vector<float> planes(np * 27);
    HIPCHK(hipMemcpy(planes.data(), src, np * 27 * 4, hipMemcpyDeviceToHost));
    for (size_t p = 0; p < np; p++)
      for (int q = 0; q < 27; q++) sh[p * 27 + q] = planes[(size_t)q * np + p];
  }
  return RTGI_OK;
}

extern "C" int rtgi_world_set_tuning(rtgi_world* w, const rtgi_tuning* t) {
  if (!w || !t) return rtgi_fail(RTGI_E_INVALID, "null argument");
  if (t->probe_tile != 1 && t->probe_tile != 2 && t->probe_tile != 4)
    return rtgi_fail(RTGI_E_INVALID, "probe_tile must be 1, 2 or 4 (got %d)", t->probe_tile);
  if (t->fused < -1 || t->fused > 1) return rtgi_fail(RTGI_E_INVALID, "fused must be -1, 0 or 1");
  HIPCHK(hipSetDevice(w->device));
  HIPCHK(world_wait(w));  // the storage order of an update in flight must not change under it
  w->tuning = *t;
  return RTGI_OK;
}

extern "C" int rtgi_world_get_tuning(const rtgi_world* w, rtgi_tuning* t) {
  if (!w || !t) return rtgi_fail(RTGI_E_INVALID, "null argument");
  *t = w->tuning;
  return RTGI_OK;
}

extern "C" int rtgi_world_destroy(rtgi_world* w) {
  if (!w) return RTGI_OK;
  (void)hipSetDevice(w->device);
  if (w->stream) (void)hipStreamSynchronize(w->stream);
  free_all(w);
  for (int i = 0; i < 5; i++)
    if (w->ev[i]) (void)hipEventDestroy(w->ev[i]);
  for (int i = 0; i < 2; i++)
    if (w->pose_copied[i]) (void)hipEventDestroy(w->pose_copied[i]);
  if (w->order_ev) (void)hipEventDestroy(w->order_ev);
  if (w->stream) (void)hipStreamDestroy(w->stream);
  delete w;
  return RTGI_OK;
}

int world_save_state(rtgi_world* w, WorldState* st) {
  HIPCHK(hipSetDevice(w->device));
  HIPCHK(world_wait(w));
  *st = WorldState{};
  if (w->sh_written) {
    st->sh.resize(sh_floats_of(w) * sizeof(float));
    HIPCHK(hipMemcpy(st->sh.data(), w->shb[w->sh_par], st->sh.size(), hipMemcpyDeviceToHost));
    for (int a = 0; a < 3; a++) st->sh_G0[a] = w->sh_G0[a];
  }
  if (w->bounce) {
    for (int i = 0; i < 2; i++) {
      st->bnvol[i].resize(BN_WORDS);
      HIPCHK(hipMemcpy(st->bnvol[i].data(), w->bnvol[i], (size_t)BN_WORDS * 4, hipMemcpyDeviceToHost));
    }
    st->bn_par = w->bn_par;
    std::memcpy(st->bn_org, w->bn_org, sizeof(st->bn_org));
    st->bn_frame = w->bn_frame;
  }
  st->valid = true;
  return RTGI_OK;
}

int world_restore_state(rtgi_world* w, const WorldState& st) {
  if (!st.valid) return RTGI_OK;
  HIPCHK(hipSetDevice(w->device));
  HIPCHK(world_wait(w));
  if (!st.sh.empty()) {
    if (st.sh.size() != sh_floats_of(w) * sizeof(float)) return rtgi_fail(RTGI_E_STATE, "saved SH of another shape");
    w->sh_par = 0;
    HIPCHK(hipMemcpy(w->shb[0], st.sh.data(), st.sh.size(), hipMemcpyHostToDevice));
    w->sh_written = true;
    for (int a = 0; a < 3; a++) w->sh_G0[a] = st.sh_G0[a];
  }
  if (w->bounce && !st.bnvol[0].empty()) {
    for (int i = 0; i < 2; i++)
      HIPCHK(hipMemcpy(w->bnvol[i], st.bnvol[i].data(), (size_t)BN_WORDS * 4, hipMemcpyHostToDevice));
    w->bn_par = st.bn_par;
    std::memcpy(w->bn_org, st.bn_org, sizeof(w->bn_org));
    w->bn_frame = st.bn_frame;
  }
  return RTGI_OK;
}

// ---- Batches ----------------------------------------------------------------------------------------------------

struct rtgi_batch {
  int device = 0;
  std::vector<rtgi_world*> worlds;
  hipStream_t stream = nullptr;
  TraceK* d_desc = nullptr;
  TraceK* h_desc[2] = {nullptr, nullptr};  // pinned staging, double-buffered
  hipEvent_t copied[2] = {nullptr, nullptr};
  int cur = 0;
  uint32_t* d_start = nullptr;
  uint32_t* d_queue = nullptr;  // the work-queue counter of the batched build
  BatchPlan plan{};
  std::vector<uint64_t> sig;  // sizes the plan was made for
  bool use_graph = false, graph_valid = false;
  // One graph of the whole update; with timing on, one per stage, launched with the stage events between them.
  hipGraph_t graph[4] = {};
  hipGraphExec_t exec[4] = {};
  hipEvent_t done = nullptr;
  bool timing = false;
  hipEvent_t ev[5] = {};
};

void plan_blocks(rtgi_world* const* worlds, int n, uint32_t* start, BatchPlan& p) {
  for (int l = 0; l < BL_COUNT; l++) start[(size_t)l * (n + 1)] = 0;
  size_t lds = 0;
  bool all_lds = true;
  p.stats = false;
  p.dyn = false;
  p.surf = false;
  p.fused = n > 0;
  for (int i = 0; i < n; i++) p.fused = p.fused && worlds[i]->tuning.fused != 0;
  for (int st = 0; st < RTGI_MAX_CASCADES; st++) {
    p.min_casc[st] = RTGI_MAX_CASCADES;
    for (int i = 0; i < n; i++)
      if (worlds[i]->N - 1 - st >= 0) p.min_casc[st] = std::min(p.min_casc[st], worlds[i]->N - 1 - st);
  }
  for (int i = 0; i < n; i++) {
    TraceK t = trace_args(worlds[i]);
    uint32_t g[BL_COUNT];
    batch_blocks(t, g);
    if (worlds[i]->rp.flags & RTGI_RC_NO_MERGE) {
      for (int st = 0; st < RTGI_MAX_CASCADES; st++) g[BL_MERGE + st] = 0;
      g[BL_SH] = 0;
    }
    for (int l = 0; l < BL_COUNT; l++) start[(size_t)l * (n + 1) + i + 1] = start[(size_t)l * (n + 1) + i] + g[l];
    size_t m = vox_mark_lds_bytes(t.vx);
    if (m == 0) all_lds = false;
    lds = std::max(lds, m);
    p.stats = p.stats || worlds[i]->stats != nullptr;
    p.dyn = p.dyn || worlds[i]->dyn_hdr != nullptr;
    p.surf = p.surf || worlds[i]->surf_on;
  }
  p.n = n;
  for (int l = 0; l < BL_COUNT; l++) {
    p.grid[l] = start[(size_t)l * (n + 1) + n];
    // About equal block counts (within 25 %) stand for about equal cost per world: then the XCD remap.
    uint32_t lo = 0xffffffffu, hi = 0;
    for (int i = 0; i < n; i++) {
      uint32_t g = start[(size_t)l * (n + 1) + i + 1] - start[(size_t)l * (n + 1) + i];
      lo = std::min(lo, g);
      hi = std::max(hi, g);
    }
    p.xcd[l] = n > 0 && lo > 0 && (uint64_t)hi * 4u <= (uint64_t)lo * 5u;
  }
  // One dynamic LDS size for the whole launch: the largest bitmap when every world's fits (each world's descriptor
  // says whether it marks in LDS).
  p.mark_lds = all_lds ? lds : 0;
}

bool plan_sh_pairs(rtgi_world* const* worlds, const TraceK* desc, int n) {
  bool any = false;
  for (int i = 0; i < n; i++) {
    if (worlds[i]->rp.flags & RTGI_RC_NO_MERGE) continue;  // no SH projection
    if (!sh_pairs(desc[i])) return false;
    any = true;
  }
  return any;
}

void batch_world_begin(rtgi_world* w, const BatchPlan& p, TraceK* out) {
  trace_begin(w);
  bounce_advance(w);
  sh_advance(w);
  *out = trace_args(w);
}

int plan_launch_stage(const BatchPlan& p, hipStream_t s, int stage) {
  int e = 0;
  if (stage == 0) {
    e = launch_vox_build_batch(p, s);
    if (!e) e = launch_dyn_build_batch(p, s);
  } else if (stage == 1) {
    e = launch_sun_batch(p, s);
    if (!e) e = launch_dyn_inject_batch(p, s);
    if (!e) e = launch_bounce_batch(p, s);
    if (!e) e = launch_relocate_batch(p, s);
    if (p.fused)
      for (int st = 0; st < RTGI_MAX_CASCADES && !e; st++) e = launch_trace_merge_batch(p, st, s);
    else if (!e)
      e = launch_trace_batch(p, s);
  } else if (stage == 2) {
    if (!p.fused)
      for (int st = 0; st < RTGI_MAX_CASCADES && !e; st++) e = launch_merge_batch(p, st, s);
  } else {
    e = launch_sh_batch(p, s);
  }
  if (e) return rtgi_fail(RTGI_E_HIP, "batched launch failed: %s", hipGetErrorString((hipError_t)e));
  return RTGI_OK;
}

int plan_launch(const BatchPlan& p, hipStream_t s, hipEvent_t* ev) {
  for (int st = 0; st < 4; st++) {
    if (ev) HIPCHK(hipEventRecord(ev[st], s));
    int e = plan_launch_stage(p, s, st);
    if (e) return e;
  }
  if (ev) HIPCHK(hipEventRecord(ev[4], s));
  return RTGI_OK;
}

namespace {

void batch_free(rtgi_batch* b) {
  for (int g = 0; g < 4; g++) {
    if (b->exec[g]) (void)hipGraphExecDestroy(b->exec[g]);
    if (b->graph[g]) (void)hipGraphDestroy(b->graph[g]);
  }
  for (int i = 0; i < 2; i++) {
    if (b->h_desc[i]) (void)hipHostFree(b->h_desc[i]);
    if (b->copied[i]) (void)hipEventDestroy(b->copied[i]);
  }
  for (int i = 0; i < 5; i++)
    if (b->ev[i]) (void)hipEventDestroy(b->ev[i]);
  if (b->d_desc) (void)hipFree(b->d_desc);
  if (b->d_start) (void)hipFree(b->d_start);
  if (b->d_queue) (void)hipFree(b->d_queue);
  if (b->done) (void)hipEventDestroy(b->done);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

// Block prefix arrays on the device (synchronous upload; only when sizes change).
int batch_plan(rtgi_batch* b, const std::vector<uint64_t>& sig) {
  const int n = (int)b->worlds.size();
  std::vector<uint32_t> start((size_t)BL_COUNT * (n + 1), 0);
  plan_blocks(b->worlds.data(), n, start.data(), b->plan);
  HIPCHK(hipMemcpy(b->d_start, start.data(), start.size() * 4, hipMemcpyHostToDevice));
  b->plan.w = b->d_desc;
  b->plan.queue = b->d_queue;
  for (int l = 0; l < BL_COUNT; l++) b->plan.start[l] = b->d_start + (size_t)l * (n + 1);
  b->sig = sig;
  b->graph_valid = false;
  return RTGI_OK;
}

}  // namespace

extern "C" int rtgi_batch_create(rtgi_world* const* worlds, int32_t n, rtgi_batch** out) {
  if (!worlds || !out || n < 1) return rtgi_fail(RTGI_E_INVALID, "rtgi_batch_create: bad argument");
  *out = nullptr;
  for (int i = 0; i < n; i++) {
    if (!worlds[i]) return rtgi_fail(RTGI_E_INVALID, "rtgi_batch_create: null world %d", i);
    if (worlds[i]->device != worlds[0]->device)
      return rtgi_fail(RTGI_E_INVALID, "rtgi_batch_create: worlds on different devices");
    for (int j = 0; j < i; j++)
      if (worlds[j] == worlds[i]) return rtgi_fail(RTGI_E_INVALID, "rtgi_batch_create: world %d repeated", i);
  }
  rtgi_batch* b = new (std::nothrow) rtgi_batch();
  if (!b) return rtgi_fail(RTGI_E_NOMEM, "out of host memory");
  b->worlds.assign(worlds, worlds + n);
  b->device = worlds[0]->device;
  int e = RTGI_OK;
  if (hipSetDevice(b->device) != hipSuccess) e = rtgi_fail(RTGI_E_HIP, "hipSetDevice failed");
  if (!e && hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
    e = rtgi_fail(RTGI_E_HIP, "hipStreamCreate failed");
  if (!e && hipMalloc((void**)&b->d_desc, sizeof(TraceK) * n) != hipSuccess) e = rtgi_fail(RTGI_E_NOMEM, "hipMalloc");
  if (!e && hipMalloc((void**)&b->d_start, sizeof(uint32_t) * BL_COUNT * (n + 1)) != hipSuccess)
    e = rtgi_fail(RTGI_E_NOMEM, "hipMalloc");
  if (!e && hipMalloc((void**)&b->d_queue, sizeof(uint32_t) * RTGI_QUEUE_WORDS) != hipSuccess)
    e = rtgi_fail(RTGI_E_NOMEM, "hipMalloc");
  for (int i = 0; i < 2 && !e; i++) {
    if (hipHostMalloc((void**)&b->h_desc[i], sizeof(TraceK) * n, hipHostMallocDefault) != hipSuccess)
      e = rtgi_fail(RTGI_E_NOMEM, "hipHostMalloc");
    if (!e && hipEventCreateWithFlags(&b->copied[i], hipEventDisableTiming) != hipSuccess)
      e = rtgi_fail(RTGI_E_HIP, "hipEventCreate");
  }
  if (!e && hipEventCreateWithFlags(&b->done, hipEventDisableTiming) != hipSuccess)
    e = rtgi_fail(RTGI_E_HIP, "hipEventCreate");
  for (int i = 0; i < 5 && !e; i++)
    if (hipEventCreate(&b->ev[i]) != hipSuccess) e = rtgi_fail(RTGI_E_HIP, "hipEventCreate");
  if (e) {
    batch_free(b);
    return e;
  }
  *out = b;
  return RTGI_OK;
}

extern "C" int rtgi_batch_set_graph(rtgi_batch* b, int32_t on) {
  if (!b) return rtgi_fail(RTGI_E_INVALID, "null batch");
  b->use_graph = on != 0;
  return RTGI_OK;
}

extern "C" int rtgi_batch_set_timing(rtgi_batch* b, int32_t on) {
  if (!b) return rtgi_fail(RTGI_E_INVALID, "null batch");
  if (b->timing != (on != 0)) b->graph_valid = false;  // timed replays use one graph per stage
  b->timing = on != 0;
  return RTGI_OK;
}

extern "C" int rtgi_batch_enqueue(rtgi_batch* b, void* stream) {
  if (!b) return rtgi_fail(RTGI_E_INVALID, "null batch");
  HIPCHK(hipSetDevice(b->device));
  hipStream_t s = stream ? (hipStream_t)stream : b->stream;
  const int n = (int)b->worlds.size();
  std::vector<uint64_t> sig;
  sig.reserve((size_t)n * 8);
  for (int i = 0; i < n; i++) {
    rtgi_world* w = b->worlds[i];
    if (!w->scene_set || !w->frame_set) return rtgi_fail(RTGI_E_STATE, "world %d has no scene or no frame", i);
    HIPCHK(hipStreamSynchronize(w->stream));  // per-world work (uploads, single updates) is done first
    sig.push_back(w->ntris);
    sig.push_back(w->ray_off[w->N]);
    sig.push_back((uint64_t)w->NL << 32 | (uint32_t)w->per);
    sig.push_back((uint64_t)w->vp.max_bricks << 32 | w->rp.flags);
    // The launch set also follows the tuning, the dynamic layer and the surfaces, which may change between enqueues.
    sig.push_back((uint64_t)(uint32_t)w->tuning.probe_tile << 32 | (uint32_t)w->tuning.fused);
    sig.push_back(w->surf_on ? 1ull : 0ull);
    sig.push_back(w->dyn_hdr ? ((uint64_t)w->dyn_first << 32 | (uint32_t)w->dyn_cap) : ~0ull);
    sig.push_back(w->pose_set ? ((uint64_t)w->objs.size() << 32 | w->pose_span) : 0ull);
  }
  if (sig != b->sig) {
    // The previous update may still read the block prefixes being replaced.
    HIPCHK(hipStreamSynchronize(s));
    int e = batch_plan(b, sig);
    if (e) return e;
  }
  // Descriptors: the staging buffer being filled was last copied two enqueues ago; wait for that copy.
  HIPCHK(hipEventSynchronize(b->copied[b->cur]));
  for (int i = 0; i < n; i++) batch_world_begin(b->worlds[i], b->plan, &b->h_desc[b->cur][i]);
  // The SH launch form follows this update's descriptors (their storage order); a change needs a new graph.
  const bool shp = plan_sh_pairs(b->worlds.data(), b->h_desc[b->cur], n);
  if (shp != b->plan.sh_pairs) {
    b->plan.sh_pairs = shp;
    b->graph_valid = false;
  }
  HIPCHK(hipMemcpyAsync(b->d_desc, b->h_desc[b->cur], sizeof(TraceK) * n, hipMemcpyHostToDevice, s));
  HIPCHK(hipEventRecord(b->copied[b->cur], s));
  b->cur ^= 1;
  if (b->use_graph) {
    const int ng = b->timing ? 4 : 1;
    if (!b->graph_valid) {
      for (int g = 0; g < 4; g++) {
        if (b->exec[g]) (void)hipGraphExecDestroy(b->exec[g]);
        if (b->graph[g]) (void)hipGraphDestroy(b->graph[g]);
        b->exec[g] = nullptr;
        b->graph[g] = nullptr;
      }
      for (int g = 0; g < ng; g++) {
        HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        int e = ng == 1 ? plan_launch(b->plan, s, nullptr) : plan_launch_stage(b->plan, s, g);
        hipError_t ce = hipStreamEndCapture(s, &b->graph[g]);
        if (e) return e;
        if (ce != hipSuccess) return rtgi_fail(RTGI_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(ce));
        size_t nodes = 0;
        HIPCHK(hipGraphGetNodes(b->graph[g], nullptr, &nodes));
        if (nodes) HIPCHK(hipGraphInstantiate(&b->exec[g], b->graph[g], nullptr, nullptr, 0));
      }
      b->graph_valid = true;
    }
    // Stage events are recorded between the stage graphs (events recorded inside a graph carry no timestamps).
    for (int g = 0; g < ng; g++) {
      if (b->timing) HIPCHK(hipEventRecord(b->ev[g], s));
      if (b->exec[g]) HIPCHK(hipGraphLaunch(b->exec[g], s));
    }
    if (b->timing) HIPCHK(hipEventRecord(b->ev[4], s));
  } else {
    int e = plan_launch(b->plan, s, b->timing ? b->ev : nullptr);
    if (e) return e;
  }
  HIPCHK(hipEventRecord(b->done, s));
  for (rtgi_world* w : b->worlds) {
    w->ext_done = b->done;
  }
  return RTGI_OK;
}

extern "C" int rtgi_batch_sync(rtgi_batch* b) {
  if (!b) return rtgi_fail(RTGI_E_INVALID, "null batch");
  HIPCHK(hipSetDevice(b->device));
  HIPCHK(hipEventSynchronize(b->done));
  return RTGI_OK;
}

extern "C" int rtgi_batch_stage_ms(rtgi_batch* b, float ms[4]) {
  if (!b || !ms) return rtgi_fail(RTGI_E_INVALID, "null argument");
  if (!b->timing) return rtgi_fail(RTGI_E_STATE, "timing is off (rtgi_batch_set_timing)");
  HIPCHK(hipEventSynchronize(b->ev[4]));
  for (int i = 0; i < 4; i++) HIPCHK(hipEventElapsedTime(&ms[i], b->ev[i], b->ev[i + 1]));
  return RTGI_OK;
}

extern "C" int rtgi_batch_destroy(rtgi_batch* b) {
  if (!b) return RTGI_OK;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->done) (void)hipEventSynchronize(b->done);
  for (rtgi_world* w : b->worlds)
    if (w->ext_done == b->done) w->ext_done = nullptr;
  batch_free(b);
  return RTGI_OK;
}
