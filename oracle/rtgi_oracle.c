/*
 * rtgi_oracle.c -- TEST INFRASTRUCTURE ONLY (the checker for librtgi; see rtgi_oracle.h).
 *
 * Plain C99, compiled with -O2 -ffp-contract=off on x86-64 (SSE: IEEE f32, correctly rounded / and sqrt).
 * Each function cites the reference WGSL it restates (paths relative to the reference repository,
 * owoDra/real-time-gi-lab). GPU atomics are taken in thread-index order; the brick allocation uses the reference's
 * ordered allocation (the one it uses when the pool overflows, voxel.ts:405-470) in every case, which is one of the
 * schedules its atomic counter allows when every brick fits.
 *
 * Departures, each documented where it occurs: no alpha test and no texture (untextured, opaque materials); no
 * dynamic layer, bounce or voxSun (the sun visibility is a frame constant); dense probe grids with relocation
 * 'none' instead of screen-registered probes; point lights are an addition (0 lights = the reference shading).
 */
#include "rtgi_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define HDR 64u
#define BW 18u /* BRICK_WORDS, voxel.ts:53 */
#define V_EMPTY 0xffffffffu
#define V_DROP 0xfffffffeu
#define V_MARK 0xfffffffdu
#define BIG 8
#define BIG_CAP (1u << 18)

typedef struct { float x, y, z; } v3;
typedef struct { int x, y, z; } iv3;

static v3 V3(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static v3 vadd(v3 a, v3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
static v3 vsub(v3 a, v3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
static v3 vmul(v3 a, v3 b) { return V3(a.x * b.x, a.y * b.y, a.z * b.z); }
static v3 vscale(v3 a, float s) { return V3(a.x * s, a.y * s, a.z * s); }
static v3 vdivs(v3 a, float s) { return V3(a.x / s, a.y / s, a.z / s); }
static v3 vneg(v3 a) { return V3(-a.x, -a.y, -a.z); }
static float vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static v3 vabs(v3 a) { return V3(fabsf(a.x), fabsf(a.y), fabsf(a.z)); }
static float fmn(float a, float b) { return a < b ? a : b; }
static float fmx(float a, float b) { return a > b ? a : b; }
static v3 vmin(v3 a, v3 b) { return V3(fmn(a.x, b.x), fmn(a.y, b.y), fmn(a.z, b.z)); }
static v3 vmax(v3 a, v3 b) { return V3(fmx(a.x, b.x), fmx(a.y, b.y), fmx(a.z, b.z)); }
static v3 vcross(v3 a, v3 b) { return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static v3 vnormalize(v3 a) { return vdivs(a, sqrtf(vdot(a, a))); }
static v3 vfloor(v3 a) { return V3(floorf(a.x), floorf(a.y), floorf(a.z)); }
static v3 vf(iv3 a) { return V3((float)a.x, (float)a.y, (float)a.z); }
static iv3 I3(int x, int y, int z) { iv3 r = {x, y, z}; return r; }
static int f2i(float f) { /* WGSL f32 -> i32 saturates */
  if (f <= -2147483648.0f) return (-2147483647 - 1);
  if (f >= 2147483520.0f) return 2147483520;
  return (int)f;
}
static iv3 vi(v3 a) { return I3(f2i(a.x), f2i(a.y), f2i(a.z)); }
static int imin(int a, int b) { return a < b ? a : b; }
static int imax(int a, int b) { return a > b ? a : b; }
static iv3 iclamp(iv3 a, iv3 lo, iv3 hi) {
  return I3(imin(imax(a.x, lo.x), hi.x), imin(imax(a.y, lo.y), hi.y), imin(imax(a.z, lo.z), hi.z));
}
static iv3 iadd(iv3 a, iv3 b) { return I3(a.x + b.x, a.y + b.y, a.z + b.z); }
static iv3 isub(iv3 a, iv3 b) { return I3(a.x - b.x, a.y - b.y, a.z - b.z); }
static iv3 imuls(iv3 a, int s) { return I3(a.x * s, a.y * s, a.z * s); }
static iv3 iadds(iv3 a, int s) { return I3(a.x + s, a.y + s, a.z + s); }
static iv3 ishr(iv3 a, int s) { return I3(a.x >> s, a.y >> s, a.z >> s); }
static int popc(uint32_t x) { return __builtin_popcount(x); }

/* ------------------------------------------------------------------------------------------------------------ */
/* Buffer layout (createVoxWorld, voxel.ts:753-760) without the bounce words and the sun word.                   */

typedef struct {
  const orc_vox* p;
  uint32_t* vox;
  uint32_t* rep;
  const float* pos;
  const float* nrm;
  const uint32_t* tri;
  const float* mat;
  uint32_t ntris;
  uint32_t ptrOff, brickOff, attrOff, plOff;
  uint32_t* big[16]; /* 0-7 voxBits per level, 8-15 markBricks per level */
} Ctx;

static void layout(const orc_vox* p, uint32_t* ptrOff, uint32_t* brickOff, uint32_t* attrOff, uint32_t* plOff) {
  uint32_t G = (uint32_t)p->G;
  *ptrOff = HDR;
  *brickOff = *ptrOff + (uint32_t)p->levels * G * G * G;
  *attrOff = *brickOff + (uint32_t)p->max_bricks * BW;
  *plOff = *attrOff + (uint32_t)p->max_voxels * 2u;
}

/* After the reference's buffer: the precision form's bf16 albedo, 2 words per attribute slot (r | g << 16,
 * b | toon << 16), written by every build and read when orc_rc.bf16 is set. */
static uint32_t albOff(const orc_vox* p) {
  uint32_t a, b, c, d;
  layout(p, &a, &b, &c, &d);
  return d + (uint32_t)p->max_voxels;
}

size_t orc_vox_words(const orc_vox* p) { return (size_t)albOff(p) + 2u * (size_t)p->max_voxels; }

/* f32 -> bf16, round to nearest even (the top 16 bits of the float, rounded); NaN stays a quiet NaN. */
static uint16_t f2bf(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
static float bf2f(uint16_t h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}
uint16_t orc_f2bf(float f) { return f2bf(f); }
float orc_bf2f(uint16_t h) { return bf2f(h); }

void orc_origins(double v0, int levels, int G, const double eye[3], int32_t out[8][3]) {
  for (int L = 0; L < levels; L++) {
    double bs = v0 * pow(2.0, L) * 8;
    for (int a = 0; a < 3; a++) out[L][a] = (int32_t)(floor(eye[a] / bs) - G / 2);
  }
}

static float voxSize(const orc_vox* p, uint32_t L) { return p->v0 * (float)(1u << L); }
static uint32_t voxPtrIndex(const Ctx* X, uint32_t L, iv3 b) {
  uint32_t G = (uint32_t)X->p->G;
  return X->ptrOff + L * G * G * G + ((uint32_t)b.z * G + (uint32_t)b.y) * G + (uint32_t)b.x;
}
static iv3 org(const orc_vox* p, uint32_t L) { return I3(p->origin[L][0], p->origin[L][1], p->origin[L][2]); }

/* axisTest / triBox, voxel.ts:231-248 (Akenine-Moller). */
static int axisTest(v3 a, v3 v0, v3 v1, v3 v2, v3 h) {
  float p0 = vdot(v0, a), p1 = vdot(v1, a), p2 = vdot(v2, a);
  float r = vdot(h, vabs(a));
  return !(fmn(fmn(p0, p1), p2) > r || fmx(fmx(p0, p1), p2) < -r);
}
static int triBox(v3 c, v3 h, v3 a, v3 b, v3 cc) {
  v3 v0 = vsub(a, c), v1 = vsub(b, c), v2 = vsub(cc, c);
  v3 mn = vmin(vmin(v0, v1), v2), mx = vmax(vmax(v0, v1), v2);
  if (mn.x > h.x || mn.y > h.y || mn.z > h.z || mx.x < -h.x || mx.y < -h.y || mx.z < -h.z) return 0;
  v3 e0 = vsub(v1, v0), e1 = vsub(v2, v1), e2 = vsub(v0, v2);
  v3 n = vcross(e0, e1);
  if (fabsf(vdot(n, v0)) > vdot(h, vabs(n))) return 0;
  return axisTest(V3(0.0f, -e0.z, e0.y), v0, v1, v2, h) && axisTest(V3(0.0f, -e1.z, e1.y), v0, v1, v2, h) &&
         axisTest(V3(0.0f, -e2.z, e2.y), v0, v1, v2, h) && axisTest(V3(e0.z, 0.0f, -e0.x), v0, v1, v2, h) &&
         axisTest(V3(e1.z, 0.0f, -e1.x), v0, v1, v2, h) && axisTest(V3(e2.z, 0.0f, -e2.x), v0, v1, v2, h) &&
         axisTest(V3(-e0.y, e0.x, 0.0f), v0, v1, v2, h) && axisTest(V3(-e1.y, e1.x, 0.0f), v0, v1, v2, h) &&
         axisTest(V3(-e2.y, e2.x, 0.0f), v0, v1, v2, h);
}
int orc_tri_box(const float c[3], const float h[3], const float a[3], const float b[3], const float cc[3]) {
  return triBox(V3(c[0], c[1], c[2]), V3(h[0], h[1], h[2]), V3(a[0], a[1], a[2]), V3(b[0], b[1], b[2]),
                V3(cc[0], cc[1], cc[2]));
}

/* closestBary, voxel.ts:250-267 (Ericson, Real-Time Collision Detection 5.1.5). */
static v3 closestBary(v3 p, v3 a, v3 b, v3 c) {
  v3 ab = vsub(b, a), ac = vsub(c, a), ap = vsub(p, a);
  float d1 = vdot(ab, ap), d2 = vdot(ac, ap);
  if (d1 <= 0.0f && d2 <= 0.0f) return V3(1.0f, 0.0f, 0.0f);
  v3 bp = vsub(p, b);
  float d3 = vdot(ab, bp), d4 = vdot(ac, bp);
  if (d3 >= 0.0f && d4 <= d3) return V3(0.0f, 1.0f, 0.0f);
  float vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
    float w = d1 / (d1 - d3);
    return V3(1.0f - w, w, 0.0f);
  }
  v3 cp = vsub(p, c);
  float d5 = vdot(ab, cp), d6 = vdot(ac, cp);
  if (d6 >= 0.0f && d5 <= d6) return V3(0.0f, 0.0f, 1.0f);
  float vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
    float w = d2 / (d2 - d6);
    return V3(1.0f - w, 0.0f, w);
  }
  float va = d3 * d6 - d5 * d4;
  if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
    float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    return V3(0.0f, 1.0f - w, w);
  }
  float dn = 1.0f / (va + vb + vc);
  float v = vb * dn, w = vc * dn;
  return V3(1.0f - v - w, v, w);
}

/* ------------------------------------------------------------------------------------------------------------ */
/* Textured surfaces: textureSampleLevel(baseColorTex, samp, uv, layer, 3 + L) with the voxel passes' sampler     */
/* (linear min/mag/mip filtering, repeat addressing, voxel.ts:793) over a mip chain built as the reference's       */
/* generateMips builds it (scene_gpu.ts:38-54); rgb decoded from sRGB before filtering (rgba8unorm-srgb).          */

static const orc_surface* g_surf; /* NULL: untextured (white layer 0 everywhere, nothing masked) */
void orc_set_surface(const orc_surface* s) { g_surf = s; }

/* IEC 61966-2-1 decoding of an encoded value c in [0, 1]. */
static double srgbDecode(double c) { return c <= 0.04045 ? c / 12.92 : pow((c + 0.055) / 1.055, 2.4); }
void orc_srgb_tables(float lut[511]) {
  for (int c = 0; c < 256; c++) lut[c] = (float)srgbDecode(c / 255.0);
  for (int k = 0; k < 255; k++) lut[256 + k] = (float)srgbDecode((k + 0.5) / 255.0);
}
static void texelAt(const orc_surface* s, uint32_t base, uint32_t w, uint32_t x, uint32_t y, float t[4]) {
  uint32_t q = s->tex[base + y * w + x];
  t[0] = s->lut[q & 255u];
  t[1] = s->lut[(q >> 8) & 255u];
  t[2] = s->lut[(q >> 16) & 255u];
  t[3] = (float)(q >> 24) / 255.0f;
}
static uint32_t mipOff(uint32_t sl, uint32_t m) {
  uint32_t off = 0;
  for (uint32_t j = 0; j < m; j++) off += 1u << (2u * (sl - j));
  return off;
}
/* Nearest code of a linear value in the sRGB encoding: the thresholds at or below it, counted one by one. */
static uint32_t srgbEncode(const float* lut, float x) {
  uint32_t c = 0;
  for (int k = 0; k < 255; k++)
    if (lut[256 + k] <= x) c++;
  return c;
}
void orc_surf_mips(const uint8_t* texels, uint32_t size, uint32_t layers, const float* lut, uint32_t* out) {
  uint32_t sl = 0;
  while ((1u << sl) < size) sl++;
  uint32_t per = mipOff(sl, sl + 1);
  orc_surface S = {out, NULL, NULL, lut, sl, sl + 1, per, layers};
  for (uint32_t l = 0; l < layers; l++) {
    for (uint32_t i = 0; i < size * size; i++) {
      const uint8_t* b = texels + ((size_t)l * size * size + i) * 4;
      out[(size_t)l * per + i] = (uint32_t)b[0] | ((uint32_t)b[1] << 8) | ((uint32_t)b[2] << 16) | ((uint32_t)b[3] << 24);
    }
    for (uint32_t m = 1; m <= sl; m++) {
      uint32_t w = 1u << (sl - m), ws = 2u * w;
      uint32_t src = l * per + mipOff(sl, m - 1), dst = l * per + mipOff(sl, m);
      for (uint32_t y = 0; y < w; y++)
        for (uint32_t x = 0; x < w; x++) {
          float t00[4], t10[4], t01[4], t11[4], r[4];
          texelAt(&S, src, ws, 2 * x, 2 * y, t00);
          texelAt(&S, src, ws, 2 * x + 1, 2 * y, t10);
          texelAt(&S, src, ws, 2 * x, 2 * y + 1, t01);
          texelAt(&S, src, ws, 2 * x + 1, 2 * y + 1, t11);
          for (int k = 0; k < 4; k++) {  /* the bilinear sample with weights 1/2, 1/2 on both axes */
            float top = t00[k] * 0.5f + t10[k] * 0.5f, bot = t01[k] * 0.5f + t11[k] * 0.5f;
            r[k] = top * 0.5f + bot * 0.5f;
          }
          uint32_t ca = (uint32_t)floorf(fminf(fmaxf(r[3], 0.0f), 1.0f) * 255.0f + 0.5f);
          out[dst + y * w + x] = srgbEncode(lut, r[0]) | (srgbEncode(lut, r[1]) << 8) | (srgbEncode(lut, r[2]) << 16) |
                                 (ca << 24);
        }
    }
  }
}
/* Repeat addressing: the texel left of coordinate x wrapped into [0, w) and the right texel's weight. */
static uint32_t wrapCoord(float x, float w, float* f) {
  if (!(fabsf(x) < 1e30f)) x = 0.0f;
  float xf = floorf(x);
  *f = x - xf;
  return (uint32_t)(xf - w * floorf(xf * (1.0f / w)));
}
void orc_surf_sample(const orc_surface* s, uint32_t layer, float u, float v, uint32_t lod, float out[4]) {
  if (layer == 0) {
    out[0] = out[1] = out[2] = out[3] = 1.0f;
    return;
  }
  uint32_t m = lod < s->mips - 1 ? lod : s->mips - 1;
  uint32_t w = 1u << (s->size_log2 - m);
  float fw = (float)w, fu, fv;
  uint32_t base = (layer - 1) * s->layer_words + mipOff(s->size_log2, m);
  uint32_t x0 = wrapCoord(u * fw - 0.5f, fw, &fu), y0 = wrapCoord(v * fw - 0.5f, fw, &fv);
  uint32_t x1 = (x0 + 1) & (w - 1), y1 = (y0 + 1) & (w - 1);
  float t00[4], t10[4], t01[4], t11[4];
  texelAt(s, base, w, x0, y0, t00);
  texelAt(s, base, w, x1, y0, t10);
  texelAt(s, base, w, x0, y1, t01);
  texelAt(s, base, w, x1, y1, t11);
  float gu = 1.0f - fu, gv = 1.0f - fv;
  for (int k = 0; k < 4; k++) {
    float top = t00[k] * gu + t10[k] * fu, bot = t01[k] * gu + t11[k] * fu;
    out[k] = top * gv + bot * fv;
  }
}
/* tex of surfaceAt / shadeSample: the material's texture times baseColorFactor at the UV of barycentrics (wa, wb, wc)
 * of vertices i0, i1, i2 (UV(a) wa + UV(b) wb + UV(c) wc), on level 3 + L. */
static void surfBaseColor(const float* mat, uint32_t m, uint32_t i0, uint32_t i1, uint32_t i2, float wa, float wb,
                          float wc, uint32_t L, float out[4]) {
  const orc_surface* s = g_surf;
  float u = s->uv[i0 * 2] * wa + s->uv[i1 * 2] * wb + s->uv[i2 * 2] * wc;
  float v = s->uv[i0 * 2 + 1] * wa + s->uv[i1 * 2 + 1] * wb + s->uv[i2 * 2 + 1] * wc;
  float t[4];
  orc_surf_sample(s, s->mext[m * 4], u, v, 3 + L, t);
  for (int k = 0; k < 4; k++) out[k] = t[k] * mat[(size_t)m * 4 + k];
}
static float surfCutoff(uint32_t m) {
  float c;
  memcpy(&c, &g_surf->mext[m * 4 + 2], 4);
  return c;
}
void orc_closest_bary(const float p[3], const float a[3], const float b[3], const float c[3], float out[3]) {
  v3 r = closestBary(V3(p[0], p[1], p[2]), V3(a[0], a[1], a[2]), V3(b[0], b[1], b[2]), V3(c[0], c[1], c[2]));
  out[0] = r.x;
  out[1] = r.y;
  out[2] = r.z;
}

typedef struct { v3 a, b, c; uint32_t i0, i1, i2, m; } Tri;
static v3 P(const Ctx* X, uint32_t i) { return V3(X->pos[i * 3], X->pos[i * 3 + 1], X->pos[i * 3 + 2]); }
static v3 N(const Ctx* X, uint32_t i) { return V3(X->nrm[i * 3], X->nrm[i * 3 + 1], X->nrm[i * 3 + 2]); }
/* loadTri, voxel.ts:271-277 (instance 0: offset 0). */
static Tri loadTri(const Ctx* X, uint32_t key) {
  Tri T;
  T.i0 = X->tri[key * 4];
  T.i1 = X->tri[key * 4 + 1];
  T.i2 = X->tri[key * 4 + 2];
  T.m = X->tri[key * 4 + 3];
  T.a = P(X, T.i0);
  T.b = P(X, T.i1);
  T.c = P(X, T.i2);
  return T;
}
static v3 regionLo(const orc_vox* p, uint32_t L) { return vscale(vscale(vf(org(p, L)), voxSize(p, L)), 8.0f); }

typedef struct { iv3 lo, n; } Range;
/* voxRange, voxel.ts:293-300. */
static Range voxRange(const Ctx* X, Tri T, uint32_t L, float cell) {
  const orc_vox* p = X->p;
  v3 lo = regionLo(p, L);
  v3 mn = vmin(vmin(T.a, T.b), T.c), mx = vmax(vmax(T.a, T.b), T.c);
  int lim = f2i((float)p->G * voxSize(p, L) * 8.0f / cell);
  Range r;
  iv3 a = iclamp(vi(vfloor(vdivs(vsub(mn, lo), cell))), I3(0, 0, 0), I3(lim, lim, lim));
  iv3 b = iclamp(iadds(vi(vfloor(vdivs(vsub(mx, lo), cell))), 1), I3(0, 0, 0), I3(lim, lim, lim));
  r.lo = a;
  r.n = I3(imax(b.x - a.x, 0), imax(b.y - a.y, 0), imax(b.z - a.z, 0));
  return r;
}
/* finerHolds / finerHoldsRange / finerBuilt / finerBuiltRange, voxel.ts:303-320. */
static int finerHolds(const orc_vox* p, uint32_t L, iv3 c) {
  if (L == 0) return 0;
  iv3 f = isub(imuls(iadd(imuls(org(p, L), 8), c), 2), imuls(org(p, L - 1), 8));
  int E = p->G * 8;
  return f.x >= 0 && f.y >= 0 && f.z >= 0 && f.x < E && f.y < E && f.z < E;
}
static int finerHoldsRange(const orc_vox* p, uint32_t L, Range r, int s) {
  return finerHolds(p, L, imuls(r.lo, s)) && finerHolds(p, L, iadds(imuls(iadd(r.lo, r.n), s), -1));
}
static int finerBuilt(const Ctx* X, uint32_t L, iv3 c) {
  if (!finerHolds(X->p, L, c)) return 0;
  if (X->vox[2] == 0u) return 1;
  iv3 f = isub(imuls(iadd(imuls(org(X->p, L), 8), c), 2), imuls(org(X->p, L - 1), 8));
  return X->vox[voxPtrIndex(X, L - 1, ishr(f, 3))] != V_DROP;
}
static int finerBuiltRange(const Ctx* X, uint32_t L, Range r) { return X->vox[2] == 0u && finerHoldsRange(X->p, L, r, 1); }

static iv3 rangeCell(Range r, int q) { return iadd(r.lo, I3(q % r.n.x, (q / r.n.x) % r.n.y, q / (r.n.x * r.n.y))); }

/* markBrick / markRange, voxel.ts:347-357. */
static void markBrick(Ctx* X, uint32_t i) {
  if (X->vox[i] == V_EMPTY) X->vox[i] = V_MARK;
}
static void markRange(Ctx* X, Tri T, uint32_t L, Range r) {
  float bs = voxSize(X->p, L) * 8.0f;
  v3 lo = regionLo(X->p, L);
  int n = r.n.x * r.n.y * r.n.z;
  for (int j = 0; j < n; j++) {
    iv3 b = rangeCell(r, j);
    if (finerHolds(X->p, L, imuls(b, 8)) && finerHolds(X->p, L, iadds(imuls(b, 8), 7))) continue;
    float hh = 0.5f * bs * 1.0001f;
    if (n == 1 || triBox(vadd(lo, vscale(vadd(vf(b), V3(0.5f, 0.5f, 0.5f)), bs)), V3(hh, hh, hh), T.a, T.b, T.c))
      markBrick(X, voxPtrIndex(X, L, b));
  }
}
/* markBricks, voxel.ts:359-376 (one invocation per triangle). */
static void markBricks(Ctx* X, uint32_t t) {
  Tri T = loadTri(X, t);
  for (uint32_t L = 0; L < (uint32_t)X->p->levels; L++) {
    Range r = voxRange(X, T, L, voxSize(X->p, L) * 8.0f);
    int n = r.n.x * r.n.y * r.n.z;
    if (n == 0 || finerHoldsRange(X->p, L, r, 8)) continue;
    if (n > 8) {
      uint32_t k = X->vox[24 + L]++;
      if (k < BIG_CAP) X->big[8 + L][k] = t; else X->vox[3]++;
      continue;
    }
    markRange(X, T, L, r);
  }
}
/* markBricksBig, voxel.ts:377-387. */
static void markBricksBig(Ctx* X) {
  for (uint32_t L = 0; L < (uint32_t)X->p->levels; L++) {
    uint32_t n = X->vox[24 + L] < BIG_CAP ? X->vox[24 + L] : BIG_CAP;
    for (uint32_t k = 0; k < n; k++) {
      Tri T = loadTri(X, X->big[8 + L][k]);
      markRange(X, T, L, voxRange(X, T, L, voxSize(X->p, L) * 8.0f));
    }
  }
}
/* markParents, voxel.ts:390-403. */
static void markParents(Ctx* X) {
  uint32_t G = (uint32_t)X->p->G, per = G * G * G, NL = (uint32_t)X->p->levels;
  for (uint32_t i = 0; i < (NL - 1) * per; i++) {
    if (X->vox[X->ptrOff + i] != V_MARK) continue;
    uint32_t L = i / per, r = i % per;
    iv3 w = iadd(org(X->p, L), I3((int)(r % G), (int)((r / G) % G), (int)(r / (G * G))));
    for (L++; L < NL; L++) {
      w = ishr(w, 1);
      iv3 b = isub(w, org(X->p, L));
      if (b.x < 0 || b.y < 0 || b.z < 0 || b.x >= (int)G || b.y >= (int)G || b.z >= (int)G) break;
      markBrick(X, voxPtrIndex(X, L, b));
    }
  }
}
/* countMarks / scanMarks / allocBricks, voxel.ts:411-470, in the ordered form (coarse levels first, pointer index
 * ascending within a level; ranks past maxBricks dropped). */
static void allocBricks(Ctx* X) {
  uint32_t G = (uint32_t)X->p->G, per = G * G * G, NL = (uint32_t)X->p->levels, mb = (uint32_t)X->p->max_bricks;
  uint32_t total = 0;
  for (uint32_t i = 0; i < NL * per; i++) total += X->vox[X->ptrOff + i] == V_MARK;
  X->vox[5] = total;
  X->vox[0] = total;
  uint32_t idx = 0;
  for (int L = (int)NL - 1; L >= 0; L--)
    for (uint32_t r = 0; r < per; r++) {
      uint32_t i = (uint32_t)L * per + r;
      if (X->vox[X->ptrOff + i] != V_MARK) continue;
      uint32_t k = idx++;
      if (k >= mb) {
        X->vox[X->ptrOff + i] = V_DROP;
        X->vox[2]++;
        continue;
      }
      X->vox[8 + L]++;
      X->vox[X->ptrOff + i] = k;
      for (uint32_t w = 0; w < 17; w++) X->vox[X->brickOff + k * BW + w] = 0;
      X->vox[X->brickOff + k * BW + 17] =
          ((uint32_t)L << 27) | ((r % G) << 18) | (((r / G) % G) << 9) | (r / (G * G));
    }
}
static uint32_t nBricks(const Ctx* X) { return X->vox[0] < (uint32_t)X->p->max_bricks ? X->vox[0] : (uint32_t)X->p->max_bricks; }

/* voxelizeOne, voxel.ts:476-490, with surfaceAt's alpha test (voxel.ts:325-333) when a surface is set. */
static void voxelizeOne(Ctx* X, Tri T, uint32_t t, uint32_t L, iv3 c, int single) {
  if (finerBuilt(X, L, c)) return;
  float v = voxSize(X->p, L);
  v3 p = vadd(regionLo(X->p, L), vscale(vadd(vf(c), V3(0.5f, 0.5f, 0.5f)), v));
  float hh = 0.5f * v * 1.0001f;
  if (!single && !triBox(p, V3(hh, hh, hh), T.a, T.b, T.c)) return;
  uint32_t ptr = X->vox[voxPtrIndex(X, L, ishr(c, 3))];
  if (ptr >= V_MARK) return;
  uint32_t bit = (uint32_t)((c.x & 7) + 8 * (c.y & 7) + 64 * (c.z & 7));
  if (g_surf && (g_surf->mext[T.m * 4 + 1] & 1u)) {  /* alpha-masked: sampled at the closest point's UV */
    v3 w = closestBary(p, T.a, T.b, T.c);
    float tex[4];
    surfBaseColor(X->mat, T.m, T.i0, T.i1, T.i2, w.x, w.y, w.z, L, tex);
    if (tex[3] < surfCutoff(T.m)) return;
  }
  X->vox[X->brickOff + ptr * BW + (bit >> 5)] |= 1u << (bit & 31u);
  uint32_t* rp = &X->rep[ptr * 512u + bit];
  if (*rp < t + 1u) *rp = t + 1u;
}
/* voxelizeSmall, voxel.ts:492-510. */
static void voxelizeSmall(Ctx* X, uint32_t t) {
  Tri T = loadTri(X, t);
  for (uint32_t L = 0; L < (uint32_t)X->p->levels; L++) {
    Range r = voxRange(X, T, L, voxSize(X->p, L));
    int n = r.n.x * r.n.y * r.n.z;
    if (n == 0 || finerBuiltRange(X, L, r)) continue;
    if (n > BIG) {
      uint32_t k = X->vox[16 + L]++;
      if (k < BIG_CAP) X->big[L][k] = t; else X->vox[3]++;
      continue;
    }
    for (int k = 0; k < n; k++) voxelizeOne(X, T, t, L, rangeCell(r, k), n == 1);
  }
}
/* voxelizeBig, voxel.ts:515-531. */
static void voxelizeBig(Ctx* X) {
  for (uint32_t L = 0; L < (uint32_t)X->p->levels; L++) {
    uint32_t n = X->vox[16 + L] < BIG_CAP ? X->vox[16 + L] : BIG_CAP;
    for (uint32_t k = 0; k < n; k++) {
      uint32_t t = X->big[L][k];
      Tri T = loadTri(X, t);
      Range r = voxRange(X, T, L, voxSize(X->p, L));
      int m = r.n.x * r.n.y * r.n.z;
      for (int j = 0; j < m; j++) voxelizeOne(X, T, t, L, rangeCell(r, j), 0);
    }
  }
}
/* reduce, voxel.ts:540-577: level L from L-1, per occupancy word of a finer brick. */
static void reduceLevel(Ctx* X, uint32_t L) {
  uint32_t nb = nBricks(X);
  for (uint32_t i = 0; i < nb * 16u; i++) {
    uint32_t ptr = i >> 4, w = i & 15u;
    uint32_t base = X->brickOff + ptr * BW;
    uint32_t code = X->vox[base + 17];
    uint32_t Lc = code >> 27;
    if (Lc + 1 != L) continue;
    uint32_t word = X->vox[base + w];
    if (word == 0) continue;
    iv3 b = I3((int)((code >> 18) & 511u), (int)((code >> 9) & 511u), (int)(code & 511u));
    iv3 c0 = iadd(imuls(b, 8), I3(0, (int)((w & 1u) * 4u), (int)(w >> 1)));
    iv3 pc0 = isub(ishr(iadd(imuls(org(X->p, Lc), 8), c0), 1), imuls(org(X->p, L), 8));
    uint32_t pptr = X->vox[voxPtrIndex(X, L, ishr(pc0, 3))];
    if (pptr >= V_MARK) continue;
    uint32_t pbit0 = (uint32_t)((pc0.x & 7) + 8 * (pc0.y & 7) + 64 * (pc0.z & 7));
    uint32_t mask = 0, lo[4] = {0, 0, 0, 0}, hi[4] = {0, 0, 0, 0};
    while (word) {
      uint32_t k = (uint32_t)__builtin_ctz(word);
      word &= word - 1u;
      uint32_t px = (k & 7u) >> 1, py = k >> 4;
      mask |= 1u << ((pbit0 + px + 8u * py) & 31u);
      uint32_t rv = X->rep[ptr * 512u + w * 32u + k];
      if (py == 0) { if (rv > lo[px]) lo[px] = rv; } else { if (rv > hi[px]) hi[px] = rv; }
    }
    X->vox[X->brickOff + pptr * BW + (pbit0 >> 5)] |= mask;
    for (uint32_t px = 0; px < 4; px++)
      for (uint32_t py = 0; py < 2; py++) {
        uint32_t rv = py == 0 ? lo[px] : hi[px];
        if (rv == 0) continue;
        uint32_t* rp = &X->rep[pptr * 512u + pbit0 + px + 8u * py];
        if (*rp < rv) *rp = rv;
      }
  }
}
/* allocVoxels, voxel.ts:579-588, bricks in slot order. */
static void allocVoxels(Ctx* X) {
  uint32_t nb = nBricks(X);
  for (uint32_t b = 0; b < nb; b++) {
    uint32_t base = X->brickOff + b * BW, n = 0;
    for (uint32_t w = 0; w < 16; w++) n += (uint32_t)popc(X->vox[base + w]);
    uint32_t first = X->vox[1];
    X->vox[1] += n;
    if (first + n > (uint32_t)X->p->max_voxels) {
      X->vox[base + 16] = V_EMPTY;
      X->vox[4]++;
      continue;
    }
    X->vox[base + 16] = first;
  }
}
/* WGSL pack4x8unorm / pack2x16snorm by the specification's formulas. */
static uint32_t pack4x8unorm(const float e[4]) {
  uint32_t r = 0;
  for (int i = 0; i < 4; i++) r |= ((uint32_t)floorf(0.5f + 255.0f * fmn(1.0f, fmx(0.0f, e[i])))) << (8 * i);
  return r;
}
static uint32_t pack2x16snorm(float x, float y) {
  int a = (int)floorf(0.5f + 32767.0f * fmn(1.0f, fmx(-1.0f, x)));
  int b = (int)floorf(0.5f + 32767.0f * fmn(1.0f, fmx(-1.0f, y)));
  return ((uint32_t)a & 0xffffu) | (((uint32_t)b & 0xffffu) << 16);
}
uint32_t orc_pack4x8unorm(const float e[4]) { return pack4x8unorm(e); }
uint32_t orc_pack2x16snorm(float x, float y) { return pack2x16snorm(x, y); }
/* octEnc, voxel.ts:584-588. */
static void octEnc(v3 n, float* px, float* py) {
  float s = fabsf(n.x) + fabsf(n.y) + fabsf(n.z);
  float x = n.x / s, y = n.y / s;
  if (n.z < 0.0f) {
    float ox = x, oy = y;
    x = (1.0f - fabsf(oy)) * (ox >= 0.0f ? 1.0f : -1.0f);
    y = (1.0f - fabsf(ox)) * (oy >= 0.0f ? 1.0f : -1.0f);
  }
  *px = x;
  *py = y;
}
void orc_oct_enc(const float n[3], float out[2]) { octEnc(V3(n[0], n[1], n[2]), &out[0], &out[1]); }
/* resolveAttr, voxel.ts:590-620, with surfaceAt(wantAttr) (voxel.ts:325-336): albedo = texture x baseColorFactor
 * (white texture when untextured), the toon flag in the albedo word's alpha. */
static void resolveAttr(Ctx* X) {
  uint32_t nb = nBricks(X);
  for (uint32_t i = 0; i < nb * 512u; i++) {
    uint32_t ptr = i >> 9, bit = i & 511u;
    uint32_t base = X->brickOff + ptr * BW, w = bit >> 5;
    uint32_t word = X->vox[base + w];
    if ((word & (1u << (bit & 31u))) == 0) continue;
    uint32_t first = X->vox[base + 16];
    if (first == V_EMPTY) continue;
    uint32_t r = (uint32_t)popc(word & ((1u << (bit & 31u)) - 1u));
    for (uint32_t j = 0; j < w; j++) r += (uint32_t)popc(X->vox[base + j]);
    uint32_t ai = first + r;
    uint32_t code = X->vox[base + 17];
    uint32_t L = code >> 27;
    iv3 b = I3((int)((code >> 18) & 511u), (int)((code >> 9) & 511u), (int)(code & 511u));
    iv3 c = iadd(imuls(b, 8), I3((int)(bit & 7u), (int)((bit >> 3) & 7u), (int)(bit >> 6)));
    v3 p = vadd(regionLo(X->p, L), vscale(vadd(vf(c), V3(0.5f, 0.5f, 0.5f)), voxSize(X->p, L)));
    Tri T = loadTri(X, X->rep[ptr * 512u + bit] - 1u);
    v3 wb = closestBary(p, T.a, T.b, T.c);
    v3 n = vadd(vadd(vscale(N(X, T.i0), wb.x), vscale(N(X, T.i1), wb.y)), vscale(N(X, T.i2), wb.z));
    if (!(vdot(n, n) > 1e-12f)) n = vcross(vsub(T.b, T.a), vsub(T.c, T.a));
    n = vnormalize(n);
    v3 q = vadd(vadd(vscale(T.a, wb.x), vscale(T.b, wb.y)), vscale(T.c, wb.z));
    const float* m = X->mat + (size_t)T.m * 4;
    float alb[4] = {1.0f * m[0], 1.0f * m[1], 1.0f * m[2], 0.0f};
    if (g_surf) {
      float tex[4];
      surfBaseColor(X->mat, T.m, T.i0, T.i1, T.i2, wb.x, wb.y, wb.z, L, tex);
      alb[0] = tex[0];
      alb[1] = tex[1];
      alb[2] = tex[2];
      alb[3] = (g_surf->mext[T.m * 4 + 1] & 4u) ? 1.0f : 0.0f;
    }
    float ox, oy;
    octEnc(n, &ox, &oy);
    X->vox[X->attrOff + ai * 2] = pack4x8unorm(alb);
    X->vox[X->attrOff + ai * 2 + 1] = pack2x16snorm(ox, oy);
    const uint32_t ao = albOff(X->p);
    X->vox[ao + ai * 2] = (uint32_t)f2bf(alb[0]) | ((uint32_t)f2bf(alb[1]) << 16);
    X->vox[ao + ai * 2 + 1] = (uint32_t)f2bf(alb[2]) | ((uint32_t)f2bf(alb[3]) << 16);
    float pl = vdot(n, vsub(q, p));
    memcpy(&X->vox[X->plOff + ai], &pl, 4);
  }
}

int orc_build(const orc_vox* p, const float* pos, const float* nrm, const uint32_t* tri, uint32_t ntris,
              const float* mat, uint32_t* vox, uint32_t* rep) {
  if (p->levels < 1 || p->levels > 8) return -1;
  Ctx X;
  memset(&X, 0, sizeof(X));
  X.p = p;
  X.vox = vox;
  X.rep = rep;
  X.pos = pos;
  X.nrm = nrm;
  X.tri = tri;
  X.mat = mat;
  X.ntris = ntris;
  layout(p, &X.ptrOff, &X.brickOff, &X.attrOff, &X.plOff);
  uint32_t cap = ntris < BIG_CAP ? ntris : BIG_CAP;
  for (int s = 0; s < 16; s++) {
    X.big[s] = (uint32_t*)malloc(sizeof(uint32_t) * (cap ? cap : 1));
    if (!X.big[s]) {
      for (int q = 0; q < s; q++) free(X.big[q]);
      return -2;
    }
  }
  /* clear, voxel.ts:338-344 */
  uint32_t G = (uint32_t)p->G, NLper = (uint32_t)p->levels * G * G * G;
  memset(vox, 0, orc_vox_words(p) * 4);
  for (uint32_t i = 0; i < NLper; i++) vox[X.ptrOff + i] = V_EMPTY;
  for (uint32_t t = 0; t < ntris; t++) markBricks(&X, t);
  markBricksBig(&X);
  if (p->levels > 1) markParents(&X);
  allocBricks(&X);
  /* clearKeys, voxel.ts:471-474 */
  memset(rep, 0, (size_t)nBricks(&X) * 512u * 4u);
  for (uint32_t t = 0; t < ntris; t++) voxelizeSmall(&X, t);
  voxelizeBig(&X);
  for (uint32_t L = 1; L < (uint32_t)p->levels; L++) reduceLevel(&X, L);
  allocVoxels(&X);
  resolveAttr(&X);
  for (int s = 0; s < 16; s++) free(X.big[s]);
  return 0;
}

int orc_canonical(const orc_vox* p, const uint32_t* vox, const uint32_t* rep, int level, uint32_t* occ_bits,
                  uint32_t* rep_out, uint32_t* albedo, uint32_t* normal, float* plane, uint8_t* brick_state) {
  uint32_t ptrOff, brickOff, attrOff, plOff;
  layout(p, &ptrOff, &brickOff, &attrOff, &plOff);
  int G = p->G, E = G * 8;
  uint32_t per = (uint32_t)(G * G * G);
  uint32_t nb = vox[0] < (uint32_t)p->max_bricks ? vox[0] : (uint32_t)p->max_bricks;
  size_t ncell = (size_t)E * E * E;
  if (occ_bits) memset(occ_bits, 0, ncell / 8);
  if (rep_out) memset(rep_out, 0, ncell * 4);
  if (albedo) memset(albedo, 0, ncell * 4);
  if (normal) memset(normal, 0, ncell * 4);
  if (plane) memset(plane, 0, ncell * 4);
  for (int bz = 0; bz < G; bz++)
    for (int by = 0; by < G; by++)
      for (int bx = 0; bx < G; bx++) {
        uint32_t bi = ((uint32_t)bz * G + by) * G + bx;
        uint32_t ptr = vox[ptrOff + (uint32_t)level * per + bi];
        uint8_t st = ptr == V_DROP ? 2 : (ptr >= V_MARK ? 0 : 1);
        if (st == 1 && ptr >= nb) return -1;
        if (st == 1 && vox[brickOff + ptr * BW + 16] == V_EMPTY) st = 3;
        if (brick_state) brick_state[bi] = st;
        if (st != 1 && st != 3) continue;
        uint32_t rank = 0;
        for (int bit = 0; bit < 512; bit++) {
          uint32_t word = vox[brickOff + ptr * BW + (bit >> 5)];
          if (!(word & (1u << (bit & 31)))) continue;
          int x = bx * 8 + (bit & 7), y = by * 8 + ((bit >> 3) & 7), z = bz * 8 + (bit >> 6);
          size_t ci = (size_t)x + (size_t)E * ((size_t)y + (size_t)E * z);
          if (occ_bits) occ_bits[ci >> 5] |= 1u << (ci & 31);
          if (rep_out) rep_out[ci] = rep[(size_t)ptr * 512 + bit];
          if (st == 1) {
            uint32_t ai = vox[brickOff + ptr * BW + 16] + rank;
            if (albedo) albedo[ci] = vox[attrOff + ai * 2];
            if (normal) normal[ci] = vox[attrOff + ai * 2 + 1];
            if (plane) memcpy(&plane[ci], &vox[plOff + ai], 4);
          }
          rank++;
        }
      }
  return 0;
}

/* ------------------------------------------------------------------------------------------------------------ */
/* Cascades.                                                                                                     */

/* binary16 conversion with round to nearest even (pack2x16float / unpack2x16float, rc.ts:309-310). */
uint16_t orc_f2h(float f) {
  uint32_t x;
  memcpy(&x, &f, 4);
  uint32_t sign = (x >> 16) & 0x8000u, a = x & 0x7fffffffu;
  if (a >= 0x7f800000u) return (uint16_t)(sign | 0x7c00u | (a > 0x7f800000u ? 0x200u : 0u));
  if (a >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u); /* >= 65520 rounds to infinity */
  if (a >= 0x38800000u) {                                  /* normal */
    uint32_t r = a - 0x38000000u;
    r += 0xfffu + ((r >> 13) & 1u);
    return (uint16_t)(sign | (r >> 13));
  }
  if (a < 0x33000000u) return (uint16_t)sign; /* below 2^-25 */
  uint32_t e = a >> 23, m = (a & 0x7fffffu) | 0x800000u, shift = 126u - e;
  uint32_t q = m >> shift, rem = m & ((1u << shift) - 1u), half = 1u << (shift - 1u);
  if (rem > half || (rem == half && (q & 1u))) q++;
  return (uint16_t)(sign | q);
}
float orc_h2f(uint16_t h) {
  uint32_t s = (uint32_t)(h >> 15) << 31, e = (h >> 10) & 31u, m = h & 1023u, bits;
  if (e == 0) {
    if (m == 0) {
      bits = s;
    } else {
      e = 113;
      while (!(m & 1024u)) {
        m <<= 1;
        e--;
      }
      bits = s | (e << 23) | ((m & 1023u) << 13);
    }
  } else if (e == 31) {
    bits = s | 0x7f800000u | (m << 13);
  } else {
    bits = s | ((e + 112u) << 23) | (m << 13);
  }
  float f;
  memcpy(&f, &bits, 4);
  return f;
}
static float q16(const orc_rc* rc, float x) { return rc->fp16 ? orc_h2f(orc_f2h(x)) : x; }

/* octEncode (frame.ts:97-103) in double; octSolidAngles (frame.ts:105-114). */
static void octEncodeD(double x, double y, double z, double* u, double* v) {
  double s = fabs(x) + fabs(y) + fabs(z);
  double px = x / s, py = y / s;
  if (z < 0) {
    double ox = px, oy = py;
    px = (1 - fabs(oy)) * (ox >= 0 ? 1 : -1);
    py = (1 - fabs(ox)) * (oy >= 0 ? 1 : -1);
  }
  *u = px * 0.5 + 0.5;
  *v = py * 0.5 + 0.5;
}
void orc_oct_solid_angles(int res, int samples, float* out) {
  double* counts = (double*)calloc((size_t)res * res, sizeof(double));
  double golden = M_PI * (3 - sqrt(5.0));
  for (int i = 0; i < samples; i++) {
    double z = 1 - (2.0 * i + 1) / samples, r = sqrt(1 - z * z), phi = golden * i, u, v;
    octEncodeD(r * cos(phi), r * sin(phi), z, &u, &v);
    int iv = (int)floor(v * res), iu = (int)floor(u * res);
    if (iv > res - 1) iv = res - 1;
    if (iu > res - 1) iu = res - 1;
    counts[iv * res + iu] += 1;
  }
  for (int i = 0; i < res * res; i++) out[i] = (float)(counts[i] / samples * 4 * M_PI);
  free(counts);
}

static int resOf(const orc_rc* rc, int c) { return 4 << (c < rc->res_cap ? c : rc->res_cap); }
static int probesOf(const orc_rc* rc, int c) { return rc->probes0 >> c; }
uint64_t orc_rays(const orc_rc* rc, int c) {
  uint64_t n = (uint64_t)probesOf(rc, c), r = (uint64_t)resOf(rc, c);
  return n * n * n * r * r;
}
static uint64_t rayOff(const orc_rc* rc, int c) {
  uint64_t o = 0;
  for (int k = 0; k < c; k++) o += orc_rays(rc, k);
  return o;
}

void orc_cascade_setup(orc_rc* rc, double v0, int levels, double s0, double k_interval, double level_map,
                       double vox_scale, const double eye[3]) {
  rc->s0 = (float)s0;
  for (int c = 0; c < rc->cascades; c++) {
    rc->b0[c] = (float)(k_interval * (pow(4.0, c) - 1) / 3);
    rc->b1[c] = (float)(k_interval * (pow(4.0, c + 1) - 1) / 3);
    /* voxLevel, rc.ts:210-213, in double; round half to even like WGSL round(). */
    double v = vox_scale * s0 * pow(2.0, level_map * c);
    double L = nearbyint(log2(v / v0));
    if (L < 0) L = 0;
    if (L > levels - 1) L = levels - 1;
    rc->vox_level[c] = (int)L;
  }
  /* Dense grid snapped to the top cascade's spacing (every grid nests in the next). */
  int top = rc->cascades - 1;
  double st = s0 * pow(2.0, top);
  int ntop = rc->probes0 >> top;
  for (int a = 0; a < 3; a++) rc->G0[a] = (int32_t)(floor(eye[a] / st) - ntop / 2) * (1 << top);
}

/* dirOf / octDecode, rc.ts:291-298, 308. */
static v3 dirOf(int di, int res) {
  float u = ((float)(di % res) + 0.5f) / (float)res, w = ((float)(di / res) + 0.5f) / (float)res;
  float fx = u * 2.0f - 1.0f, fy = w * 2.0f - 1.0f;
  v3 n = V3(fx, fy, 1.0f - fabsf(fx) - fabsf(fy));
  float t = fmx(-n.z, 0.0f);
  n.x += n.x >= 0.0f ? -t : t;
  n.y += n.y >= 0.0f ? -t : t;
  return vnormalize(n);
}

typedef struct { const orc_vox* p; const uint32_t* vox; uint32_t ptrOff, brickOff, attrOff, plOff, alOff; } View;

static uint32_t vPtr(const View* W, uint32_t L, iv3 b) {
  uint32_t G = (uint32_t)W->p->G;
  return W->vox[W->ptrOff + L * G * G * G + ((uint32_t)b.z * G + (uint32_t)b.y) * G + (uint32_t)b.x];
}
/* voxRank, voxel.ts:81-87; a brick without attribute slots has none (returns V_EMPTY). */
static uint32_t voxRank(const View* W, uint32_t brick, uint32_t bit) {
  uint32_t base = W->brickOff + brick * BW, w = bit >> 5;
  if (W->vox[base + 16] == V_EMPTY) return V_EMPTY;
  uint32_t r = (uint32_t)popc(W->vox[base + w] & ((1u << (bit & 31u)) - 1u));
  for (uint32_t i = 0; i < w; i++) r += (uint32_t)popc(W->vox[base + i]);
  return W->vox[base + 16] + r;
}
/* voxAt, voxel.ts:90-107. Returns 1 with (brick, bit, level, cell) for an occupied voxel. */
static int voxAt(const View* W, uint32_t L, iv3 c, uint32_t* brick, uint32_t* bit, uint32_t* lev, iv3* cell) {
  uint32_t l = L;
  iv3 cc = c;
  int E = W->p->G * 8;
  for (;;) {
    uint32_t ptr = vPtr(W, l, ishr(cc, 3));
    if (ptr != V_DROP) {
      if (ptr >= V_MARK) return 0;
      uint32_t b = (uint32_t)((cc.x & 7) + 8 * (cc.y & 7) + 64 * (cc.z & 7));
      if ((W->vox[W->brickOff + ptr * BW + (b >> 5)] & (1u << (b & 31u))) == 0u) return 0;
      *brick = ptr; *bit = b; *lev = l; *cell = cc;
      return 1;
    }
    if (l + 1u >= (uint32_t)W->p->levels) return 0;
    iv3 w = ishr(iadd(imuls(org(W->p, l), 8), cc), 1);
    l++;
    cc = isub(w, imuls(org(W->p, l), 8));
    if (cc.x < 0 || cc.y < 0 || cc.z < 0 || cc.x >= E || cc.y >= E || cc.z >= E) return 0;
  }
}

#define PI_F ((float)3.14159265)

static int bnInside(iv3 c) { return c.x >= 0 && c.y >= 0 && c.z >= 0 && c.x < ORC_BN_DX && c.y < ORC_BN_DY && c.z < ORC_BN_DZ; }
static const uint16_t* bnTexel(const uint16_t* vol, iv3 c, int a) {
  return vol + ((size_t)(c.x + ORC_BN_DX * (c.y + ORC_BN_DY * c.z)) * 6 + (size_t)a) * 4;
}
static float bnReach(const uint16_t* vol, iv3 c, int a) { return orc_h2f(bnTexel(vol, c, a)[3]); }
static v3 bnRgb(const uint16_t* vol, iv3 c, int a) {
  const uint16_t* t = bnTexel(vol, c, a);
  return V3(orc_h2f(t[0]), orc_h2f(t[1]), orc_h2f(t[2]));
}
/* WGSL smoothstep. */
static float smoothStep(float lo, float hi, float x) {
  float t = fmn(fmx((x - lo) / (hi - lo), 0.0f), 1.0f);
  return t * t * (3.0f - 2.0f * t);
}
/* bnRead (bounce.ts:49-75), tol = half a level-0 voxel. */
static v3 bnRead(const uint16_t* vol, const int32_t org3[3], float tol, v3 q, v3 p, v3 n) {
  iv3 o = I3(org3[0], org3[1], org3[2]);
  v3 g = vsub(vdivs(q, ORC_BN_CELL), vf(o));
  iv3 bb = vi(vfloor(g));
  v3 fr = vsub(g, vf(bb));
  v3 n2 = vmul(n, n);
  v3 e = V3(0, 0, 0);
  float ws = 0.0f;
  for (int k = 0; k < 8; k++) {
    iv3 off = I3(k & 1, (k >> 1) & 1, (k >> 2) & 1);
    iv3 cc = iadd(bb, off);
    if (!bnInside(cc)) continue;
    v3 pp = vscale(vf(iadd(o, cc)), ORC_BN_CELL);
    v3 u = vsub(q, pp);
    if (!(bnReach(vol, cc, 0) > 0.0f)) continue;
    if (fabsf(u.x) > bnReach(vol, cc, u.x >= 0.0f ? 0 : 1) + tol || fabsf(u.y) > bnReach(vol, cc, u.y >= 0.0f ? 2 : 3) + tol ||
        fabsf(u.z) > bnReach(vol, cc, u.z >= 0.0f ? 4 : 5) + tol)
      continue;
    v3 t3 = V3(off.x ? fr.x : 1.0f - fr.x, off.y ? fr.y : 1.0f - fr.y, off.z ? fr.z : 1.0f - fr.z);
    float w = t3.x * t3.y * t3.z * smoothStep(-0.05f, 0.25f, vdot(vsub(pp, p), n) / ORC_BN_CELL);
    if (w <= 0.0f) continue;
    v3 A = bnRgb(vol, cc, n.x >= 0.0f ? 0 : 1), B = bnRgb(vol, cc, n.y >= 0.0f ? 2 : 3), C = bnRgb(vol, cc, n.z >= 0.0f ? 4 : 5);
    v3 sum = vadd(vadd(vscale(A, n2.x), vscale(B, n2.y)), vscale(C, n2.z));
    e = vadd(e, vscale(sum, w));
    ws += w;
  }
  return ws > 1e-6f ? vdivs(e, ws) : V3(0, 0, 0);
}

/* dynE of the bounce volume (bounce.ts:267-298): the irradiance a dynamic voxel at p with the projected areas P
 * receives: per axis face, the mean over the probes around p that are active, see p within their reach (+ tol) and
 * lie on that face's side (smoothstep weight) of their texel on that axis; the faces weighted by P. */
static v3 dynE(const uint16_t* vol, const int32_t org3[3], float tol, v3 p, const float P[6]) {
  iv3 o = I3(org3[0], org3[1], org3[2]);
  v3 g = vsub(vdivs(p, ORC_BN_CELL), vf(o));
  iv3 bb = vi(vfloor(g));
  v3 fr = vsub(g, vf(bb));
  v3 e[6];
  float ws[6];
  for (int a = 0; a < 6; a++) {
    e[a] = V3(0, 0, 0);
    ws[a] = 0.0f;
  }
  for (int k = 0; k < 8; k++) {
    iv3 off = I3(k & 1, (k >> 1) & 1, (k >> 2) & 1);
    iv3 cc = iadd(bb, off);
    if (!bnInside(cc)) continue;
    v3 pp = vscale(vf(iadd(o, cc)), ORC_BN_CELL);
    v3 u = vsub(p, pp);
    if (!(bnReach(vol, cc, 0) > 0.0f)) continue;
    if (fabsf(u.x) > bnReach(vol, cc, u.x >= 0.0f ? 0 : 1) + tol || fabsf(u.y) > bnReach(vol, cc, u.y >= 0.0f ? 2 : 3) + tol ||
        fabsf(u.z) > bnReach(vol, cc, u.z >= 0.0f ? 4 : 5) + tol)
      continue;
    v3 t3 = V3(off.x ? fr.x : 1.0f - fr.x, off.y ? fr.y : 1.0f - fr.y, off.z ? fr.z : 1.0f - fr.z);
    float wt = t3.x * t3.y * t3.z;
    for (int a = 0; a < 6; a++) {
      if (P[a] <= 0.0f) continue;
      float sg = (a & 1) ? -1.0f : 1.0f;
      v3 n = V3(a / 2 == 0 ? sg : 0.0f, a / 2 == 1 ? sg : 0.0f, a / 2 == 2 ? sg : 0.0f);
      float w = wt * smoothStep(-0.05f, 0.25f, vdot(vsub(pp, p), n) / ORC_BN_CELL);
      if (w <= 0.0f) continue;
      e[a] = vadd(e[a], vscale(bnRgb(vol, cc, a), w));
      ws[a] += w;
    }
  }
  v3 s = V3(0, 0, 0);
  float ps = 0.0f;
  for (int a = 0; a < 6; a++) {
    if (ws[a] > 1e-6f) s = vadd(s, vdivs(vscale(e[a], P[a]), ws[a]));
    ps += P[a];
  }
  return vdivs(s, fmx(ps, 1e-9f));
}

/* shadeHitV (rc.ts:108-116) at a voxel hit with the voxel's attributes (voxel.ts:150-156, 185-192), no bounce,
 * the frame's sun visibility; plus the point lights of this path (none: the reference shading). */
static v3 shadeVoxel(const View* W, const orc_rc* rc, uint32_t ai, v3 d, v3 entry, uint32_t br, uint32_t bit) {
  uint32_t pa = 0, pn = 0;
  if (ai != V_EMPTY) {
    pa = W->vox[W->attrOff + ai * 2];
    pn = W->vox[W->attrOff + ai * 2 + 1];
  }
  v3 alb = V3((float)(pa & 255u) / 255.0f, (float)((pa >> 8) & 255u) / 255.0f, (float)((pa >> 16) & 255u) / 255.0f);
  if (rc->bf16) { /* precision form: the albedo as bf16 */
    uint32_t w0 = ai != V_EMPTY ? W->vox[W->alOff + ai * 2] : 0u, w1 = ai != V_EMPTY ? W->vox[W->alOff + ai * 2 + 1] : 0u;
    alb = V3(bf2f((uint16_t)(w0 & 0xffffu)), bf2f((uint16_t)(w0 >> 16)), bf2f((uint16_t)(w1 & 0xffffu)));
  }
  float nx = fmx((float)(int16_t)(uint16_t)(pn & 0xffffu) / 32767.0f, -1.0f);
  float ny = fmx((float)(int16_t)(uint16_t)(pn >> 16) / 32767.0f, -1.0f);
  v3 n0 = V3(nx, ny, 1.0f - fabsf(nx) - fabsf(ny));
  float q = fmx(-n0.z, 0.0f);
  n0.x += n0.x >= 0.0f ? -q : q;
  n0.y += n0.y >= 0.0f ? -q : q;
  v3 n = vdot(n0, n0) > 0.01f ? n0 : vneg(d);
  if (vdot(n, d) > 0.0f) n = vneg(n);
  v3 e = V3(0, 0, 0);
  if (rc->bn_prev && ai != V_EMPTY) {
    /* voxBounce of bounce.ts:96-99: the previous volume read at the shading point, on the side the ray comes from. */
    uint32_t code = W->vox[W->brickOff + br * BW + 17];
    uint32_t L = code >> 27;
    iv3 bc = I3((int)((code >> 18) & 511u), (int)((code >> 9) & 511u), (int)(code & 511u));
    iv3 c = iadd(imuls(bc, 8), I3((int)(bit & 7u), (int)((bit >> 3) & 7u), (int)(bit >> 6)));
    float v = voxSize(W->p, L);
    v3 pc = vadd(vscale(vscale(vf(org(W->p, L)), v), 8.0f), vscale(vadd(vf(c), V3(0.5f, 0.5f, 0.5f)), v));
    v3 p = vsub(entry, vscale(d, 0.25f * v));
    v3 nb;
    if (vdot(n0, n0) <= 0.01f) {
      nb = vneg(d);
    } else {
      float sd = vdot(vsub(p, pc), n0);
      float side = fabsf(sd) > 0.05f * v ? (sd > 0.0f ? 1.0f : -1.0f) : (vdot(n0, d) > 0.0f ? -1.0f : 1.0f);
      nb = vscale(n0, side);
    }
    e = bnRead(rc->bn_prev, rc->bn_prev_org, 0.5f * W->p->v0, p, p, nb);
  } else if (rc->bounce && ai != V_EMPTY) {
    /* voxBounce / voxSide (rc.ts:122-126, 198-202) at a point a quarter voxel back from the entry (voxShade). */
    uint32_t code = W->vox[W->brickOff + br * BW + 17];
    uint32_t L = code >> 27;
    iv3 bc = I3((int)((code >> 18) & 511u), (int)((code >> 9) & 511u), (int)(code & 511u));
    iv3 c = iadd(imuls(bc, 8), I3((int)(bit & 7u), (int)((bit >> 3) & 7u), (int)(bit >> 6)));
    float v = voxSize(W->p, L);
    v3 pc = vadd(vscale(vscale(vf(org(W->p, L)), v), 8.0f), vscale(vadd(vf(c), V3(0.5f, 0.5f, 0.5f)), v));
    v3 p = vsub(entry, vscale(d, 0.25f * v));
    float sd = vdot(vsub(p, pc), n0);
    int back = fabsf(sd) > 0.05f * v ? !(sd > 0.0f) : vdot(n0, d) > 0.0f;
    const uint32_t* B = rc->bounce + (size_t)ai * 4 + (back ? 2 : 0);
    if (rc->bf16)
      e = V3(bf2f((uint16_t)(B[0] & 0xffffu)), bf2f((uint16_t)(B[0] >> 16)), bf2f((uint16_t)(B[1] & 0xffffu)));
    else
      e = V3(orc_h2f((uint16_t)(B[0] & 0xffffu)), orc_h2f((uint16_t)(B[0] >> 16)), orc_h2f((uint16_t)(B[1] & 0xffffu)));
  }
  v3 b = vmul(vdivs(alb, PI_F), e);
  v3 sd = V3(rc->sun_dir[0], rc->sun_dir[1], rc->sun_dir[2]);
  float ndl = fmx(vdot(n, sd), 0.0f);
  v3 out = b;
  float vis = (rc->sunvis && ai != V_EMPTY) ? rc->sunvis[ai] : rc->sun_vis;
  if (ndl > 0.0f)
    out = vadd(b, vscale(vscale(vmul(vdivs(alb, PI_F), V3(rc->sun_e[0], rc->sun_e[1], rc->sun_e[2])), ndl), vis));
  for (int k = 0; k < rc->n_point; k++) {
    v3 l = vsub(V3(rc->point_pos[k][0], rc->point_pos[k][1], rc->point_pos[k][2]), entry);
    float r2 = vdot(l, l);
    float nl = fmx(vdot(n, l) / sqrtf(r2), 0.0f);
    if (nl > 0.0f)
      out = vadd(out, vscale(vmul(vdivs(alb, PI_F), V3(rc->point_i[k][0], rc->point_i[k][1], rc->point_i[k][2])), nl / r2));
  }
  return out;
}

static uint32_t hitId(const orc_vox* p, uint32_t L, iv3 c) {
  int E = p->G * 8;
  return (L << 29) | (uint32_t)(c.x + E * (c.y + E * c.z));
}

/* traceVox, voxel.ts:121-202. out: rgb, T. */
static void traceVoxT(const View* W, const orc_rc* rc, v3 o, v3 d, float t0, float t1, float tSkip, uint32_t L,
                      float out[4], uint32_t* hid, uint32_t* steps, float* tin) {
  const orc_vox* p = W->p;
  *hid = 0xffffffffu;
  *steps = 0;
  out[0] = out[1] = out[2] = 0.0f;
  out[3] = 1.0f;
  float v = voxSize(p, L);
  int G = p->G;
  v3 lo = vscale(vscale(vf(org(p, L)), v), 8.0f);
  float span = (float)G * v * 8.0f;
  v3 hi = vadd(lo, V3(span, span, span));
  v3 dd = V3(fabsf(d.x) < 1e-9f ? 1e-9f : d.x, fabsf(d.y) < 1e-9f ? 1e-9f : d.y, fabsf(d.z) < 1e-9f ? 1e-9f : d.z);
  v3 inv = V3(1.0f / dd.x, 1.0f / dd.y, 1.0f / dd.z);
  v3 ta = vmul(vsub(lo, o), inv), tb = vmul(vsub(hi, o), inv);
  v3 tn = vmin(ta, tb), tf = vmax(ta, tb);
  float t = fmx(fmx(fmx(tn.x, tn.y), tn.z), t0);
  float tEnd = fmn(fmn(fmn(tf.x, tf.y), tf.z), t1);
  if (t >= tEnd) return;
  iv3 st = I3(dd.x > 0 ? 1 : -1, dd.y > 0 ? 1 : -1, dd.z > 0 ? 1 : -1);
  v3 sp = V3(st.x > 0 ? 1.0f : 0.0f, st.y > 0 ? 1.0f : 0.0f, st.z > 0 ? 1.0f : 0.0f);
  int nVox = G * 8;
  iv3 c = iclamp(vi(vfloor(vdivs(vsub(vadd(o, vscale(d, t)), lo), v))), I3(0, 0, 0), I3(nVox - 1, nVox - 1, nVox - 1));
  v3 tMax = vmul(vsub(vadd(vscale(vadd(vf(c), sp), v), lo), o), inv);
  v3 tDelta = vabs(vscale(inv, v));
  float tIn = t;
  for (int it = 0; it < 1024; it++) {
    (*steps)++; /* voxSteps (voxel.ts:143) */
    iv3 b = ishr(c, 3);
    uint32_t ptr = vPtr(W, L, b);
    if (ptr == V_DROP) {
      uint32_t br, bit, lev;
      iv3 cell;
      if (voxAt(W, L, c, &br, &bit, &lev, &cell) && tIn >= tSkip) {
        *hid = hitId(p, lev, cell);
        if (tin) *tin = tIn;
        v3 r = shadeVoxel(W, rc, voxRank(W, br, bit), d, vadd(o, vscale(d, tIn)), br, bit);
        out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = 0.0f;
        return;
      }
      float tx = fmn(fmn(tMax.x, tMax.y), tMax.z);
      if (tx >= tEnd) break;
      tIn = tx;
      if (tMax.x <= tMax.y && tMax.x <= tMax.z) { c.x += st.x; tMax.x += tDelta.x; }
      else if (tMax.y <= tMax.z) { c.y += st.y; tMax.y += tDelta.y; }
      else { c.z += st.z; tMax.z += tDelta.z; }
      if (c.x < 0 || c.y < 0 || c.z < 0 || c.x >= nVox || c.y >= nVox || c.z >= nVox) break;
      continue;
    }
    if (ptr >= V_MARK) {
      v3 bt = vmul(vsub(vadd(vscale(vadd(vf(imuls(b, 8)), vscale(sp, 8.0f)), v), lo), o), inv);
      float te = fmn(fmn(bt.x, bt.y), bt.z);
      if (te >= tEnd) break;
      iv3 nb = b;
      if (bt.x <= bt.y && bt.x <= bt.z) nb.x += st.x; else if (bt.y <= bt.z) nb.y += st.y; else nb.z += st.z;
      if (nb.x < 0 || nb.y < 0 || nb.z < 0 || nb.x >= G || nb.y >= G || nb.z >= G) break;
      t = te;
      tIn = te;
      c = iclamp(vi(vfloor(vdivs(vsub(vadd(o, vscale(d, t)), lo), v))), imuls(nb, 8), iadds(imuls(nb, 8), 7));
      tMax = vmul(vsub(vadd(vscale(vadd(vf(c), sp), v), lo), o), inv);
      continue;
    }
    uint32_t bit = (uint32_t)((c.x & 7) + 8 * (c.y & 7) + 64 * (c.z & 7));
    uint32_t occ = W->vox[W->brickOff + ptr * BW + (bit >> 5)];
    float tExit = fmn(fmn(tMax.x, tMax.y), tMax.z);
    if ((occ & (1u << (bit & 31u))) != 0u && tIn >= tSkip) {
      *hid = hitId(p, L, c);
      if (tin) *tin = tIn;
      v3 r = shadeVoxel(W, rc, voxRank(W, ptr, bit), d, vadd(o, vscale(d, tIn)), ptr, bit);
      out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = 0.0f;
      return;
    }
    if (tExit >= tEnd) break;
    tIn = tExit;
    if (tMax.x <= tMax.y && tMax.x <= tMax.z) { c.x += st.x; tMax.x += tDelta.x; }
    else if (tMax.y <= tMax.z) { c.y += st.y; tMax.y += tDelta.y; }
    else { c.z += st.z; tMax.z += tDelta.z; }
    if (c.x < 0 || c.y < 0 || c.z < 0 || c.x >= nVox || c.y >= nVox || c.z >= nVox) break;
  }
}

static void traceVox(const View* W, const orc_rc* rc, v3 o, v3 d, float t0, float t1, float tSkip, uint32_t L,
                     float out[4], uint32_t* hid, uint32_t* steps) {
  traceVoxT(W, rc, o, d, t0, t1, tSkip, L, out, hid, steps, NULL);
}

/* voxInside / voxLevelFor, voxel.ts:108-119. */
static uint32_t voxLevelFor(const orc_vox* p, uint32_t L, v3 e) {
  uint32_t l = L, n = (uint32_t)p->levels;
  for (;;) {
    if (!(l + 1u < n)) break;
    float bs = voxSize(p, l) * 8.0f;
    v3 lo = vscale(vf(org(p, l)), bs);
    float span = (float)p->G * bs;
    v3 hi = vadd(lo, V3(span, span, span));
    int in = e.x > lo.x && e.y > lo.y && e.z > lo.z && e.x < hi.x && e.y < hi.y && e.z < hi.z;
    if (in) break;
    l++;
  }
  return l;
}

static int cascadeOf(const orc_rc* rc, uint64_t ray, uint64_t* local) {
  uint64_t off = 0;
  for (int c = 0; c < rc->cascades; c++) {
    uint64_t n = orc_rays(rc, c);
    if (ray < off + n) { *local = ray - off; return c; }
    off += n;
  }
  return -1;
}

/* trace (rc.ts:600-622) -> traceWorld (rc.ts:216-245) with the probe at its grid point, normal 0. */
/* One ray of cascade c from o along d (traceWorld, rc.ts:216-222): interval [b0, b1] s0, level voxLevelFor at the
 * interval's end, start voxOffset voxels along the probe normal n (relocated probes only; 0 on the grid). */
/* ---- Dynamic coverage layer (dynlayer.ts): build and read ---- */
#define DYN_W 13u
#define DYN_EMPTY_ 0xffffffffu
#define DYN_DROPPED_ 0xfffffffdu
#define DYN_FIX_ 65536.0f

static uint32_t dynNW(const orc_vox* p) { return (uint32_t)(p->levels * p->G * p->G * p->G) / 32u; }

int orc_dyn_init(const orc_vox* p, orc_dyn* d) {
  size_t per = (size_t)p->G * p->G * p->G, NLper = per * (size_t)p->levels;
  memset(d->hdr, 0, 64 * 4);
  d->hdr[32] = d->hdr[33] = d->hdr[34] = 0xffffffffu;
  for (size_t i = 0; i < NLper; i++) d->ptr[i] = DYN_EMPTY_;
  memset(d->bits, 0, 2 * (NLper / 32) * 4);
  memset(d->meta, 0, (size_t)d->cap * 4);
  memset(d->mask, 0, (size_t)d->cap * 16 * 4);
  memset(d->data, 0, (size_t)d->cap * 512 * DYN_W * 4);
  return 0;
}

static v3 lo3(const orc_vox* p, uint32_t L) { return vscale(vscale(vf(org(p, L)), voxSize(p, L)), 8.0f); }
/* splatLevel (dynlayer.ts:240-247) */
static int splatLevel(const orc_vox* p, v3 mn, v3 mx) {
  for (int L = 0; L < p->levels; L++) {
    float v = voxSize(p, (uint32_t)L);
    v3 lo = lo3(p, (uint32_t)L);
    float span = (float)p->G * v * 8.0f;
    v3 hi = vadd(lo, V3(span, span, span));
    v3 a = vsub(mn, V3(2.0f * v, 2.0f * v, 2.0f * v)), b = vadd(mx, V3(2.0f * v, 2.0f * v, 2.0f * v));
    if (a.x > lo.x && a.y > lo.y && a.z > lo.z && b.x < hi.x && b.y < hi.y && b.z < hi.z) return L;
  }
  return -1;
}
static void dynWant(const orc_vox* p, uint32_t* cur, int L, iv3 b) {
  int G = p->G;
  if (b.x < 0 || b.y < 0 || b.z < 0 || b.x >= G || b.y >= G || b.z >= G) return;
  uint32_t per = (uint32_t)(G * G * G);
  uint32_t g = (uint32_t)(p->levels - 1 - L) * per + ((uint32_t)b.z * G + (uint32_t)b.y) * G + (uint32_t)b.x;
  cur[g >> 5] |= 1u << (g & 31u);
}
static uint32_t fixq(float x) {
  float r = rintf(x * DYN_FIX_);
  return (uint32_t)(r < 4294967040.0f ? r : 4294967040.0f);
}
static void splatSample(const orc_vox* p, orc_dyn* D, int L, v3 pt, v3 n, float area, v3 rad, v3 alb, int albw) {
  float v = voxSize(p, (uint32_t)L);
  v3 g = vsub(vdivs(vsub(pt, lo3(p, (uint32_t)L)), v), V3(0.5f, 0.5f, 0.5f));
  iv3 c0 = vi(vfloor(g));
  v3 f = vsub(g, vf(c0));
  float sv = area / (v * v);
  v3 pn = V3(fmx(n.x, 0.0f), fmx(n.y, 0.0f), fmx(n.z, 0.0f)), nn = V3(fmx(-n.x, 0.0f), fmx(-n.y, 0.0f), fmx(-n.z, 0.0f));
  int E = p->G * 8, G = p->G;
  uint32_t per = (uint32_t)(G * G * G);
  for (int q = 0; q < 8; q++) {
    iv3 o = I3(q & 1, (q >> 1) & 1, q >> 2);
    iv3 c = iadd(c0, o);
    if (c.x < 0 || c.y < 0 || c.z < 0 || c.x >= E || c.y >= E || c.z >= E) continue;
    v3 tw = V3(o.x ? f.x : 1.0f - f.x, o.y ? f.y : 1.0f - f.y, o.z ? f.z : 1.0f - f.z);
    float w = tw.x * tw.y * tw.z * sv;
    if (w <= 0.0f) continue;
    iv3 b = ishr(c, 3);
    uint32_t ptr = D->ptr[(uint32_t)L * per + ((uint32_t)b.z * G + (uint32_t)b.y) * G + (uint32_t)b.x];
    if (ptr >= (uint32_t)D->cap) continue;
    uint32_t vi_ = (uint32_t)((c.x & 7) + 8 * (c.y & 7) + 64 * (c.z & 7));
    D->mask[ptr * 16 + (vi_ >> 5)] |= 1u << (vi_ & 31u);
    uint32_t* dd = D->data + ((size_t)ptr * 512 + vi_) * DYN_W;
    float x[13] = {w * pn.x, w * nn.x, w * pn.y, w * nn.y, w * pn.z, w * nn.z, w, w * rad.x, w * rad.y, w * rad.z,
                   w * alb.x, w * alb.y, w * alb.z};
    for (int k = 0; k < (albw ? 13 : 10); k++) dd[k] += fixq(x[k]);
  }
}

int orc_dyn_build(const orc_vox* p, const orc_rc* rc, const float* pos, const float* nrm, const uint32_t* tri,
                  const float* mat, orc_dyn* D) {
  const int G = p->G, NL = p->levels;
  const uint32_t per = (uint32_t)(G * G * G), nw = dynNW(p), cap = (uint32_t)D->cap;
  uint32_t* h = D->hdr;
  uint32_t* cur = D->bits + (h[10] & 1u) * nw;
  uint32_t* old = D->bits + (1u - (h[10] & 1u)) * nw;
  if (D->end_tri <= D->first_tri) return 0;
  /* mark */
  for (uint32_t t = D->first_tri; t < D->end_tri; t++) {
    const uint32_t* q = tri + (size_t)t * 4;
    v3 a = V3(pos[q[0] * 3], pos[q[0] * 3 + 1], pos[q[0] * 3 + 2]);
    v3 b = V3(pos[q[1] * 3], pos[q[1] * 3 + 1], pos[q[1] * 3 + 2]);
    v3 c = V3(pos[q[2] * 3], pos[q[2] * 3 + 1], pos[q[2] * 3 + 2]);
    v3 mn = vmin(vmin(a, b), c), mx = vmax(vmax(a, b), c);
    int Ls = splatLevel(p, mn, mx);
    if (Ls < 0) {
      h[38]++;
      continue;
    }
    float v = voxSize(p, (uint32_t)Ls), bs = v * 8.0f;
    v3 lo = lo3(p, (uint32_t)Ls);
    iv3 g0 = vi(vfloor(vdivs(vsub(vsub(mn, V3(v, v, v)), lo), bs)));
    iv3 g1 = vi(vfloor(vdivs(vsub(vadd(mx, V3(v, v, v)), lo), bs)));
    for (int z = g0.z; z <= g1.z; z++)
      for (int y = g0.y; y <= g1.y; y++)
        for (int x = g0.x; x <= g1.x; x++) dynWant(p, cur, Ls, I3(x, y, z));
    float gv = voxSize(p, (uint32_t)(NL - 1));
    v3 blo = vfloor(vscale(vsub(mn, V3(gv, gv, gv)), 100.0f));
    v3 bh = vadd(mx, V3(gv, gv, gv));
    v3 bhi = V3(ceilf(bh.x * 100.0f), ceilf(bh.y * 100.0f), ceilf(bh.z * 100.0f));
    uint32_t lo_[3] = {(uint32_t)(f2i(blo.x) + 1073741824), (uint32_t)(f2i(blo.y) + 1073741824), (uint32_t)(f2i(blo.z) + 1073741824)};
    uint32_t hi_[3] = {(uint32_t)(f2i(bhi.x) + 1073741824), (uint32_t)(f2i(bhi.y) + 1073741824), (uint32_t)(f2i(bhi.z) + 1073741824)};
    for (int k = 0; k < 3; k++) {
      if (lo_[k] < h[32 + k]) h[32 + k] = lo_[k];
      if (hi_[k] > h[35 + k]) h[35 + k] = hi_[k];
    }
  }
  /* parents: bit order is coarse first, so walking words from the finest level up visits a word after every bit
   * that finer bricks add to it; each wanted brick adds all its ancestors anyway. */
  for (uint32_t w = per / 32u; w < nw; w++) {
    uint32_t m = cur[w];
    while (m) {
      uint32_t bit = (uint32_t)__builtin_ctz(m);
      m &= m - 1u;
      uint32_t g = w * 32u + bit, lr = g / per;
      int L = NL - 1 - (int)lr;
      uint32_t r = g - lr * per;
      iv3 b = I3((int)(r % G), (int)((r / G) % G), (int)(r / (G * G)));
      while (L + 1 < NL) {
        iv3 o0 = org(p, (uint32_t)L), o1 = org(p, (uint32_t)(L + 1));
        b = ishr(isub(imuls(iadd(o0, b), 4), imuls(o1, 8)), 3);
        L++;
        dynWant(p, cur, L, b);
      }
    }
  }
  /* alloc + finalize */
  uint32_t first = 0;
  for (uint32_t w = 0; w < nw; w++) {
    uint32_t cw = cur[w], ow = old[w];
    uint32_t c = (uint32_t)__builtin_popcount(cw);
    uint32_t excl = first;
    if ((w * 32u) % per == 0u) h[16u + (uint32_t)NL - 1u - w * 32u / per] = excl;
    if (w == nw - 1u) h[0] = excl + c;
    uint32_t m = cw | ow;
    while (m) {
      uint32_t bit = (uint32_t)__builtin_ctz(m);
      m &= m - 1u;
      uint32_t g = w * 32u + bit, lr = g / per;
      uint32_t pi = ((uint32_t)NL - 1u - lr) * per + g - lr * per;
      if ((cw >> bit) & 1u) {
        uint32_t rank = excl + (uint32_t)__builtin_popcount(cw & ((1u << bit) - 1u));
        if (rank < cap) {
          D->ptr[pi] = rank;
          D->meta[rank] = pi;
        } else {
          D->ptr[pi] = DYN_DROPPED_;
        }
      } else {
        D->ptr[pi] = DYN_EMPTY_;
      }
    }
    old[w] = 0u;
    first += c;
  }
  {
    uint32_t total = h[0], al = total < cap ? total : cap;
    h[8] = al;
    h[1] = total - al;
    for (int L = 0; L < NL; L++) {
      uint32_t e = L > 0 ? h[16 + L - 1] : total;
      uint32_t s0 = h[16 + L] < cap ? h[16 + L] : cap;
      h[16 + L] = s0;
      h[24 + L] = (e < cap ? e : cap) - s0;
    }
    for (int a = 0; a < 3; a++) {
      h[2 + a] = h[32 + a];
      h[32 + a] = 0xffffffffu;
      h[5 + a] = h[35 + a];
      h[35 + a] = 0u;
    }
    h[9] = h[38];
    h[38] = 0u;
    h[10] = h[10] + 1u;
  }
  /* clear */
  memset(D->data, 0, (size_t)h[8] * 512 * DYN_W * 4);
  memset(D->mask, 0, (size_t)h[8] * 16 * 4);
  /* splat */
  v3 sd = V3(rc->sun_dir[0], rc->sun_dir[1], rc->sun_dir[2]), se = V3(rc->sun_e[0], rc->sun_e[1], rc->sun_e[2]);
  for (uint32_t t = D->first_tri; t < D->end_tri; t++) {
    const uint32_t* q = tri + (size_t)t * 4;
    v3 a = V3(pos[q[0] * 3], pos[q[0] * 3 + 1], pos[q[0] * 3 + 2]);
    v3 b = V3(pos[q[1] * 3], pos[q[1] * 3 + 1], pos[q[1] * 3 + 2]);
    v3 c = V3(pos[q[2] * 3], pos[q[2] * 3 + 1], pos[q[2] * 3 + 2]);
    int L = splatLevel(p, vmin(vmin(a, b), c), vmax(vmax(a, b), c));
    if (L < 0) continue;
    v3 cr = vcross(vsub(b, a), vsub(c, a));
    float area = 0.5f * sqrtf(vdot(cr, cr));
    if (area <= 0.0f) continue;
    v3 e0 = vsub(b, a), e1 = vsub(c, b), e2 = vsub(a, c);
    float e = fmx(fmx(sqrtf(vdot(e0, e0)), sqrtf(vdot(e1, e1))), sqrtf(vdot(e2, e2)));
    float mf = ceilf(e / voxSize(p, (uint32_t)L));
    mf = mf > 1.0f ? mf : 1.0f;
    mf = mf < 16.0f ? mf : 16.0f;
    uint32_t m = (uint32_t)mf;
    float fm = (float)m, sa = area / (fm * fm);
    const float* M = mat + (size_t)q[3] * 4;
    const int masked = g_surf && (g_surf->mext[q[3] * 4 + 1] & 1u);
    v3 n0 = V3(nrm[q[0] * 3], nrm[q[0] * 3 + 1], nrm[q[0] * 3 + 2]);
    v3 n1 = V3(nrm[q[1] * 3], nrm[q[1] * 3 + 1], nrm[q[1] * 3 + 2]);
    v3 n2 = V3(nrm[q[2] * 3], nrm[q[2] * 3 + 1], nrm[q[2] * 3 + 2]);
    for (uint32_t u = 0; u < m; u++)
      for (uint32_t v_ = 0; v_ + u < m; v_++)
        for (uint32_t flip = 0; flip < 2; flip++) {
          if (flip == 1 && u + v_ + 1 >= m) continue;
          float off = flip == 1 ? 2.0f / 3.0f : 1.0f / 3.0f;
          float wb = ((float)u + off) / fm, wc = ((float)v_ + off) / fm, wa = 1.0f - wb - wc;
          v3 pt = vadd(vadd(vscale(a, wa), vscale(b, wb)), vscale(c, wc));
          v3 nv = vadd(vadd(vscale(n0, wa), vscale(n1, wb)), vscale(n2, wc));
          nv = vnormalize(vdot(nv, nv) > 1e-12f ? nv : cr);
          /* shadeSample's tex (dynlayer.ts:451-452): the alpha test drops the sample, its rgb is the albedo */
          v3 alb = V3(1.0f * M[0], 1.0f * M[1], 1.0f * M[2]);
          if (g_surf) {
            float tex[4];
            surfBaseColor(mat, q[3], q[0], q[1], q[2], wa, wb, wc, (uint32_t)L, tex);
            if (masked && tex[3] < surfCutoff(q[3])) continue;
            alb = V3(tex[0], tex[1], tex[2]);
          }
          float ndl = fmx(vdot(nv, sd), 0.0f);
          v3 rad = V3(0, 0, 0);
          if (ndl > 0.0f) rad = vscale(vscale(vmul(vdivs(alb, PI_F), se), ndl), rc->sun_vis);
          splatSample(p, D, L, pt, nv, sa, rad, alb, rc->objects);
        }
  }
  /* resolve, fine to coarse */
  for (int L = 1; L < NL; L++) {
    uint32_t cnt = h[24 + L], s0 = h[16 + L];
    for (uint32_t i = 0; i < cnt * 512u; i++) {
      uint32_t slot = s0 + (i >> 9), bit = i & 511u;
      uint32_t pi = D->meta[slot], r = pi % per;
      iv3 c = iadd(imuls(I3((int)(r % G), (int)((r / G) % G), (int)(r / (G * G))), 8),
                   I3((int)(bit & 7u), (int)((bit >> 3) & 7u), (int)(bit >> 6)));
      uint32_t own = (D->mask[slot * 16 + (bit >> 5)] >> (bit & 31u)) & 1u;
      iv3 gc = isub(imuls(iadd(imuls(org(p, (uint32_t)L), 8), c), 2), imuls(org(p, (uint32_t)(L - 1)), 8));
      uint32_t cptr = cap;
      int E = G * 8;
      if (gc.x >= 0 && gc.y >= 0 && gc.z >= 0 && gc.x < E && gc.y < E && gc.z < E) {
        iv3 b = ishr(gc, 3);
        cptr = D->ptr[(uint32_t)(L - 1) * per + ((uint32_t)b.z * G + (uint32_t)b.y) * G + (uint32_t)b.x];
      }
      uint32_t kids = 0;
      if (cptr < cap) {
        for (int q = 0; q < 8; q++) {
          iv3 l = iadd(I3(gc.x & 7, gc.y & 7, gc.z & 7), I3(q & 1, (q >> 1) & 1, q >> 2));
          uint32_t vv = (uint32_t)(l.x + 8 * l.y + 64 * l.z);
          if ((D->mask[cptr * 16 + (vv >> 5)] >> (vv & 31u)) & 1u) kids |= 1u << q;
        }
      }
      if (own == 0u && kids == 0u) continue;
      uint32_t* dd = D->data + ((size_t)slot * 512 + bit) * DYN_W;
      float acc[DYN_W];
      for (uint32_t k = 0; k < DYN_W; k++) acc[k] = (float)dd[k] / DYN_FIX_;
      for (int q = 0; q < 8; q++) {
        if (!((kids >> q) & 1u)) continue;
        iv3 l = iadd(I3(gc.x & 7, gc.y & 7, gc.z & 7), I3(q & 1, (q >> 1) & 1, q >> 2));
        const uint32_t* cb = D->data + ((size_t)cptr * 512 + (uint32_t)(l.x + 8 * l.y + 64 * l.z)) * DYN_W;
        for (uint32_t k = 0; k < DYN_W; k++) {
          uint32_t u = cb[k];
          float x;
          if (L == 1) x = (float)u / DYN_FIX_;
          else memcpy(&x, &u, 4);
          acc[k] += 0.25f * x;
        }
      }
      for (uint32_t k = 0; k < DYN_W; k++) memcpy(&dd[k], &acc[k], 4);
      if (own == 0u) D->mask[slot * 16 + (bit >> 5)] |= 1u << (bit & 31u);
    }
  }
  return 0;
}

/* log2 of a positive normal float for the level blend of traceDyn (dynlayer.ts:140, WGSL log2): x = m 2^e with m in
 * [sqrt(1/2), sqrt(2)], ln m = 2 atanh((m - 1) / (m + 1)) by its series to s^9, in plain float operations in a fixed
 * order so that the device's restatement gives the same bits. */
float orc_log2(float x) {
  uint32_t u;
  memcpy(&u, &x, 4);
  int e = (int)(u >> 23) - 127;
  uint32_t mu = (u & 0x007fffffu) | 0x3f800000u;
  float m;
  memcpy(&m, &mu, 4);
  if (m > 1.41421356f) {
    m = m * 0.5f;
    e += 1;
  }
  float s = (m - 1.0f) / (m + 1.0f), s2 = s * s;
  float p = 1.0f / 9.0f;
  p = p * s2 + 1.0f / 7.0f;
  p = p * s2 + 1.0f / 5.0f;
  p = p * s2 + 1.0f / 3.0f;
  p = p * s2 + 1.0f;
  return (float)e + ((2.0f * s) * p) * 1.44269504f;
}

/* dynFind / dynWord / dynRead (dynlayer.ts:81-116) */
static uint32_t dynFind(const orc_vox* p, const orc_dyn* D, uint32_t L, v3 pt, uint32_t* lev) {
  uint32_t l = L;
  int G = p->G;
  uint32_t per = (uint32_t)(G * G * G);
  for (;;) {
    float v = voxSize(p, l);
    iv3 c = vi(vfloor(vdivs(vsub(pt, vscale(vscale(vf(org(p, l)), v), 8.0f)), v)));
    *lev = l;
    if (c.x < 0 || c.y < 0 || c.z < 0 || c.x >= G * 8 || c.y >= G * 8 || c.z >= G * 8) return DYN_EMPTY_;
    iv3 b = ishr(c, 3);
    uint32_t ptr = D->ptr[l * per + ((uint32_t)b.z * G + (uint32_t)b.y) * G + (uint32_t)b.x];
    if (ptr == DYN_DROPPED_ && l + 1u < (uint32_t)p->levels) {
      l++;
      continue;
    }
    if (ptr >= (uint32_t)D->cap) return DYN_EMPTY_;
    return (ptr * 512u + (uint32_t)((c.x & 7) + 8 * (c.y & 7) + 64 * (c.z & 7))) * DYN_W;
  }
}
static float dynWord(const orc_dyn* D, uint32_t base, uint32_t k, uint32_t L) {
  uint32_t u = D->data[base + k];
  if (L == 0u) return (float)u / DYN_FIX_;
  float x;
  memcpy(&x, &u, 4);
  return x;
}
static void dynRead(const orc_dyn* D, uint32_t base, uint32_t L, v3 d, float out[4]) {
  out[0] = out[1] = out[2] = out[3] = 0.0f;
  if (base == DYN_EMPTY_) return;
  float a = dynWord(D, base, 6, L);
  if (a <= 0.0f) return;
  uint32_t sx = d.x > 0.0f ? 1u : 0u, sy = d.y > 0.0f ? 1u : 0u, sz = d.z > 0.0f ? 1u : 0u;
  float front = fabsf(d.x) * dynWord(D, base, sx, L) + fabsf(d.y) * dynWord(D, base, 2 + sy, L) +
                fabsf(d.z) * dynWord(D, base, 4 + sz, L);
  out[0] = dynWord(D, base, 7, L) / a;
  out[1] = dynWord(D, base, 8, L) / a;
  out[2] = dynWord(D, base, 9, L) / a;
  out[3] = fmn(front, 1.0f);
}
/* dynInject (dynlayer.ts:166-190) with the bounce volume's dynE: each voxel of the allocated bricks with coverage
 * adds its area x albedo / pi x the irradiance of the previous volume at its centre, weighted over its six projected
 * areas; level 0 in fixed point (round half to even), the coarser levels as floats. */
int orc_dyn_inject(const orc_vox* p, const orc_dyn* D, const uint16_t* vol, const int32_t vorg[3]) {
  const int G = p->G;
  const uint32_t per = (uint32_t)(G * G * G), n = D->hdr[8];
  const float tol = 0.5f * p->v0;
  for (uint32_t slot = 0; slot < n; slot++) {
    const uint32_t pi = D->meta[slot], L = pi / per, r = pi % per;
    for (uint32_t bit = 0; bit < 512u; bit++) {
      const uint32_t base = (slot * 512u + bit) * DYN_W;
      if (dynWord(D, base, 6, L) <= 0.0f) continue;
      float P[6];
      for (uint32_t k = 0; k < 6; k++) P[k] = dynWord(D, base, k, L);
      iv3 c = iadd(imuls(I3((int)(r % G), (int)((r / G) % G), (int)(r / (G * G))), 8),
                   I3((int)(bit & 7u), (int)((bit >> 3) & 7u), (int)(bit >> 6)));
      float v = voxSize(p, L);
      v3 pc = vadd(vscale(vscale(vf(org(p, L)), v), 8.0f), vscale(vadd(vf(c), V3(0.5f, 0.5f, 0.5f)), v));
      v3 E = dynE(vol, vorg, tol, pc, P);
      v3 alb = V3(dynWord(D, base, 10, L), dynWord(D, base, 11, L), dynWord(D, base, 12, L));
      v3 add = vmul(vdivs(alb, PI_F), E);
      float ad[3] = {add.x, add.y, add.z};
      for (uint32_t k = 0; k < 3; k++) {
        uint32_t* w = D->data + base + 7 + k;
        if (L == 0) {
          *w += (uint32_t)fmn(rintf(ad[k] * DYN_FIX_), 4294967040.0f); /* clamped as the splat's words are */
        } else {
          float x;
          memcpy(&x, w, 4);
          x += ad[k];
          memcpy(w, &x, 4);
        }
      }
    }
  }
  return 0;
}

/* traceDyn (dynlayer.ts:118-160) */
static void traceDyn(const orc_vox* p, const orc_dyn* D, v3 o, v3 d, float t0, float t1, float kc, float out[4]) {
  out[0] = out[1] = out[2] = out[3] = 0.0f;
  if (D->hdr[0] == 0u || t1 <= t0) return;
  v3 bmin = vscale(vsub(V3((float)D->hdr[2], (float)D->hdr[3], (float)D->hdr[4]), V3(1073741824.0f, 1073741824.0f, 1073741824.0f)), 0.01f);
  v3 bmax = vscale(vsub(V3((float)D->hdr[5], (float)D->hdr[6], (float)D->hdr[7]), V3(1073741824.0f, 1073741824.0f, 1073741824.0f)), 0.01f);
  v3 dd = V3(fabsf(d.x) < 1e-9f ? 1e-9f : d.x, fabsf(d.y) < 1e-9f ? 1e-9f : d.y, fabsf(d.z) < 1e-9f ? 1e-9f : d.z);
  v3 ta = V3((bmin.x - o.x) / dd.x, (bmin.y - o.y) / dd.y, (bmin.z - o.z) / dd.z);
  v3 tb = V3((bmax.x - o.x) / dd.x, (bmax.y - o.y) / dd.y, (bmax.z - o.z) / dd.z);
  v3 tn = vmin(ta, tb), tf = vmax(ta, tb);
  float ba = fmx(fmx(fmx(tn.x, tn.y), tn.z), t0);
  float bb = fmn(fmn(fmn(tf.x, tf.y), tf.z), t1);
  if (ba >= bb) return;
  float v0 = p->v0;
  int nl = p->levels;
  v3 rad = V3(0, 0, 0);
  float cov = 0.0f, t = t0;
  for (int i = 0; i < 128; i++) {
    if (t >= bb || cov >= 1.0f) break;
    float w = fmx(kc * t, v0);
    float te = fmn(t + 0.5f * w, t1);
    if (te > ba) {
      float len = te - t;
      v3 pt = vadd(o, vscale(d, 0.5f * (t + te)));
      float lf = fmn(fmx(orc_log2(w / v0), 0.0f), (float)(nl - 1));
      uint32_t L0 = (uint32_t)floorf(lf);
      float f = lf - (float)L0;
      uint32_t Lw = voxLevelFor(p, L0, pt);
      if (Lw != L0) {
        L0 = Lw;
        f = 0.0f;
      }
      uint32_t la = 0, lb = 0;
      float sa[4], sb[4];
      /* the level comes out of dynFind, so it is called before dynRead reads it */
      uint32_t fa = dynFind(p, D, L0, pt, &la);
      dynRead(D, fa, la, d, sa);
      float aa = sa[3] * len / voxSize(p, la);
      v3 ra = vscale(V3(sa[0], sa[1], sa[2]), aa);
      if (f > 0.0f && (int)L0 + 1 < nl) {
        uint32_t fb = dynFind(p, D, L0 + 1, pt, &lb);
        dynRead(D, fb, lb, d, sb);
        float ab = sb[3] * len / voxSize(p, lb);
        v3 rb = vscale(V3(sb[0], sb[1], sb[2]), ab);
        ra = vadd(vscale(ra, 1.0f - f), vscale(rb, f));
        aa = aa * (1.0f - f) + ab * f;
      }
      if (aa > 0.0f) {
        float ac = fmn(aa, 1.0f - cov);
        rad = vadd(rad, vscale(ra, ac / aa));
        cov += ac;
      }
    }
    t = te;
  }
  out[0] = rad.x;
  out[1] = rad.y;
  out[2] = rad.z;
  out[3] = cov;
}

static void traceFromN(const View* W, const orc_rc* rc, int c, v3 o, v3 n, float out[4], uint32_t* hid,
                       uint32_t* steps, v3 d);
static void traceFrom(const View* W, const orc_rc* rc, int c, v3 o, v3 d, float out[4], uint32_t* hid,
                      uint32_t* steps) {
  traceFromN(W, rc, c, o, V3(0, 0, 0), out, hid, steps, d);
}
static void traceFromN(const View* W, const orc_rc* rc, int c, v3 o, v3 n, float out[4], uint32_t* hid,
                       uint32_t* steps, v3 d) {
  float t0 = rc->b0[c] * rc->s0, t1 = rc->b1[c] * rc->s0;
  uint32_t L = voxLevelFor(W->p, (uint32_t)rc->vox_level[c], vadd(o, vscale(d, t1)));
  float v = voxSize(W->p, L);
  v3 ot = rc->reloc ? vadd(o, vscale(n, rc->vox_offset * v)) : o;
  float tin = 0.0f;
  traceVoxT(W, rc, ot, d, t0, t1, rc->vox_skip * v, L, out, hid, steps, &tin);
  if (rc->dyn) {
    /* the dynamic layer in front of the static hit (traceWorld, rc.ts:224-231) */
    float tS = out[3] < 0.5f ? vdot(vsub(vadd(ot, vscale(d, tin)), o), d) : t1;
    float kc = 3.5449077f / (float)resOf(rc, c);
    float dl[4];
    traceDyn(W->p, rc->dyn, o, d, fmx(t0, W->p->v0), fmn(t1, tS), kc, dl);
    if (dl[3] > 0.0f) {
      for (int k = 0; k < 3; k++) out[k] = dl[k] + (1.0f - dl[3]) * out[k];
      out[3] = (1.0f - dl[3]) * out[3];
    }
  }
}

static void traceRay(const View* W, const orc_rc* rc, uint64_t ray, float out[4], uint32_t* hid, uint32_t* steps) {
  uint64_t r = 0;
  int c = cascadeOf(rc, ray, &r);
  int res = resOf(rc, c), n = probesOf(rc, c);
  uint64_t dirs = (uint64_t)res * res;
  uint64_t pi = r / dirs;
  int di = (int)(r % dirs);
  iv3 g = I3((rc->G0[0] >> c) + (int)(pi % n), (rc->G0[1] >> c) + (int)((pi / n) % n), (rc->G0[2] >> c) + (int)(pi / ((uint64_t)n * n)));
  float s = rc->s0 * (float)(1u << c);
  if (rc->reloc) {
    const float* P = rc->ppos[c] + pi * 3;
    const float* Nn = rc->pnrm[c] + pi * 3;
    traceFromN(W, rc, c, V3(P[0], P[1], P[2]), V3(Nn[0], Nn[1], Nn[2]), out, hid, steps, dirOf(di, res));
    return;
  }
  traceFrom(W, rc, c, vscale(vf(g), s), dirOf(di, res), out, hid, steps);
}

static void makeView(View* W, const orc_vox* p, const uint32_t* vox) {
  W->p = p;
  W->vox = vox;
  layout(p, &W->ptrOff, &W->brickOff, &W->attrOff, &W->plOff);
  W->alOff = albOff(p);
}

int orc_trace(const orc_vox* p, const uint32_t* vox, const orc_rc* rc, uint64_t r0, uint64_t r1, float* rgba,
              uint32_t* hits, uint32_t* steps, int threads) {
  View W;
  makeView(&W, p, vox);
#ifdef _OPENMP
  if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 4096) num_threads(threads)
#endif
  for (int64_t q = (int64_t)r0; q < (int64_t)r1; q++) {
    float out[4];
    uint32_t h, st;
    traceRay(&W, rc, (uint64_t)q, out, &h, &st);
    if (steps) steps[q - (int64_t)r0] = st;
    for (int k = 0; k < 4; k++) out[k] = q16(rc, out[k]); /* packRad of the trace output (rc.ts:621) */
    if (rgba) memcpy(rgba + (size_t)(q - (int64_t)r0) * 4, out, 16);
    if (hits) hits[q - (int64_t)r0] = h;
  }
  (void)threads;
  return 0;
}

int orc_trace_rays(const orc_vox* p, const uint32_t* vox, const orc_rc* rc, int c, const float* o, const float* d,
                   uint64_t n, float* rgba, uint32_t* hits, int threads) {
  if (c < 0 || c >= rc->cascades) return -1;
  View W;
  makeView(&W, p, vox);
#ifdef _OPENMP
  if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads)
#endif
  for (int64_t i = 0; i < (int64_t)n; i++) {
    float out[4];
    uint32_t h, st;
    traceFrom(&W, rc, c, V3(o[3 * i], o[3 * i + 1], o[3 * i + 2]), V3(d[3 * i], d[3 * i + 1], d[3 * i + 2]), out, &h,
              &st);
    for (int k = 0; k < 4; k++) out[k] = q16(rc, out[k]);
    if (rgba) memcpy(rgba + (size_t)i * 4, out, 16);
    if (hits) hits[i] = h;
  }
  (void)threads;
  return 0;
}

/* hash32 (rc.ts:248-250). */
static uint32_t hash32(uint32_t x) {
  uint32_t h = x;
  h ^= h >> 16;
  h *= 0x7feb352du;
  h ^= h >> 15;
  h *= 0x846ca68bu;
  h ^= h >> 16;
  return h;
}
/* histHash (rc.ts:781-783): the key of a grid point and LOD (u32 wrap of negative coordinates). */
static uint32_t histHash(int gx, int gy, int gz, int lod) {
  return hash32(((uint32_t)gx * 73856093u) ^ ((uint32_t)gy * 19349663u) ^ ((uint32_t)gz * 83492791u) ^
                ((uint32_t)lod * 2654435761u));
}

int orc_hist_init(orc_hist* h, int32_t cap) {
  uint32_t H = 1;
  while (H < (uint32_t)cap * 2u) H <<= 1; /* hashSize = nextPow2(cap * 2) (rc.ts:1029) */
  h->cap = cap;
  h->H = H;
  h->par = 0;
  h->frames = 0;
  h->misses = 0;
  h->limit = 64;
  return 0;
}

/* histBlend (rc.ts:784-814) of probe pi (grid point g, LOD 0) into v (27 coefficients; the reference's v[27..31] are
 * 0 and stay 0): look the key up in the previous parity's hash (linear probing, at most h->limit slots (the
 * reference's 64; 0: no limit), stop at an empty slot); when found, v = mix(prev, v, HIST_W) = prev * (1 - 0.5) + v * 0.5 per coefficient (WGSL mix); then record
 * the probe's grid point and insert it into the current parity's hash (compare-exchange on empty slots, at most
 * h->limit).
 * The stored SH (sh_out, written by the caller after packing) is what the next frame reads. Returns 1 when a probe
 * of the previous frame held the key but the 64-slot search missed it (counted, to show the dense lookup's
 * equivalence). */
static int histBlend(orc_hist* h, uint64_t pi, const int32_t g[3], float v[27]) {
  const uint32_t H = h->H, cur = (uint32_t)h->par, prv = 1u - cur;
  const uint32_t h0 = histHash(g[0], g[1], g[2], 0) & (H - 1u);
  const uint32_t lim = h->limit > 0 ? (uint32_t)h->limit : H;
  uint32_t hh = h0;
  int64_t found = -1;
  for (uint32_t i = 0; i < lim; i++) {
    uint32_t sl = h->hash[prv * H + hh];
    if (sl == 0u) break;
    uint64_t q = sl - 1u;
    const int32_t* o = h->grid + ((uint64_t)prv * h->cap + q) * 4;
    if (o[0] == g[0] && o[1] == g[1] && o[2] == g[2] && o[3] == 0) {
      found = (int64_t)q;
      break;
    }
    hh = (hh + 1u) & (H - 1u);
  }
  int missed = 0;
  if (found >= 0) {
    const float* f = h->sh + ((uint64_t)prv * h->cap + (uint64_t)found) * 32;
    for (int k = 0; k < 27; k++) v[k] = f[k] * (1.0f - 0.5f) + v[k] * 0.5f;
  } else if (h->frames > 0) {
    /* did the previous frame's dense grid hold this point? */
    int in = 1;
    for (int a = 0; a < 3; a++) in = in && g[a] >= h->G0[prv][a] && g[a] < h->G0[prv][a] + h->n;
    missed = in;
  }
  int32_t* og = h->grid + ((uint64_t)cur * h->cap + pi) * 4;
  og[0] = g[0];
  og[1] = g[1];
  og[2] = g[2];
  og[3] = 0;
  hh = h0;
  for (uint32_t i = 0; i < lim; i++) {
    if (h->hash[cur * H + hh] == 0u) {
      h->hash[cur * H + hh] = (uint32_t)pi + 1u;
      break;
    }
    hh = (hh + 1u) & (H - 1u);
  }
  return missed;
}

int orc_merge_sh(const orc_rc* rc, float* rgba, float* sh) { return orc_merge_sh_hist(rc, rgba, sh, NULL); }

/* upperAvg / mergeWorld (rc.ts:681-708) with the dense grid's parents, then projectSH (rc.ts:745-774) with the c0
 * history (clearHist + histBlend, rc.ts:775-818, 1213) when hist is given. */
int orc_merge_sh_hist(const orc_rc* rc, float* rgba, float* sh, orc_hist* hist) {
  int Nc = rc->cascades;
  for (int c = Nc - 1; c >= 0; c--) {
    int res = resOf(rc, c), n = probesOf(rc, c);
    uint64_t off = rayOff(rc, c), nr = orc_rays(rc, c), dirs = (uint64_t)res * res;
    for (uint64_t t = 0; t < nr; t++) {
      float* v = rgba + (off + t) * 4;
      v3 up = V3(rc->sky[0], rc->sky[1], rc->sky[2]);
      if (c + 1 < Nc && v[3] > 0.0f) {
        int ur = resOf(rc, c + 1), un = probesOf(rc, c + 1);
        uint64_t uoff = rayOff(rc, c + 1);
        uint64_t pi = t / dirs;
        int di = (int)(t % dirs);
        int u0 = di % res, w0 = di / res;
        iv3 g = I3((int)(pi % n), (int)((pi / n) % n), (int)(pi / ((uint64_t)n * n)));
        v3 acc = V3(0, 0, 0);
        float ws = 0.0f;
        for (int k = 0; k < 8; k++) {
          iv3 o = I3(k & 1, (k >> 1) & 1, (k >> 2) & 1);
          /* linkParents (rc.ts:540-570): an odd axis may take the next upper point; absent outside the grid. */
          if ((o.x > 0 && !(g.x & 1)) || (o.y > 0 && !(g.y & 1)) || (o.z > 0 && !(g.z & 1))) continue;
          iv3 q = iadd(ishr(g, 1), o);
          if (q.x >= un || q.y >= un || q.z >= un) continue;
          if (rc->reloc && !((rc->pvis[c][pi] >> k) & 1u)) continue; /* hidden by voxels (rc.ts:559-566) */
          uint64_t qi = (uint64_t)q.x + (uint64_t)un * ((uint64_t)q.y + (uint64_t)un * q.z);
          const float* U = rgba + (uoff + qi * (uint64_t)ur * ur) * 4;
          v3 a;
          if (ur == res) {
            const float* x = U + ((uint64_t)w0 * ur + u0) * 4;
            a = V3(x[0], x[1], x[2]);
          } else {
            const float* b = U + ((uint64_t)w0 * 2 * ur + (uint64_t)u0 * 2) * 4;
            const float* b1 = b + 4;
            const float* b2 = b + (uint64_t)ur * 4;
            const float* b3 = b2 + 4;
            a = vscale(vadd(vadd(vadd(V3(b[0], b[1], b[2]), V3(b1[0], b1[1], b1[2])), V3(b2[0], b2[1], b2[2])), V3(b3[0], b3[1], b3[2])), 0.25f);
          }
          acc = vadd(acc, a);
          ws += 1.0f;
        }
        up = ws > 0.0f ? vdivs(acc, fmx(ws, 1.0f)) : V3(0, 0, 0);
      }
      /* packRad of the merged value (rc.ts:707) */
      v[0] = q16(rc, v[0] + v[3] * up.x);
      v[1] = q16(rc, v[1] + v[3] * up.y);
      v[2] = q16(rc, v[2] + v[3] * up.z);
      v[3] = 1.0f;
    }
  }
  if (sh) {
    int n = probesOf(rc, 0);
    uint64_t np = (uint64_t)n * n * n;
    if (hist) {
      if ((uint64_t)hist->cap != np) return -1;
      /* encode: the parity flips (rc.ts:1170); clearHist empties the new parity's hash (rc.ts:815-818, 1213) */
      hist->par = 1 - hist->par;
      memset(hist->hash + (size_t)hist->par * hist->H, 0, (size_t)hist->H * sizeof(uint32_t));
    }
    const float a0 = (float)3.14159265, a1 = (float)(2.0 * 3.14159265 / 3.0), a2 = (float)(3.14159265 / 4.0);
    for (uint64_t pi = 0; pi < np; pi++) {
      v3 cf[9];
      for (int k = 0; k < 9; k++) cf[k] = V3(0, 0, 0);
      for (int d = 0; d < 16; d++) {
        const float* x = rgba + (pi * 16 + d) * 4;
        v3 r = vscale(V3(x[0], x[1], x[2]), rc->omega0[d]);
        v3 nn = dirOf(d, 4);
        float y[9] = {0.282095f, 0.488603f * nn.y, 0.488603f * nn.z, 0.488603f * nn.x, 1.092548f * nn.x * nn.y,
                      1.092548f * nn.y * nn.z, 0.315392f * (3.0f * nn.z * nn.z - 1.0f), 1.092548f * nn.x * nn.z,
                      0.546274f * (nn.x * nn.x - nn.y * nn.y)};
        for (int k = 0; k < 9; k++) cf[k] = vadd(cf[k], vscale(r, y[k]));
      }
      float v[27];
      for (int k = 0; k < 9; k++) {
        float a = k == 0 ? a0 : (k < 4 ? a1 : a2);
        v[k * 3] = cf[k].x * a;
        v[k * 3 + 1] = cf[k].y * a;
        v[k * 3 + 2] = cf[k].z * a;
      }
      if (hist) {
        int32_t g[3] = {rc->G0[0] + (int32_t)(pi % n), rc->G0[1] + (int32_t)((pi / n) % n),
                        rc->G0[2] + (int32_t)(pi / ((uint64_t)n * n))};
        hist->misses += (uint64_t)histBlend(hist, pi, g, v);
      }
      for (int k = 0; k < 27; k++) sh[pi * 27 + k] = q16(rc, v[k]); /* pack2x16float (rc.ts:770-771) */
      /* the history holds the packed words of this frame's SH (rc.ts:772): their unpacked values */
      if (hist) memcpy(hist->sh + ((uint64_t)hist->par * hist->cap + pi) * 32, sh + pi * 27, 27 * sizeof(float));
    }
    if (hist) {
      for (int a = 0; a < 3; a++) hist->G0[hist->par][a] = rc->G0[a];
      hist->n = n;
      hist->frames++;
    }
  }
  return 0;
}

/* evalSH (rc.ts:819-829) of probe pi of sh ([P0][27]): coefficients times shBasis(n) in order, clamped at 0. */
static v3 evalSH(const float* sh, uint64_t pi, v3 n) {
  float y[9] = {0.282095f, 0.488603f * n.y, 0.488603f * n.z, 0.488603f * n.x, 1.092548f * n.x * n.y,
                1.092548f * n.y * n.z, 0.315392f * (3.0f * n.z * n.z - 1.0f), 1.092548f * n.x * n.z,
                0.546274f * (n.x * n.x - n.y * n.y)};
  const float* v = sh + pi * 27;
  v3 e = V3(0, 0, 0);
  for (int k = 0; k < 9; k++) e = vadd(e, vscale(V3(v[k * 3], v[k * 3 + 1], v[k * 3 + 2]), y[k]));
  return V3(fmaxf(e.x, 0.0f), fmaxf(e.y, 0.0f), fmaxf(e.z, 0.0f));
}

/* segOccluded (rc.ts:845-869): min(ceil(|b - a| / v), 16) samples at the midpoints; one in an occupied voxel of
 * level L (voxAt: a dropped brick read coarser) hides the segment; samples outside the window are empty. */
static int segOccluded(const View* W, v3 a, v3 b, uint32_t L) {
  const orc_vox* p = W->p;
  v3 d = vsub(b, a);
  float v = voxSize(p, L);
  float kf = ceilf(sqrtf(vdot(d, d)) / v);
  uint32_t k = (uint32_t)(kf < 16.0f ? kf : 16.0f);
  v3 lo = vscale(vscale(vf(org(p, L)), v), 8.0f);
  int E = p->G * 8;
  for (uint32_t i = 0; i < k; i++) {
    v3 q = vadd(a, vscale(d, ((float)i + 0.5f) / (float)k));
    iv3 c = vi(vfloor(vdivs(vsub(q, lo), v)));
    if (c.x < 0 || c.y < 0 || c.z < 0 || c.x >= E || c.y >= E || c.z >= E) continue;
    uint32_t br, bit, lev;
    iv3 cell;
    if (voxAt(W, L, c, &br, &bit, &lev, &cell)) return 1;
  }
  return 0;
}

/* The c0 probe at corner k of the cell b around x: its gather weight and index (gatherWorld, rc.ts:886-895). */
static int gatherCorner(const orc_rc* rc, v3 x, v3 n, iv3 b, v3 f, int k, v3* pr, float* w, uint64_t* pi) {
  iv3 o = I3(k & 1, (k >> 1) & 1, (k >> 2) & 1);
  iv3 g = iadd(b, o);
  int N = rc->probes0;
  iv3 gl = isub(g, I3(rc->G0[0], rc->G0[1], rc->G0[2]));
  if (gl.x < 0 || gl.y < 0 || gl.z < 0 || gl.x >= N || gl.y >= N || gl.z >= N) return 0;
  v3 tw = V3(o.x ? f.x : 1.0f - f.x, o.y ? f.y : 1.0f - f.y, o.z ? f.z : 1.0f - f.z);
  *pi = (uint64_t)gl.x + (uint64_t)N * ((uint64_t)gl.y + (uint64_t)N * (uint64_t)gl.z);
  if (rc->reloc) {
    const float* P = rc->ppos[0] + *pi * 3;
    *pr = V3(P[0], P[1], P[2]);
  } else {
    *pr = vscale(vf(g), rc->s0);
  }
  v3 dv = vsub(*pr, x);
  float side = vdot(n, dv) / fmaxf(sqrtf(vdot(dv, dv)), 1e-5f);
  *w = (tw.x * tw.y * tw.z + 1e-4f) * ((side + 1.0f) * (side + 1.0f) * 0.25f + 0.02f);
  return 1;
}

/* The irradiance of the field described by rc (c0 grid, probe positions) at x for normal nn. */
static v3 gatherPoint(const View* W, const orc_rc* rc, const float* sh, v3 x, v3 nn, int occlusion) {
  const orc_vox* p = W->p;
  v3 gp = vdivs(x, rc->s0);
  v3 fl = vfloor(gp);
  iv3 b = vi(fl);
  v3 f = vsub(gp, vf(b));
  uint32_t Lo = voxLevelFor(p, (uint32_t)rc->vox_level[0], x);
  float vLo = voxSize(p, Lo);
  v3 xo1 = vadd(x, vscale(nn, 1.0f * vLo));
  v3 e = V3(0, 0, 0);
  float ws = 0.0f;
  uint32_t hidden = 0;
  for (int k = 0; k < 8; k++) {
    v3 pr;
    float w;
    uint64_t pi;
    if (!gatherCorner(rc, x, nn, b, f, k, &pr, &w, &pi)) continue;
    if (occlusion && segOccluded(W, xo1, pr, Lo)) {
      hidden |= 1u << k;
      continue;
    }
    e = vadd(e, vscale(evalSH(sh, pi, nn), w));
    ws += w;
  }
  if (ws <= 0.0f && hidden) {
    v3 xo2 = vadd(x, vscale(nn, 2.0f * vLo));
    for (int k = 0; k < 8; k++) {
      if (!((hidden >> k) & 1u)) continue;
      v3 pr;
      float w;
      uint64_t pi;
      gatherCorner(rc, x, nn, b, f, k, &pr, &w, &pi);
      if (segOccluded(W, xo2, pr, Lo)) continue;
      e = vadd(e, vscale(evalSH(sh, pi, nn), w));
      ws += w;
    }
  }
  return ws > 0.0f ? vdivs(e, ws) : V3(0, 0, 0);
}

int orc_gather(const orc_vox* p, const uint32_t* vox, const orc_rc* rc, const float* sh, const float* pos,
               const float* nrm, uint64_t n, int occlusion, float* irr) {
  View W;
  makeView(&W, p, vox);
  for (uint64_t i = 0; i < n; i++) {
    v3 out = gatherPoint(&W, rc, sh, V3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]),
                         V3(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]), occlusion);
    irr[3 * i] = out.x;
    irr[3 * i + 1] = out.y;
    irr[3 * i + 2] = out.z;
  }
  return 0;
}

static int voxCell(const View* W, uint32_t L, v3 p, uint32_t* br, uint32_t* bit);

void orc_bn_origin(const float eye[3], int32_t o[3]) {
  const int dims[3] = {ORC_BN_DX, ORC_BN_DY, ORC_BN_DZ};
  for (int a = 0; a < 3; a++) o[a] = (int32_t)floorf(eye[a] / ORC_BN_CELL) - dims[a] / 2;
}

/* sfDir (bounce.ts:155-160), evaluated in double and rounded to f32. */
static v3 sfDir(int i, int n) {
  double z = 1.0 - (2.0 * i + 1.0) / n;
  double r = sqrt(fmax(0.0, 1.0 - z * z));
  double f = i * 0.618033988749;
  double phi = 2.0 * 3.14159265 * (f - floor(f));
  return V3((float)(r * cos(phi)), (float)(r * sin(phi)), (float)z);
}

/* bnTrace (bounce.ts:164-175): rgb and the hit's entry distance (sky and 1e9 without a hit). */
static void bnTrace(const View* W, const orc_rc* rc, v3 o, v3 d, float tMax, float out[4]) {
  const orc_vox* p = W->p;
  float b0 = 0.0f;
  for (uint32_t k = 0; k < (uint32_t)p->levels && b0 * ORC_BN_CELL < tMax; k++) {
    float b1 = b0 * 4.0f + 1.0f;
    uint32_t L = voxLevelFor(p, k, vadd(o, vscale(vscale(d, b1), ORC_BN_CELL)));
    float r[4], tin = 0.0f;
    uint32_t hid, st;
    traceVoxT(W, rc, o, d, b0 * ORC_BN_CELL, fmn(b1 * ORC_BN_CELL, tMax), 0.0f, L, r, &hid, &st, &tin);
    if (r[3] < 0.5f) {
      v3 entry = vadd(o, vscale(d, tin));
      out[0] = r[0]; out[1] = r[1]; out[2] = r[2];
      out[3] = vdot(vsub(entry, o), d);
      return;
    }
    b0 = b1;
  }
  out[0] = rc->sky[0]; out[1] = rc->sky[1]; out[2] = rc->sky[2];
  out[3] = 1e9f;
}

int orc_bounce_volume(const orc_vox* p, const uint32_t* vox, const orc_rc* rc0, const uint16_t* prev,
                      const int32_t prev_org[3], uint32_t frame, const int32_t o3[3], uint16_t* cur) {
  View W;
  makeView(&W, p, vox);
  orc_rc rc = *rc0;
  rc.bounce = NULL;
  rc.bn_prev = prev;
  for (int a = 0; a < 3; a++) rc.bn_prev_org[a] = prev ? prev_org[a] : 0;
  /* levelOf (bounce.ts:110) */
  double lv = nearbyint(log2((double)(ORC_BN_CELL / p->v0)));
  uint32_t Lc = (uint32_t)fmin(fmax(lv, 0.0), (double)(p->levels - 1));
  v3 dirs[ORC_BN_RAYS];
  for (int r = 0; r < ORC_BN_RAYS; r++) dirs[r] = sfDir(r, ORC_BN_RAYS);
  const int n = ORC_BN_DX * ORC_BN_DY * ORC_BN_DZ;
  for (int i = 0; i < n; i++) {
    iv3 c = I3(i % ORC_BN_DX, (i / ORC_BN_DX) % ORC_BN_DY, i / (ORC_BN_DX * ORC_BN_DY));
    iv3 wp = iadd(I3(o3[0], o3[1], o3[2]), c);
    v3 pt = vscale(vf(wp), ORC_BN_CELL);
    uint16_t* out = cur + (size_t)i * 24;
    int near = 0;
    uint32_t br, bit;
    for (int k = 0; k < 8 && !near; k++) {
      v3 corner = vsub(V3((float)(k & 1), (float)((k >> 1) & 1), (float)(k >> 2)), V3(0.5f, 0.5f, 0.5f));
      near = voxCell(&W, Lc, vadd(pt, vscale(corner, ORC_BN_CELL)), &br, &bit);
    }
    if (!near || voxCell(&W, 0, pt, &br, &bit)) {
      memset(out, 0, 24 * sizeof(uint16_t));
      continue;
    }
    if (ORC_BN_DIV > 1 && ((uint32_t)(wp.x + wp.y + wp.z + 1048576) + frame) % ORC_BN_DIV != 0u && prev) {
      iv3 pc = isub(wp, I3(prev_org[0], prev_org[1], prev_org[2]));
      if (bnInside(pc) && bnReach(prev, pc, 0) > 0.0f) {
        memcpy(out, bnTexel(prev, pc, 0), 24 * sizeof(uint16_t));
        continue;
      }
    }
    float rad[ORC_BN_RAYS][4], reach[6];
    for (int r = 0; r < ORC_BN_RAYS; r++) {
      bnTrace(&W, &rc, pt, dirs[r], 1e9f, rad[r]);
      if (rc.objects && rc.dyn) {
        /* bnTraceDyn (bounce.ts:173-178): the layer in front of the static result, read with one ray's cone */
        float dl[4];
        traceDyn(p, rc.dyn, pt, dirs[r], 0.0f, fmn(rad[r][3], 1e4f), sqrtf(4.0f * PI_F / (float)ORC_BN_RAYS), dl);
        for (int k = 0; k < 3; k++) rad[r][k] = dl[k] + (1.0f - dl[3]) * rad[r][k];
      }
    }
    for (int a = 0; a < 6; a++) {
      float sg = (a & 1) ? -1.0f : 1.0f;
      v3 d = V3(a / 2 == 0 ? sg : 0.0f, a / 2 == 1 ? sg : 0.0f, a / 2 == 2 ? sg : 0.0f);
      float r[4];
      bnTrace(&W, &rc, pt, d, 1.5f * ORC_BN_CELL, r);
      reach[a] = fmx(fmn(r[3], 1.5f * ORC_BN_CELL), 1e-3f);
    }
    for (int a = 0; a < 6; a++) {
      float sg = (a & 1) ? -1.0f : 1.0f;
      v3 e = V3(0, 0, 0);
      float ws = 0.0f;
      for (int r = 0; r < ORC_BN_RAYS; r++) {
        float comp = a / 2 == 0 ? dirs[r].x : (a / 2 == 1 ? dirs[r].y : dirs[r].z);
        float w = fmx(sg * comp, 0.0f);
        e = vadd(e, vscale(V3(rad[r][0], rad[r][1], rad[r][2]), w));
        ws += w;
      }
      float m = fmx(ws, 1e-6f);
      out[a * 4 + 0] = orc_f2h(PI_F * e.x / m);
      out[a * 4 + 1] = orc_f2h(PI_F * e.y / m);
      out[a * 4 + 2] = orc_f2h(PI_F * e.z / m);
      out[a * 4 + 3] = orc_f2h(reach[a]);
    }
  }
  return 0;
}

int orc_bounce_inject(const orc_vox* p, const uint32_t* vox, const uint16_t* vol, const int32_t vorg[3],
                      uint32_t* bounce, int bf16) {
  View W;
  makeView(&W, p, vox);
  const float tol = 0.5f * p->v0;
  uint32_t nb = W.vox[0] < (uint32_t)p->max_bricks ? W.vox[0] : (uint32_t)p->max_bricks;
  for (uint32_t ptr = 0; ptr < nb; ptr++) {
    uint32_t base = W.brickOff + ptr * BW;
    for (uint32_t bit = 0; bit < 512; bit++) {
      if ((W.vox[base + (bit >> 5)] & (1u << (bit & 31u))) == 0u) continue;
      uint32_t ai = voxRank(&W, ptr, bit);
      if (ai == V_EMPTY) continue;
      uint32_t code = W.vox[base + 17];
      uint32_t L = code >> 27;
      iv3 b = I3((int)((code >> 18) & 511u), (int)((code >> 9) & 511u), (int)(code & 511u));
      iv3 c = iadd(imuls(b, 8), I3((int)(bit & 7u), (int)((bit >> 3) & 7u), (int)(bit >> 6)));
      float v = voxSize(p, L);
      v3 pc = vadd(vscale(vscale(vf(org(p, L)), v), 8.0f), vscale(vadd(vf(c), V3(0.5f, 0.5f, 0.5f)), v));
      uint32_t pn = W.vox[W.attrOff + ai * 2 + 1];
      float ex = fmx((float)(int16_t)(uint16_t)(pn & 0xffffu) / 32767.0f, -1.0f);
      float ey = fmx((float)(int16_t)(uint16_t)(pn >> 16) / 32767.0f, -1.0f);
      v3 n = V3(ex, ey, 1.0f - fabsf(ex) - fabsf(ey));
      float q0 = fmx(-n.z, 0.0f);
      n.x += n.x >= 0.0f ? -q0 : q0;
      n.y += n.y >= 0.0f ? -q0 : q0;
      v3 nf = vdot(n, n) > 0.01f ? vnormalize(n) : V3(0.0f, 1.0f, 0.0f);
      float off = L == 0 ? 0.0f : 0.5f * v;
      v3 ef = bnRead(vol, vorg, tol, vadd(pc, vscale(nf, off)), pc, nf);
      v3 eb = bnRead(vol, vorg, tol, vadd(pc, vscale(vneg(nf), off)), pc, vneg(nf));
      uint32_t* o = bounce + (size_t)ai * 4;
      uint16_t (*enc)(float) = bf16 ? f2bf : orc_f2h; /* precision form: the injected light as bf16 */
      o[0] = (uint32_t)enc(ef.x) | ((uint32_t)enc(ef.y) << 16);
      o[1] = (uint32_t)enc(ef.z) | ((uint32_t)enc(0.0f) << 16);
      o[2] = (uint32_t)enc(eb.x) | ((uint32_t)enc(eb.y) << 16);
      o[3] = (uint32_t)enc(eb.z) | ((uint32_t)enc(0.0f) << 16);
    }
  }
  return 0;
}

/* A shadow ray from x toward the sun: traced at level L, then from each level's window exit on at the next coarser
 * level; occupied voxels entered within tSkip are ignored. 1 if it leaves the clipmap unhit. */
static float sunRay(const View* W, const orc_rc* rc, v3 x, v3 d, uint32_t L, float tSkip) {
  const orc_vox* p = W->p;
  v3 dd = V3(fabsf(d.x) < 1e-9f ? 1e-9f : d.x, fabsf(d.y) < 1e-9f ? 1e-9f : d.y, fabsf(d.z) < 1e-9f ? 1e-9f : d.z);
  v3 inv = V3(1.0f / dd.x, 1.0f / dd.y, 1.0f / dd.z);
  float t = 0.0f;
  for (uint32_t l = L; l < (uint32_t)p->levels; l++) {
    float out[4];
    uint32_t hid, st;
    traceVox(W, rc, x, d, t, 1e30f, tSkip, l, out, &hid, &st);
    if (hid != 0xffffffffu) return 0.0f;
    float v = voxSize(p, l);
    v3 lo = vscale(vscale(vf(org(p, l)), v), 8.0f);
    float span = (float)p->G * v * 8.0f;
    v3 hi = vadd(lo, V3(span, span, span));
    v3 ta = vmul(vsub(lo, x), inv), tb = vmul(vsub(hi, x), inv);
    v3 tf = vmax(ta, tb);
    t = fmx(t, fmn(fmn(tf.x, tf.y), tf.z));
  }
  return 1.0f;
}

int orc_sun(const orc_vox* p, const uint32_t* vox, const orc_rc* rc, float* sunvis) {
  View W;
  makeView(&W, p, vox);
  uint32_t nb = W.vox[0] < (uint32_t)p->max_bricks ? W.vox[0] : (uint32_t)p->max_bricks;
  v3 sd = V3(rc->sun_dir[0], rc->sun_dir[1], rc->sun_dir[2]);
  for (uint32_t ptr = 0; ptr < nb; ptr++) {
    uint32_t base = W.brickOff + ptr * BW;
    for (uint32_t bit = 0; bit < 512; bit++) {
      if ((W.vox[base + (bit >> 5)] & (1u << (bit & 31u))) == 0u) continue;
      uint32_t ai = voxRank(&W, ptr, bit);
      if (ai == V_EMPTY) continue;
      uint32_t code = W.vox[base + 17];
      uint32_t L = code >> 27;
      iv3 b = I3((int)((code >> 18) & 511u), (int)((code >> 9) & 511u), (int)(code & 511u));
      iv3 c = iadd(imuls(b, 8), I3((int)(bit & 7u), (int)((bit >> 3) & 7u), (int)(bit >> 6)));
      float v = voxSize(p, L);
      v3 pc = vadd(vscale(vscale(vf(org(p, L)), v), 8.0f), vscale(vadd(vf(c), V3(0.5f, 0.5f, 0.5f)), v));
      uint32_t pn = W.vox[W.attrOff + ai * 2 + 1];
      float ex = fmx((float)(int16_t)(uint16_t)(pn & 0xffffu) / 32767.0f, -1.0f);
      float ey = fmx((float)(int16_t)(uint16_t)(pn >> 16) / 32767.0f, -1.0f);
      v3 n = V3(ex, ey, 1.0f - fabsf(ex) - fabsf(ey));
      float q0 = fmx(-n.z, 0.0f);
      n.x += n.x >= 0.0f ? -q0 : q0;
      n.y += n.y >= 0.0f ? -q0 : q0;
      v3 q = pc;
      if (vdot(n, n) > 1e-4f) {
        n = vnormalize(n);
        float pl;
        memcpy(&pl, &W.vox[W.plOff + ai], 4);
        q = vadd(pc, vscale(n, pl));
      } else {
        n = V3(0.0f, 1.0f, 0.0f);
      }
      v3 ns = vdot(n, sd) < 0.0f ? vneg(n) : n;
      v3 t1 = vnormalize(vcross(ns, fabsf(ns.y) > 0.9f ? V3(1.0f, 0.0f, 0.0f) : V3(0.0f, 1.0f, 0.0f)));
      v3 t2 = vcross(ns, t1);
      float h = 0.25f * v;
      v3 o = vadd(q, vscale(ns, 0.25f * v + 0.03f));
      v3 pp = vscale(vadd(t1, t2), h), pm = vscale(vsub(t1, t2), h);
      float vis = 0.25f * (sunRay(&W, rc, vadd(o, pp), sd, L, v) + sunRay(&W, rc, vadd(o, pm), sd, L, v) +
                           sunRay(&W, rc, vsub(o, pp), sd, L, v) + sunRay(&W, rc, vsub(o, pm), sd, L, v));
      sunvis[ai] = vis;
    }
  }
  return 0;
}

/* voxCell (rc.ts:379-386): the occupied voxel of level L at p (voxAt: a dropped brick read coarser), or 0 when
 * the cell is empty or outside the window. */
static int voxCell(const View* W, uint32_t L, v3 p, uint32_t* br, uint32_t* bit) {
  const orc_vox* pp = W->p;
  float v = voxSize(pp, L);
  iv3 c = vi(vfloor(vdivs(vsub(p, vscale(vscale(vf(org(pp, L)), v), 8.0f)), v)));
  int E = pp->G * 8;
  if (c.x < 0 || c.y < 0 || c.z < 0 || c.x >= E || c.y >= E || c.z >= E) return 0;
  uint32_t lev;
  iv3 cell;
  return voxAt(W, L, c, br, bit, &lev, &cell);
}
/* voxStoredNormal (rc.ts:387-393); a brick without attribute slots gives (0, 0, 1). */
static v3 storedNormal(const View* W, uint32_t br, uint32_t bit) {
  uint32_t ai = voxRank(W, br, bit);
  uint32_t pn = ai == V_EMPTY ? 0u : W->vox[W->attrOff + ai * 2 + 1];
  float nx = fmaxf((float)(int16_t)(uint16_t)(pn & 0xffffu) / 32767.0f, -1.0f);
  float ny = fmaxf((float)(int16_t)(uint16_t)(pn >> 16) / 32767.0f, -1.0f);
  v3 n = V3(nx, ny, 1.0f - fabsf(nx) - fabsf(ny));
  float q = fmx(-n.z, 0.0f);
  n.x += n.x >= 0.0f ? -q : q;
  n.y += n.y >= 0.0f ? -q : q;
  return n;
}

int orc_relocate(const orc_vox* p, const uint32_t* vox, const orc_rc* rc, int c, float* pos, float* nrm) {
  View W;
  makeView(&W, p, vox);
  int n = probesOf(rc, c);
  float s = rc->s0 * (float)(1u << c);
  for (int64_t pi = 0; pi < (int64_t)n * n * n; pi++) {
    iv3 g = I3((rc->G0[0] >> c) + (int)(pi % n), (rc->G0[1] >> c) + (int)((pi / n) % n),
               (rc->G0[2] >> c) + (int)(pi / ((int64_t)n * n)));
    v3 pp = vscale(vf(g), s);
    uint32_t L = voxLevelFor(p, (uint32_t)rc->vox_level[c], pp);
    uint32_t Lg = voxLevelFor(p, L + 1 < (uint32_t)p->levels ? L + 1 : (uint32_t)p->levels - 1, pp);
    float vg = voxSize(p, Lg);
    /* The 4^3 cube of the next level around the grid point (cubeOff, rc.ts:395). */
    v3 grad = V3(0, 0, 0), nsum = V3(0, 0, 0);
    float cnt = 0.0f;
    for (int k = 0; k < 64; k++) {
      v3 off = vscale(vsub(V3((float)(k & 3), (float)((k >> 2) & 3), (float)(k >> 4)), V3(1.5f, 1.5f, 1.5f)), vg);
      uint32_t br, bit;
      if (!voxCell(&W, Lg, vadd(pp, off), &br, &bit)) continue;
      grad = vsub(grad, off);
      cnt += 1.0f;
      nsum = vadd(nsum, storedNormal(&W, br, bit));
    }
    int flag = cnt < 1.0f ? 0 : (sqrtf(vdot(grad, grad)) < 0.1f * cnt * vg ? 1 : 2);
    v3 dir = flag == 1 ? nsum : grad;
    if (flag != 0) {
      if (sqrtf(vdot(dir, dir)) < 1e-4f) flag = 0;
      else dir = vnormalize(dir);
    }
    v3 r = pp, nn = V3(0, 0, 0);
    if (flag != 0) {
      nn = dir;
      float back = rc->lift * s;
      float stp = fmx(0.5f * voxSize(p, L), 1.5f * s / 16.0f);
      for (int i = 0; i <= 16; i++) {
        float di = (float)i * stp;
        if (di > 1.5f * s) break;
        v3 q = vadd(pp, vscale(dir, di));
        uint32_t br, bit;
        if (!voxCell(&W, L, q, &br, &bit) && !voxCell(&W, L, vsub(q, vscale(dir, back)), &br, &bit)) {
          r = q;
          break;
        }
      }
    }
    memcpy(pos + pi * 3, &r, 12);
    memcpy(nrm + pi * 3, &nn, 12);
  }
  return 0;
}

int orc_link(const orc_vox* p, const uint32_t* vox, const orc_rc* rc, int c, const float* pos_c, const float* pos_u,
             uint32_t* vis) {
  if (c + 1 >= rc->cascades) return -1;
  View W;
  makeView(&W, p, vox);
  int n = probesOf(rc, c), un = probesOf(rc, c + 1);
  for (int64_t pi = 0; pi < (int64_t)n * n * n; pi++) {
    iv3 g = I3((int)(pi % n), (int)((pi / n) % n), (int)(pi / ((int64_t)n * n)));
    v3 a = V3(pos_c[pi * 3], pos_c[pi * 3 + 1], pos_c[pi * 3 + 2]);
    uint32_t Lv = voxLevelFor(p, (uint32_t)rc->vox_level[c], a);
    uint32_t m = 0;
    for (int k = 0; k < 8; k++) {
      iv3 o = I3(k & 1, (k >> 1) & 1, (k >> 2) & 1);
      if ((o.x > 0 && !(g.x & 1)) || (o.y > 0 && !(g.y & 1)) || (o.z > 0 && !(g.z & 1))) continue;
      iv3 q = iadd(ishr(g, 1), o);
      if (q.x >= un || q.y >= un || q.z >= un) continue;
      uint64_t qi = (uint64_t)q.x + (uint64_t)un * ((uint64_t)q.y + (uint64_t)un * q.z);
      v3 b = V3(pos_u[qi * 3], pos_u[qi * 3 + 1], pos_u[qi * 3 + 2]);
      if (!segOccluded(&W, a, b, Lv)) m |= 1u << k;
    }
    vis[pi] = m;
  }
  return 0;
}

/* ------------------------------------------------------------------------------------------------------------ */
/* FP64 oracle: the same algorithm in double precision (directions, ray-voxel traversal, shading, merge, SH), on   */
/* the same occupancy and attributes.                                                                            */

typedef struct { double x, y, z; } d3;
static d3 D3(double x, double y, double z) { d3 r = {x, y, z}; return r; }
static double ddot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

static d3 dirOfD(int di, int res) {
  double u = ((di % res) + 0.5) / res, w = ((di / res) + 0.5) / res;
  double fx = u * 2 - 1, fy = w * 2 - 1;
  d3 n = D3(fx, fy, 1 - fabs(fx) - fabs(fy));
  double t = n.z < 0 ? -n.z : 0;
  n.x += n.x >= 0 ? -t : t;
  n.y += n.y >= 0 ? -t : t;
  double l = sqrt(ddot(n, n));
  return D3(n.x / l, n.y / l, n.z / l);
}

static d3 shadeVoxelD(const View* W, const orc_rc* rc, uint32_t ai, d3 d, d3 entry) {
  uint32_t pa = 0, pn = 0;
  if (ai != V_EMPTY) {
    pa = W->vox[W->attrOff + ai * 2];
    pn = W->vox[W->attrOff + ai * 2 + 1];
  }
  d3 alb = D3((pa & 255u) / 255.0, ((pa >> 8) & 255u) / 255.0, ((pa >> 16) & 255u) / 255.0);
  if (rc->bf16) { /* the stored bf16 albedo, widened exactly */
    uint32_t w0 = ai != V_EMPTY ? W->vox[W->alOff + ai * 2] : 0u, w1 = ai != V_EMPTY ? W->vox[W->alOff + ai * 2 + 1] : 0u;
    alb = D3(bf2f((uint16_t)(w0 & 0xffffu)), bf2f((uint16_t)(w0 >> 16)), bf2f((uint16_t)(w1 & 0xffffu)));
  }
  double nx = fmax((int16_t)(uint16_t)(pn & 0xffffu) / 32767.0, -1.0), ny = fmax((int16_t)(uint16_t)(pn >> 16) / 32767.0, -1.0);
  d3 n = D3(nx, ny, 1 - fabs(nx) - fabs(ny));
  double q = n.z < 0 ? -n.z : 0;
  n.x += n.x >= 0 ? -q : q;
  n.y += n.y >= 0 ? -q : q;
  if (!(ddot(n, n) > 0.01)) n = D3(-d.x, -d.y, -d.z);
  if (ddot(n, d) > 0) n = D3(-n.x, -n.y, -n.z);
  double ndl = fmax(n.x * rc->sun_dir[0] + n.y * rc->sun_dir[1] + n.z * rc->sun_dir[2], 0.0);
  d3 out = D3(0, 0, 0);
  if (ndl > 0) {
    double k = ndl * rc->sun_vis / M_PI;
    out = D3(alb.x * rc->sun_e[0] * k, alb.y * rc->sun_e[1] * k, alb.z * rc->sun_e[2] * k);
  }
  for (int i = 0; i < rc->n_point; i++) {
    d3 l = D3(rc->point_pos[i][0] - entry.x, rc->point_pos[i][1] - entry.y, rc->point_pos[i][2] - entry.z);
    double r2 = ddot(l, l), nl = ddot(n, l) / sqrt(r2);
    if (nl > 0) {
      double k = nl / r2 / M_PI;
      out.x += alb.x * rc->point_i[i][0] * k;
      out.y += alb.y * rc->point_i[i][1] * k;
      out.z += alb.z * rc->point_i[i][2] * k;
    }
  }
  return out;
}

/* Two values closer than eps: their order may differ between f32 and f64 evaluation. Exact equality counts too:
 * values that tie exactly in double (a ray through a voxel edge or corner) need not tie in f32. */
static int near(double a, double b, double eps) { return fabs(a - b) < eps; }

/* Voxel-by-voxel traversal of level L in double; the first occupied voxel entered at t >= tSkip. */
static void traceVoxD(const View* W, const orc_rc* rc, d3 o, d3 d, double t0, double t1, double tSkip, uint32_t L,
                      double eps, double out[4], uint32_t* hid, uint8_t* tie) {
  const orc_vox* p = W->p;
  out[0] = out[1] = out[2] = 0;
  out[3] = 1;
  *hid = 0xffffffffu;
  *tie = 0;
  double v = p->v0 * (double)(1u << L);
  int G = p->G, nVox = G * 8;
  double lo[3] = {p->origin[L][0] * v * 8, p->origin[L][1] * v * 8, p->origin[L][2] * v * 8};
  double dv[3] = {d.x, d.y, d.z}, ov[3] = {o.x, o.y, o.z};
  double dd[3], inv[3], t = t0, tEnd = t1;
  for (int a = 0; a < 3; a++) {
    dd[a] = fabs(dv[a]) < 1e-9 ? 1e-9 : dv[a];
    inv[a] = 1.0 / dd[a];
    double ta = (lo[a] - ov[a]) * inv[a], tb = (lo[a] + nVox * v - ov[a]) * inv[a];
    double tn = ta < tb ? ta : tb, tf = ta < tb ? tb : ta;
    if (tn > t) t = tn;
    if (tf < tEnd) tEnd = tf;
  }
  double scale = fabs(t1) + 1.0;
  if (near(t, tEnd, eps * scale)) *tie = 1;
  if (t >= tEnd) return;
  int c[3], st[3];
  double tMax[3], tDelta[3];
  for (int a = 0; a < 3; a++) {
    st[a] = dd[a] > 0 ? 1 : -1;
    double x = (ov[a] + dv[a] * t - lo[a]) / v;
    int ci = (int)floor(x);
    /* A start on a voxel plane is exact in both precisions when the interval starts at the probe (t0 = 0). */
    if (t != 0 && near(x, floor(x + 0.5), eps * (fabs(x) + 1)) && fabs(dv[a]) >= 1e-9) *tie = 1;
    c[a] = ci < 0 ? 0 : (ci > nVox - 1 ? nVox - 1 : ci);
    tMax[a] = ((c[a] + (st[a] > 0 ? 1 : 0)) * v + lo[a] - ov[a]) * inv[a];
    tDelta[a] = fabs(v * inv[a]);
  }
  double tIn = t;
  for (int it = 0; it < 4 * nVox; it++) {
    uint32_t br, bit, lev;
    iv3 cell;
    if (voxAt(W, L, I3(c[0], c[1], c[2]), &br, &bit, &lev, &cell)) {
      if (near(tIn, tSkip, eps * scale)) *tie = 1;
      if (tIn >= tSkip) {
        *hid = hitId(p, lev, cell);
        d3 e = D3(o.x + d.x * tIn, o.y + d.y * tIn, o.z + d.z * tIn);
        d3 r = shadeVoxelD(W, rc, voxRank(W, br, bit), d, e);
        out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = 0;
        return;
      }
    }
    int a = tMax[0] <= tMax[1] && tMax[0] <= tMax[2] ? 0 : (tMax[1] <= tMax[2] ? 1 : 2);
    double tx = tMax[a];
    for (int b2 = 0; b2 < 3; b2++)
      if (b2 != a && near(tMax[b2], tx, eps * scale)) *tie = 1;
    if (near(tx, tEnd, eps * scale)) *tie = 1;
    if (tx >= tEnd) break;
    tIn = tx;
    c[a] += st[a];
    tMax[a] += tDelta[a];
    if (c[a] < 0 || c[a] >= nVox) break;
  }
}

/* One ray of cascade c in double: the level (the smallest level >= voxLevel whose window strictly contains the
 * interval's end), then the traversal. tie: as traceVoxD, or the end lies within eps of a window face. */
static void traceRayD(const View* W, const orc_rc* rc, int c, d3 o, d3 d, double eps, double out[4], uint32_t* h,
                      uint8_t* tie) {
  const orc_vox* p = W->p;
  double t0 = (double)rc->b0[c] * rc->s0, t1 = (double)rc->b1[c] * rc->s0;
  d3 e = D3(o.x + d.x * t1, o.y + d.y * t1, o.z + d.z * t1);
  uint32_t L = (uint32_t)rc->vox_level[c];
  uint8_t tl = 0;
  while (L + 1 < (uint32_t)p->levels) {
    double bs = p->v0 * (double)(1u << L) * 8;
    double ev[3] = {e.x, e.y, e.z};
    int in = 1;
    for (int a = 0; a < 3; a++) {
      double lo = p->origin[L][a] * bs, hi = lo + p->G * bs;
      if (near(ev[a], lo, eps * (fabs(lo) + 1)) || near(ev[a], hi, eps * (fabs(hi) + 1))) tl = 1;
      if (!(ev[a] > lo && ev[a] < hi)) in = 0;
    }
    if (in) break;
    L++;
  }
  double v = p->v0 * (double)(1u << L);
  uint8_t tt;
  traceVoxD(W, rc, o, d, t0, t1, (double)rc->vox_skip * v, L, eps, out, h, &tt);
  *tie = (uint8_t)(tt | tl);
}

int orc_trace_f64(const orc_vox* p, const uint32_t* vox, const orc_rc* rc, uint64_t r0, uint64_t r1, double* rgbt,
                  uint32_t* hits, uint8_t* tie, double eps, int threads) {
  View W;
  makeView(&W, p, vox);
#ifdef _OPENMP
  if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 4096) num_threads(threads)
#endif
  for (int64_t q = (int64_t)r0; q < (int64_t)r1; q++) {
    uint64_t r = 0;
    int c = cascadeOf(rc, (uint64_t)q, &r);
    int res = resOf(rc, c), n = probesOf(rc, c);
    uint64_t dirs = (uint64_t)res * res, pi = r / dirs;
    int di = (int)(r % dirs);
    double s = (double)rc->s0 * (double)(1u << c);
    d3 o = D3(((rc->G0[0] >> c) + (int)(pi % n)) * s, ((rc->G0[1] >> c) + (int)((pi / n) % n)) * s,
              ((rc->G0[2] >> c) + (int)(pi / ((uint64_t)n * n))) * s);
    double out[4];
    uint32_t h;
    uint8_t tt;
    traceRayD(&W, rc, c, o, dirOfD(di, res), eps, out, &h, &tt);
    size_t i = (size_t)(q - (int64_t)r0);
    if (rgbt) memcpy(rgbt + i * 4, out, 32);
    if (hits) hits[i] = h;
    if (tie) tie[i] = tt;
  }
  (void)threads;
  return 0;
}

/* The radiance in double of a ray that enters the voxel of hit id `id` (level << 29 | window cell): entered where the
 * ray crosses into that voxel's box (the larger of the three slab entries), shaded in double; a miss (NO_HIT) returns
 * (0, 0, 0, 1). */
static void shadeIdD(const View* W, const orc_rc* rc, d3 o, d3 d, uint32_t id, double out[4]) {
  out[0] = out[1] = out[2] = 0;
  out[3] = 1;
  if (id == 0xffffffffu) return;
  const orc_vox* p = W->p;
  uint32_t L = id >> 29, idx = id & ((1u << 29) - 1u), E = (uint32_t)p->G * 8u;
  iv3 cell = I3((int)(idx % E), (int)((idx / E) % E), (int)(idx / (E * E)));
  double v = p->v0 * (double)(1u << L);
  double lo[3] = {(p->origin[L][0] * 8.0 + cell.x) * v, (p->origin[L][1] * 8.0 + cell.y) * v,
                  (p->origin[L][2] * 8.0 + cell.z) * v};
  double ov[3] = {o.x, o.y, o.z}, dv[3] = {d.x, d.y, d.z}, t = -1e300;
  for (int a = 0; a < 3; a++) {
    double dd = fabs(dv[a]) < 1e-9 ? 1e-9 : dv[a];
    double ta = (lo[a] - ov[a]) / dd, tb = (lo[a] + v - ov[a]) / dd;
    double tn = ta < tb ? ta : tb;
    if (tn > t) t = tn;
  }
  uint32_t br, bit, lev;
  iv3 cc;
  uint32_t ai = voxAt(W, L, cell, &br, &bit, &lev, &cc) ? voxRank(W, br, bit) : V_EMPTY;
  d3 r = shadeVoxelD(W, rc, ai, d, D3(o.x + d.x * t, o.y + d.y * t, o.z + d.z * t));
  out[0] = r.x;
  out[1] = r.y;
  out[2] = r.z;
  out[3] = 0;
}

int orc_trace_f64_take(const orc_vox* p, const uint32_t* vox, const orc_rc* rc, uint64_t r0, uint64_t r1,
                       const uint32_t* hits32, double* rgbt, uint8_t* taken, double eps, int threads) {
  View W;
  makeView(&W, p, vox);
#ifdef _OPENMP
  if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 4096) num_threads(threads)
#endif
  for (int64_t q = (int64_t)r0; q < (int64_t)r1; q++) {
    uint64_t r = 0;
    int c = cascadeOf(rc, (uint64_t)q, &r);
    int res = resOf(rc, c), n = probesOf(rc, c);
    uint64_t dirs = (uint64_t)res * res, pi = r / dirs;
    int di = (int)(r % dirs);
    double s = (double)rc->s0 * (double)(1u << c);
    d3 o = D3(((rc->G0[0] >> c) + (int)(pi % n)) * s, ((rc->G0[1] >> c) + (int)((pi / n) % n)) * s,
              ((rc->G0[2] >> c) + (int)(pi / ((uint64_t)n * n))) * s);
    d3 d = dirOfD(di, res);
    double out[4];
    uint32_t h;
    uint8_t tt;
    traceRayD(&W, rc, c, o, d, eps, out, &h, &tt);
    size_t i = (size_t)(q - (int64_t)r0);
    uint8_t take = (uint8_t)(tt || h != hits32[i]);
    if (take) shadeIdD(&W, rc, o, d, hits32[i], out);
    memcpy(rgbt + i * 4, out, 32);
    if (taken) taken[i] = take;
  }
  (void)threads;
  return 0;
}

int orc_trace_rays_f64(const orc_vox* p, const uint32_t* vox, const orc_rc* rc, int c, const double* o,
                       const double* d, uint64_t n, double* rgbt, uint32_t* hits, uint8_t* tie, double eps,
                       int threads) {
  if (c < 0 || c >= rc->cascades) return -1;
  View W;
  makeView(&W, p, vox);
#ifdef _OPENMP
  if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads)
#endif
  for (int64_t i = 0; i < (int64_t)n; i++) {
    double out[4];
    uint32_t h;
    uint8_t tt;
    traceRayD(&W, rc, c, D3(o[3 * i], o[3 * i + 1], o[3 * i + 2]), D3(d[3 * i], d[3 * i + 1], d[3 * i + 2]), eps, out,
              &h, &tt);
    if (rgbt) memcpy(rgbt + (size_t)i * 4, out, 32);
    if (hits) hits[i] = h;
    if (tie) tie[i] = tt;
  }
  (void)threads;
  return 0;
}

int orc_merge_sh_f64(const orc_rc* rc, double* rgbt, double* sh) {
  int Nc = rc->cascades;
  for (int c = Nc - 1; c >= 0; c--) {
    int res = resOf(rc, c), n = probesOf(rc, c);
    uint64_t off = rayOff(rc, c), nr = orc_rays(rc, c), dirs = (uint64_t)res * res;
    for (uint64_t t = 0; t < nr; t++) {
      double* v = rgbt + (off + t) * 4;
      double up[3] = {rc->sky[0], rc->sky[1], rc->sky[2]};
      if (c + 1 < Nc && v[3] > 0) {
        int ur = resOf(rc, c + 1), un = probesOf(rc, c + 1);
        uint64_t uoff = rayOff(rc, c + 1), pi = t / dirs;
        int di = (int)(t % dirs), u0 = di % res, w0 = di / res;
        int g[3] = {(int)(pi % n), (int)((pi / n) % n), (int)(pi / ((uint64_t)n * n))};
        double acc[3] = {0, 0, 0}, ws = 0;
        for (int k = 0; k < 8; k++) {
          int o[3] = {k & 1, (k >> 1) & 1, (k >> 2) & 1}, q[3], skip = 0;
          for (int a = 0; a < 3; a++) {
            if (o[a] && !(g[a] & 1)) skip = 1;
            q[a] = (g[a] >> 1) + o[a];
            if (q[a] >= un) skip = 1;
          }
          if (skip) continue;
          uint64_t qi = (uint64_t)q[0] + (uint64_t)un * ((uint64_t)q[1] + (uint64_t)un * q[2]);
          const double* U = rgbt + (uoff + qi * (uint64_t)ur * ur) * 4;
          for (int ch = 0; ch < 3; ch++) {
            if (ur == res) acc[ch] += U[((uint64_t)w0 * ur + u0) * 4 + ch];
            else {
              const double* b = U + ((uint64_t)w0 * 2 * ur + (uint64_t)u0 * 2) * 4;
              acc[ch] += 0.25 * (b[ch] + b[4 + ch] + b[(uint64_t)ur * 4 + ch] + b[(uint64_t)ur * 4 + 4 + ch]);
            }
          }
          ws += 1;
        }
        for (int ch = 0; ch < 3; ch++) up[ch] = ws > 0 ? acc[ch] / ws : 0;
      }
      for (int ch = 0; ch < 3; ch++) v[ch] += v[3] * up[ch];
      v[3] = 1;
    }
  }
  if (sh) {
    int n = probesOf(rc, 0);
    uint64_t np = (uint64_t)n * n * n;
    for (uint64_t pi = 0; pi < np; pi++) {
      double cf[27] = {0};
      for (int d = 0; d < 16; d++) {
        const double* x = rgbt + (pi * 16 + d) * 4;
        d3 nn = dirOfD(d, 4);
        double y[9] = {0.282095, 0.488603 * nn.y, 0.488603 * nn.z, 0.488603 * nn.x, 1.092548 * nn.x * nn.y,
                       1.092548 * nn.y * nn.z, 0.315392 * (3 * nn.z * nn.z - 1), 1.092548 * nn.x * nn.z,
                       0.546274 * (nn.x * nn.x - nn.y * nn.y)};
        for (int k = 0; k < 9; k++)
          for (int ch = 0; ch < 3; ch++) cf[k * 3 + ch] += x[ch] * rc->omega0[d] * y[k];
      }
      for (int k = 0; k < 9; k++) {
        double a = k == 0 ? M_PI : (k < 4 ? 2 * M_PI / 3 : M_PI / 4);
        for (int ch = 0; ch < 3; ch++) sh[pi * 27 + k * 3 + ch] = cf[k * 3 + ch] * a;
      }
    }
  }
  return 0;
}

/* posedVertices, scenes.ts:80-102: per vertex x = rest - pivot in double (JavaScript numbers), the turn about +y
 * (c, s) and the offset, each component stored into a Float32Array (rounded once). Deformations (scenes.ts:56-78)
 * are not restated. */
void orc_pose(uint32_t n, const uint32_t* first, const uint32_t* count, const double* pivot, const double* pose,
              const float* rest_pos, const float* rest_nrm, float* pos, float* nrm) {
  if (!n) return;
  const uint32_t f0 = first[0];
  for (uint32_t i = 0; i < n; i++) {
    const double* pv = pivot + 3 * i;
    const double c = pose[5 * i], s = pose[5 * i + 1];
    const double* off = pose + 5 * i + 2;
    for (uint32_t v = first[i]; v < first[i] + count[i]; v++) {
      const float* r = rest_pos + (size_t)(v - f0) * 3;
      const float* rn = rest_nrm + (size_t)(v - f0) * 3;
      double x = r[0] - pv[0], y = r[1] - pv[1], z = r[2] - pv[2];
      double nx = rn[0], ny = rn[1], nz = rn[2];
      pos[(size_t)v * 3] = (float)(pv[0] + c * x + s * z + off[0]);
      pos[(size_t)v * 3 + 1] = (float)(pv[1] + y + off[1]);
      pos[(size_t)v * 3 + 2] = (float)(pv[2] - s * x + c * z + off[2]);
      nrm[(size_t)v * 3] = (float)(c * nx + s * nz);
      nrm[(size_t)v * 3 + 1] = (float)ny;
      nrm[(size_t)v * 3 + 2] = (float)(-s * nx + c * nz);
    }
  }
}
