// Base-colour texture sampling of the voxel passes (surfaceAt, src/core/voxel.ts:325-336; shadeSample,
// src/core/dynlayer.ts:451-452) on gfx950: textureSampleLevel(baseColorTex, samp, uv, layer, 3 + L) with the voxel
// passes' sampler (linear min/mag/mip filtering, repeat addressing, voxel.ts:793). An integer level selects one mip,
// so a sample is the bilinear mean of the four texels around uv on level min(3 + L, last): rgb decoded from sRGB
// before filtering (rgba8unorm-srgb, scene_gpu.ts:66), weights (1 - f, f) per axis, the x lerps first. Every value
// is one f32 expression in a fixed order (the sources build with -ffp-contract=off), so the oracle's restatement
// gives the same bits.
#pragma once

#include "rtgi_internal.h"
#include "vmath.h"

// First texel of mip m in a layer of 2^s x 2^s texels: the sizes of the levels above it summed.
__device__ __forceinline__ uint32_t surf_mip_off(uint32_t s, uint32_t m) {
  uint32_t off = 0u;
  for (uint32_t j = 0; j < m; j++) off += 1u << (2u * (s - j));
  return off;
}

// Texel (x, y) of a level of edge w at texel offset base: rgb decoded, alpha c / 255.
__device__ __forceinline__ float4 surf_texel(const SurfK& s, uint32_t base, uint32_t w, uint32_t x, uint32_t y) {
  const uint32_t p = s.tex[base + y * w + x];
  return make_float4(s.lut[p & 255u], s.lut[(p >> 8) & 255u], s.lut[(p >> 16) & 255u], (float)(p >> 24) / 255.0f);
}

// Repeat addressing of the texel coordinate x on a level of edge w (a power of two): the texel left of x, wrapped
// into [0, w), and the weight of the one right of it. Non-finite or huge coordinates sample at 0.
__device__ __forceinline__ uint32_t surf_wrap(float x, float w, float& f) {
  if (!(fabsf(x) < 1e30f)) x = 0.0f;
  const float xf = floorf(x);
  f = x - xf;
  const float m = xf - w * floorf(xf * (1.0f / w));  // exact: w is a power of two and xf an integer
  return (uint32_t)m;
}

// Bilinear sample of layer `layer` (>= 1: texel layer layer - 1; 0: the white default layer, exactly 1) at uv on
// mip level min(lod, last).
__device__ __forceinline__ float4 surf_sample(const SurfK& s, uint32_t layer, float u, float v, uint32_t lod) {
  if (layer == 0u) return make_float4(1.0f, 1.0f, 1.0f, 1.0f);
  const uint32_t m = min(lod, s.mips - 1u);
  const uint32_t wl = s.size_log2 - m, w = 1u << wl;
  const float fw = (float)w;
  const uint32_t base = (layer - 1u) * s.layer_words + surf_mip_off(s.size_log2, m);
  float fu, fv;
  const uint32_t x0 = surf_wrap(u * fw - 0.5f, fw, fu), y0 = surf_wrap(v * fw - 0.5f, fw, fv);
  const uint32_t x1 = (x0 + 1u) & (w - 1u), y1 = (y0 + 1u) & (w - 1u);
  const float4 t00 = surf_texel(s, base, w, x0, y0), t10 = surf_texel(s, base, w, x1, y0);
  const float4 t01 = surf_texel(s, base, w, x0, y1), t11 = surf_texel(s, base, w, x1, y1);
  const float gu = 1.0f - fu, gv = 1.0f - fv;
  const float4 top = make_float4(t00.x * gu + t10.x * fu, t00.y * gu + t10.y * fu, t00.z * gu + t10.z * fu,
                                 t00.w * gu + t10.w * fu);
  const float4 bot = make_float4(t01.x * gu + t11.x * fu, t01.y * gu + t11.y * fu, t01.z * gu + t11.z * fu,
                                 t01.w * gu + t11.w * fu);
  return make_float4(top.x * gv + bot.x * fv, top.y * gv + bot.y * fv, top.z * gv + bot.z * fv,
                     top.w * gv + bot.w * fv);
}

// UV at barycentrics w of a triangle's vertices i0, i1, i2 (UV(a) w.x + UV(b) w.y + UV(c) w.z, left to right).
__device__ __forceinline__ void surf_uv(const SurfK& s, uint32_t i0, uint32_t i1, uint32_t i2, f3 w, float& u,
                                        float& v) {
  const float2 a = reinterpret_cast<const float2*>(s.uv)[i0], b = reinterpret_cast<const float2*>(s.uv)[i1],
               c = reinterpret_cast<const float2*>(s.uv)[i2];
  u = a.x * w.x + b.x * w.y + c.x * w.z;
  v = a.y * w.x + b.y * w.y + c.y * w.z;
}

// The material's texture times baseColorFactor at uv on voxel level L (tex in surfaceAt / shadeSample).
__device__ __forceinline__ float4 surf_base_color(const SurfK& s, const float* mat, uint32_t mi, float u, float v,
                                                  int L) {
  const float4 f = reinterpret_cast<const float4*>(mat)[mi];
  const float4 t = surf_sample(s, s.mext[mi * 4u], u, v, 3u + (uint32_t)L);
  return make_float4(t.x * f.x, t.y * f.y, t.z * f.z, t.w * f.w);
}
__device__ __forceinline__ uint32_t surf_flags(const SurfK& s, uint32_t mi) { return s.mext[mi * 4u + 1u]; }
__device__ __forceinline__ float surf_cutoff(const SurfK& s, uint32_t mi) { return __uint_as_float(s.mext[mi * 4u + 2u]); }

// generateMips (src/core/scene_gpu.ts:38-54) of one level: texel (x, y) of level m is the bilinear sample of level
// m - 1 at the texel's centre, which falls on the corner of texels 2x..2x+1, 2y..2y+1 (weights 1/2 per axis): the
// mean in linear space, stored as the nearest code: rgb in the sRGB encoding (the number of encode thresholds at or
// below the value), alpha floor(255 a + 0.5).
__device__ __forceinline__ uint32_t surf_encode_srgb(const float* lut, float x) {
  uint32_t lo = 0u, hi = 255u;  // code = number of thresholds <= x (binary search over the 255 thresholds)
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (lut[256u + mid] <= x)
      lo = mid + 1u;
    else
      hi = mid;
  }
  return lo;
}
__device__ __forceinline__ uint32_t surf_mip_texel(const SurfK& s, uint32_t layer0, uint32_t m, uint32_t x,
                                                   uint32_t y) {
  const uint32_t ws = 1u << (s.size_log2 - m + 1u);
  const uint32_t src = layer0 * s.layer_words + surf_mip_off(s.size_log2, m - 1u);
  const float4 t00 = surf_texel(s, src, ws, 2u * x, 2u * y), t10 = surf_texel(s, src, ws, 2u * x + 1u, 2u * y);
  const float4 t01 = surf_texel(s, src, ws, 2u * x, 2u * y + 1u), t11 = surf_texel(s, src, ws, 2u * x + 1u, 2u * y + 1u);
  const float h = 0.5f;
  const float r = (t00.x * h + t10.x * h) * h + (t01.x * h + t11.x * h) * h;
  const float g = (t00.y * h + t10.y * h) * h + (t01.y * h + t11.y * h) * h;
  const float b = (t00.z * h + t10.z * h) * h + (t01.z * h + t11.z * h) * h;
  const float a = (t00.w * h + t10.w * h) * h + (t01.w * h + t11.w * h) * h;
  const uint32_t ca = (uint32_t)floorf(fminf(fmaxf(a, 0.0f), 1.0f) * 255.0f + 0.5f);
  return surf_encode_srgb(s.lut, r) | (surf_encode_srgb(s.lut, g) << 8) | (surf_encode_srgb(s.lut, b) << 16) |
         (ca << 24);
}
