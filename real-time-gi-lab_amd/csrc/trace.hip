// Radiance-cascade probe-interval trace, cascade merge and SH projection on gfx950.
//
//   k_trace   one thread per (cascade, probe, direction) ray over all cascades in a single launch: octahedral texel
//             direction, interval [b_c, b_{c+1}] s0, voxel level from the cone footprint, hierarchical DDA through
//             the clipmap (traceVox, src/core/voxel.ts:121-202), Lambert hit shading (shadeHitV, src/core/rc.ts:108-116)
//             -> rgba f32 (radiance, transmittance) and optionally the first-hit voxel id.
//   k_merge   one launch per cascade, N-1 down to 0: rad_c += T_c * mean over the grid parents of the upper 2x2
//             direction average (mergeWorld/upperAvg, src/core/rc.ts:681-708).
//   k_sh      c0 probes: solid-angle-weighted L2 SH, cosine-convolved (projectSH, src/core/rc.ts:745-774).
//
// The DDA restates the reference's f32 operation sequence exactly (compiled with -ffp-contract=off; IEEE division
// and square root), so first-hit voxels are bit-identical to a CPU restatement of the same WGSL.
#include <hip/hip_runtime.h>

#include "rtgi_internal.h"
#include "vmath.h"

namespace {

constexpr int BLK = 256;
// PI of the WGSL (an abstract-float constant): f32(3.14159265) where it meets f32 values, and its products with
// literals folded in double before rounding to f32 (src/core/rc.ts:145, 765).
constexpr float PI_F = (float)3.14159265;
constexpr float SH_A0 = (float)3.14159265;
constexpr float SH_A1 = (float)(2.0 * 3.14159265 / 3.0);
constexpr float SH_A2 = (float)(3.14159265 / 4.0);

struct VoxHit {
  uint32_t brick, bit;
  int level;
  i3 cell;
};

__device__ __forceinline__ float vox_size(const VoxK& k, int L) { return k.v0 * (float)(1u << L); }
__device__ __forceinline__ i3 origin(const VoxK& k, int L) { return i3{k.origin[L][0], k.origin[L][1], k.origin[L][2]}; }
__device__ __forceinline__ uint32_t ptr_index(const VoxK& k, int L, i3 b) {
  return (uint32_t)L * (uint32_t)k.per + ((uint32_t)b.z * (uint32_t)k.G + (uint32_t)b.y) * (uint32_t)k.G + (uint32_t)b.x;
}
__device__ __forceinline__ uint32_t bit_of(i3 c) { return (uint32_t)((c.x & 7) + 8 * (c.y & 7) + 64 * (c.z & 7)); }

// voxAt (src/core/voxel.ts:90-107): the occupied voxel at cell c of level L, a dropped brick read one level coarser.
__device__ bool vox_at(const VoxK& k, int L, i3 c, VoxHit& h) {
  int l = L;
  i3 cc = c;
  for (;;) {
    uint32_t ptr = k.ptr[ptr_index(k, l, shr3(cc, 3))];
    if (ptr != PTR_DROP) {
      if (ptr >= PTR_MARK) return false;
      uint32_t bit = bit_of(cc);
      if ((k.occ[ptr * 16u + (bit >> 5)] & (1u << (bit & 31u))) == 0u) return false;
      h = VoxHit{ptr, bit, l, cc};
      return true;
    }
    if (l + 1 >= k.NL) return false;
    i3 w = shr3(origin(k, l) * 8 + cc, 1);
    l++;
    cc = w - origin(k, l) * 8;
    if (!inside3(cc, k.G * 8)) return false;
  }
}

// voxRank (src/core/voxel.ts:81-87): attribute slot of a voxel, ATTR_NONE when its brick got no slots.
__device__ __forceinline__ uint32_t vox_rank(const VoxK& k, uint32_t brick, uint32_t bit) {
  uint32_t first = k.battr[brick];
  if (first == ATTR_NONE) return ATTR_NONE;
  const uint32_t* o = k.occ + (size_t)brick * 16u;
  uint32_t w = bit >> 5;
  uint32_t r = __popc(o[w] & ((1u << (bit & 31u)) - 1u));
  for (uint32_t i = 0; i < w; i++) r += __popc(o[i]);
  return first + r;
}

__device__ __forceinline__ float unorm8_to_f(uint32_t b) { return (float)b / 255.0f; }
__device__ __forceinline__ float snorm16_to_f(uint32_t h) {
  return fmaxf((float)(int16_t)(uint16_t)h / 32767.0f, -1.0f);
}

// Hit radiance of an occupied voxel: attributes, oct-decoded (unnormalised) normal, shadeHitV with no bounce
// (src/core/rc.ts:108-116, 136-140; voxel.ts:150-156, 185-192). Point lights are an addition of this path (0 lights:
// the reference's sun-only shading).
__device__ f3 shade_voxel(const TraceK& a, uint32_t ai, f3 d, f3 entry) {
  uint32_t pa = 0u, pn = 0u;
  if (ai != ATTR_NONE) {
    pa = a.vx.attr[ai * 2u];
    pn = a.vx.attr[ai * 2u + 1u];
  }
  f3 alb = mk3(unorm8_to_f(pa & 255u), unorm8_to_f((pa >> 8) & 255u), unorm8_to_f((pa >> 16) & 255u));
  float nx = snorm16_to_f(pn & 0xffffu), ny = snorm16_to_f(pn >> 16);
  f3 n0 = mk3(nx, ny, 1.0f - fabsf(nx) - fabsf(ny));
  float q = fmaxf(-n0.z, 0.0f);
  n0.x += n0.x >= 0.0f ? -q : q;
  n0.y += n0.y >= 0.0f ? -q : q;
  // shadeHitV
  f3 n = dot3(n0, n0) > 0.01f ? n0 : -d;
  if (dot3(n, d) > 0.0f) n = -n;
  f3 b = alb / PI_F * 0.0f;
  f3 sd = mk3(a.sun_dir[0], a.sun_dir[1], a.sun_dir[2]);
  float ndl = fmaxf(dot3(n, sd), 0.0f);
  f3 out = b;
  if (ndl > 0.0f) out = b + alb / PI_F * mk3(a.sun_e[0], a.sun_e[1], a.sun_e[2]) * ndl * a.sun_vis;
  for (int k = 0; k < a.n_point; k++) {
    f3 l = mk3(a.point_pos[k][0], a.point_pos[k][1], a.point_pos[k][2]) - entry;
    float r2 = dot3(l, l);
    float nl = fmaxf(dot3(n, l) / sqrtf(r2), 0.0f);
    if (nl > 0.0f) out = out + alb / PI_F * mk3(a.point_i[k][0], a.point_i[k][1], a.point_i[k][2]) * (nl / r2);
  }
  return out;
}

__device__ __forceinline__ uint32_t hit_id(const VoxK& k, int L, i3 c) {
  return ((uint32_t)L << 29) | (uint32_t)(c.x + k.E * (c.y + k.E * c.z));
}

// traceVox (src/core/voxel.ts:121-202): hierarchical DDA over level L; empty bricks are skipped whole.
__device__ float4 trace_vox(const TraceK& a, f3 o, f3 d, float t0, float t1, float tSkip, int L, uint32_t& hid) {
  const VoxK& k = a.vx;
  hid = RTGI_NO_HIT;
  float v = vox_size(k, L);
  int G = k.G;
  f3 lo = tof3(origin(k, L)) * v * 8.0f;
  f3 hi = lo + splat3((float)G * v * 8.0f);
  f3 dd = mk3(fabsf(d.x) < 1e-9f ? 1e-9f : d.x, fabsf(d.y) < 1e-9f ? 1e-9f : d.y, fabsf(d.z) < 1e-9f ? 1e-9f : d.z);
  f3 inv = mk3(1.0f / dd.x, 1.0f / dd.y, 1.0f / dd.z);
  f3 ta = (lo - o) * inv, tb = (hi - o) * inv;
  f3 tn = min3(ta, tb), tf = max3(ta, tb);
  float t = fmaxf(fmaxf(fmaxf(tn.x, tn.y), tn.z), t0);
  float tEnd = fminf(fminf(fminf(tf.x, tf.y), tf.z), t1);
  if (t >= tEnd) return make_float4(0.0f, 0.0f, 0.0f, 1.0f);
  i3 st{dd.x > 0.0f ? 1 : -1, dd.y > 0.0f ? 1 : -1, dd.z > 0.0f ? 1 : -1};
  f3 sp = mk3(st.x > 0 ? 1.0f : 0.0f, st.y > 0 ? 1.0f : 0.0f, st.z > 0 ? 1.0f : 0.0f);
  int nVox = G * 8;
  i3 zero{0, 0, 0}, top{nVox - 1, nVox - 1, nVox - 1};
  i3 c = clampi3(f2i3(floor3((o + d * t - lo) / v)), zero, top);
  f3 tMax = ((tof3(c) + sp) * v + lo - o) * inv;
  f3 tDelta = abs3(inv * v);
  float tIn = t;
  for (int it = 0; it < 1024; it++) {
    i3 b = shr3(c, 3);
    uint32_t ptr = k.ptr[ptr_index(k, L, b)];
    if (ptr == PTR_DROP) {
      VoxHit h;
      if (vox_at(k, L, c, h) && tIn >= tSkip) {
        hid = hit_id(k, h.level, h.cell);
        f3 rgb = shade_voxel(a, vox_rank(k, h.brick, h.bit), d, o + d * tIn);
        return make_float4(rgb.x, rgb.y, rgb.z, 0.0f);
      }
      float tx = fminf(fminf(tMax.x, tMax.y), tMax.z);
      if (tx >= tEnd) break;
      tIn = tx;
      if (tMax.x <= tMax.y && tMax.x <= tMax.z) {
        c.x += st.x;
        tMax.x += tDelta.x;
      } else if (tMax.y <= tMax.z) {
        c.y += st.y;
        tMax.y += tDelta.y;
      } else {
        c.z += st.z;
        tMax.z += tDelta.z;
      }
      if (!inside3(c, nVox)) break;
      continue;
    }
    if (ptr >= PTR_MARK) {
      f3 bt = ((tof3(b * 8) + sp * 8.0f) * v + lo - o) * inv;
      float te = fminf(fminf(bt.x, bt.y), bt.z);
      if (te >= tEnd) break;
      i3 nb = b;
      if (bt.x <= bt.y && bt.x <= bt.z)
        nb.x += st.x;
      else if (bt.y <= bt.z)
        nb.y += st.y;
      else
        nb.z += st.z;
      if (!inside3(nb, G)) break;
      t = te;
      tIn = te;
      c = clampi3(f2i3(floor3((o + d * t - lo) / v)), nb * 8, nb * 8 + 7);
      tMax = ((tof3(c) + sp) * v + lo - o) * inv;
      continue;
    }
    uint32_t bit = bit_of(c);
    uint32_t occ = k.occ[ptr * 16u + (bit >> 5)];
    float tExit = fminf(fminf(tMax.x, tMax.y), tMax.z);
    if ((occ & (1u << (bit & 31u))) != 0u && tIn >= tSkip) {
      hid = hit_id(k, L, c);
      f3 rgb = shade_voxel(a, vox_rank(k, ptr, bit), d, o + d * tIn);
      return make_float4(rgb.x, rgb.y, rgb.z, 0.0f);
    }
    if (tExit >= tEnd) break;
    tIn = tExit;
    if (tMax.x <= tMax.y && tMax.x <= tMax.z) {
      c.x += st.x;
      tMax.x += tDelta.x;
    } else if (tMax.y <= tMax.z) {
      c.y += st.y;
      tMax.y += tDelta.y;
    } else {
      c.z += st.z;
      tMax.z += tDelta.z;
    }
    if (!inside3(c, nVox)) break;
  }
  return make_float4(0.0f, 0.0f, 0.0f, 1.0f);
}

// octDecode of a texel centre (dirOf, src/core/rc.ts:291-298, 308).
__device__ __forceinline__ f3 dir_of(int di, int res) {
  float u = ((float)(di % res) + 0.5f) / (float)res, w = ((float)(di / res) + 0.5f) / (float)res;
  float fx = u * 2.0f - 1.0f, fy = w * 2.0f - 1.0f;
  f3 n = mk3(fx, fy, 1.0f - fabsf(fx) - fabsf(fy));
  float t = fmaxf(-n.z, 0.0f);
  n.x += n.x >= 0.0f ? -t : t;
  n.y += n.y >= 0.0f ? -t : t;
  return normalize3(n);
}

// voxInside / voxLevelFor (src/core/voxel.ts:108-119).
__device__ __forceinline__ int vox_level_for(const VoxK& k, int L, f3 e) {
  int l = L;
  while (l + 1 < k.NL) {
    float bs = vox_size(k, l) * 8.0f;
    f3 lo = tof3(origin(k, l)) * bs;
    f3 hi = lo + splat3((float)k.G * bs);
    bool in = e.x > lo.x && e.y > lo.y && e.z > lo.z && e.x < hi.x && e.y < hi.y && e.z < hi.z;
    if (in) break;
    l++;
  }
  return l;
}

__global__ void __launch_bounds__(BLK) k_trace(TraceK a) {
  uint64_t ray = (uint64_t)blockIdx.x * BLK + threadIdx.x;
  if (ray >= a.total_rays) return;
  int c = 0;
  while (c + 1 < a.N && ray >= a.c[c + 1].ray_off) c++;
  const CascK& C = a.c[c];
  uint32_t r = (uint32_t)(ray - C.ray_off);
  uint32_t dirs = (uint32_t)(C.res * C.res);
  uint32_t pi = r / dirs, di = r % dirs;
  uint32_t n = (uint32_t)C.n;
  i3 g{a.G0[0] >> c, a.G0[1] >> c, a.G0[2] >> c};
  g = g + i3{(int)(pi % n), (int)((pi / n) % n), (int)(pi / (n * n))};
  float s = a.s0 * (float)(1u << c);
  f3 o = tof3(g) * s;
  f3 d = dir_of((int)di, C.res);
  float t0 = C.b0 * a.s0, t1 = C.b1 * a.s0;
  int L = vox_level_for(a.vx, C.vox_level, o + d * t1);
  float v = vox_size(a.vx, L);
  uint32_t hid;
  float4 out = trace_vox(a, o, d, t0, t1, a.vox_skip * v, L, hid);
  reinterpret_cast<float4*>(a.rad)[ray] = out;
  if (a.hits) a.hits[ray] = hid;
}

// upperAvg (src/core/rc.ts:681-686) over the f32 cascade.
__device__ __forceinline__ f3 upper_avg(const float4* rad, const CascK& U, uint32_t q, uint32_t u, uint32_t w,
                                        uint32_t lr) {
  uint32_t ur = (uint32_t)U.res;
  uint64_t base0 = U.ray_off + (uint64_t)q * ur * ur;
  if (ur == lr) {
    float4 x = rad[base0 + w * ur + u];
    return mk3(x.x, x.y, x.z);
  }
  uint64_t base = base0 + w * 2u * ur + u * 2u;
  float4 p0 = rad[base], p1 = rad[base + 1], p2 = rad[base + ur], p3 = rad[base + ur + 1];
  return (mk3(p0.x, p0.y, p0.z) + mk3(p1.x, p1.y, p1.z) + mk3(p2.x, p2.y, p2.z) + mk3(p3.x, p3.y, p3.z)) * 0.25f;
}

// mergeWorld (src/core/rc.ts:688-708) on the dense grid: the parents of probe g are the c+1 grid points
// (g >> 1) + o, o in {0,1}^3 with o = 0 on even axes (linkParents, rc.ts:540-570), those outside the grid absent.
__global__ void __launch_bounds__(BLK) k_merge(TraceK a, int c) {
  const CascK& C = a.c[c];
  uint64_t r = (uint64_t)blockIdx.x * BLK + threadIdx.x;
  if (r >= C.rays) return;
  float4* rad = reinterpret_cast<float4*>(a.rad);
  float4 v = rad[C.ray_off + r];
  f3 up = mk3(a.sky[0], a.sky[1], a.sky[2]);
  if (c + 1 < a.N && v.w > 0.0f) {
    const CascK& U = a.c[c + 1];
    uint32_t dirs = (uint32_t)(C.res * C.res);
    uint32_t pi = (uint32_t)(r / dirs), di = (uint32_t)(r % dirs);
    uint32_t n = (uint32_t)C.n, un = (uint32_t)U.n;
    i3 g{(int)(pi % n), (int)((pi / n) % n), (int)(pi / (n * n))};
    uint32_t u0 = di % (uint32_t)C.res, w0 = di / (uint32_t)C.res;
    f3 acc = splat3(0.0f);
    float ws = 0.0f;
    for (int kk = 0; kk < 8; kk++) {
      i3 o{kk & 1, (kk >> 1) & 1, (kk >> 2) & 1};
      if ((o.x && !(g.x & 1)) || (o.y && !(g.y & 1)) || (o.z && !(g.z & 1))) continue;
      i3 q = shr3(g, 1) + o;
      if (q.x >= (int)un || q.y >= (int)un || q.z >= (int)un) continue;
      uint32_t qi = (uint32_t)q.x + un * ((uint32_t)q.y + un * (uint32_t)q.z);
      acc = acc + upper_avg(rad, U, qi, u0, w0, (uint32_t)C.res);
      ws += 1.0f;
    }
    up = ws > 0.0f ? acc / fmaxf(ws, 1.0f) : splat3(0.0f);
  }
  rad[C.ray_off + r] = make_float4(v.x + v.w * up.x, v.y + v.w * up.y, v.z + v.w * up.z, 1.0f);
}

// projectSH (src/core/rc.ts:745-774): 27 f32 per c0 probe, stored as 27 planes of probes0^3 floats.
__global__ void __launch_bounds__(BLK) k_sh(TraceK a) {
  const CascK& C = a.c[0];
  uint32_t np = (uint32_t)C.n * (uint32_t)C.n * (uint32_t)C.n;
  uint32_t pi = blockIdx.x * BLK + threadIdx.x;
  if (pi >= np) return;
  const float4* rad = reinterpret_cast<const float4*>(a.rad);
  f3 cf[9];
  for (int q = 0; q < 9; q++) cf[q] = splat3(0.0f);
  for (int d = 0; d < 16; d++) {
    float4 x = rad[(uint64_t)pi * 16u + d];
    f3 r = mk3(x.x, x.y, x.z) * a.omega0[d];
    f3 n = dir_of(d, 4);
    float y[9] = {0.282095f,
                  0.488603f * n.y,
                  0.488603f * n.z,
                  0.488603f * n.x,
                  1.092548f * n.x * n.y,
                  1.092548f * n.y * n.z,
                  0.315392f * (3.0f * n.z * n.z - 1.0f),
                  1.092548f * n.x * n.z,
                  0.546274f * (n.x * n.x - n.y * n.y)};
    for (int q = 0; q < 9; q++) cf[q] = cf[q] + r * y[q];
  }
  for (int q = 0; q < 9; q++) {
    float s = q == 0 ? SH_A0 : (q < 4 ? SH_A1 : SH_A2);
    a.sh[(size_t)(q * 3 + 0) * np + pi] = cf[q].x * s;
    a.sh[(size_t)(q * 3 + 1) * np + pi] = cf[q].y * s;
    a.sh[(size_t)(q * 3 + 2) * np + pi] = cf[q].z * s;
  }
}

int grid_for(uint64_t n) {
  uint64_t g = (n + BLK - 1) / BLK;
  return (int)(g < 1 ? 1 : g);
}

}  // namespace

int launch_trace(const TraceK& k, void* stream) {
  if (k.total_rays) k_trace<<<grid_for(k.total_rays), BLK, 0, (hipStream_t)stream>>>(k);
  return (int)hipGetLastError();
}

int launch_merge(const TraceK& k, int c, void* stream) {
  if (k.c[c].rays) k_merge<<<grid_for(k.c[c].rays), BLK, 0, (hipStream_t)stream>>>(k, c);
  return (int)hipGetLastError();
}

int launch_sh(const TraceK& k, void* stream) {
  uint64_t np = (uint64_t)k.c[0].n * k.c[0].n * k.c[0].n;
  k_sh<<<grid_for(np), BLK, 0, (hipStream_t)stream>>>(k);
  return (int)hipGetLastError();
}
