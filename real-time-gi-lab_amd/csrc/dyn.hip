// Dynamic coverage layer build on gfx950 (src/core/dynlayer.ts, BUILD_WGSL): every update, the dynamic triangles
// [first_tri, end_tri) are splatted into a separate layer of coverage and mean radiance, prefiltered from fine to
// coarse levels; the cascade trace reads it in front of the static hit (trace_dyn in trace.hip).
//
//   k_dl_mark     per dynamic triangle: splat level (the finest level whose window holds its box grown by two
//                 voxels), the bricks its box grown by one voxel touches there (bits of this update's bitset, via
//                 an LDS copy when it fits), the bounds of everything a read can see, reduced per wavefront
//                                                                                       (mark, dynlayer.ts:255-279)
//   k_dl_parents  per bitset word below the coarsest level: each wanted brick wants its ancestors (parents, :281-303)
//   k_dl_alloc    one workgroup: slots in bit order (coarse levels first, then brick index), dropped past the pool;
//                 bricks wanted last update and not now emptied; per-level ranges, bounds published, parity flipped
//                                                                       (count, alloc, finalize, :318-384)
//   k_dl_clear    the allocated bricks' words and voxel masks                                  (clear, :385-390)
//   k_dl_splat    per dynamic triangle: m^2 sub-triangle samples shaded like a voxel hit, added with trilinear
//                 weights over the 8 nearest voxel centres as fixed-point sums (exact: order-independent); 16
//                 triangles per workgroup, 4 lanes each, through an LDS window of voxels
//                                                                       (splat, splatTri, splatSample, :392-540)
//   k_dl_resolve  level L = 1..: own splats + 1/4 of each child with a value, as floats    (resolve, :542-580)
//
// Sample shading uses the frame's sun visibility (the reference taps its raster shadow map) and the base colour of
// surfaceAt's sampler (texture x baseColorFactor at the sample's UV, mip 3 + L; white when untextured), skipping
// samples of alpha-masked materials below their cutoff.
#include <hip/hip_runtime.h>

#include "batch.h"
#include "rtgi_internal.h"
#include "surface.h"
#include "vmath.h"

namespace {

constexpr int BLK = 256;
constexpr float PI_F = 3.14159265f;

struct Tri {
  f3 a, b, c;
  uint32_t i0, i1, i2, m;
};
__device__ __forceinline__ f3 ldp(const float* p, uint32_t i) { return mk3(p[i * 3u], p[i * 3u + 1u], p[i * 3u + 2u]); }
__device__ __forceinline__ Tri load_tri(const VoxK& k, uint32_t t) {
  const uint4 q = reinterpret_cast<const uint4*>(k.tri)[t];
  return Tri{ldp(k.pos, q.x), ldp(k.pos, q.y), ldp(k.pos, q.z), q.x, q.y, q.z, q.w};
}
__device__ __forceinline__ float vox_size(const VoxK& k, int L) { return k.v0 * (float)(1u << L); }
__device__ __forceinline__ f3 region_lo(const VoxK& k, int L) {
  return tof3(i3{k.origin[L][0], k.origin[L][1], k.origin[L][2]}) * vox_size(k, L) * 8.0f;
}
__device__ __forceinline__ f3 min3v(f3 a, f3 b) { return mk3(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)); }
__device__ __forceinline__ f3 max3v(f3 a, f3 b) { return mk3(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)); }
__device__ __forceinline__ uint32_t* cur_bits(const DynK& D) { return D.bits + (D.hdr[10] & 1u) * D.nw; }
__device__ __forceinline__ uint32_t* prev_bits(const DynK& D) { return D.bits + (1u - (D.hdr[10] & 1u)) * D.nw; }

// splatLevel (dynlayer.ts:240-247): the finest level whose window holds the box grown by two voxels, or -1.
__device__ __forceinline__ int splat_level(const VoxK& k, f3 mn, f3 mx) {
  for (int L = 0; L < k.NL; L++) {
    float v = vox_size(k, L);
    f3 lo = region_lo(k, L);
    f3 hi = lo + splat3((float)k.G * v * 8.0f);
    f3 a = mn - splat3(2.0f * v), b = mx + splat3(2.0f * v);
    if (a.x > lo.x && a.y > lo.y && a.z > lo.z && b.x < hi.x && b.y < hi.y && b.z < hi.z) return L;
  }
  return -1;
}
// want (dynlayer.ts:248-254): brick b of level L is wanted this update.
__device__ __forceinline__ void want(const VoxK& k, uint32_t* cur, int L, i3 b) {
  const int G = k.G;
  if (b.x < 0 || b.y < 0 || b.z < 0 || b.x >= G || b.y >= G || b.z >= G) return;
  uint32_t g = (uint32_t)(k.NL - 1 - L) * (uint32_t)k.per + ((uint32_t)b.z * G + (uint32_t)b.y) * G + (uint32_t)b.x;
  atomicOr(&cur[g >> 5], 1u << (g & 31u));
}

__device__ __forceinline__ uint32_t wave_min(uint32_t x) {
  for (int s = 32; s > 0; s >>= 1) x = min(x, (uint32_t)__shfl_xor((int)x, s));
  return x;
}
__device__ __forceinline__ uint32_t wave_max(uint32_t x) {
  for (int s = 32; s > 0; s >>= 1) x = max(x, (uint32_t)__shfl_xor((int)x, s));
  return x;
}

// The block's rounds are uniform (i0 < n), so the wavefront reductions of the bounds see every lane; one lane per
// wavefront then takes the min / max into the header (atomic min and max commute, so the result is the per-triangle
// atomics' result). The wanted bits gather in a bitset in LDS (when the world's bitset fits MARK_LDS words) and
// are ORed into the layer's bitset once per block: the triangles of a block want the same few bricks.
constexpr uint32_t MARK_LDS = 4096;
__device__ __forceinline__ void mark_body(const VoxK& k, uint32_t* lbits, uint32_t lb, uint32_t nblk) {
  const DynK& D = k.dyn;
  const bool in_lds = D.nw <= MARK_LDS;
  uint32_t* cur = in_lds ? lbits : cur_bits(D);
  const uint32_t n = D.end_tri - D.first_tri;
  if (in_lds) {
    for (uint32_t w = threadIdx.x; w < D.nw; w += BLK) lbits[w] = 0u;
    __syncthreads();
  }
  for (uint32_t i0 = lb * BLK; i0 < n; i0 += nblk * BLK) {
    const uint32_t i = i0 + threadIdx.x;
    uint32_t lo3[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, hi3[3] = {0u, 0u, 0u};
    uint32_t outside = 0u;
    if (i < n) {
      Tri T = load_tri(k, D.first_tri + i);
      f3 mn = min3v(min3v(T.a, T.b), T.c), mx = max3v(max3v(T.a, T.b), T.c);
      int Ls = splat_level(k, mn, mx);
      if (Ls < 0) {
        outside = 1u;
      } else {
        float v = vox_size(k, Ls), bs = v * 8.0f;
        f3 lo = region_lo(k, Ls);
        f3 q0 = (mn - splat3(v) - lo) / bs, q1 = (mx + splat3(v) - lo) / bs;
        i3 g0 = f2i3(floor3(q0)), g1 = f2i3(floor3(q1));
        for (int z = g0.z; z <= g1.z; z++)
          for (int y = g0.y; y <= g1.y; y++)
            for (int x = g0.x; x <= g1.x; x++) want(k, cur, Ls, i3{x, y, z});
        // Bounds of everything a read can see: the box grown by the coarsest voxel (cm, offset by 2^30).
        float gv = vox_size(k, k.NL - 1);
        f3 blo = floor3((mn - splat3(gv)) * 100.0f), bhi = mx + splat3(gv);
        bhi = mk3(ceilf(bhi.x * 100.0f), ceilf(bhi.y * 100.0f), ceilf(bhi.z * 100.0f));
        lo3[0] = (uint32_t)(f2i(blo.x) + 1073741824);
        lo3[1] = (uint32_t)(f2i(blo.y) + 1073741824);
        lo3[2] = (uint32_t)(f2i(blo.z) + 1073741824);
        hi3[0] = (uint32_t)(f2i(bhi.x) + 1073741824);
        hi3[1] = (uint32_t)(f2i(bhi.y) + 1073741824);
        hi3[2] = (uint32_t)(f2i(bhi.z) + 1073741824);
      }
    }
    const uint64_t out_mask = __ballot(outside != 0u);
    for (int a = 0; a < 3; a++) {
      lo3[a] = wave_min(lo3[a]);
      hi3[a] = wave_max(hi3[a]);
    }
    if ((threadIdx.x & 63u) == 0u) {
      if (out_mask) atomicAdd(&D.hdr[38], (uint32_t)__popcll(out_mask));
      if (lo3[0] != 0xffffffffu) {
        for (int a = 0; a < 3; a++) {
          atomicMin(&D.hdr[32 + a], lo3[a]);
          atomicMax(&D.hdr[35 + a], hi3[a]);
        }
      }
    }
  }
  if (in_lds) {
    __syncthreads();
    uint32_t* g = cur_bits(D);
    for (uint32_t w = threadIdx.x; w < D.nw; w += BLK)
      if (lbits[w]) atomicOr(&g[w], lbits[w]);
  }
}

__device__ __forceinline__ void parents_body(const VoxK& k, uint32_t lb, uint32_t nblk) {
  const DynK& D = k.dyn;
  uint32_t* cur = cur_bits(D);
  const uint32_t G = (uint32_t)k.G, G3 = (uint32_t)k.per, n = (uint32_t)k.NL;
  for (uint32_t w = lb * BLK + threadIdx.x + G3 / 32u; w < D.nw; w += nblk * BLK) {  // the coarsest level is first
    uint32_t m = cur[w];
    while (m) {
      uint32_t bit = __builtin_ctz(m);
      m &= m - 1u;
      uint32_t g = w * 32u + bit, lr = g / G3;
      int L = (int)(n - 1u - lr);
      uint32_t r = g - lr * G3;
      i3 b{(int)(r % G), (int)((r / G) % G), (int)(r / (G * G))};
      while (L + 1 < k.NL) {
        i3 o0{k.origin[L][0], k.origin[L][1], k.origin[L][2]}, o1{k.origin[L + 1][0], k.origin[L + 1][1], k.origin[L + 1][2]};
        b = shr3((o0 + b) * 4 - o1 * 8, 3);  // ((o0 + b) * 8 / 2 - o1 * 8) >> 3: the 8 voxels map to 4 of L + 1
        L++;
        want(k, cur, L, b);
      }
    }
  }
}

// Exclusive prefix sum of x over the 1024 threads of the block; *total gets the sum.
__device__ __forceinline__ uint32_t block_scan(uint32_t x, uint32_t* wsum, uint32_t* total) {
  const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
  uint32_t incl = x;
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(incl, o, 64);
    if (lane >= (uint32_t)o) incl += y;
  }
  if (lane == 63u) wsum[wv] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t s = 0;
    for (uint32_t i = 0; i < blockDim.x / 64u; i++) {
      uint32_t t = wsum[i];
      wsum[i] = s;
      s += t;
    }
    wsum[16] = s;
  }
  __syncthreads();
  uint32_t r = wsum[wv] + incl - x;
  *total = wsum[16];
  __syncthreads();
  return r;
}

__device__ __forceinline__ void alloc_body(const VoxK& k) {
  __shared__ uint32_t wsum[17];
  const DynK& D = k.dyn;
  const uint32_t G3 = (uint32_t)k.per, n = (uint32_t)k.NL, nw = D.nw, cap = (uint32_t)D.cap;
  uint32_t* cur = cur_bits(D);
  uint32_t* old = prev_bits(D);
  uint32_t first = 0;
  for (uint32_t w0 = 0; w0 < nw; w0 += blockDim.x) {
    const uint32_t w = w0 + threadIdx.x;
    uint32_t cw = 0, ow = 0;
    if (w < nw) {
      cw = cur[w];
      ow = old[w];
    }
    const uint32_t c = __popc(cw);
    uint32_t round;
    const uint32_t excl = first + block_scan(c, wsum, &round);
    if (w < nw) {
      if ((w * 32u) % G3 == 0u) D.hdr[16u + n - 1u - w * 32u / G3] = excl;
      if (w == nw - 1u) D.hdr[0] = excl + c;
      uint32_t m = cw | ow;
      while (m) {
        uint32_t bit = __builtin_ctz(m);
        m &= m - 1u;
        uint32_t g = w * 32u + bit, lr = g / G3;
        uint32_t pi = (n - 1u - lr) * G3 + g - lr * G3;
        if ((cw >> bit) & 1u) {
          uint32_t rank = excl + __popc(cw & ((1u << bit) - 1u));
          if (rank < cap) {
            D.ptr[pi] = rank;
            D.meta[rank] = pi;
          } else {
            D.ptr[pi] = DYN_DROPPED;
          }
        } else {
          D.ptr[pi] = DYN_EMPTY;
        }
      }
      old[w] = 0u;
    }
    first += round;
  }
  __syncthreads();
  if (threadIdx.x == 0) {  // finalize (dynlayer.ts:362-384)
    uint32_t* h = D.hdr;
    const uint32_t total = h[0], al = min(total, cap);
    h[8] = al;
    h[1] = total - al;
    for (uint32_t L = 0; L < n; L++) {
      uint32_t e = L > 0u ? h[16u + L - 1u] : total;
      uint32_t s0 = min(h[16u + L], cap);
      h[16u + L] = s0;
      h[24u + L] = min(e, cap) - s0;
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
}

__device__ __forceinline__ void clear_body(const VoxK& k, uint32_t lb, uint32_t nblk) {
  const DynK& D = k.dyn;
  const uint32_t al = D.hdr[8];
  // The allocated slots are [0, al): one contiguous range of 512 x 13 words (whole 16-byte units) per slot.
  uint4* d4 = reinterpret_cast<uint4*>(D.data);
  const uint32_t nq = al * (512u * DYN_WORDS / 4u);
  for (uint32_t i = lb * BLK + threadIdx.x; i < nq; i += nblk * BLK) d4[i] = make_uint4(0u, 0u, 0u, 0u);
  for (uint32_t i = lb * BLK + threadIdx.x; i < al * 16u; i += nblk * BLK) D.mask[i] = 0u;
}

// u32(round(x * FIX)) of WGSL for x >= 0 (round half to even; saturating as WGSL's conversion).
__device__ __forceinline__ uint32_t fix(float x) { return (uint32_t)fminf(rintf(x * DYN_FIX), 4294967040.0f); }

// The splat window (dynlayer.ts:398-540): a workgroup of SPLAT_TPW consecutive dynamic triangles (one wavefront,
// SPLAT_LPT lanes per triangle)
// accumulates its samples' fixed-point adds in workgroup memory over the voxel box of its triangles at their finest
// splat level, then adds each touched voxel's words to the layer once. Neighbouring triangles hit the same voxels,
// so this replaces contended device atomics with LDS atomics. Corners outside the window (other levels, or a box
// larger than SPLAT_WIN voxels) add straight to the layer. Fixed-point sums are exact, so the words are the same.
constexpr uint32_t SPLAT_WG = 64;
constexpr uint32_t SPLAT_LPT = 4;                     // lanes per triangle, over its samples
constexpr uint32_t SPLAT_TPW = SPLAT_WG / SPLAT_LPT;  // triangles per workgroup
constexpr uint32_t SPLAT_WIN = 512;  // voxels: 512 x 13 words = 26 KiB of LDS
constexpr uint32_t SPLAT_W = 13;     // words splatted: 10, or 13 with the albedo words (token obj)

struct SplatWin {
  uint32_t acc[SPLAT_WIN * SPLAT_W];
  uint32_t touched[SPLAT_WIN / 32];
};

__device__ __forceinline__ int wave_min_i(int x) {
  for (int s = 32; s > 0; s >>= 1) x = min(x, __shfl_xor(x, s));
  return x;
}
__device__ __forceinline__ int wave_max_i(int x) {
  for (int s = 32; s > 0; s >>= 1) x = max(x, __shfl_xor(x, s));
  return x;
}

// splatSample (dynlayer.ts:404-438): trilinear weights over the 8 voxel centres nearest to p (level L), each adding
// its projected areas (front and back per axis), area and area x radiance in units of v^2; into the window when the
// corner lies in it (level wl, voxels wo .. wo + wd - 1), else into the layer.
__device__ void splat_sample(const VoxK& k, SplatWin& W, int wl, i3 wo, i3 wd, int L, f3 p, f3 n, float area,
                             f3 rad, f3 alb, uint32_t nw) {
  const DynK& D = k.dyn;
  float v = vox_size(k, L);
  f3 g = (p - region_lo(k, L)) / v - splat3(0.5f);
  i3 c0 = f2i3(floor3(g));
  f3 f = g - tof3(c0);
  float s = area / (v * v);
  f3 pn = mk3(fmaxf(n.x, 0.0f), fmaxf(n.y, 0.0f), fmaxf(n.z, 0.0f));
  f3 nn = mk3(fmaxf(-n.x, 0.0f), fmaxf(-n.y, 0.0f), fmaxf(-n.z, 0.0f));
  const int E = k.G * 8;
  for (int q = 0; q < 8; q++) {
    i3 o{q & 1, (q >> 1) & 1, q >> 2};
    i3 c = c0 + o;
    if (c.x < 0 || c.y < 0 || c.z < 0 || c.x >= E || c.y >= E || c.z >= E) continue;
    f3 tw = mk3(o.x ? f.x : 1.0f - f.x, o.y ? f.y : 1.0f - f.y, o.z ? f.z : 1.0f - f.z);
    float w = tw.x * tw.y * tw.z * s;
    if (w <= 0.0f) continue;
    const uint32_t x[SPLAT_W] = {fix(w * pn.x),  fix(w * nn.x),  fix(w * pn.y),  fix(w * nn.y), fix(w * pn.z),
                                 fix(w * nn.z),  fix(w),         fix(w * rad.x), fix(w * rad.y), fix(w * rad.z),
                                 fix(w * alb.x), fix(w * alb.y), fix(w * alb.z)};
    i3 r = c - wo;
    if (L == wl && r.x >= 0 && r.y >= 0 && r.z >= 0 && r.x < wd.x && r.y < wd.y && r.z < wd.z) {
      uint32_t j = (uint32_t)(r.x + wd.x * (r.y + wd.y * r.z));
      atomicOr(&W.touched[j >> 5], 1u << (j & 31u));
      for (uint32_t q2 = 0; q2 < nw; q2++)
        if (x[q2]) atomicAdd(&W.acc[j * SPLAT_W + q2], x[q2]);
      continue;
    }
    i3 b = shr3(c, 3);
    uint32_t ptr = D.ptr[(uint32_t)L * (uint32_t)k.per + ((uint32_t)b.z * k.G + (uint32_t)b.y) * k.G + (uint32_t)b.x];
    if (ptr >= (uint32_t)D.cap) continue;
    uint32_t vi = (uint32_t)((c.x & 7) + 8 * (c.y & 7) + 64 * (c.z & 7));
    atomicOr(&D.mask[ptr * 16u + (vi >> 5)], 1u << (vi & 31u));
    uint32_t* d = D.data + ((size_t)ptr * 512u + vi) * DYN_WORDS;
    for (uint32_t q2 = 0; q2 < nw; q2++)
      if (x[q2]) atomicAdd(&d[q2], x[q2]);
  }
}

// splatTri (dynlayer.ts:461-492): m^2 equal sub-triangles (m = ceil(longest edge / voxel) in 1..16), each a sample at
// its centroid with its area; shadeSample (:440-452) with the frame's sun visibility. The rounds
// of the loop are uniform over the wavefront (i0 < n), so the window reductions see every lane.
__device__ __forceinline__ void splat_body(const VoxK& k, const TraceK& t, SplatWin& W, uint32_t lb, uint32_t nblk) {
  const DynK& D = k.dyn;
  const uint32_t n = D.end_tri - D.first_tri, lane = threadIdx.x, sub = lane % SPLAT_LPT;
  const uint32_t nw = t.objects ? 13u : 10u;  // the albedo words with token obj (DYN_ALB), else they stay 0
  const f3 sd = mk3(t.sun_dir[0], t.sun_dir[1], t.sun_dir[2]), se = mk3(t.sun_e[0], t.sun_e[1], t.sun_e[2]);
  for (uint32_t i0 = lb * SPLAT_TPW; i0 < n; i0 += nblk * SPLAT_TPW) {
    const uint32_t i = i0 + lane / SPLAT_LPT;
    Tri T{};
    int L = -1;
    f3 mn = splat3(0.0f), mx = splat3(0.0f);
    if (i < n) {
      T = load_tri(k, D.first_tri + i);
      mn = min3v(min3v(T.a, T.b), T.c);
      mx = max3v(max3v(T.a, T.b), T.c);
      L = splat_level(k, mn, mx);
    }
    // The window: the box of the triangles at the finest splat level among them, the voxels their samples'
    // trilinear corners can reach (floor(x / v - 0.5) .. + 1 over each triangle's box).
    const int wl = wave_min_i(L < 0 ? 0x7fffffff : L);
    i3 c0{0x7fffffff, 0x7fffffff, 0x7fffffff}, c1{-0x7fffffff, -0x7fffffff, -0x7fffffff};
    if (L >= 0 && L == wl) {
      float v = vox_size(k, wl);
      f3 lo = region_lo(k, wl);
      c0 = f2i3(floor3((mn - lo) / v - splat3(0.5f)));
      c1 = f2i3(floor3((mx - lo) / v - splat3(0.5f))) + i3{1, 1, 1};
    }
    const i3 wo{wave_min_i(c0.x), wave_min_i(c0.y), wave_min_i(c0.z)};
    const i3 whi{wave_max_i(c1.x), wave_max_i(c1.y), wave_max_i(c1.z)};
    i3 wd{0, 0, 0};
    if (wl != 0x7fffffff) {
      const i3 e = whi - wo + i3{1, 1, 1};
      if ((uint32_t)e.x * (uint32_t)e.y <= SPLAT_WIN)
        wd = i3{e.x, e.y, min(e.z, (int)(SPLAT_WIN / ((uint32_t)e.x * (uint32_t)e.y)))};
    }
    const uint32_t vol = (uint32_t)(wd.x * wd.y * wd.z);
    for (uint32_t j = lane; j < vol * SPLAT_W; j += SPLAT_WG) W.acc[j] = 0u;  // all 13 words: the flush reads nw
    for (uint32_t j = lane; j < SPLAT_WIN / 32; j += SPLAT_WG) W.touched[j] = 0u;
    __syncthreads();
    if (L >= 0) {
      f3 cr = cross3(T.b - T.a, T.c - T.a);
      float area = 0.5f * sqrtf(dot3(cr, cr));
      if (area > 0.0f) {
        f3 e0 = T.b - T.a, e1 = T.c - T.b, e2 = T.a - T.c;
        float e = fmaxf(fmaxf(sqrtf(dot3(e0, e0)), sqrtf(dot3(e1, e1))), sqrtf(dot3(e2, e2)));
        uint32_t m = (uint32_t)fminf(fmaxf(ceilf(e / vox_size(k, L)), 1.0f), 16.0f);
        float fm = (float)m;
        float sa = area / (fm * fm);
        const float4 mat = reinterpret_cast<const float4*>(k.mat)[T.m];
        const bool masked = k.surf.tex && (surf_flags(k.surf, T.m) & RTGI_SURF_ALPHA_MASK);
        const f3 n0 = ldp(k.nrm, T.i0), n1 = ldp(k.nrm, T.i1), n2 = ldp(k.nrm, T.i2);
        for (uint32_t s = sub; s < m * m; s += SPLAT_LPT) {
          // Sample s in the order u, then v < m - u, then flip (flip 1 only while u + v + 1 < m): row u holds
          // 2 (m - u) - 1 samples, alternating flip 0 and 1 from v = 0.
          uint32_t u = 0u, r = s;
          while (r >= 2u * (m - u) - 1u) {
            r -= 2u * (m - u) - 1u;
            u++;
          }
          const uint32_t v = r >> 1, flip = r & 1u;
          float off = flip == 1u ? 2.0f / 3.0f : 1.0f / 3.0f;
          float wb = ((float)u + off) / fm, wc = ((float)v + off) / fm, wa = 1.0f - wb - wc;
          f3 p = T.a * wa + T.b * wb + T.c * wc;
          f3 nv = n0 * wa + n1 * wb + n2 * wc;
          nv = normalize3(dot3(nv, nv) > 1e-12f ? nv : cr);
          // shadeSample's tex (dynlayer.ts:451-452): the alpha test drops the sample, its rgb is the albedo.
          f3 alb = mk3(1.0f * mat.x, 1.0f * mat.y, 1.0f * mat.z);
          if (k.surf.tex) {
            float su, sv;
            surf_uv(k.surf, T.i0, T.i1, T.i2, mk3(wa, wb, wc), su, sv);
            const float4 tx = surf_base_color(k.surf, k.mat, T.m, su, sv, L);
            if (masked && tx.w < surf_cutoff(k.surf, T.m)) continue;
            alb = mk3(tx.x, tx.y, tx.z);
          }
          float ndl = fmaxf(dot3(nv, sd), 0.0f);
          f3 rad = splat3(0.0f);
          if (ndl > 0.0f) rad = alb / PI_F * se * ndl * t.sun_vis;
          splat_sample(k, W, wl, wo, wd, L, p, nv, sa, rad, alb, nw);
        }
      }
    }
    __syncthreads();
    // Add the window to the layer: every touched voxel (a corner with weight > 0 reached it, as a direct add would
    // have set its mask bit) of a claimed brick.
    for (uint32_t j = lane; j < vol; j += SPLAT_WG) {
      if (!((W.touched[j >> 5] >> (j & 31u)) & 1u)) continue;
      const i3 c = wo + i3{(int)(j % (uint32_t)wd.x), (int)((j / (uint32_t)wd.x) % (uint32_t)wd.y),
                           (int)(j / (uint32_t)(wd.x * wd.y))};
      const i3 b = shr3(c, 3);
      uint32_t ptr = D.ptr[(uint32_t)wl * (uint32_t)k.per + ((uint32_t)b.z * k.G + (uint32_t)b.y) * k.G + (uint32_t)b.x];
      if (ptr >= (uint32_t)D.cap) continue;
      uint32_t vi = (uint32_t)((c.x & 7) + 8 * (c.y & 7) + 64 * (c.z & 7));
      atomicOr(&D.mask[ptr * 16u + (vi >> 5)], 1u << (vi & 31u));
      uint32_t* d = D.data + ((size_t)ptr * 512u + vi) * DYN_WORDS;
      for (uint32_t q = 0; q < nw; q++) {
        const uint32_t x = W.acc[j * SPLAT_W + q];
        if (x) atomicAdd(&d[q], x);
      }
    }
    __syncthreads();
  }
}

// resolve (dynlayer.ts:542-580) of level L >= 1.
__device__ __forceinline__ void resolve_body(const VoxK& k, int L, uint32_t lb, uint32_t nblk) {
  const DynK& D = k.dyn;
  const uint32_t cnt = D.hdr[24u + L], s0 = D.hdr[16u + L];
  const uint32_t G = (uint32_t)k.G, G3 = (uint32_t)k.per;
  for (uint32_t i = lb * BLK + threadIdx.x; i < cnt * 512u; i += nblk * BLK) {
    const uint32_t slot = s0 + (i >> 9), bit = i & 511u;
    const uint32_t pi = D.meta[slot], r = pi % G3;
    i3 c = i3{(int)(r % G), (int)((r / G) % G), (int)(r / (G * G))} * 8 +
           i3{(int)(bit & 7u), (int)((bit >> 3) & 7u), (int)(bit >> 6)};
    const uint32_t own = (D.mask[slot * 16u + (bit >> 5)] >> (bit & 31u)) & 1u;
    i3 o0{k.origin[L][0], k.origin[L][1], k.origin[L][2]}, o1{k.origin[L - 1][0], k.origin[L - 1][1], k.origin[L - 1][2]};
    i3 gc = (o0 * 8 + c) * 2 - o1 * 8;
    uint32_t cptr = (uint32_t)D.cap;
    const int E = k.G * 8;
    if (gc.x >= 0 && gc.y >= 0 && gc.z >= 0 && gc.x < E && gc.y < E && gc.z < E) {
      i3 b = shr3(gc, 3);
      cptr = D.ptr[(uint32_t)(L - 1) * G3 + ((uint32_t)b.z * G + (uint32_t)b.y) * G + (uint32_t)b.x];
    }
    uint32_t kids = 0;
    if (cptr < (uint32_t)D.cap) {
      for (int q = 0; q < 8; q++) {
        i3 l = i3{gc.x & 7, gc.y & 7, gc.z & 7} + i3{q & 1, (q >> 1) & 1, q >> 2};
        uint32_t v = (uint32_t)(l.x + 8 * l.y + 64 * l.z);
        if ((D.mask[cptr * 16u + (v >> 5)] >> (v & 31u)) & 1u) kids |= 1u << q;
      }
    }
    if (own == 0u && kids == 0u) continue;
    uint32_t* d = D.data + ((size_t)slot * 512u + bit) * DYN_WORDS;
    float acc[DYN_WORDS];
    for (uint32_t q = 0; q < DYN_WORDS; q++) acc[q] = (float)d[q] / DYN_FIX;
    for (int q = 0; q < 8; q++) {
      if (!((kids >> q) & 1u)) continue;
      i3 l = i3{gc.x & 7, gc.y & 7, gc.z & 7} + i3{q & 1, (q >> 1) & 1, q >> 2};
      const uint32_t* cb = D.data + ((size_t)cptr * 512u + (uint32_t)(l.x + 8 * l.y + 64 * l.z)) * DYN_WORDS;
      for (uint32_t w = 0; w < DYN_WORDS; w++) {
        uint32_t u = cb[w];
        float x = L == 1 ? (float)u / DYN_FIX : __uint_as_float(u);
        acc[w] += 0.25f * x;
      }
    }
    for (uint32_t q = 0; q < DYN_WORDS; q++) d[q] = __float_as_uint(acc[q]);
    if (own == 0u) atomicOr(&D.mask[slot * 16u + (bit >> 5)], 1u << (bit & 31u));
  }
}

__global__ void __launch_bounds__(BLK) k_dl_mark(VoxK k) {
  __shared__ uint32_t lbits[MARK_LDS];
  mark_body(k, lbits, blockIdx.x, gridDim.x);
}
__global__ void __launch_bounds__(BLK) k_dl_parents(VoxK k) { parents_body(k, blockIdx.x, gridDim.x); }
__global__ void __launch_bounds__(1024) k_dl_alloc(VoxK k) { alloc_body(k); }
__global__ void __launch_bounds__(BLK) k_dl_clear(VoxK k) { clear_body(k, blockIdx.x, gridDim.x); }
__global__ void __launch_bounds__(SPLAT_WG) k_dl_splat(TraceK t) {
  __shared__ SplatWin W;
  splat_body(t.vx, t, W, blockIdx.x, gridDim.x);
}
__global__ void __launch_bounds__(BLK) k_dl_resolve(VoxK k, int L) { resolve_body(k, L, blockIdx.x, gridDim.x); }

#define DL_BATCHED(name, body)                                              \
  __global__ void __launch_bounds__(BLK) name(BatchK b) {                   \
    const uint32_t blk = batch_block(b);                  \
    int w = batch_world(b, blk);                                            \
    body(b.w[w].vx, blk - b.start[w], b.start[w + 1] - b.start[w]);         \
  }
__global__ void __launch_bounds__(BLK) k_dl_mark_b(BatchK b) {
  __shared__ uint32_t lbits[MARK_LDS];
  const uint32_t blk = batch_block(b);
  int w = batch_world(b, blk);
  mark_body(b.w[w].vx, lbits, blk - b.start[w], b.start[w + 1] - b.start[w]);
}
DL_BATCHED(k_dl_parents_b, parents_body)
DL_BATCHED(k_dl_clear_b, clear_body)
__global__ void __launch_bounds__(1024) k_dl_alloc_b(BatchK b) {
  int w = batch_world(b, batch_block(b));
  alloc_body(b.w[w].vx);
}
__global__ void __launch_bounds__(SPLAT_WG) k_dl_splat_b(BatchK b) {
  __shared__ SplatWin W;
  const uint32_t blk = batch_block(b);
  int w = batch_world(b, blk);
  splat_body(b.w[w].vx, b.w[w], W, blk - b.start[w], b.start[w + 1] - b.start[w]);
}
__global__ void __launch_bounds__(BLK) k_dl_resolve_b(BatchK b) {
  const uint32_t blk = batch_block(b);
  int w = batch_world(b, blk);
  resolve_body(b.w[w].vx, b.arg, blk - b.start[w], b.start[w + 1] - b.start[w]);
}

}  // namespace

static BatchK dbk(const BatchPlan& p, int l, int arg = 0) { return BatchK{p.w, p.start[l], p.n, arg, p.xcd[l] ? 1 : 0}; }

int launch_dyn_build(const VoxK& k, void* stream) {
  if (!k.dyn.hdr || k.dyn.end_tri <= k.dyn.first_tri) return 0;
  hipStream_t s = (hipStream_t)stream;
  TraceK t{};
  t.vx = k;
  uint32_t g[BL_COUNT];
  batch_blocks(t, g);
  k_dl_mark<<<g[BL_DL_MARK], BLK, 0, s>>>(k);
  if (g[BL_DL_PARENTS]) k_dl_parents<<<g[BL_DL_PARENTS], BLK, 0, s>>>(k);
  k_dl_alloc<<<1, 1024, 0, s>>>(k);
  k_dl_clear<<<g[BL_DL_CLEAR], BLK, 0, s>>>(k);
  return (int)hipGetLastError();
}

// The splat shades with the frame (sun), so it takes the trace arguments; the resolve follows it.
int launch_dyn_splat(const TraceK& t, void* stream) {
  const VoxK& k = t.vx;
  if (!k.dyn.hdr || k.dyn.end_tri <= k.dyn.first_tri) return 0;
  hipStream_t s = (hipStream_t)stream;
  uint32_t g[BL_COUNT];
  batch_blocks(t, g);
  k_dl_splat<<<g[BL_DL_SPLAT], SPLAT_WG, 0, s>>>(t);
  for (int L = 1; L < k.NL; L++) k_dl_resolve<<<g[BL_DL_RESOLVE + L - 1], BLK, 0, s>>>(k, L);
  return (int)hipGetLastError();
}

int launch_dyn_build_batch(const BatchPlan& p, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (p.grid[BL_DL_MARK]) k_dl_mark_b<<<p.grid[BL_DL_MARK], BLK, 0, s>>>(dbk(p, BL_DL_MARK));
  if (p.grid[BL_DL_PARENTS]) k_dl_parents_b<<<p.grid[BL_DL_PARENTS], BLK, 0, s>>>(dbk(p, BL_DL_PARENTS));
  if (p.grid[BL_DL_ALLOC]) k_dl_alloc_b<<<p.grid[BL_DL_ALLOC], 1024, 0, s>>>(dbk(p, BL_DL_ALLOC));
  if (p.grid[BL_DL_CLEAR]) k_dl_clear_b<<<p.grid[BL_DL_CLEAR], BLK, 0, s>>>(dbk(p, BL_DL_CLEAR));
  if (p.grid[BL_DL_SPLAT]) k_dl_splat_b<<<p.grid[BL_DL_SPLAT], SPLAT_WG, 0, s>>>(dbk(p, BL_DL_SPLAT));
  for (int L = 1; L < RTGI_MAX_LEVELS; L++)
    if (p.grid[BL_DL_RESOLVE + L - 1])
      k_dl_resolve_b<<<p.grid[BL_DL_RESOLVE + L - 1], BLK, 0, s>>>(dbk(p, BL_DL_RESOLVE + L - 1, L));
  return (int)hipGetLastError();
}
