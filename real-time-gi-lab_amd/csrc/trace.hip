// Radiance-cascade probe-interval trace, cascade merge and SH projection on gfx950.
//
//   k_trace        one thread per (cascade, direction, probe) ray over all cascades in one launch. Rays of a cascade
//                  are ordered direction-major (ray = di * probes + probe), so a wavefront marches 64 neighbouring
//                  probes along one direction: their DDA steps stay in lockstep and their rgba writes coalesce.
//                  Per ray: interval [b_c, b_{c+1}] s0, voxel level from the cone footprint, hierarchical DDA
//                  through the clipmap (traceVox, src/core/voxel.ts:121-202), Lambert hit shading (shadeHitV,
//                  src/core/rc.ts:108-116) -> rgba f32 (radiance, transmittance), optionally the first-hit voxel id.
//   k_merge_up     cascade c >= 1, launched N-1 .. 1: rad_c += T_c * mean over the grid parents of the upper 2x2
//                  direction average (mergeWorld/upperAvg, src/core/rc.ts:681-708); a wavefront holds 16 probes x
//                  one 2x2 block of directions and also stores the block's average once, so the next cascade's
//                  merge reads 8 parent values per ray, not 32.
//   k_merge0       cascade 0, thread per ray.
//   k_trace_merge  the trace and merge of one cascade in one launch (cascades N-1 .. 0): the traced value is merged
//                  in registers and stored once, which saves the merge's read and write of every ray.
//   k_sh           c0 probes: solid-angle-weighted L2 SH, cosine-convolved (projectSH, src/core/rc.ts:745-774).
//   k_trace_rays   ray queries: arbitrary rays traced as a cascade's rays (error measurements).
//   *_b            the same bodies for a batch of worlds: a launch covers every world's blocks (block -> world by a
//                  binary search over per-launch block prefixes); per-world parameters come from descriptors.
// Shading is deferred out of the DDA loop: a lane that hits leaves the loop with its voxel, and the wavefront shades
// all its hits once after the loop instead of once per loop iteration in which some lane hit.
//
// The DDA restates the reference's f32 operation sequence exactly (compiled with -ffp-contract=off; IEEE division
// and square root; x / v becomes x * (1/v) only when v is a power of two, where both are the same value), so
// first-hit voxels are bit-identical to a CPU restatement of the same WGSL.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <mutex>

#include "rtgi_internal.h"
#include "batch.h"
#include "vmath.h"

namespace {

constexpr int BLK = 256;
// PI of the WGSL (an abstract-float constant): f32(3.14159265) where it meets f32 values, and its products with
// literals folded in double before rounding to f32 (src/core/rc.ts:145, 765).
constexpr float PI_F = (float)3.14159265;
constexpr float SH_A0 = (float)3.14159265;
constexpr float SH_A1 = (float)(2.0 * 3.14159265 / 3.0);
constexpr float SH_A2 = (float)(3.14159265 / 4.0);

// Radiance storage: f32 rgba, or (precision form) 4 halves rounded to nearest even, as packRad / unpackRad
// (src/core/rc.ts:309-310). rad_store returns the value a later read sees (the rounded one in f16 mode).
__device__ __forceinline__ uint32_t h2(float x, float y) {
  return (uint32_t)__half_as_ushort(__float2half_rn(x)) | ((uint32_t)__half_as_ushort(__float2half_rn(y)) << 16);
}
__device__ __forceinline__ float hlo(uint32_t p) { return __half2float(__ushort_as_half((unsigned short)(p & 0xffffu))); }
__device__ __forceinline__ float hhi(uint32_t p) { return __half2float(__ushort_as_half((unsigned short)(p >> 16))); }
__device__ __forceinline__ float4 rad_load(const TraceK& a, size_t i) {
  if (a.fp16) {
    uint2 p = reinterpret_cast<const uint2*>(a.rad)[i];
    return make_float4(hlo(p.x), hhi(p.x), hlo(p.y), hhi(p.y));
  }
  return reinterpret_cast<const float4*>(a.rad)[i];
}
__device__ __forceinline__ float4 rad_store(const TraceK& a, size_t i, float4 v) {
  if (a.fp16) {
    uint2 p = make_uint2(h2(v.x, v.y), h2(v.z, v.w));
    reinterpret_cast<uint2*>(a.rad)[i] = p;
    return make_float4(hlo(p.x), hhi(p.x), hlo(p.y), hhi(p.y));
  }
  reinterpret_cast<float4*>(a.rad)[i] = v;
  return v;
}

// Buffer-resource loads for the DDA's per-step reads: the base sits in scalar registers (one per world, uniform over
// a workgroup) and a lane supplies a 32-bit byte offset, so a read costs no 64-bit address arithmetic; offsets past
// the resource's size read 0 (never the case here: indexes are in range by construction).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ uint32_t buf_u32(__amdgpu_buffer_rsrc_t r, uint32_t byte_off) {
  return (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(r, (int)byte_off, 0, 0);
}

__device__ __forceinline__ float4* ppos4(const TraceK& a) { return reinterpret_cast<float4*>(a.ppos); }
__device__ __forceinline__ float4* pnrm4(const TraceK& a) { return reinterpret_cast<float4*>(a.pnrm); }

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

// voxRank (src/core/voxel.ts:81-87): attribute slot of a voxel, ATTR_NONE when its brick got no slots. The per-word
// first slots (written with the slot allocation) hold the brick's first slot plus the popcounts of the earlier words.
__device__ __forceinline__ uint32_t vox_rank(const VoxK& k, uint32_t brick, uint32_t bit) {
  uint32_t w = bit >> 5;
  uint32_t first = k.wslot[brick * 16u + w];
  if (first == ATTR_NONE) return ATTR_NONE;
  return first + __popc(k.occ[brick * 16u + w] & ((1u << (bit & 31u)) - 1u));
}

__device__ __forceinline__ float unorm8_to_f(uint32_t b) { return div255((float)b); }
__device__ __forceinline__ float snorm16_to_f(uint32_t h) { return fmaxf(div32767((float)(int16_t)(uint16_t)h), -1.0f); }

// Hit radiance of an occupied voxel: attributes, oct-decoded (unnormalised) normal, shadeHitV with no bounce
// (src/core/rc.ts:108-116, 136-140; voxel.ts:150-156, 185-192). Point lights are an addition of this path (0 lights:
// the reference's sun-only shading).
__device__ f3 shade_voxel(const TraceK& a, uint32_t ai, f3 d, f3 entry, uint32_t brick, uint32_t bit) {
  uint32_t pa = 0u, pn = 0u;
  if (ai != ATTR_NONE) {
    pa = a.vx.attr[ai * 2u];
    pn = a.vx.attr[ai * 2u + 1u];
  }
  f3 alb = mk3(unorm8_to_f(pa & 255u), unorm8_to_f((pa >> 8) & 255u), unorm8_to_f((pa >> 16) & 255u));
  if (a.vx.alb16) {  // precision form: the bf16 albedo
    const uint2 q = ai != ATTR_NONE ? reinterpret_cast<const uint2*>(a.vx.alb16)[ai] : make_uint2(0u, 0u);
    alb = mk3(bf2f(q.x), bf2f(q.x >> 16), bf2f(q.y));
  }
  float nx = snorm16_to_f(pn & 0xffffu), ny = snorm16_to_f(pn >> 16);
  f3 n0 = mk3(nx, ny, 1.0f - fabsf(nx) - fabsf(ny));
  float q = fmaxf(-n0.z, 0.0f);
  n0.x += n0.x >= 0.0f ? -q : q;
  n0.y += n0.y >= 0.0f ? -q : q;
  // shadeHitV
  f3 n = dot3(n0, n0) > 0.01f ? n0 : -d;
  if (dot3(n, d) > 0.0f) n = -n;
  f3 e = splat3(0.0f);
  if (a.bounce && ai != ATTR_NONE) {
    // voxBounce (rc.ts:198-202): the side of the voxel the ray comes from, judged at a point a quarter voxel back
    // from the entry (voxShade) against the voxel centre (voxSide, rc.ts:122-126).
    uint32_t code = a.vx.bcode[brick];
    int L = (int)(code >> 27);
    i3 bc{(int)((code >> 18) & 511u), (int)((code >> 9) & 511u), (int)(code & 511u)};
    i3 c = bc * 8 + i3{(int)(bit & 7u), (int)((bit >> 3) & 7u), (int)(bit >> 6)};
    float v = vox_size(a.vx, L);
    f3 pc = tof3(origin(a.vx, L)) * v * 8.0f + (tof3(c) + splat3(0.5f)) * v;
    f3 p = entry - d * (0.25f * v);
    float sd = dot3(p - pc, n0);
    bool back = fabsf(sd) > 0.05f * v ? !(sd > 0.0f) : dot3(n0, d) > 0.0f;
    const uint32_t* B = a.bounce + (size_t)ai * 4u + (back ? 2u : 0u);
    uint32_t w0 = B[0], w1 = B[1];
    e = a.bf16 ? mk3(bf2f(w0), bf2f(w0 >> 16), bf2f(w1)) : mk3(hlo(w0), hhi(w0), hlo(w1));
  }
  f3 b = alb / PI_F * e;
  f3 sd = mk3(a.sun_dir[0], a.sun_dir[1], a.sun_dir[2]);
  float ndl = fmaxf(dot3(n, sd), 0.0f);
  f3 out = b;
  float vis = (a.sunvis && ai != ATTR_NONE) ? a.sunvis[ai] : a.sun_vis;
  if (ndl > 0.0f) out = b + alb / PI_F * mk3(a.sun_e[0], a.sun_e[1], a.sun_e[2]) * ndl * vis;
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

// floor((p - lo) / v) per component; a multiplication by 1/v when v is a power of two (the same value).
__device__ __forceinline__ i3 cell_at(f3 q, float v, float iv, bool pow2) {
  return pow2 ? f2i3(floor3(q * iv)) : f2i3(floor3(q / v));
}

struct Hit {
  bool hit;
  uint32_t brick, bit;  // occupied voxel (brick slot, bit)
  float tIn;            // entry distance
  uint32_t id;          // level << 29 | window cell
  uint32_t steps;       // loop iterations (the reference's voxSteps), counted only by trace_vox<true>
};

// traceVox (src/core/voxel.ts:121-202): hierarchical DDA over level L; empty bricks are skipped whole. d and inv come
// from the direction table (dd = d with |d| < 1e-9 components set to 1e-9, inv = 1 / dd). Returns the first occupied
// voxel entered at tIn >= tSkip (shading happens after the loop).
template <bool ST>
__device__ Hit trace_vox(const TraceK& a, f3 o, f3 d, f3 inv, float t0, float t1, float tSkip, int L) {
  const VoxK& k = a.vx;
  Hit res{false, 0u, 0u, 0.0f, RTGI_NO_HIT, 0u};
  const bool p2 = k.pow2 != 0;
  float v = vox_size(k, L);
  float iv = k.inv_v0 * inv_pow2(L);
  int G = k.G;
  f3 lo = tof3(origin(k, L)) * v * 8.0f;
  f3 hi = lo + splat3((float)G * v * 8.0f);
  f3 ta = (lo - o) * inv, tb = (hi - o) * inv;
  f3 tn = min3(ta, tb), tf = max3(ta, tb);
  float t = fmaxf(fmaxf(fmaxf(tn.x, tn.y), tn.z), t0);
  float tEnd = fminf(fminf(fminf(tf.x, tf.y), tf.z), t1);
  if (t >= tEnd) return res;
  // sign(dd): inv has the sign of dd.
  i3 st{inv.x > 0.0f ? 1 : -1, inv.y > 0.0f ? 1 : -1, inv.z > 0.0f ? 1 : -1};
  f3 sp = mk3(st.x > 0 ? 1.0f : 0.0f, st.y > 0 ? 1.0f : 0.0f, st.z > 0 ? 1.0f : 0.0f);
  int nVox = G * 8;
  i3 zero{0, 0, 0}, top{nVox - 1, nVox - 1, nVox - 1};
  i3 c = clampi3(cell_at(o + d * t - lo, v, iv, p2), zero, top);
  f3 tMax = ((tof3(c) + sp) * v + lo - o) * inv;
  f3 tDelta = abs3(inv * v);
  float tIn = t;
  const __amdgpu_buffer_rsrc_t ptr_r = buf_rsrc(k.ptr, (uint32_t)(k.NL * k.per) * 4u);
  const __amdgpu_buffer_rsrc_t occ_r = buf_rsrc(k.occ, (uint32_t)k.max_bricks * 64u);
  const uint32_t ptr_base = (uint32_t)L * (uint32_t)k.per * 4u;
  uint32_t ptr = 0;
  i3 lc{-8, -8, -8};  // a voxel of the brick whose pointer was read last (none yet: every XOR below is negative)
  int it;
  // Steps and brick jumps select with conditional moves rather than branches (the same f32 operations: each
  // candidate sum is the one the branch would form), so lanes on different axes do not split the wavefront.
  for (it = 0; it < 1024; it++) {
    i3 b = shr3(c, 3);
    // The pointer of a brick is read once per brick visited: c left that brick when some coordinate differs above
    // its low 3 bits (c >= 0 inside the window).
    if ((uint32_t)((c.x ^ lc.x) | (c.y ^ lc.y) | (c.z ^ lc.z)) > 7u) {
      uint32_t bi = ((uint32_t)b.z * (uint32_t)G + (uint32_t)b.y) * (uint32_t)G + (uint32_t)b.x;
      ptr = buf_u32(ptr_r, ptr_base + bi * 4u);
      lc = c;
    }
    if (ptr >= PTR_MARK && ptr != PTR_DROP) {  // empty or marked-unallocated brick: jump to its exit
      f3 bt = ((tof3(b * 8) + sp * 8.0f) * v + lo - o) * inv;
      float te = fminf(fminf(bt.x, bt.y), bt.z);
      if (te >= tEnd) break;
      bool bx = bt.x <= bt.y && bt.x <= bt.z, by = !bx && bt.y <= bt.z, bz = !bx && !by;
      i3 nb{b.x + (bx ? st.x : 0), b.y + (by ? st.y : 0), b.z + (bz ? st.z : 0)};
      if (!inside3(nb, G)) break;
      t = te;
      tIn = te;
      // The next iteration reads nb's pointer; read it here. When nb is empty too, that iteration jumps again, and a
      // jump reads only the brick of c, t's successor and the ray: the entry cell and tMax of nb are never used, so
      // they are formed only where the chain of empty bricks ends (the same values the step-by-step loop forms).
      ptr = buf_u32(ptr_r, ptr_base + (((uint32_t)nb.z * (uint32_t)G + (uint32_t)nb.y) * (uint32_t)G + (uint32_t)nb.x) * 4u);
      if (ptr >= PTR_MARK && ptr != PTR_DROP) {
        c = nb * 8;
        lc = c;
        continue;
      }
      c = clampi3(cell_at(o + d * t - lo, v, iv, p2), nb * 8, nb * 8 + 7);
      tMax = ((tof3(c) + sp) * v + lo - o) * inv;
      lc = c;
      continue;
    }
    bool occupied;
    uint32_t hbrick, hbit;
    int hlev = L;
    i3 hcell = c;
    if (ptr == PTR_DROP) {  // dropped brick: voxel by voxel, each read at the coarser level that holds it
      VoxHit h;
      occupied = vox_at(k, L, c, h);
      hbrick = h.brick;
      hbit = h.bit;
      hlev = h.level;
      hcell = h.cell;
    } else {
      hbrick = ptr;
      hbit = bit_of(c);
      occupied = (buf_u32(occ_r, ptr * 64u + (hbit >> 5) * 4u) & (1u << (hbit & 31u))) != 0u;
    }
    if (occupied && tIn >= tSkip) {  // the hit id is formed once, here
      res = Hit{true, hbrick, hbit, tIn, hit_id(k, hlev, hcell), ST ? (uint32_t)it + 1u : 0u};
      return res;
    }
    float tExit = fminf(fminf(tMax.x, tMax.y), tMax.z);
    if (tExit >= tEnd) break;
    tIn = tExit;
    bool sx = tMax.x <= tMax.y && tMax.x <= tMax.z, sy = !sx && tMax.y <= tMax.z, sz = !sx && !sy;
    c = i3{c.x + (sx ? st.x : 0), c.y + (sy ? st.y : 0), c.z + (sz ? st.z : 0)};
    f3 nt = tMax + tDelta;
    tMax = mk3(sx ? nt.x : tMax.x, sy ? nt.y : tMax.y, sz ? nt.z : tMax.z);
    if (!inside3(c, nVox)) break;
  }
  if (ST) res.steps = (uint32_t)min(it + 1, 1024);
  return res;
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

// ---- Dynamic layer read (traceDyn, src/core/dynlayer.ts:104-160) ----

// dynFind: the voxel at p of level L, or of the first coarser level whose brick was not dropped; x = its first data
// word (DYN_EMPTY: empty brick), y = the level read.
__device__ __forceinline__ uint2 dyn_find(const VoxK& k, int L, f3 p) {
  const DynK& D = k.dyn;
  int l = L;
  for (;;) {
    float v = vox_size(k, l);
    i3 c = cell_at(p - tof3(origin(k, l)) * v * 8.0f, v, k.inv_v0 * inv_pow2(l), k.pow2 != 0);
    if (!inside3(c, k.G * 8)) return make_uint2(DYN_EMPTY, (uint32_t)l);
    i3 b = shr3(c, 3);
    uint32_t ptr = D.ptr[(uint32_t)l * (uint32_t)k.per + ((uint32_t)b.z * k.G + (uint32_t)b.y) * k.G + (uint32_t)b.x];
    if (ptr == DYN_DROPPED && l + 1 < k.NL) {
      l++;
      continue;
    }
    if (ptr >= (uint32_t)D.cap) return make_uint2(DYN_EMPTY, (uint32_t)l);
    return make_uint2((ptr * 512u + bit_of(c)) * DYN_WORDS, (uint32_t)l);
  }
}
// dynWord: level 0 holds fixed point, coarser levels floats.
__device__ __forceinline__ float dyn_word(const DynK& D, uint32_t base, uint32_t q, uint32_t L) {
  uint32_t u = D.data[base + q];
  return L == 0u ? (float)u / DYN_FIX : __uint_as_float(u);
}
// dynRead: mean radiance and coverage per voxel length for direction d (the three faces against d).
__device__ __forceinline__ float4 dyn_read(const DynK& D, uint2 f, f3 d) {
  if (f.x == DYN_EMPTY) return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  float a = dyn_word(D, f.x, 6, f.y);
  if (a <= 0.0f) return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  uint32_t sx = d.x > 0.0f ? 1u : 0u, sy = d.y > 0.0f ? 1u : 0u, sz = d.z > 0.0f ? 1u : 0u;
  float front = fabsf(d.x) * dyn_word(D, f.x, sx, f.y) + fabsf(d.y) * dyn_word(D, f.x, 2u + sy, f.y) +
                fabsf(d.z) * dyn_word(D, f.x, 4u + sz, f.y);
  f3 r = mk3(dyn_word(D, f.x, 7, f.y) / a, dyn_word(D, f.x, 8, f.y) / a, dyn_word(D, f.x, 9, f.y) / a);
  return make_float4(r.x, r.y, r.z, fminf(front, 1.0f));
}
// traceDyn: covered radiance and coverage of [t0, t1] along d from o, marched in steps of half the cone width
// w = kc t (not below v0), read at the two levels around w and interpolated; coverage adds front to back up to 1.
__device__ float4 trace_dyn(const TraceK& a, f3 o, f3 d, float t0, float t1, float kc) {
  const VoxK& k = a.vx;
  const DynK& D = k.dyn;
  if (D.hdr[0] == 0u || t1 <= t0) return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  f3 bmin = (mk3((float)D.hdr[2], (float)D.hdr[3], (float)D.hdr[4]) - splat3(1073741824.0f)) * 0.01f;
  f3 bmax = (mk3((float)D.hdr[5], (float)D.hdr[6], (float)D.hdr[7]) - splat3(1073741824.0f)) * 0.01f;
  f3 dd = mk3(fabsf(d.x) < 1e-9f ? 1e-9f : d.x, fabsf(d.y) < 1e-9f ? 1e-9f : d.y, fabsf(d.z) < 1e-9f ? 1e-9f : d.z);
  f3 ta = (bmin - o) / dd, tb = (bmax - o) / dd;
  f3 tn = min3(ta, tb), tf = max3(ta, tb);
  float ba = fmaxf(fmaxf(fmaxf(tn.x, tn.y), tn.z), t0);
  float bb = fminf(fminf(fminf(tf.x, tf.y), tf.z), t1);
  if (ba >= bb) return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const float v0 = k.v0;
  const int nl = k.NL;
  f3 rad = splat3(0.0f);
  float cov = 0.0f, t = t0;
  for (int i = 0; i < 128; i++) {
    if (t >= bb || cov >= 1.0f) break;
    float w = fmaxf(kc * t, v0);
    float te = fminf(t + 0.5f * w, t1);
    if (te > ba) {
      float len = te - t;
      f3 p = o + d * (0.5f * (t + te));
      // w = v0 gives log2(1) = 0 exactly: the series is skipped there
      float lf = w > v0 ? fminf(fmaxf(log2_series(w / v0), 0.0f), (float)(nl - 1)) : 0.0f;
      int L0 = (int)floorf(lf);
      float f = lf - (float)L0;
      int Lw = vox_level_for(k, L0, p);
      if (Lw != L0) {
        L0 = Lw;
        f = 0.0f;
      }
      uint2 fa = dyn_find(k, L0, p);
      float4 sa = dyn_read(D, fa, d);
      float aa = sa.w > 0.0f ? sa.w * len / vox_size(k, (int)fa.y) : 0.0f;  // 0 x len / v = 0: no division
      f3 ra = mk3(sa.x, sa.y, sa.z) * aa;
      if (f > 0.0f && L0 + 1 < nl) {
        uint2 fb = dyn_find(k, L0 + 1, p);
        float4 sb = dyn_read(D, fb, d);
        float ab = sb.w > 0.0f ? sb.w * len / vox_size(k, (int)fb.y) : 0.0f;
        f3 rb = mk3(sb.x, sb.y, sb.z) * ab;
        ra = ra * (1.0f - f) + rb * f;  // mix
        aa = aa * (1.0f - f) + ab * f;
      }
      if (aa > 0.0f) {
        float ac = fminf(aa, 1.0f - cov);
        rad = rad + ra * (ac / aa);
        cov += ac;
      }
    }
    t = te;
  }
  return make_float4(rad.x, rad.y, rad.z, cov);
}
// The dynamic layer in front of a cascade ray's static result r (traceWorld, src/core/rc.ts:224-231): read from
// max(t0, v0) to the static hit's entry distance (or t1), at the width of the direction bin's cone.
__device__ __forceinline__ float4 with_dyn(const TraceK& a, int c, f3 o, f3 d, float t0, float t1, const Hit& h, f3 ot,
                                           float4 r) {
  if (!a.vx.dyn.hdr) return r;
  float tS = h.hit ? dot3(ot + d * h.tIn - o, d) : t1;
  float kc = 3.5449077f / (float)a.c[c].res;
  float4 dl = trace_dyn(a, o, d, fmaxf(t0, a.vx.v0), fminf(t1, tS), kc);
  if (dl.w > 0.0f)
    r = make_float4(dl.x + (1.0f - dl.w) * r.x, dl.y + (1.0f - dl.w) * r.y, dl.z + (1.0f - dl.w) * r.z,
                    (1.0f - dl.w) * r.w);
  return r;
}

__device__ __forceinline__ int cascade_of(const TraceK& a, uint64_t ray) {
  int c = 0;
  while (c + 1 < a.N && ray >= a.c[c + 1].ray_off) c++;
  return c;
}

// Probe grid coordinates of probe index pi of a cascade (storage order: rows, or blocks of 2^tile_log2 per axis; with
// 4^3 blocks a wavefront of 64 consecutive probes holds one compact block of parallel rays).
__device__ __forceinline__ i3 probe_xyz(const CascK& C, uint32_t pi) {
  if (C.tile_log2 > 0) {
    int T = C.tile_log2, nt = C.n_log2 - T;
    uint32_t mT = (1u << T) - 1u, mt = (1u << nt) - 1u, t = pi >> (3 * T);
    return i3{(int)(((t & mt) << T) | (pi & mT)), (int)((((t >> nt) & mt) << T) | ((pi >> T) & mT)),
              (int)(((t >> (2 * nt)) << T) | ((pi >> (2 * T)) & mT))};
  }
  if (C.n_log2 >= 0) {
    uint32_t m = (uint32_t)C.n - 1u;
    return i3{(int)(pi & m), (int)((pi >> C.n_log2) & m), (int)(pi >> (2 * C.n_log2))};
  }
  uint32_t n = (uint32_t)C.n;
  return i3{(int)(pi % n), (int)((pi / n) % n), (int)(pi / (n * n))};
}
// (direction, probe) of ray r of a cascade (direction-major).
__device__ __forceinline__ void split_ray(const CascK& C, uint32_t r, uint32_t& di, uint32_t& pi) {
  if (C.n_log2 >= 0) {
    di = r >> (3 * C.n_log2);
    pi = r & (C.np - 1u);
  } else {
    di = r / C.np;
    pi = r % C.np;
  }
}

// traceWorld (src/core/rc.ts:216-222) for probe (grid point g, device index pi) of cascade c along d: interval
// [b0, b1] s0, level voxLevelFor at the interval's end, start voxOffset voxels along the probe normal (relocated
// probes; the grid points have normal 0), skip voxSkip voxels. ot: the traversal's origin (hits enter at ot + d t).
template <bool ST>
__device__ __forceinline__ void trace_probe_ray(const TraceK& a, int c, i3 g, uint32_t pi, f3 d, f3 inv, Hit& h,
                                                f3& ot, f3& op) {
  const CascK& C = a.c[c];
  f3 o, n = splat3(0.0f);
  if (a.reloc) {
    float4 P = ppos4(a)[C.probe_off + pi], N = pnrm4(a)[C.probe_off + pi];
    o = mk3(P.x, P.y, P.z);
    n = mk3(N.x, N.y, N.z);
  } else {
    o = tof3(g) * (a.s0 * (float)(1u << c));
  }
  float t0 = C.b0 * a.s0, t1 = C.b1 * a.s0;
  int L = vox_level_for(a.vx, C.vox_level, o + d * t1);
  float v = vox_size(a.vx, L);
  ot = a.reloc ? o + n * (a.vox_offset * v) : o;
  op = o;
  h = trace_vox<ST>(a, ot, d, inv, t0, t1, a.vox_skip * v, L);
}

template <bool ST, bool DY>
__device__ __forceinline__ void trace_body(const TraceK& a, uint32_t lb, uint32_t nblk, uint32_t tx = threadIdx.x) {
  uint64_t ray = (uint64_t)lb * BLK + tx;
  if (ray >= a.total_rays) return;
  int c = cascade_of(a, ray);
  const CascK& C = a.c[c];
  uint32_t di, pi;
  split_ray(C, (uint32_t)(ray - C.ray_off), di, pi);
  i3 g = i3{a.G0[0] >> c, a.G0[1] >> c, a.G0[2] >> c} + probe_xyz(C, pi);
  const DirEntry& de = a.dirs[C.dir_off + di];
  f3 d = mk3(de.d[0], de.d[1], de.d[2]);
  f3 inv = mk3(de.inv[0], de.inv[1], de.inv[2]);
  Hit h;
  f3 ot;
  f3 op;
  trace_probe_ray<ST>(a, c, g, pi, d, inv, h, ot, op);
  float4 out = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
  if (h.hit) {
    f3 rgb = shade_voxel(a, vox_rank(a.vx, h.brick, h.bit), d, ot + d * h.tIn, h.brick, h.bit);
    out = make_float4(rgb.x, rgb.y, rgb.z, 0.0f);
  }
  if (DY) out = with_dyn(a, c, op, d, C.b0 * a.s0, C.b1 * a.s0, h, ot, out);
  rad_store(a, ray, out);
  if (a.hits) a.hits[ray] = h.id;
  if (ST && a.vx.stats) {
    // counts[24 + c], [32 + c], [64 + c] of traceWorld (src/core/rc.ts:232-235), summed per wavefront when all its
    // lanes trace the same cascade.
    unsigned long long* S = a.vx.stats;
    int c0 = __shfl(c, 0);
    bool full = __ballot(1) == ~0ull;  // the reduction reads every lane: only for a wavefront with no lane returned
    if (full && __all(c == c0)) {
      unsigned long long st = h.steps, cap = h.steps >= 1024u ? 1ull : 0ull, n = 1ull;
      for (int m = 32; m >= 1; m >>= 1) {
        st += __shfl_xor(st, m);
        cap += __shfl_xor(cap, m);
        n += __shfl_xor(n, m);
      }
      if (__lane_id() == 0) {
        atomicAdd(&S[c], n);
        atomicAdd(&S[RTGI_MAX_CASCADES + c], st);
        if (cap) atomicAdd(&S[2 * RTGI_MAX_CASCADES + c], cap);
      }
    } else {
      atomicAdd(&S[c], 1ull);
      atomicAdd(&S[RTGI_MAX_CASCADES + c], (unsigned long long)h.steps);
      if (h.steps >= 1024u) atomicAdd(&S[2 * RTGI_MAX_CASCADES + c], 1ull);
    }
  }
}
// Ray queries: n rays from origins o along unit directions d (3 floats each) traced as rays of cascade c (its
// interval, level choice and skip), shaded as the trace does; rgba is rounded as packRad stores it in the precision
// form. inv = 1 / dd with dd = d where |d| >= 1e-9, else 1e-9 (traceVox, src/core/voxel.ts:128-129).
__global__ void __launch_bounds__(BLK) k_trace_rays(TraceK a, int c, const float* o, const float* d, uint32_t n,
                                                    float4* out, uint32_t* hits) {
  uint32_t i = blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  const CascK& C = a.c[c];
  f3 O = mk3(o[3 * i], o[3 * i + 1], o[3 * i + 2]);
  f3 D = mk3(d[3 * i], d[3 * i + 1], d[3 * i + 2]);
  f3 dd = mk3(fabsf(D.x) < 1e-9f ? 1e-9f : D.x, fabsf(D.y) < 1e-9f ? 1e-9f : D.y, fabsf(D.z) < 1e-9f ? 1e-9f : D.z);
  f3 inv = mk3(1.0f / dd.x, 1.0f / dd.y, 1.0f / dd.z);
  float t0 = C.b0 * a.s0, t1 = C.b1 * a.s0;
  int L = vox_level_for(a.vx, C.vox_level, O + D * t1);
  float v = vox_size(a.vx, L);
  Hit h = trace_vox<false>(a, O, D, inv, t0, t1, a.vox_skip * v, L);
  float4 r = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
  if (h.hit) {
    f3 rgb = shade_voxel(a, vox_rank(a.vx, h.brick, h.bit), D, O + D * h.tIn, h.brick, h.bit);
    r = make_float4(rgb.x, rgb.y, rgb.z, 0.0f);
  }
  r = with_dyn(a, c, O, D, t0, t1, h, O, r);
  if (a.fp16) r = make_float4(hlo(h2(r.x, r.y)), hhi(h2(r.x, r.y)), hlo(h2(r.z, r.w)), hhi(h2(r.z, r.w)));
  out[i] = r;
  hits[i] = h.id;
}

// ST: count the tracer statistics (a separate instantiation, so the counting costs the plain trace nothing).
// DY: some world has a dynamic layer (a separate instantiation: the read costs registers the plain trace keeps).
// One wavefront per workgroup (workgroup h traces quarter h & 3 of block h >> 2): a workgroup of four would hold its
// slots until its slowest ray ends.
template <bool ST, bool DY>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(8, 8))) k_trace(TraceK a) {
  trace_body<ST, DY>(a, blockIdx.x >> 2, 0, (blockIdx.x & 3u) * 64u + threadIdx.x);
}
template <bool ST, bool DY>
__global__ void __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(8, 8))) k_trace_b(BatchK b) {
  const uint32_t blk = batch_block(b);
  int w = batch_world(b, blk);
  trace_body<ST, DY>(b.w[w], blk - b.start[w], b.start[w + 1] - b.start[w]);
}

__device__ __forceinline__ f3 rgb_of(float4 x) { return mk3(x.x, x.y, x.z); }

// Parents of probe g of cascade c < N-1 on the dense grid (linkParents, src/core/rc.ts:540-570): the c+1 grid points
// (g >> 1) + o, o in {0,1}^3 with o = 0 on even axes, those outside the grid absent. Bit kk of the mask is parent
// o = (kk & 1, kk >> 1 & 1, kk >> 2 & 1). Per axis one or two parents exist, so their count is 1, 2, 4 or 8.
__device__ __forceinline__ uint32_t parent_mask(const CascK& U, i3 g) {
  uint32_t m = 0;
  for (int kk = 0; kk < 8; kk++) {
    i3 o{kk & 1, (kk >> 1) & 1, (kk >> 2) & 1};
    i3 q = shr3(g, 1) + o;
    bool ok = !((o.x && !(g.x & 1)) || (o.y && !(g.y & 1)) || (o.z && !(g.z & 1))) && q.x < U.n && q.y < U.n &&
              q.z < U.n;
    m |= (ok ? 1u : 0u) << kk;
  }
  return m;
}
// Probe index of grid point q (the inverse of probe_xyz).
__device__ __forceinline__ uint32_t probe_index(const CascK& C, i3 q) {
  if (C.tile_log2 > 0) {
    int T = C.tile_log2, nt = C.n_log2 - T;
    uint32_t mT = (1u << T) - 1u;
    uint32_t t = ((uint32_t)q.x >> T) | (((uint32_t)q.y >> T) << nt) | (((uint32_t)q.z >> T) << (2 * nt));
    uint32_t r = ((uint32_t)q.x & mT) | (((uint32_t)q.y & mT) << T) | (((uint32_t)q.z & mT) << (2 * T));
    return (t << (3 * T)) | r;
  }
  return C.n_log2 >= 0 ? (uint32_t)q.x | ((uint32_t)q.y << C.n_log2) | ((uint32_t)q.z << (2 * C.n_log2))
                       : (uint32_t)q.x + (uint32_t)C.n * ((uint32_t)q.y + (uint32_t)C.n * (uint32_t)q.z);
}
__device__ __forceinline__ uint32_t parent_index(const CascK& U, i3 g, int kk) {
  return probe_index(U, shr3(g, 1) + i3{kk & 1, (kk >> 1) & 1, (kk >> 2) & 1});
}

// Parent mean for one (probe g, direction di) of cascade c < N-1 (mergeWorld, src/core/rc.ts:693-706): sum over the
// existing parents in kk order, divided by their count. avg holds cascade c+1's upperAvg at c's directions. Every
// parent value is loaded before any is used, so a thread's loads are in flight together. The count is a power of two,
// for which acc * (1 / n) is the correctly rounded acc / n.
// With relocation, parents hidden by voxels are left out (linkParents, rc.ts:559-566); their count is then any of
// 0..8, so the mean divides (0 parents: 0, as mergeWorld's select gives).
// The common case (rows of a power-of-two grid, no relocation) reads the eight candidates unconditionally through a
// buffer resource over the cascade's averages (a candidate past its end reads 0) and zeroes the absent ones by
// select: adding +0 for an absent parent leaves the running sum exactly as skipping it does (the sum starts at +0 and
// never is -0), so the result is the same sum in kk order.
__device__ __forceinline__ float3 avg_rgb(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0);
  return make_float3(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]));
}
__device__ __forceinline__ f3 upper_mean(const TraceK& a, int c, i3 g, uint32_t pi, uint32_t di) {
  const CascK& U = a.c[c + 1];
  if (U.tile_log2 == 0 && U.n_log2 >= 0 && !a.reloc) {
    const int nl = U.n_log2;
    const i3 b = shr3(g, 1);
    const bool hx = (g.x & 1) && b.x + 1 < U.n, hy = (g.y & 1) && b.y + 1 < U.n, hz = (g.z & 1) && b.z + 1 < U.n;
    // U's averages: U.np entries per direction of cascade c (res_c^2 of them), float4 each (below 4 GiB: the world
    // layout refuses larger)
    const uint32_t res = (uint32_t)a.c[c].res;
    const __amdgpu_buffer_rsrc_t r =
        buf_rsrc(reinterpret_cast<const float4*>(a.avg) + U.avg_off, U.np * res * res * 16u);
    const uint32_t o = ((di << (3 * nl)) + ((uint32_t)b.x | ((uint32_t)b.y << nl) | ((uint32_t)b.z << (2 * nl)))) * 16u;
    const uint32_t oy = o + (16u << nl), oz = o + (16u << (2 * nl)), oyz = oy + (16u << (2 * nl));
    const float3 x000 = avg_rgb(r, o), x100 = avg_rgb(r, o + 16u), x010 = avg_rgb(r, oy), x110 = avg_rgb(r, oy + 16u),
                 x001 = avg_rgb(r, oz), x101 = avg_rgb(r, oz + 16u), x011 = avg_rgb(r, oyz),
                 x111 = avg_rgb(r, oyz + 16u);
    const float3 z = make_float3(0.0f, 0.0f, 0.0f);
    f3 acc = splat3(0.0f) + mk3(x000.x, x000.y, x000.z);
    const float3 p100 = hx ? x100 : z, p010 = hy ? x010 : z, p110 = hx && hy ? x110 : z, p001 = hz ? x001 : z,
                 p101 = hx && hz ? x101 : z, p011 = hy && hz ? x011 : z, p111 = hx && hy && hz ? x111 : z;
    acc = acc + mk3(p100.x, p100.y, p100.z);
    acc = acc + mk3(p010.x, p010.y, p010.z);
    acc = acc + mk3(p110.x, p110.y, p110.z);
    acc = acc + mk3(p001.x, p001.y, p001.z);
    acc = acc + mk3(p101.x, p101.y, p101.z);
    acc = acc + mk3(p011.x, p011.y, p011.z);
    acc = acc + mk3(p111.x, p111.y, p111.z);
    // 1 / (parent count) = 2^-(hx + hy + hz), exact
    return acc * inv_pow2((hx ? 1 : 0) + (hy ? 1 : 0) + (hz ? 1 : 0));
  }
  const float4* A = reinterpret_cast<const float4*>(a.avg) + U.avg_off + (size_t)di * U.np;
  uint32_t m;
  float4 x[8];
  if (U.tile_log2 == 0 && U.n_log2 >= 0) {
    // Rows of a power-of-two grid: the parents are the base point g >> 1 plus unit steps, and the mask is the AND of
    // one pattern per axis (bits with o = 0 on that axis: 0x55, 0x33, 0x0f; all 8 where the axis has a second parent).
    i3 b = shr3(g, 1);
    const int nl = U.n_log2;
    bool hx = (g.x & 1) && b.x + 1 < U.n, hy = (g.y & 1) && b.y + 1 < U.n, hz = (g.z & 1) && b.z + 1 < U.n;
    m = (hx ? 0xffu : 0x55u) & (hy ? 0xffu : 0x33u) & (hz ? 0xffu : 0x0fu);
    if (a.reloc) m &= a.pvis[a.c[c].probe_off + pi];
    uint32_t i0 = (uint32_t)b.x | ((uint32_t)b.y << nl) | ((uint32_t)b.z << (2 * nl));
    for (int kk = 0; kk < 8; kk++) {
      uint32_t i = i0 + (uint32_t)(kk & 1) + ((uint32_t)((kk >> 1) & 1) << nl) + ((uint32_t)(kk >> 2) << (2 * nl));
      x[kk] = ((m >> kk) & 1u) ? A[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
  } else {
    m = parent_mask(U, g);
    if (a.reloc) m &= a.pvis[a.c[c].probe_off + pi];
    for (int kk = 0; kk < 8; kk++)
      x[kk] = ((m >> kk) & 1u) ? A[parent_index(U, g, kk)] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
  f3 acc = splat3(0.0f);
  for (int kk = 0; kk < 8; kk++)
    if ((m >> kk) & 1u) acc = acc + mk3(x[kk].x, x[kk].y, x[kk].z);
  if (a.reloc) return m ? acc / fmaxf((float)__popc(m), 1.0f) : splat3(0.0f);
  return acc * (1.0f / (float)__popc(m));
}

// rad_c = v.rgb + v.a * up, a := 1, with up = sky at the last cascade or where v.a = 0 (mergeWorld, rc.ts:690-707).
// The parent mean is read before v.a is known (it is discarded where unused), so the loads do not wait for v.
__device__ __forceinline__ float4 merged(const TraceK& a, int c, i3 g, uint32_t di, float4 v, f3 um) {
  f3 up = (c + 1 < a.N && v.w > 0.0f) ? um : mk3(a.sky[0], a.sky[1], a.sky[2]);
  return make_float4(v.x + v.w * up.x, v.y + v.w * up.y, v.z + v.w * up.z, 1.0f);
}
__device__ __forceinline__ float4 merge_ray(const TraceK& a, int c, i3 g, uint32_t pi, uint32_t di, size_t i) {
  f3 um = c + 1 < a.N ? upper_mean(a, c, g, pi, di) : splat3(0.0f);
  return rad_store(a, i, merged(a, c, g, di, rad_load(a, i), um));
}

// Cascade c >= 1, thread per ray. A wavefront holds 16 probes x the 4 directions of one 2x2 block of c's directions
// under texel dl of cascade c-1's resolution (lane = 16 j + probe, j = 2x2 position); it merges its rays and lane j = 0
// stores the block's average ((p00 + p10 + p01 + p11) * 0.25, upperAvg's order, rc.ts:681-688) for cascade c-1's
// merge. Where the resolution is capped (rc.ts:682) the block is one texel: thread per ray, average = merged value.
__device__ __forceinline__ void merge_up_body(const TraceK& a, int c, uint32_t lb, uint32_t nblk) {
  const CascK& C = a.c[c];
  const CascK& D = a.c[c - 1];
  uint64_t t = (uint64_t)lb * BLK + threadIdx.x;
  float4* AV = reinterpret_cast<float4*>(a.avg) + C.avg_off;
  if (C.res == D.res) {
    if (t >= C.rays) return;
    uint32_t di, pi;
    split_ray(C, (uint32_t)t, di, pi);
    float4 m = merge_ray(a, c, probe_xyz(C, pi), pi, di, C.ray_off + t);
    AV[(size_t)di * C.np + pi] = make_float4(m.x, m.y, m.z, 0.0f);
    return;
  }
  uint32_t npb = ((uint32_t)C.np + 15u) >> 4;
  uint64_t grp = t >> 6;
  if (grp >= (uint64_t)D.res * D.res * npb) return;  // uniform over the wavefront
  uint32_t lane = (uint32_t)t & 63u, j = lane >> 4;
  uint32_t dl = (uint32_t)(grp / npb), pi = (uint32_t)(grp % npb) * 16u + (lane & 15u);
  bool ok = pi < (uint32_t)C.np;
  uint32_t lu = dl & ((uint32_t)D.res - 1u), lw = dl >> D.res_log2;
  uint32_t di = ((2u * lw + (j >> 1)) << C.res_log2) + 2u * lu + (j & 1u);
  float4 m = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (ok) m = merge_ray(a, c, probe_xyz(C, pi), pi, di, C.ray_off + (size_t)di * C.np + pi);
  uint32_t l0 = lane & 15u;
  f3 p00 = mk3(__shfl(m.x, l0), __shfl(m.y, l0), __shfl(m.z, l0));
  f3 p10 = mk3(__shfl(m.x, l0 + 16), __shfl(m.y, l0 + 16), __shfl(m.z, l0 + 16));
  f3 p01 = mk3(__shfl(m.x, l0 + 32), __shfl(m.y, l0 + 32), __shfl(m.z, l0 + 32));
  f3 p11 = mk3(__shfl(m.x, l0 + 48), __shfl(m.y, l0 + 48), __shfl(m.z, l0 + 48));
  if (ok && j == 0) {
    f3 avg = (p00 + p10 + p01 + p11) * 0.25f;
    AV[(size_t)dl * C.np + pi] = make_float4(avg.x, avg.y, avg.z, 0.0f);
  }
}
__global__ void __launch_bounds__(BLK) k_merge_up(TraceK a, int c) { merge_up_body(a, c, blockIdx.x, gridDim.x); }

// Cascade 0: thread per ray.
__device__ __forceinline__ void merge0_body(const TraceK& a, uint32_t lb, uint32_t nblk) {
  const CascK& C = a.c[0];
  uint64_t r = (uint64_t)lb * BLK + threadIdx.x;
  if (r >= C.rays) return;
  uint32_t di, pi;
  split_ray(C, (uint32_t)r, di, pi);
  merge_ray(a, 0, probe_xyz(C, pi), pi, di, C.ray_off + r);
}
__global__ void __launch_bounds__(BLK) k_merge0(TraceK a) { merge0_body(a, blockIdx.x, gridDim.x); }

// shBasis (src/core/rc.ts:745-749) of c0 texel d, from the direction table.
__device__ __forceinline__ void sh_basis3(float nx, float ny, float nz, float y[9]) {
  y[0] = 0.282095f;
  y[1] = 0.488603f * ny;
  y[2] = 0.488603f * nz;
  y[3] = 0.488603f * nx;
  y[4] = 1.092548f * nx * ny;
  y[5] = 1.092548f * ny * nz;
  y[6] = 0.315392f * (3.0f * nz * nz - 1.0f);
  y[7] = 1.092548f * nx * nz;
  y[8] = 0.546274f * (nx * nx - ny * ny);
}
__device__ __forceinline__ void sh_basis(const DirEntry& de, float y[9]) { sh_basis3(de.d[0], de.d[1], de.d[2], y); }

// Grid-row index x + n (y + n z) of local grid point q of a cascade with n probes per axis.
__device__ __forceinline__ uint32_t grid_row(const CascK& C, i3 q) {
  return (uint32_t)q.x + (uint32_t)C.n * ((uint32_t)q.y + (uint32_t)C.n * (uint32_t)q.z);
}

// Row of the c0 history that probe gl (local grid point of this update) blends with: the same world grid point in
// the previous update's grid (histBlend's key, rc.ts:784-800; the dense grid holds every point once, so the lookup is
// the index itself), or ~0u when that grid did not hold it or there is no history.
__device__ __forceinline__ uint32_t hist_row(const TraceK& a, i3 gl) {
  if (!a.sh_prev) return ~0u;
  i3 q = gl + i3{a.hist_d[0], a.hist_d[1], a.hist_d[2]};
  return inside3(q, a.c[0].n) ? grid_row(a.c[0], q) : ~0u;
}

// One SH value (plane k = 3 * coefficient + channel) of probe row: the scaled coefficient v, blended with the history
// value at hrow (mix(prev, v, HIST_W = 0.5) = prev * (1 - 0.5) + v * 0.5, rc.ts:801-804), stored as f32 or as a half
// (pack2x16float per value, rc.ts:770-771).
__device__ __forceinline__ float sh_get(const float* sh, int fp16, size_t j) {
  return fp16 ? __half2float(__ushort_as_half(reinterpret_cast<const unsigned short*>(sh)[j])) : sh[j];
}
__device__ __forceinline__ void sh_put(const TraceK& a, uint32_t np, uint32_t row, uint32_t hrow, int k, float v) {
  const size_t i = (size_t)k * np + row;
  if (hrow != ~0u) v = sh_get(a.sh_prev, a.fp16, (size_t)k * np + hrow) * (1.0f - 0.5f) + v * 0.5f;
  if (a.fp16)
    reinterpret_cast<unsigned short*>(a.sh)[i] = __half_as_ushort(__float2half_rn(v));
  else
    a.sh[i] = v;
}
__device__ __forceinline__ float sh_scale(int q) { return q == 0 ? SH_A0 : (q < 4 ? SH_A1 : SH_A2); }

// projectSH (src/core/rc.ts:752-774) of already merged c0 radiance, thread per probe in grid-row order: 27 values per
// probe in 27 planes, blended with the history (histBlend, rc.ts:767).
__device__ __forceinline__ void sh_body(const TraceK& a, uint32_t lb, uint32_t nblk) {
  const CascK& C = a.c[0];
  uint32_t row = lb * BLK + threadIdx.x;
  if (row >= C.np) return;
  const uint32_t n = (uint32_t)C.n;
  const i3 gl{(int)(row % n), (int)((row / n) % n), (int)(row / (n * n))};
  const uint32_t pi = probe_index(C, gl);  // device order of the radiance
  f3 cf[9];
  for (int q = 0; q < 9; q++) cf[q] = splat3(0.0f);
  for (uint32_t d = 0; d < 16; d++) {
    f3 r = rgb_of(rad_load(a, C.ray_off + (size_t)d * C.np + pi)) * a.omega0[d];
    float y[9];
    sh_basis(a.dirs[C.dir_off + d], y);
    for (int q = 0; q < 9; q++) cf[q] = cf[q] + r * y[q];
  }
  float v[27];
  for (int q = 0; q < 9; q++) {
    const float s = sh_scale(q);
    v[3 * q] = cf[q].x * s;
    v[3 * q + 1] = cf[q].y * s;
    v[3 * q + 2] = cf[q].z * s;
  }
  // The history's 27 values are read before any store (the stores might alias them as far as the compiler knows),
  // so the reads are in flight together.
  const uint32_t hrow = hist_row(a, gl);
  if (hrow != ~0u) {
    float p[27];
    for (int k = 0; k < 27; k++) p[k] = sh_get(a.sh_prev, a.fp16, (size_t)k * C.np + hrow);
    for (int k = 0; k < 27; k++) v[k] = p[k] * (1.0f - 0.5f) + v[k] * 0.5f;
  }
  for (int k = 0; k < 27; k++) sh_put(a, C.np, row, ~0u, k, v[k]);
}
// Where the radiance and SH are f16 in grid rows of even length (sh_pairs), a thread projects two neighbouring
// probes: rows r and r + 1 (r even) are grid points (x, y, z) and (x + 1, y, z), their radiance of one direction is
// one 16-byte read and their SH value of one plane one 4-byte write, each probe getting the same operations as in
// sh_body. The launches keep one thread per probe (the block counts are planned before an update's descriptors are
// formed, and its storage order decides the form); in pair form the upper half of the threads has no work. Where
// every world of a launch takes pairs, the pair-only kernel runs (78 VGPRs against the combined kernel's 138: twice
// the wavefronts per SIMD; 32 mixed FP16 worlds' SH 0.65 -> 0.41 ms).
__device__ __forceinline__ void sh_pair_body(const TraceK& a, uint32_t lb) {
  const CascK& C = a.c[0];
  const uint32_t row = 2u * (lb * BLK + threadIdx.x);
  if (row >= C.np) return;
  const uint32_t n = (uint32_t)C.n;
  const i3 gl{(int)(row % n), (int)((row / n) % n), (int)(row / (n * n))};
  const i3 gl1{gl.x + 1, gl.y, gl.z};
  const uint4* __restrict__ R = reinterpret_cast<const uint4*>(a.rad);
  f3 c0[9], c1[9];
  for (int q = 0; q < 9; q++) c0[q] = c1[q] = splat3(0.0f);
  for (uint32_t d = 0; d < 16; d++) {
    const uint4 p = R[(C.ray_off + (size_t)d * C.np + row) >> 1];
    const float w = a.omega0[d];
    const f3 r0 = mk3(hlo(p.x), hhi(p.x), hlo(p.y)) * w, r1 = mk3(hlo(p.z), hhi(p.z), hlo(p.w)) * w;
    float y[9];
    sh_basis(a.dirs[C.dir_off + d], y);
    for (int q = 0; q < 9; q++) {
      c0[q] = c0[q] + r0 * y[q];
      c1[q] = c1[q] + r1 * y[q];
    }
  }
  const unsigned short* __restrict__ H = reinterpret_cast<const unsigned short*>(a.sh_prev);
  uint32_t* __restrict__ S = reinterpret_cast<uint32_t*>(a.sh);
  const uint32_t h0 = hist_row(a, gl), h1 = hist_row(a, gl1);
  for (int q = 0; q < 9; q++) {
    const float s = sh_scale(q);
    const float v0[3] = {c0[q].x * s, c0[q].y * s, c0[q].z * s}, v1[3] = {c1[q].x * s, c1[q].y * s, c1[q].z * s};
    for (int ch = 0; ch < 3; ch++) {
      const size_t k = (size_t)(3 * q + ch) * C.np;
      float x0 = v0[ch], x1 = v1[ch];
      if (h0 != ~0u) x0 = __half2float(__ushort_as_half(H[k + h0])) * (1.0f - 0.5f) + x0 * 0.5f;
      if (h1 != ~0u) x1 = __half2float(__ushort_as_half(H[k + h1])) * (1.0f - 0.5f) + x1 * 0.5f;
      S[(k + row) >> 1] = (uint32_t)__half_as_ushort(__float2half_rn(x0)) |
                          ((uint32_t)__half_as_ushort(__float2half_rn(x1)) << 16);
    }
  }
}
__global__ void __launch_bounds__(BLK) k_sh(TraceK a) { sh_body(a, blockIdx.x, gridDim.x); }
__global__ void __launch_bounds__(BLK) k_sh_p(TraceK a) { sh_pair_body(a, blockIdx.x); }
__global__ void __launch_bounds__(BLK) k_sh_pb(BatchK b) {
  const uint32_t blk = batch_block(b);
  int w = batch_world(b, blk);
  sh_pair_body(b.w[w], blk - b.start[w]);
}
__global__ void __launch_bounds__(BLK) k_sh_b(BatchK b) {
  const uint32_t blk = batch_block(b);
  int w = batch_world(b, blk);
  if (sh_pairs(b.w[w]))
    sh_pair_body(b.w[w], blk - b.start[w]);
  else
    sh_body(b.w[w], blk - b.start[w], b.start[w + 1] - b.start[w]);
}

// Merge step of a batch: each world merges its cascade N-1-step (c > 0 with the 2x2 averages, c = 0 per ray).
__global__ void __launch_bounds__(BLK) k_merge_step_b(BatchK b) {
  const uint32_t blk = batch_block(b);
  int w = batch_world(b, blk);
  const TraceK& a = b.w[w];
  uint32_t lb = blk - b.start[w], nb = b.start[w + 1] - b.start[w];
  int c = a.N - 1 - b.arg;
  if (c > 0)
    merge_up_body(a, c, lb, nb);
  else
    merge0_body(a, lb, nb);
}

// Trace and merge of cascade c in one pass (cascades are launched N-1 .. 0): each thread traces one ray, rounds it
// as packRad stores it, merges it with the parent mean of cascade c+1's averages (merged in the previous launch) and
// stores the merged value once; the traced values are never stored unmerged, which saves the merge pass's read and
// write of every ray. Where cascade c's directions double those of c-1, a wavefront holds 16 probes x the 4
// directions of one 2x2 block under a texel of c-1 (lane = 16 j + probe, j the block position, the layout of
// merge_up_body) and the block's average for c-1's merge comes from the other lanes; otherwise a thread takes one ray
// (direction-major). Results are identical to the trace followed by the merge steps (the same operations on the same
// values).
__device__ __forceinline__ float4 rad_round(const TraceK& a, float4 v) {
  if (!a.fp16) return v;
  uint32_t p0 = h2(v.x, v.y), p1 = h2(v.z, v.w);
  return make_float4(hlo(p0), hhi(p0), hlo(p1), hhi(p1));
}

// projectSH of a c0 block of 16 probes x 16 directions (sh_fuse). Before the trace, sh0_setup stages the basis and
// solid angles of the 16 directions in LDS; after it, each thread gives its merged radiance (p = probe of the block,
// di = direction) and the block forms value k = 3 q + channel of probe p per thread (432 per block), summed over the
// directions in order as sh_body does (the same operations on the same stored values).
struct Sh0Lds {
  float rad[16][16][3];  // [probe][direction][rgb]
  float y[16][9];        // shBasis of the 16 c0 directions
  float om[16];          // their solid angles
};
__device__ __forceinline__ void sh0_setup(const TraceK& a, Sh0Lds& s) {
  if (threadIdx.x < 16u) {
    sh_basis(a.dirs[a.c[0].dir_off + threadIdx.x], s.y[threadIdx.x]);
    s.om[threadIdx.x] = a.omega0[threadIdx.x];
  }
}
__device__ __forceinline__ void sh0_block(const TraceK& a, Sh0Lds& s, uint32_t lb, uint32_t p, uint32_t di, float mx,
                                          float my, float mz) {
  const CascK& C = a.c[0];
  s.rad[p][di][0] = mx;
  s.rad[p][di][1] = my;
  s.rad[p][di][2] = mz;
  __syncthreads();
#pragma unroll 1
  for (uint32_t o = threadIdx.x; o < 16u * 27u; o += BLK) {
    const uint32_t pb = o / 27u, k = o % 27u, q = k / 3u, ch = k % 3u;
    const uint32_t pp = lb * 16u + pb;
    if (pp >= C.np) continue;
    float cf = 0.0f;
    for (uint32_t d = 0; d < 16; d++) cf = cf + (s.rad[pb][d][ch] * s.om[d]) * s.y[d][q];
    const i3 gl = probe_xyz(C, pp);
    sh_put(a, C.np, grid_row(C, gl), hist_row(a, gl), (int)k, cf * sh_scale((int)q));
  }
}

// With sh_fuse, cascade 0's launch also projects the SH: a block holds 16 probes x all 16 c0 directions (wavefront w
// = 2x2 block (w & 1, w >> 1) of the 4x4 texels, lane = 16 j + probe as above) and ends with sh0_block, so the c0
// radiance is not read back. SH: the launch holds some world's cascade 0 (a separate instantiation).
// tx: the thread's index in its block of BLK threads (threadIdx.x, or the block's quarter times 64 plus the lane where
// a workgroup is one wavefront).
template <bool ST, bool DY, bool SH>
__device__ __forceinline__ void trace_merge_body(const TraceK& a, int c, uint32_t lb, uint32_t nblk,
                                                 uint32_t tx = threadIdx.x) {
  __shared__ Sh0Lds s_sh[1];  // referenced only by the SH instantiation
  const CascK& C = a.c[c];
  const bool quad = !SH && !a.no_merge && c > 0 && C.res != a.c[c - 1].res;
  const bool sh0 = SH;  // the SH instantiation runs cascade 0 of worlds with sh_fuse only
  uint64_t t = (uint64_t)lb * BLK + tx;
  uint32_t lane = threadIdx.x & 63u, di = 0, pi = 0, dl = 0, j = 0;
  bool ok;
  if (sh0) {
    const uint32_t wv = threadIdx.x >> 6;
    j = lane >> 4;
    pi = lb * 16u + (lane & 15u);
    ok = pi < C.np;
    di = ((2u * (wv >> 1) + (j >> 1)) << 2) + 2u * (wv & 1u) + (j & 1u);
    if (SH) sh0_setup(a, s_sh[0]);
  } else if (quad) {
    const CascK& D = a.c[c - 1];
    uint32_t npb = (C.np + 15u) >> 4;
    uint64_t grp = t >> 6;
    if (grp >= (uint64_t)D.res * D.res * npb) return;  // uniform over the wavefront
    j = lane >> 4;
    dl = (uint32_t)(grp / npb);
    pi = (uint32_t)(grp % npb) * 16u + (lane & 15u);
    ok = pi < C.np;
    uint32_t lu = dl & ((uint32_t)D.res - 1u), lw = dl >> D.res_log2;
    di = ((2u * lw + (j >> 1)) << C.res_log2) + 2u * lu + (j & 1u);
  } else {
    if ((t & ~63ull) >= C.rays) return;  // uniform over the wavefront
    ok = t < C.rays;
    if (ok) split_ray(C, (uint32_t)t, di, pi);
  }
  float4 m = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  uint32_t steps = 0;
  if (ok) {
    i3 g = i3{a.G0[0] >> c, a.G0[1] >> c, a.G0[2] >> c} + probe_xyz(C, pi);
    const DirEntry& de = a.dirs[C.dir_off + di];
    f3 d = mk3(de.d[0], de.d[1], de.d[2]);
    f3 inv = mk3(de.inv[0], de.inv[1], de.inv[2]);
    Hit h;
    f3 ot;
    f3 op;
    trace_probe_ray<ST>(a, c, g, pi, d, inv, h, ot, op);
    float4 out = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    if (h.hit) {
      f3 rgb = shade_voxel(a, vox_rank(a.vx, h.brick, h.bit), d, ot + d * h.tIn, h.brick, h.bit);
      out = make_float4(rgb.x, rgb.y, rgb.z, 0.0f);
    }
    if (DY) out = with_dyn(a, c, op, d, C.b0 * a.s0, C.b1 * a.s0, h, ot, out);
    size_t ray = C.ray_off + (size_t)di * C.np + pi;
    if (a.hits) a.hits[ray] = h.id;
    steps = h.steps;
    if (a.no_merge) {
      m = rad_store(a, ray, out);
    } else {
      f3 um = c + 1 < a.N ? upper_mean(a, c, probe_xyz(C, pi), pi, di) : splat3(0.0f);
      m = rad_store(a, ray, merged(a, c, g, di, rad_round(a, out), um));
    }
  }
  if (ST && a.vx.stats) {
    // counts[24 + c], [32 + c], [64 + c] of traceWorld (src/core/rc.ts:232-235), summed per wavefront (every lane
    // of a wavefront that did not return reaches this point).
    unsigned long long st = steps, cap = steps >= 1024u ? 1ull : 0ull, n = ok ? 1ull : 0ull;
    for (int k = 32; k >= 1; k >>= 1) {
      st += __shfl_xor(st, k);
      cap += __shfl_xor(cap, k);
      n += __shfl_xor(n, k);
    }
    if (lane == 0) {
      unsigned long long* S = a.vx.stats;
      atomicAdd(&S[c], n);
      atomicAdd(&S[RTGI_MAX_CASCADES + c], st);
      if (cap) atomicAdd(&S[2 * RTGI_MAX_CASCADES + c], cap);
    }
  }
  if (SH && sh0) {
    sh0_block(a, s_sh[0], lb, lane & 15u, di, m.x, m.y, m.z);
    return;
  }
  if (a.no_merge || c == 0) return;
  float4* AV = reinterpret_cast<float4*>(a.avg) + C.avg_off;
  if (!quad) {  // resolution capped: the average of a one-texel block is the merged value
    if (ok) AV[(size_t)di * C.np + pi] = make_float4(m.x, m.y, m.z, 0.0f);
    return;
  }
  uint32_t l0 = lane & 15u;
  f3 p00 = mk3(__shfl(m.x, l0), __shfl(m.y, l0), __shfl(m.z, l0));
  f3 p10 = mk3(__shfl(m.x, l0 + 16), __shfl(m.y, l0 + 16), __shfl(m.z, l0 + 16));
  f3 p01 = mk3(__shfl(m.x, l0 + 32), __shfl(m.y, l0 + 32), __shfl(m.z, l0 + 32));
  f3 p11 = mk3(__shfl(m.x, l0 + 48), __shfl(m.y, l0 + 48), __shfl(m.z, l0 + 48));
  if (ok && j == 0) {
    f3 avg = (p00 + p10 + p01 + p11) * 0.25f;
    AV[(size_t)dl * C.np + pi] = make_float4(avg.x, avg.y, avg.z, 0.0f);
  }
}
// Eight wavefronts per SIMD (at most 64 VGPRs): the trace is latency-bound, and occupancy pays more than registers.
template <bool ST, bool DY, bool SH>
__global__ void __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(8, 8))) k_trace_merge(TraceK a, int c) {
  trace_merge_body<ST, DY, SH>(a, c, blockIdx.x, gridDim.x);
}
// Step of a batch: each world traces and merges its cascade N-1-step. Batches project the SH in their own launch
// (k_sh_b): with the projection, this kernel needs 80 VGPRs and spills at 8 wavefronts per SIMD.
// With a work queue (queue: the step's counters, given where the launch's worlds have about equal block counts),
// workgroups take their blocks from it (queue_block), else by batch_block (hardware order). The descriptors and the
// counters come as restrict parameters: the compiler then knows that the counters' atomics do not write the
// descriptors and keeps reading them with scalar loads (with both behind one struct, every descriptor field read
// after the atomic became a vector load and the trace spilled).
template <bool ST, bool DY>
__global__ void __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(8, 8)))
k_trace_merge_b(BatchK b, const TraceK* __restrict__ ws, uint32_t* __restrict__ queue) {
  const uint32_t blk = queue ? queue_block(queue) : batch_block(b);
  if (blk == ~0u) return;
  int w = batch_world(b, blk);
  const TraceK& a = ws[w];
  trace_merge_body<ST, DY, false>(a, a.N - 1 - b.arg, blk - b.start[w], b.start[w + 1] - b.start[w]);
}
// The same step without a work queue, one wavefront per workgroup: workgroup h runs quarter h & 3 of block h >> 2.
// A workgroup of four wavefronts holds its slots until its slowest wavefront ends, so that the next workgroup waits
// for four free slots on one CU; single wavefronts leave and enter one at a time (64 city worlds: c0 step 4.27 ->
// 3.95 ms; 32 mixed worlds' trace 3.85 -> 3.54 ms). With a queue each workgroup pays a device-scope atomic, and per
// wavefront that cost more than it saved (c1 step 3.30 -> 3.68 ms), so queued steps keep four wavefronts.
template <bool ST, bool DY>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(8, 8)))
k_trace_merge_b1(BatchK b, const TraceK* __restrict__ ws) {
  const uint32_t h = batch_block(b), blk = h >> 2;
  int w = batch_world(b, blk);
  const TraceK& a = ws[w];
  trace_merge_body<ST, DY, false>(a, a.N - 1 - b.arg, blk - b.start[w], b.start[w + 1] - b.start[w],
                                  (h & 3u) * 64u + threadIdx.x);
}

// ---- Gather: irradiance of the field at surface points (gatherWorld, src/core/rc.ts:871-927) ----

// evalSH (src/core/rc.ts:819-829) of c0 probe row q (grid order): coefficients times shBasis(n) in coefficient order,
// clamped at 0. The precision form's f16 coefficients are widened, as unpack2x16float does.
__device__ f3 eval_sh(const TraceK& a, uint32_t q, f3 n) {
  uint32_t np = a.c[0].np;
  float y[9];
  sh_basis3(n.x, n.y, n.z, y);
  f3 e = splat3(0.0f);
  for (int k = 0; k < 9; k++) {
    float v[3];
    for (int ch = 0; ch < 3; ch++) {
      size_t j = (size_t)(k * 3 + ch) * np + q;
      v[ch] = a.fp16 ? __half2float(__ushort_as_half(reinterpret_cast<const unsigned short*>(a.sh)[j])) : a.sh[j];
    }
    e = e + mk3(v[0], v[1], v[2]) * y[k];
  }
  return mk3(fmaxf(e.x, 0.0f), fmaxf(e.y, 0.0f), fmaxf(e.z, 0.0f));
}

// segOccluded (src/core/rc.ts:845-869): min(ceil(|b - a| / v), 16) samples at the segment's midpoints; true if one
// lies in an occupied voxel of level L (a dropped brick read one level coarser). Samples outside the window count
// as empty.
__device__ bool seg_occluded(const VoxK& k, f3 pa, f3 pb, int L) {
  f3 d = pb - pa;
  float v = vox_size(k, L);
  float iv = k.inv_v0 * inv_pow2(L);
  uint32_t n = (uint32_t)fminf(ceilf(sqrtf(dot3(d, d)) / v), 16.0f);
  f3 lo = tof3(origin(k, L)) * v * 8.0f;
  int nVox = k.G * 8;
  for (uint32_t i = 0; i < n; i++) {
    f3 p = pa + d * (((float)i + 0.5f) / (float)n);
    i3 c = cell_at(p - lo, v, iv, k.pow2 != 0);
    if (!inside3(c, nVox)) continue;
    uint32_t ptr = k.ptr[ptr_index(k, L, shr3(c, 3))];
    if (ptr == PTR_DROP) {
      VoxHit h;
      if (vox_at(k, L, c, h)) return true;
      continue;
    }
    if (ptr >= PTR_MARK) continue;
    uint32_t bit = bit_of(c);
    if ((k.occ[ptr * 16u + (bit >> 5)] & (1u << (bit & 31u))) != 0u) return true;
  }
  return false;
}

// The 8 c0 probes around x with their gather weights: trilinear weight + 1e-4 times ((side + 1)^2 / 4 + 0.02),
// side = cos(n, probe - x) (gatherWorld, rc.ts:886-895). Corner k: o = (k & 1, k >> 1 & 1, k >> 2 & 1); absent
// (outside the dense grid) -> false. q: the probe's SH row (grid order).
__device__ __forceinline__ bool gather_corner(const TraceK& a, const int32_t* G0, f3 x, f3 n, i3 b, f3 f, int k,
                                              f3& pr, float& w, uint32_t& q) {
  const CascK& C = a.c[0];
  i3 o{k & 1, (k >> 1) & 1, (k >> 2) & 1};
  i3 g = b + o;
  i3 gl = g - i3{G0[0], G0[1], G0[2]};
  if (!inside3(gl, C.n)) return false;
  f3 tw = mk3(o.x ? f.x : 1.0f - f.x, o.y ? f.y : 1.0f - f.y, o.z ? f.z : 1.0f - f.z);
  q = grid_row(C, gl);
  if (a.reloc) {
    float4 P = ppos4(a)[C.probe_off + probe_index(C, gl)];
    pr = mk3(P.x, P.y, P.z);
  } else {
    pr = tof3(g) * a.s0;
  }
  f3 dv = pr - x;
  float side = dot3(n, dv) / fmaxf(sqrtf(dot3(dv, dv)), 1e-5f);
  w = (tw.x * tw.y * tw.z + 1e-4f) * ((side + 1.0f) * (side + 1.0f) * 0.25f + 0.02f);
  return true;
}

// The irradiance of the field (c0 grid at G0) at x for normal nn.
__device__ f3 gather_point(const TraceK& a, const int32_t* G0, f3 x, f3 nn, bool occl) {
  f3 gp = x / a.s0;
  i3 b = f2i3(floor3(gp));
  f3 f = gp - tof3(b);
  // Occlusion tests start one voxel off the surface along the normal, at the level of the c0 traces.
  int Lo = vox_level_for(a.vx, a.c[0].vox_level, x);
  float vLo = vox_size(a.vx, Lo);
  f3 xo1 = x + nn * (1.0f * vLo);
  f3 e = splat3(0.0f);
  float ws = 0.0f;
  uint32_t hidden = 0;
  for (int k = 0; k < 8; k++) {
    f3 pr;
    float w;
    uint32_t q;
    if (!gather_corner(a, G0, x, nn, b, f, k, pr, w, q)) continue;
    if (occl && seg_occluded(a.vx, xo1, pr, Lo)) {
      hidden |= 1u << k;
      continue;
    }
    e = e + eval_sh(a, q, nn) * w;
    ws += w;
  }
  // Every probe hidden (the offset start fell into a neighbouring face): retest from two voxels out.
  if (ws <= 0.0f && hidden != 0u) {
    f3 xo2 = x + nn * (2.0f * vLo);
    for (int k = 0; k < 8; k++) {
      if (!((hidden >> k) & 1u)) continue;
      f3 pr;
      float w;
      uint32_t q;
      gather_corner(a, G0, x, nn, b, f, k, pr, w, q);
      if (seg_occluded(a.vx, xo2, pr, Lo)) continue;
      e = e + eval_sh(a, q, nn) * w;
      ws += w;
    }
  }
  return ws > 0.0f ? e / ws : splat3(0.0f);
}

__global__ void __launch_bounds__(BLK) k_gather(TraceK a, const float* pos, const float* nrm, uint32_t n, int occl,
                                                float* irr) {
  uint32_t i = blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  f3 x = mk3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
  f3 nn = mk3(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]);
  f3 out = gather_point(a, a.G0, x, nn, occl != 0);
  irr[3 * i] = out.x;
  irr[3 * i + 1] = out.y;
  irr[3 * i + 2] = out.z;
}

// ---- Voxel relocation (relocateVox, src/core/rc.ts:413-472) and parent visibility (linkParents, :559-566) ----

// voxCell (rc.ts:379-386): the occupied voxel of level L at p (a dropped brick read coarser), or false when the
// cell is empty or outside the window.
__device__ __forceinline__ bool vox_cell(const VoxK& k, int L, f3 p, VoxHit& h) {
  float v = vox_size(k, L);
  i3 c = cell_at(p - tof3(origin(k, L)) * v * 8.0f, v, k.inv_v0 * inv_pow2(L), k.pow2 != 0);
  if (!inside3(c, k.G * 8)) return false;
  return vox_at(k, L, c, h);
}
// voxStoredNormal (rc.ts:387-393): the oct-decoded stored normal (unnormalised); a brick without attribute slots
// (voxel-pool overflow, where the reference reads past its array) gives (0, 0, 1).
__device__ __forceinline__ f3 stored_normal(const VoxK& k, const VoxHit& h) {
  uint32_t ai = vox_rank(k, h.brick, h.bit);
  uint32_t pn = ai == ATTR_NONE ? 0u : k.attr[ai * 2u + 1u];
  float nx = snorm16_to_f(pn & 0xffffu), ny = snorm16_to_f(pn >> 16);
  f3 n = mk3(nx, ny, 1.0f - fabsf(nx) - fabsf(ny));
  float q = fmaxf(-n.z, 0.0f);
  n.x += n.x >= 0.0f ? -q : q;
  n.y += n.y >= 0.0f ? -q : q;
  return n;
}

// Probe (grid point g) of cascade c: the normal points away from the occupied voxels of the next level in the 4^3
// cube around the grid point (their stored normals decide when that is ambiguous); the probe then moves along it,
// in steps of max(v/2, 1.5 s/16) up to 1.5 s, to the first point whose voxel and the voxel lift * s behind it are
// empty. One thread per probe, every sum in the reference's serial order.
__device__ void relocate_probe(const TraceK& a, int c, uint32_t pi, float4& pos, float4& nrm) {
  const VoxK& k = a.vx;
  const CascK& C = a.c[c];
  i3 g = i3{a.G0[0] >> c, a.G0[1] >> c, a.G0[2] >> c} + probe_xyz(C, pi);
  float s = a.s0 * (float)(1u << c);
  f3 p = tof3(g) * s;
  int L = vox_level_for(k, C.vox_level, p);
  int Lg = vox_level_for(k, min(L + 1, k.NL - 1), p);
  float vg = vox_size(k, Lg);
  f3 grad = splat3(0.0f);
  float cnt = 0.0f;
  uint64_t occm = 0;
  for (int q = 0; q < 64; q++) {
    f3 off = (mk3((float)(q & 3), (float)((q >> 2) & 3), (float)(q >> 4)) - splat3(1.5f)) * vg;
    VoxHit h;
    if (!vox_cell(k, Lg, p + off, h)) continue;
    grad = grad - off;
    cnt += 1.0f;
    occm |= 1ull << q;
  }
  int flag = cnt < 1.0f ? 0 : (sqrtf(dot3(grad, grad)) < 0.1f * cnt * vg ? 1 : 2);
  f3 dir = grad;
  if (flag == 1) {
    f3 ns = splat3(0.0f);
    for (int q = 0; q < 64; q++) {
      if (!((occm >> q) & 1ull)) continue;
      f3 off = (mk3((float)(q & 3), (float)((q >> 2) & 3), (float)(q >> 4)) - splat3(1.5f)) * vg;
      VoxHit h;
      vox_cell(k, Lg, p + off, h);
      ns = ns + stored_normal(k, h);
    }
    dir = ns;
  }
  if (flag != 0) {
    if (sqrtf(dot3(dir, dir)) < 1e-4f)
      flag = 0;
    else
      dir = normalize3(dir);
  }
  if (flag == 0) {
    pos = make_float4(p.x, p.y, p.z, a.s0);
    nrm = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    return;
  }
  float back = a.lift * s;
  float stp = fmaxf(0.5f * vox_size(k, L), 1.5f * s / 16.0f);
  f3 r = p;
  for (int i = 0; i <= 16; i++) {
    float di = (float)i * stp;
    if (di > 1.5f * s) break;
    f3 q = p + dir * di;
    VoxHit h;
    if (!vox_cell(k, L, q, h) && !vox_cell(k, L, q - dir * back, h)) {
      r = q;
      break;
    }
  }
  pos = make_float4(r.x, r.y, r.z, a.s0);
  nrm = make_float4(dir.x, dir.y, dir.z, 0.0f);
}

// ---- Traced sun visibility of the occupied voxels (voxSun, src/core/voxel.ts:631-673) ----

// 1 if a shadow ray from x toward the sun leaves the clipmap without entering an occupied voxel: traced at level L,
// then from where it leaves each level's window on at the next coarser level; occupied voxels entered within tSkip
// of x are ignored (the voxel the point lies on).
__device__ float sun_ray(const TraceK& a, f3 x, f3 d, f3 inv, int L, float tSkip) {
  const VoxK& k = a.vx;
  float t = 0.0f;
  for (int l = L; l < k.NL; l++) {
    Hit h = trace_vox<false>(a, x, d, inv, t, 1e30f, tSkip, l);
    if (h.hit) return 0.0f;
    float v = vox_size(k, l);
    f3 lo = tof3(origin(k, l)) * v * 8.0f;
    f3 hi = lo + splat3((float)k.G * v * 8.0f);
    f3 ta = (lo - x) * inv, tb = (hi - x) * inv;
    f3 tf = max3(ta, tb);
    t = fmaxf(t, fminf(fminf(tf.x, tf.y), tf.z));
  }
  return 1.0f;
}

// The voxel at attribute slot ai (slot list written by the resolve), or false when the slot holds no voxel of this
// update (a slot past the voxel-pool limit keeps an older entry: its brick must be allocated, the bit occupied and
// the rank equal to the slot).
__device__ __forceinline__ bool slot_voxel(const VoxK& k, uint32_t ai, uint32_t nb, uint32_t& ptr, uint32_t& bit) {
  uint32_t e = k.vslot[ai];
  ptr = e >> 9;
  bit = e & 511u;
  if (ptr >= nb) return false;
  if ((k.occ[ptr * 16u + (bit >> 5)] & (1u << (bit & 31u))) == 0u) return false;
  return vox_rank(k, ptr, bit) == ai;
}

// The voxel's sun visibility: the mean of 2 x 2 shadow rays from points a quarter voxel along two tangents of the
// representative's plane, (0.25 voxel + 3 cm) off it on its sunward side, as voxSun places its shadow-map taps.
// Four lanes per occupied voxel (dense over the attribute slots), one tap each; the quad sums the taps in tap order.
__device__ __forceinline__ void sun_body(const TraceK& a, uint32_t lb, uint32_t nblk, uint32_t tx = threadIdx.x) {
  const VoxK& k = a.vx;
  const uint32_t nb = min(k.hdr[H_MARKED], (uint32_t)k.max_bricks);
  const uint32_t nv = min(k.hdr[H_VOXELS], (uint32_t)k.max_voxels);
  f3 sd = mk3(a.sun_dir[0], a.sun_dir[1], a.sun_dir[2]);
  f3 dd = mk3(fabsf(sd.x) < 1e-9f ? 1e-9f : sd.x, fabsf(sd.y) < 1e-9f ? 1e-9f : sd.y,
              fabsf(sd.z) < 1e-9f ? 1e-9f : sd.z);
  f3 inv = mk3(1.0f / dd.x, 1.0f / dd.y, 1.0f / dd.z);
  // The stride is a multiple of 4 and so is the bound: a voxel's four lanes are always in the loop together.
  for (uint64_t t = (uint64_t)lb * BLK + tx; t < (uint64_t)nv * 4u; t += (uint64_t)nblk * BLK) {
    const uint32_t ai = (uint32_t)(t >> 2), tap = (uint32_t)(t & 3u);
    uint32_t ptr, bit;
    if (!slot_voxel(k, ai, nb, ptr, bit)) continue;  // the same decision for the quad's four lanes
    uint32_t code = k.bcode[ptr];
    int L = (int)(code >> 27);
    i3 b{(int)((code >> 18) & 511u), (int)((code >> 9) & 511u), (int)(code & 511u)};
    i3 c = b * 8 + i3{(int)(bit & 7u), (int)((bit >> 3) & 7u), (int)(bit >> 6)};
    float v = vox_size(k, L);
    f3 pc = tof3(origin(k, L)) * v * 8.0f + (tof3(c) + splat3(0.5f)) * v;
    uint32_t pn = k.attr[ai * 2u + 1u];
    float ex = snorm16_to_f(pn & 0xffffu), ey = snorm16_to_f(pn >> 16);
    f3 n = mk3(ex, ey, 1.0f - fabsf(ex) - fabsf(ey));
    float q0 = fmaxf(-n.z, 0.0f);
    n.x += n.x >= 0.0f ? -q0 : q0;
    n.y += n.y >= 0.0f ? -q0 : q0;
    f3 q = pc;
    if (dot3(n, n) > 1e-4f) {
      n = normalize3(n);
      q = pc + n * k.plane[ai];
    } else {
      n = mk3(0.0f, 1.0f, 0.0f);
    }
    f3 ns = dot3(n, sd) < 0.0f ? -n : n;
    f3 t1 = normalize3(cross3(ns, fabsf(ns.y) > 0.9f ? mk3(1.0f, 0.0f, 0.0f) : mk3(0.0f, 1.0f, 0.0f)));
    f3 t2 = cross3(ns, t1);
    float h = 0.25f * v;
    f3 o = q + ns * (0.25f * v + 0.03f);
    f3 x = tap == 0u ? o + (t1 + t2) * h : tap == 1u ? o + (t1 - t2) * h : tap == 2u ? o - (t1 + t2) * h
                                                                                    : o - (t1 - t2) * h;
    float r = sun_ray(a, x, sd, inv, L, v);
    float r0 = __shfl(r, 0, 4), r1 = __shfl(r, 1, 4), r2 = __shfl(r, 2, 4), r3 = __shfl(r, 3, 4);
    if (tap == 0u) a.sunvis[ai] = 0.25f * (r0 + r1 + r2 + r3);
  }
}
__global__ void __launch_bounds__(BLK) k_sun(TraceK a) { sun_body(a, blockIdx.x, gridDim.x); }
// Batched: one wavefront per workgroup (workgroup h runs quarter h & 3 of block h >> 2), as the batched trace.
__global__ void __launch_bounds__(64) k_sun_b(BatchK b) {
  const uint32_t h = batch_block(b), blk = h >> 2;
  int w = batch_world(b, blk);
  sun_body(b.w[w], blk - b.start[w], b.start[w + 1] - b.start[w], (h & 3u) * 64u + threadIdx.x);
}

// ---- Bounce volume (src/core/bounce.ts): update (bnClassify / bnUpdate) and injection into the voxels (bnInject) ----

static_assert(BLK == 32 * (int)BN_PROBES_PER_BLOCK, "bounce volume blocks are 8 probes x 32 lanes");
static_assert(BN_RAYS + 6u <= 32u, "a probe's rays fit its 32 lanes");
static_assert(BN_PROBES % BN_PROBES_PER_BLOCK == 0u, "whole blocks of probes");

__device__ __forceinline__ uint32_t bn_probe(i3 c) { return (uint32_t)(c.x + BN_DX * (c.y + BN_DY * c.z)); }
__device__ __forceinline__ bool bn_inside(i3 c) {
  return c.x >= 0 && c.y >= 0 && c.z >= 0 && c.x < BN_DX && c.y < BN_DY && c.z < BN_DZ;
}
__device__ __forceinline__ f3 bn_rgb(uint2 t) { return mk3(hlo(t.x), hhi(t.x), hlo(t.y)); }
__device__ __forceinline__ float bn_reach(uint2 t) { return hhi(t.y); }

// WGSL smoothstep: t = clamp((x - lo) / (hi - lo), 0, 1); t * t * (3 - 2 t).
__device__ __forceinline__ float smooth_step(float lo, float hi, float x) {
  float t = fminf(fmaxf((x - lo) / (hi - lo), 0.0f), 1.0f);
  return t * t * (3.0f - 2.0f * t);
}

// bnRead (bounce.ts:49-75): irradiance on the side n of a surface at p, read at q from the volume vol (origin org):
// trilinear over the 8 probes around q, counting a probe when it is active, q lies within the reach it sees along
// its axes towards q (plus tol = half a level-0 voxel), and it lies on the side n of p (smoothstep weight).
__device__ f3 bn_read(const uint32_t* vol, const int32_t* org, float tol, f3 q, f3 p, f3 n) {
  const uint2* V = reinterpret_cast<const uint2*>(vol);
  i3 o{org[0], org[1], org[2]};
  f3 g = mk3(q.x / BN_CELL, q.y / BN_CELL, q.z / BN_CELL) - tof3(o);
  i3 bb = f2i3(floor3(g));
  f3 fr = g - tof3(bb);
  f3 n2 = n * n;
  f3 e = splat3(0.0f);
  float ws = 0.0f;
  for (int k = 0; k < 8; k++) {
    i3 off{k & 1, (k >> 1) & 1, (k >> 2) & 1};
    i3 cc = bb + off;
    if (!bn_inside(cc)) continue;
    f3 pp = tof3(o + cc) * BN_CELL;
    f3 u = q - pp;
    const uint2* t = V + (size_t)bn_probe(cc) * 6u;
    if (!(bn_reach(t[0]) > 0.0f)) continue;
    if (fabsf(u.x) > bn_reach(t[u.x >= 0.0f ? 0 : 1]) + tol || fabsf(u.y) > bn_reach(t[u.y >= 0.0f ? 2 : 3]) + tol ||
        fabsf(u.z) > bn_reach(t[u.z >= 0.0f ? 4 : 5]) + tol)
      continue;
    f3 t3 = mk3(off.x ? fr.x : 1.0f - fr.x, off.y ? fr.y : 1.0f - fr.y, off.z ? fr.z : 1.0f - fr.z);
    float w = t3.x * t3.y * t3.z * smooth_step(-0.05f, 0.25f, dot3(pp - p, n) / BN_CELL);
    if (w <= 0.0f) continue;
    f3 A = bn_rgb(t[n.x >= 0.0f ? 0 : 1]), B = bn_rgb(t[n.y >= 0.0f ? 2 : 3]), C = bn_rgb(t[n.z >= 0.0f ? 4 : 5]);
    f3 s = mk3(n2.x * A.x + n2.y * B.x + n2.z * C.x, n2.x * A.y + n2.y * B.y + n2.z * C.y,
               n2.x * A.z + n2.y * B.z + n2.z * C.z);
    e = e + s * w;
    ws += w;
  }
  return ws > 1e-6f ? mk3(e.x / ws, e.y / ws, e.z / ws) : splat3(0.0f);
}

// dynE (bounce.ts:267-298): the irradiance a dynamic voxel at p with the projected areas P (+x, -x, +y, -y, +z, -z)
// receives from the volume vol: per face a, the mean over the 8 probes around p that are active, see p within
// their reach (+ tol) and lie on that face's side of p (smoothstep weight) of the probe's texel a; the faces weighted
// by P.
__device__ f3 dyn_e(const uint32_t* vol, const int32_t* org, float tol, f3 p, const float P[6]) {
  const uint2* V = reinterpret_cast<const uint2*>(vol);
  i3 o{org[0], org[1], org[2]};
  f3 g = mk3(p.x / BN_CELL, p.y / BN_CELL, p.z / BN_CELL) - tof3(o);
  i3 bb = f2i3(floor3(g));
  f3 fr = g - tof3(bb);
  f3 e[6];
  float ws[6];
  for (int q = 0; q < 6; q++) {
    e[q] = splat3(0.0f);
    ws[q] = 0.0f;
  }
  for (int k = 0; k < 8; k++) {
    i3 off{k & 1, (k >> 1) & 1, (k >> 2) & 1};
    i3 cc = bb + off;
    if (!bn_inside(cc)) continue;
    f3 pp = tof3(o + cc) * BN_CELL;
    f3 u = p - pp;
    const uint2* t = V + (size_t)bn_probe(cc) * 6u;
    if (!(bn_reach(t[0]) > 0.0f)) continue;
    if (fabsf(u.x) > bn_reach(t[u.x >= 0.0f ? 0 : 1]) + tol || fabsf(u.y) > bn_reach(t[u.y >= 0.0f ? 2 : 3]) + tol ||
        fabsf(u.z) > bn_reach(t[u.z >= 0.0f ? 4 : 5]) + tol)
      continue;
    f3 t3 = mk3(off.x ? fr.x : 1.0f - fr.x, off.y ? fr.y : 1.0f - fr.y, off.z ? fr.z : 1.0f - fr.z);
    const float wt = t3.x * t3.y * t3.z;
    for (int q = 0; q < 6; q++) {
      if (P[q] <= 0.0f) continue;
      const float sg = (q & 1) ? -1.0f : 1.0f;
      f3 n = mk3(q / 2 == 0 ? sg : 0.0f, q / 2 == 1 ? sg : 0.0f, q / 2 == 2 ? sg : 0.0f);
      float w = wt * smooth_step(-0.05f, 0.25f, dot3(pp - p, n) / BN_CELL);
      if (w <= 0.0f) continue;
      e[q] = e[q] + bn_rgb(t[q]) * w;
      ws[q] += w;
    }
  }
  f3 s = splat3(0.0f);
  float ps = 0.0f;
  for (int q = 0; q < 6; q++) {
    if (ws[q] > 1e-6f) {
      f3 m = e[q] * P[q];
      s = s + mk3(m.x / ws[q], m.y / ws[q], m.z / ws[q]);
    }
    ps += P[q];
  }
  const float pm = fmaxf(ps, 1e-9f);
  return mk3(s.x / pm, s.y / pm, s.z / pm);
}

// dynInject (dynlayer.ts:166-190) with dynE (token obj): every voxel of the layer's allocated bricks with coverage
// adds its albedo / pi x the previous bounce volume's irradiance at its centre (the albedo words already hold the
// area weight) to its radiance words; level 0 in fixed point (round half to even), the coarser levels as floats.
// One thread per voxel: each voxel's words are its own.
__device__ __forceinline__ void dyn_inject_body(const TraceK& a, uint32_t lb, uint32_t nblk) {
  if (!a.objects) return;
  const VoxK& k = a.vx;
  const DynK& D = k.dyn;
  const uint32_t n = D.hdr[8], G = (uint32_t)k.G, G3 = (uint32_t)k.per;
  const float tol = 0.5f * k.v0;
  for (uint32_t i = lb * BLK + threadIdx.x; i < n * 512u; i += nblk * BLK) {
    const uint32_t slot = i >> 9, bit = i & 511u;
    const uint32_t pi = D.meta[slot], L = pi / G3, r = pi % G3;
    uint32_t* d = D.data + ((size_t)slot * 512u + bit) * DYN_WORDS;
    float wv[DYN_WORDS];
    for (uint32_t q = 0; q < DYN_WORDS; q++) wv[q] = L == 0u ? (float)d[q] / DYN_FIX : __uint_as_float(d[q]);
    if (wv[6] <= 0.0f) continue;
    i3 c = i3{(int)(r % G), (int)((r / G) % G), (int)(r / (G * G))} * 8 +
           i3{(int)(bit & 7u), (int)((bit >> 3) & 7u), (int)(bit >> 6)};
    const float v = k.v0 * (float)(1u << L);
    f3 o = mk3((float)k.origin[L][0], (float)k.origin[L][1], (float)k.origin[L][2]);
    f3 pc = o * v * 8.0f + (tof3(c) + splat3(0.5f)) * v;
    f3 E = dyn_e(a.bn_prev, a.bn_prev_org, tol, pc, wv);
    const float add[3] = {wv[10] / PI_F * E.x, wv[11] / PI_F * E.y, wv[12] / PI_F * E.z};
    for (uint32_t q = 0; q < 3; q++) {
      if (L == 0u)
        d[7u + q] += (uint32_t)fminf(rintf(add[q] * DYN_FIX), 4294967040.0f);  // clamped as the splat's words are
      else
        d[7u + q] = __float_as_uint(wv[7u + q] + add[q]);
    }
  }
}
__global__ void __launch_bounds__(BLK) k_dl_inject(TraceK a) { dyn_inject_body(a, blockIdx.x, gridDim.x); }
__global__ void __launch_bounds__(BLK) k_dl_inject_b(BatchK b) {
  const uint32_t blk = batch_block(b);
  int w = batch_world(b, blk);
  dyn_inject_body(b.w[w], blk - b.start[w], b.start[w + 1] - b.start[w]);
}

// A voxel hit of a bounce-volume ray (voxShade with the voxBounce of bounce.ts:96-99): the bounce term reads the
// previous update's volume at a point a quarter voxel back from the entry, on the side the ray comes from (voxSide,
// src/core/rc.ts:122-126); a voxel without a normal is read facing the ray. Otherwise the cascades' shading.
__device__ f3 shade_voxel_vol(const TraceK& a, uint32_t ai, f3 d, f3 entry, uint32_t brick, uint32_t bit) {
  uint32_t pa = 0u, pn = 0u;
  if (ai != ATTR_NONE) {
    pa = a.vx.attr[ai * 2u];
    pn = a.vx.attr[ai * 2u + 1u];
  }
  f3 alb = mk3(unorm8_to_f(pa & 255u), unorm8_to_f((pa >> 8) & 255u), unorm8_to_f((pa >> 16) & 255u));
  if (a.vx.alb16) {  // precision form: the bf16 albedo
    const uint2 q = ai != ATTR_NONE ? reinterpret_cast<const uint2*>(a.vx.alb16)[ai] : make_uint2(0u, 0u);
    alb = mk3(bf2f(q.x), bf2f(q.x >> 16), bf2f(q.y));
  }
  float nx = snorm16_to_f(pn & 0xffffu), ny = snorm16_to_f(pn >> 16);
  f3 n0 = mk3(nx, ny, 1.0f - fabsf(nx) - fabsf(ny));
  float q = fmaxf(-n0.z, 0.0f);
  n0.x += n0.x >= 0.0f ? -q : q;
  n0.y += n0.y >= 0.0f ? -q : q;
  f3 n = dot3(n0, n0) > 0.01f ? n0 : -d;
  if (dot3(n, d) > 0.0f) n = -n;
  f3 e = splat3(0.0f);
  if (ai != ATTR_NONE) {
    uint32_t code = a.vx.bcode[brick];
    int L = (int)(code >> 27);
    i3 bc{(int)((code >> 18) & 511u), (int)((code >> 9) & 511u), (int)(code & 511u)};
    i3 c = bc * 8 + i3{(int)(bit & 7u), (int)((bit >> 3) & 7u), (int)(bit >> 6)};
    float v = vox_size(a.vx, L);
    f3 pc = tof3(origin(a.vx, L)) * v * 8.0f + (tof3(c) + splat3(0.5f)) * v;
    f3 p = entry - d * (0.25f * v);
    f3 nb;
    if (dot3(n0, n0) <= 0.01f) {
      nb = -d;
    } else {
      float sd = dot3(p - pc, n0);
      float side = fabsf(sd) > 0.05f * v ? (sd > 0.0f ? 1.0f : -1.0f) : (dot3(n0, d) > 0.0f ? -1.0f : 1.0f);
      nb = n0 * side;
    }
    e = bn_read(a.bn_prev, a.bn_prev_org, 0.5f * a.vx.v0, p, p, nb);
  }
  f3 b = alb / PI_F * e;
  f3 sd = mk3(a.sun_dir[0], a.sun_dir[1], a.sun_dir[2]);
  float ndl = fmaxf(dot3(n, sd), 0.0f);
  f3 out = b;
  float vis = (a.sunvis && ai != ATTR_NONE) ? a.sunvis[ai] : a.sun_vis;
  if (ndl > 0.0f) out = b + alb / PI_F * mk3(a.sun_e[0], a.sun_e[1], a.sun_e[2]) * ndl * vis;
  for (int k = 0; k < a.n_point; k++) {
    f3 l = mk3(a.point_pos[k][0], a.point_pos[k][1], a.point_pos[k][2]) - entry;
    float r2 = dot3(l, l);
    float nl = fmaxf(dot3(n, l) / sqrtf(r2), 0.0f);
    if (nl > 0.0f) out = out + alb / PI_F * mk3(a.point_i[k][0], a.point_i[k][1], a.point_i[k][2]) * (nl / r2);
  }
  return out;
}

// bnTrace (bounce.ts:164-175): segments [b_k, b_k+1] * BN_CELL (b_k+1 = 4 b_k + 1) on voxel level k (or the first
// level holding the segment's end), nothing skipped, up to tMax. Returns rgb and the entry distance of the hit, or
// the sky and 1e9. shade = false leaves rgb 0 (the axis rays use only the distance).
__device__ float4 bn_trace(const TraceK& a, f3 o, f3 d, f3 inv, float tMax, bool shade) {
  const VoxK& k = a.vx;
  float b0 = 0.0f;
  for (int kk = 0; kk < k.NL && b0 * BN_CELL < tMax; kk++) {
    float b1 = b0 * 4.0f + 1.0f;
    int L = vox_level_for(k, kk, o + d * b1 * BN_CELL);
    Hit h = trace_vox<false>(a, o, d, inv, b0 * BN_CELL, fminf(b1 * BN_CELL, tMax), 0.0f, L);
    if (h.hit) {
      f3 entry = o + d * h.tIn;
      f3 rgb = shade ? shade_voxel_vol(a, vox_rank(k, h.brick, h.bit), d, entry, h.brick, h.bit) : splat3(0.0f);
      return make_float4(rgb.x, rgb.y, rgb.z, dot3(entry - o, d));
    }
    b0 = b1;
  }
  return make_float4(a.sky[0], a.sky[1], a.sky[2], 1e9f);
}

// bnClassify + bnUpdate (bounce.ts:117-144, 182-208), 8 probes per 256-thread block, 32 lanes per probe. A probe is
// active when an occupied voxel of level levelOf(BN_CELL) lies at one of the 8 points half a cell around it and
// none of level 0 at the probe itself; an inactive probe is cleared. An active probe whose turn it is not (every
// BN_DIV-th update by world position) keeps the previous volume's texels when that volume holds it active;
// otherwise lanes [0, BN_RAYS) trace the irradiance rays, lanes [BN_RAYS, BN_RAYS + 6) the axis rays (free reach,
// at most 1.5 cells), and lanes 0..5 form the ambient cube: axis a holds pi * sum(w r) / sum(w), w = max(a . dir, 0),
// summed in ray order. Texels are rounded to halves (RNE).
__device__ __forceinline__ void bnvol_body(const TraceK& a, uint32_t lb) {
  __shared__ float s_rad[BN_PROBES_PER_BLOCK][BN_RAYS][3];
  __shared__ float s_reach[BN_PROBES_PER_BLOCK][6];
  const VoxK& k = a.vx;
  const uint32_t g = threadIdx.x >> 5, li = threadIdx.x & 31u;
  const uint32_t i = lb * BN_PROBES_PER_BLOCK + g;
  i3 c{(int)(i % (uint32_t)BN_DX), (int)((i / (uint32_t)BN_DX) % (uint32_t)BN_DY), (int)(i / (uint32_t)(BN_DX * BN_DY))};
  i3 wp{a.bn_org[0] + c.x, a.bn_org[1] + c.y, a.bn_org[2] + c.z};
  f3 p = tof3(wp) * BN_CELL;
  bool occ = false;
  VoxHit vh;
  if (li < 8u) {
    f3 corner = mk3((float)(li & 1u), (float)((li >> 1) & 1u), (float)(li >> 2)) - splat3(0.5f);
    occ = vox_cell(k, a.bn_level, p + corner * BN_CELL, vh);
  } else if (li == 8u) {
    occ = vox_cell(k, 0, p, vh);
  }
  uint32_t gm = (uint32_t)(__ballot(occ) >> (threadIdx.x & 32u));
  int mode = ((gm & 0xffu) != 0u && ((gm >> 8) & 1u) == 0u) ? 2 : 0;  // 0 clear, 1 keep, 2 trace
  i3 pc{0, 0, 0};
  const uint2* prev = reinterpret_cast<const uint2*>(a.bn_prev);
  if (mode == 2 && BN_DIV > 1u && ((uint32_t)(wp.x + wp.y + wp.z + 1048576) + a.bn_frame) % BN_DIV != 0u) {
    pc = wp - i3{a.bn_prev_org[0], a.bn_prev_org[1], a.bn_prev_org[2]};
    if (bn_inside(pc) && bn_reach(prev[(size_t)bn_probe(pc) * 6u]) > 0.0f) mode = 1;
  }
  if (mode == 2) {
    if (li < BN_RAYS) {
      const DirEntry& de = a.dirs[a.bn_dir_off + li];
      const f3 dir = mk3(de.d[0], de.d[1], de.d[2]);
      float4 r = bn_trace(a, p, dir, mk3(de.inv[0], de.inv[1], de.inv[2]), 1e9f, true);
      if (a.objects) {  // bnTraceDyn (bounce.ts:173-178): the layer in front of the static result, one ray's cone
        float4 dl = trace_dyn(a, p, dir, 0.0f, fminf(r.w, 1e4f), sqrtf(4.0f * PI_F / (float)BN_RAYS));
        r.x = dl.x + (1.0f - dl.w) * r.x;
        r.y = dl.y + (1.0f - dl.w) * r.y;
        r.z = dl.z + (1.0f - dl.w) * r.z;
      }
      s_rad[g][li][0] = r.x;
      s_rad[g][li][1] = r.y;
      s_rad[g][li][2] = r.z;
    } else if (li < BN_RAYS + 6u) {
      uint32_t ax = li - BN_RAYS;
      float sg = (ax & 1u) ? -1.0f : 1.0f;
      f3 d = mk3(ax / 2u == 0u ? sg : 0.0f, ax / 2u == 1u ? sg : 0.0f, ax / 2u == 2u ? sg : 0.0f);
      f3 inv = mk3(1.0f / (d.x == 0.0f ? 1e-9f : d.x), 1.0f / (d.y == 0.0f ? 1e-9f : d.y),
                   1.0f / (d.z == 0.0f ? 1e-9f : d.z));
      float4 r = bn_trace(a, p, d, inv, 1.5f * BN_CELL, false);
      s_reach[g][ax] = fmaxf(fminf(r.w, 1.5f * BN_CELL), 1e-3f);
    }
  }
  __syncthreads();
  if (li < 6u) {
    uint2* out = reinterpret_cast<uint2*>(a.bn_cur) + (size_t)i * 6u + li;
    if (mode == 0) {
      *out = make_uint2(0u, 0u);
    } else if (mode == 1) {
      *out = prev[(size_t)bn_probe(pc) * 6u + li];
    } else {
      float sg = (li & 1u) ? -1.0f : 1.0f;
      f3 e = splat3(0.0f);
      float ws = 0.0f;
      for (uint32_t r = 0; r < BN_RAYS; r++) {
        float w = fmaxf(sg * a.dirs[a.bn_dir_off + r].d[li >> 1], 0.0f);
        e = e + mk3(s_rad[g][r][0], s_rad[g][r][1], s_rad[g][r][2]) * w;
        ws += w;
      }
      float m = fmaxf(ws, 1e-6f);
      *out = make_uint2(h2(PI_F * e.x / m, PI_F * e.y / m), h2(PI_F * e.z / m, s_reach[g][li]));
    }
  }
}
__global__ void __launch_bounds__(BLK) k_bnvol(TraceK a) { bnvol_body(a, blockIdx.x); }
__global__ void __launch_bounds__(BLK) k_bnvol_b(BatchK b) {
  const uint32_t blk = batch_block(b);
  int w = batch_world(b, blk);
  bnvol_body(b.w[w], blk - b.start[w]);
}

// bnInject (bounce.ts:224-262): every occupied voxel stores this update's volume irradiance on its two sides (along
// and against its normal; no normal: +y): level 0 read at the voxel centre, coarser levels half a voxel out towards
// the side. Four words per attribute slot: front rg, b, back rg, b (halves).
__device__ __forceinline__ void bounce_body(const TraceK& a, uint32_t lb, uint32_t nblk) {
  const VoxK& k = a.vx;
  const uint32_t nb = min(k.hdr[H_MARKED], (uint32_t)k.max_bricks);
  const uint32_t nv = min(k.hdr[H_VOXELS], (uint32_t)k.max_voxels);
  const float tol = 0.5f * k.v0;
  for (uint32_t ai = lb * BLK + threadIdx.x; ai < nv; ai += nblk * BLK) {  // dense over the attribute slots
    uint32_t ptr, bit;
    if (!slot_voxel(k, ai, nb, ptr, bit)) continue;
    uint32_t code = k.bcode[ptr];
    int L = (int)(code >> 27);
    i3 bc{(int)((code >> 18) & 511u), (int)((code >> 9) & 511u), (int)(code & 511u)};
    i3 c = bc * 8 + i3{(int)(bit & 7u), (int)((bit >> 3) & 7u), (int)(bit >> 6)};
    float v = vox_size(k, L);
    f3 p = tof3(origin(k, L)) * v * 8.0f + (tof3(c) + splat3(0.5f)) * v;
    uint32_t pn = k.attr[ai * 2u + 1u];
    float ex = snorm16_to_f(pn & 0xffffu), ey = snorm16_to_f(pn >> 16);
    f3 n = mk3(ex, ey, 1.0f - fabsf(ex) - fabsf(ey));
    float q0 = fmaxf(-n.z, 0.0f);
    n.x += n.x >= 0.0f ? -q0 : q0;
    n.y += n.y >= 0.0f ? -q0 : q0;
    f3 nf = dot3(n, n) > 0.01f ? normalize3(n) : mk3(0.0f, 1.0f, 0.0f);
    float off = L == 0 ? 0.0f : 0.5f * v;
    f3 ef = bn_read(a.bn_cur, a.bn_org, tol, p + nf * off, p, nf);
    f3 eb = bn_read(a.bn_cur, a.bn_org, tol, p + (-nf) * off, p, -nf);
    if (a.bf16)  // precision form: the injected light as bf16
      reinterpret_cast<uint4*>(a.bounce)[ai] = make_uint4(f2bf(ef.x) | (f2bf(ef.y) << 16), f2bf(ef.z) | (f2bf(0.0f) << 16),
                                                          f2bf(eb.x) | (f2bf(eb.y) << 16), f2bf(eb.z) | (f2bf(0.0f) << 16));
    else
      reinterpret_cast<uint4*>(a.bounce)[ai] = make_uint4(h2(ef.x, ef.y), h2(ef.z, 0.0f), h2(eb.x, eb.y), h2(eb.z, 0.0f));
  }
}
__global__ void __launch_bounds__(BLK) k_bounce(TraceK a) { bounce_body(a, blockIdx.x, gridDim.x); }
__global__ void __launch_bounds__(BLK) k_bounce_b(BatchK b) {
  const uint32_t blk = batch_block(b);
  int w = batch_world(b, blk);
  bounce_body(b.w[w], blk - b.start[w], b.start[w + 1] - b.start[w]);
}

__device__ __forceinline__ int probe_cascade(const TraceK& a, uint64_t t) {
  int c = 0;
  while (c + 1 < a.N && t >= a.c[c + 1].probe_off) c++;
  return c;
}

__device__ __forceinline__ void relocate_body(const TraceK& a, uint32_t lb, uint32_t nblk, uint32_t tx = threadIdx.x) {
  uint64_t t = (uint64_t)lb * BLK + tx;
  if (t >= a.total_probes) return;
  int c = probe_cascade(a, t);
  relocate_probe(a, c, (uint32_t)(t - a.c[c].probe_off), ppos4(a)[t], pnrm4(a)[t]);
}

// Parents of probe pi of cascade c < N-1 that exist and whose segment from the probe is free of occupied voxels at
// the probe's trace level (voxOcc samples of linkParents = segOccluded's).
__device__ __forceinline__ void link_body(const TraceK& a, uint32_t lb, uint32_t nblk, uint32_t tx = threadIdx.x) {
  uint64_t t = (uint64_t)lb * BLK + tx;
  if (a.N < 2 || t >= a.c[a.N - 1].probe_off) return;
  int c = probe_cascade(a, t);
  const CascK& C = a.c[c];
  const CascK& U = a.c[c + 1];
  uint32_t pi = (uint32_t)(t - C.probe_off);
  i3 g = probe_xyz(C, pi);
  float4 P = ppos4(a)[t];
  f3 pa = mk3(P.x, P.y, P.z);
  int Lv = vox_level_for(a.vx, C.vox_level, pa);
  uint32_t m = parent_mask(U, g), vis = 0;
  for (int kk = 0; kk < 8; kk++) {
    if (!((m >> kk) & 1u)) continue;
    float4 Q = ppos4(a)[U.probe_off + parent_index(U, g, kk)];
    if (!seg_occluded(a.vx, pa, mk3(Q.x, Q.y, Q.z), Lv)) vis |= 1u << kk;
  }
  a.pvis[t] = vis;
}
__global__ void __launch_bounds__(BLK) k_relocate(TraceK a) { relocate_body(a, blockIdx.x, gridDim.x); }
__global__ void __launch_bounds__(BLK) k_link(TraceK a) { link_body(a, blockIdx.x, gridDim.x); }
__global__ void __launch_bounds__(64) k_relocate_b(BatchK b) {
  const uint32_t h = batch_block(b), blk = h >> 2;
  int w = batch_world(b, blk);
  relocate_body(b.w[w], blk - b.start[w], b.start[w + 1] - b.start[w], (h & 3u) * 64u + threadIdx.x);
}
__global__ void __launch_bounds__(64) k_link_b(BatchK b) {
  const uint32_t h = batch_block(b), blk = h >> 2;
  int w = batch_world(b, blk);
  link_body(b.w[w], blk - b.start[w], b.start[w + 1] - b.start[w], (h & 3u) * 64u + threadIdx.x);
}

int grid_for(uint64_t n) {
  uint64_t g = (n + BLK - 1) / BLK;
  return (int)(g < 1 ? 1 : g);
}

}  // namespace

int launch_trace(const TraceK& k, void* stream) {
  const bool dy = k.vx.dyn.hdr != nullptr;
  if (k.total_rays) {
    const uint32_t g = grid_for(k.total_rays) * 4u;
    if (k.vx.stats)
      (dy ? k_trace<true, true> : k_trace<true, false>)<<<g, 64, 0, (hipStream_t)stream>>>(k);
    else
      (dy ? k_trace<false, true> : k_trace<false, false>)<<<g, 64, 0, (hipStream_t)stream>>>(k);
  }
  return (int)hipGetLastError();
}

uint64_t merge_threads(const TraceK& k, int c) {
  if (c == 0 || k.c[c].res == k.c[c - 1].res) return k.c[c].rays;
  return (uint64_t)k.c[c - 1].res * k.c[c - 1].res * ((k.c[c].np + 15u) >> 4) * 64u;
}

int launch_trace_rays(const TraceK& k, int c, const float* o, const float* d, uint32_t n, float* rgba,
                      uint32_t* hits, void* stream) {
  if (n) k_trace_rays<<<grid_for(n), BLK, 0, (hipStream_t)stream>>>(k, c, o, d, n, (float4*)rgba, hits);
  return (int)hipGetLastError();
}

uint64_t trace_merge_threads(const TraceK& k, int c) {
  if (k.no_merge) return k.c[c].rays;
  if (c == 0 && k.sh_fuse) return (uint64_t)((k.c[0].np + 15u) >> 4) * BLK;  // 16 probes x 16 directions per block
  return merge_threads(k, c);
}

using TraceMergeFn = void (*)(TraceK, int);
using TraceMergeBatchFn = void (*)(BatchK, const TraceK*, uint32_t*);
template <bool SH>
TraceMergeFn trace_merge_fn(bool st, bool dy) {
  return st ? (dy ? k_trace_merge<true, true, SH> : k_trace_merge<true, false, SH>)
            : (dy ? k_trace_merge<false, true, SH> : k_trace_merge<false, false, SH>);
}
TraceMergeBatchFn trace_merge_batch_fn(bool st, bool dy) {
  return st ? (dy ? k_trace_merge_b<true, true> : k_trace_merge_b<true, false>)
            : (dy ? k_trace_merge_b<false, true> : k_trace_merge_b<false, false>);
}
using TraceMergeBatch1Fn = void (*)(BatchK, const TraceK*);
TraceMergeBatch1Fn trace_merge_batch1_fn(bool st, bool dy) {
  return st ? (dy ? k_trace_merge_b1<true, true> : k_trace_merge_b1<true, false>)
            : (dy ? k_trace_merge_b1<false, true> : k_trace_merge_b1<false, false>);
}

int launch_trace_merge(const TraceK& k, int c, void* stream) {
  const bool dy = k.vx.dyn.hdr != nullptr;
  uint64_t n = trace_merge_threads(k, c);
  if (!n) return 0;
  const bool sh = c == 0 && k.sh_fuse && !k.no_merge;
  TraceMergeFn f = sh ? trace_merge_fn<true>(k.vx.stats, dy) : trace_merge_fn<false>(k.vx.stats, dy);
  f<<<grid_for(n), BLK, 0, (hipStream_t)stream>>>(k, c);
  return (int)hipGetLastError();
}

// Staging of the allocation-free gather: [pos 3 x max][nrm 3 x max][irr 3 x max] floats. Each GPU holds its own copy
// from the library's load to the end of the process; every call on a GPU shares it (calls that overlap on one GPU
// race). In host code `scratch` only names the symbol: kernels get the per-device address from hipGetSymbolAddress.
static __device__ __align__(256) float scratch[(size_t)RTGI_GATHER_SCRATCH_POINTS * 9];

int gather_scratch(int device, float** out) {
  static std::mutex mu;
  static float* ptr[RTGI_SCRATCH_MAX_DEVICES] = {};
  if (device < 0 || device >= RTGI_SCRATCH_MAX_DEVICES) return (int)hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(mu);
  if (!ptr[device]) {
    int cur = 0;
    hipError_t e = hipGetDevice(&cur);
    if (e == hipSuccess && cur != device) e = hipSetDevice(device);  // the symbol's address on this device
    float* p = nullptr;
    if (e == hipSuccess) e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    if (e != hipSuccess) return (int)e;
    ptr[device] = p;
  }
  *out = ptr[device];
  return 0;
}

int launch_gather(const TraceK& k, const float* pos, const float* nrm, uint32_t n, int occl, float* irr, void* stream) {
  if (n) k_gather<<<grid_for(n), BLK, 0, (hipStream_t)stream>>>(k, pos, nrm, n, occl, irr);
  return (int)hipGetLastError();
}

int launch_sun(const TraceK& k, void* stream) {
  if (!k.sunvis) return 0;
  uint32_t g[BL_COUNT];
  batch_blocks(k, g);
  k_sun<<<g[BL_SUN], BLK, 0, (hipStream_t)stream>>>(k);
  return (int)hipGetLastError();
}

int launch_bounce(const TraceK& k, void* stream) {
  if (!k.bounce) return 0;
  uint32_t g[BL_COUNT];
  batch_blocks(k, g);
  k_bnvol<<<g[BL_BNVOL], BLK, 0, (hipStream_t)stream>>>(k);
  k_bounce<<<g[BL_BOUNCE], BLK, 0, (hipStream_t)stream>>>(k);
  return (int)hipGetLastError();
}

int launch_dyn_inject(const TraceK& k, void* stream) {
  if (!k.objects) return 0;
  uint32_t g[BL_COUNT];
  batch_blocks(k, g);
  if (g[BL_DL_INJECT]) k_dl_inject<<<g[BL_DL_INJECT], BLK, 0, (hipStream_t)stream>>>(k);
  return (int)hipGetLastError();
}

int launch_relocate(const TraceK& k, void* stream) {
  if (!k.reloc || !k.total_probes) return 0;
  k_relocate<<<grid_for(k.total_probes), BLK, 0, (hipStream_t)stream>>>(k);
  if (k.N > 1) k_link<<<grid_for(k.c[k.N - 1].probe_off), BLK, 0, (hipStream_t)stream>>>(k);
  return (int)hipGetLastError();
}

int launch_merge(const TraceK& k, int c, void* stream) {
  if (c > 0) {
    k_merge_up<<<grid_for(merge_threads(k, c)), BLK, 0, (hipStream_t)stream>>>(k, c);
  } else if (k.c[0].rays) {
    k_merge0<<<grid_for(k.c[0].rays), BLK, 0, (hipStream_t)stream>>>(k);
  }
  return (int)hipGetLastError();
}

int launch_sh(const TraceK& k, void* stream) {
  (sh_pairs(k) ? k_sh_p : k_sh)<<<grid_for(k.c[0].np), BLK, 0, (hipStream_t)stream>>>(k);
  return (int)hipGetLastError();
}

static BatchK bk(const BatchPlan& p, int l, int arg = 0) { return BatchK{p.w, p.start[l], p.n, arg, p.xcd[l] ? 1 : 0}; }

int launch_sun_batch(const BatchPlan& p, void* stream) {
  if (p.grid[BL_SUN]) k_sun_b<<<p.grid[BL_SUN] * 4u, 64, 0, (hipStream_t)stream>>>(bk(p, BL_SUN));
  return (int)hipGetLastError();
}

int launch_bounce_batch(const BatchPlan& p, void* stream) {
  if (p.grid[BL_BNVOL]) k_bnvol_b<<<p.grid[BL_BNVOL], BLK, 0, (hipStream_t)stream>>>(bk(p, BL_BNVOL));
  if (p.grid[BL_BOUNCE]) k_bounce_b<<<p.grid[BL_BOUNCE], BLK, 0, (hipStream_t)stream>>>(bk(p, BL_BOUNCE));
  return (int)hipGetLastError();
}

int launch_dyn_inject_batch(const BatchPlan& p, void* stream) {
  if (p.grid[BL_DL_INJECT]) k_dl_inject_b<<<p.grid[BL_DL_INJECT], BLK, 0, (hipStream_t)stream>>>(bk(p, BL_DL_INJECT));
  return (int)hipGetLastError();
}

int launch_relocate_batch(const BatchPlan& p, void* stream) {
  if (p.grid[BL_RELOC]) k_relocate_b<<<p.grid[BL_RELOC] * 4u, 64, 0, (hipStream_t)stream>>>(bk(p, BL_RELOC));
  if (p.grid[BL_LINK]) k_link_b<<<p.grid[BL_LINK] * 4u, 64, 0, (hipStream_t)stream>>>(bk(p, BL_LINK));
  return (int)hipGetLastError();
}

int launch_trace_batch(const BatchPlan& p, void* stream) {
  if (p.grid[BL_TRACE]) {
    if (p.stats)
      (p.dyn ? k_trace_b<true, true> : k_trace_b<true, false>)<<<p.grid[BL_TRACE], BLK, 0, (hipStream_t)stream>>>(bk(p, BL_TRACE));
    else
      (p.dyn ? k_trace_b<false, true> : k_trace_b<false, false>)<<<p.grid[BL_TRACE], BLK, 0, (hipStream_t)stream>>>(bk(p, BL_TRACE));
  }
  return (int)hipGetLastError();
}

int launch_merge_batch(const BatchPlan& p, int step, void* stream) {
  if (p.grid[BL_MERGE + step])
    k_merge_step_b<<<p.grid[BL_MERGE + step], BLK, 0, (hipStream_t)stream>>>(bk(p, BL_MERGE + step, step));
  return (int)hipGetLastError();
}

int launch_trace_merge_batch(const BatchPlan& p, int step, void* stream) {
  const int l = BL_TM + step;
  // Only the coarse cascades (3 and up) take the queue: few blocks per world of long rays, whose cost varies most
  // between worlds. The finer steps run in the static order with one wavefront per workgroup, which there beats the
  // queue's atomic per workgroup (64 city worlds, static against queued: c0 3.91 against 4.27 ms in four-wavefront
  // workgroups, c1 3.01 against 3.30, c2 2.64 against 2.82; c3 2.34 against 2.01).
  uint32_t* queue = p.queue && p.xcd[l] && p.min_casc[step] >= 3 ? p.queue + VQ_SET * (size_t)(1 + step) : nullptr;
  if (p.grid[l] && queue)
    trace_merge_batch_fn(p.stats, p.dyn)<<<p.grid[l], BLK, 0, (hipStream_t)stream>>>(bk(p, l, step), p.w, queue);
  else if (p.grid[l])
    trace_merge_batch1_fn(p.stats, p.dyn)<<<p.grid[l] * 4u, 64, 0, (hipStream_t)stream>>>(bk(p, l, step), p.w);
  return (int)hipGetLastError();
}

int launch_sh_batch(const BatchPlan& p, void* stream) {
  if (p.grid[BL_SH]) (p.sh_pairs ? k_sh_pb : k_sh_b)<<<p.grid[BL_SH], BLK, 0, (hipStream_t)stream>>>(bk(p, BL_SH));
  return (int)hipGetLastError();
}
