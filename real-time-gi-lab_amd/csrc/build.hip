// Clipmap voxel rebuild on gfx950: every triangle is conservatively voxelised into an L-level camera-centred clipmap
// of 8^3-voxel bricks (the StrataBrick build of src/core/voxel.ts:205-620).
//
// Pipeline (one stream, no host round trip; device-side counts bound every grid-stride loop instead of indirect
// dispatch):
//   k_clear        header and pointer table                               (clear, voxel.ts:338-344)
//   k_mark         per triangle and level: overlapped bricks, marked in   (markBricks, voxel.ts:347-375)
//                  an LDS bitmap of all pointers and flushed with plain
//                  stores (no atomics); ranges of more than 8 bricks go
//                  to a list with one wavefront-aggregated atomic per wave
//   k_mark_big     one wavefront per listed (triangle, level), LDS bitmap (markBricksBig, voxel.ts:377-387)
//   k_mark_parents ancestors of marked bricks                             (markParents, voxel.ts:390-403)
//   k_alloc_bricks one workgroup, one scan: ordered allocation,         (countMarks/scanMarks/allocBricks,
//                  coarse levels first, finest dropped first               voxel.ts:405-470)
//   k_init_bricks  zero occupancy and representatives of the pool         (clearKeys, voxel.ts:471-474)
//   k_vox_small    per triangle and level: SAT per voxel (the wavefront's (voxBits, voxel.ts:476-512)
//                  cells laid end to end over its lanes); bit and
//                  triangle + 1 merged per workgroup in an LDS hash table,
//                  then one atomicOr / atomicMax per unique voxel
//   k_vox_big      listed (triangle, level) entries, up to 64 per         (voxBitsBig, voxel.ts:514-534)
//                  wavefront, their cells laid end to end over its lanes
//   k_reduce(L)    coarse level L from L-1, gathered per parent voxel     (reduce, voxel.ts:540-577)
//                  (single writer per word: no atomics)
//   k_alloc_voxels one workgroup, scan rounds: attribute slots by         (allocVoxels, voxel.ts:579-588)
//                  prefix sum in slot order
//   k_resolve      per occupied voxel (laid end to end over the lanes     (resolveAttr, voxel.ts:590-620)
//                  by popcount scans): albedo, oct normal, plane
//
// Device-scope atomics bypass the XCD-local L2 on this part, so the build avoids them where the result does not
// need them (marks, reduce); the voxel updates that remain return nothing, so no thread waits on them.
// Allocation is always the ordered one: the reference uses it when the pool overflows and an atomic counter
// otherwise (any order is valid then); a single order makes the raw layout deterministic as well as the canonical
// content, and voxel-pool overflow then drops the same bricks every run.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "rtgi_internal.h"
#include "batch.h"
#include "surface.h"
#include "vmath.h"

namespace {

constexpr int BLK = 256;

struct Tri {
  f3 a, b, c;
  uint32_t i0, i1, i2, m;
};

__device__ __forceinline__ f3 ldp(const float* p, uint32_t i) { return mk3(p[i * 3u], p[i * 3u + 1u], p[i * 3u + 2u]); }

__device__ __forceinline__ Tri load_tri(const VoxK& k, uint32_t t) {
  const uint4 q = reinterpret_cast<const uint4*>(k.tri)[t];
  Tri T;
  T.i0 = q.x;
  T.i1 = q.y;
  T.i2 = q.z;
  T.m = q.w;
  T.a = ldp(k.pos, q.x);
  T.b = ldp(k.pos, q.y);
  T.c = ldp(k.pos, q.z);
  return T;
}

__device__ __forceinline__ float vox_size(const VoxK& k, int L) { return k.v0 * (float)(1u << L); }
__device__ __forceinline__ i3 origin(const VoxK& k, int L) { return i3{k.origin[L][0], k.origin[L][1], k.origin[L][2]}; }
__device__ __forceinline__ f3 region_lo(const VoxK& k, int L) { return tof3(origin(k, L)) * vox_size(k, L) * 8.0f; }
__device__ __forceinline__ uint32_t ptr_index(const VoxK& k, int L, i3 b) {
  return (uint32_t)L * (uint32_t)k.per + ((uint32_t)b.z * (uint32_t)k.G + (uint32_t)b.y) * (uint32_t)k.G + (uint32_t)b.x;
}
// x / cell, as a multiplication when cell is a power of two (then the two are the same f32 value).
__device__ __forceinline__ float div_cell(float x, float cell, float inv, bool pow2) { return pow2 ? x * inv : x / cell; }

// Akenine-Moller triangle / box overlap (axisTest and triBox, src/core/voxel.ts:231-248).
__device__ __forceinline__ bool axis_test(f3 a, f3 v0, f3 v1, f3 v2, f3 h) {
  float p0 = dot3(v0, a), p1 = dot3(v1, a), p2 = dot3(v2, a);
  float r = dot3(h, abs3(a));
  return !(fminf(fminf(p0, p1), p2) > r || fmaxf(fmaxf(p0, p1), p2) < -r);
}
__device__ bool tri_box(f3 c, f3 h, f3 a, f3 b, f3 cc) {
  f3 v0 = a - c, v1 = b - c, v2 = cc - c;
  f3 mn = min3(min3(v0, v1), v2), mx = max3(max3(v0, v1), v2);
  if (mn.x > h.x || mn.y > h.y || mn.z > h.z || mx.x < -h.x || mx.y < -h.y || mx.z < -h.z) return false;
  f3 e0 = v1 - v0, e1 = v2 - v1, e2 = v0 - v2;
  f3 n = cross3(e0, e1);
  if (fabsf(dot3(n, v0)) > dot3(h, abs3(n))) return false;
  return axis_test(mk3(0.0f, -e0.z, e0.y), v0, v1, v2, h) && axis_test(mk3(0.0f, -e1.z, e1.y), v0, v1, v2, h) &&
         axis_test(mk3(0.0f, -e2.z, e2.y), v0, v1, v2, h) && axis_test(mk3(e0.z, 0.0f, -e0.x), v0, v1, v2, h) &&
         axis_test(mk3(e1.z, 0.0f, -e1.x), v0, v1, v2, h) && axis_test(mk3(e2.z, 0.0f, -e2.x), v0, v1, v2, h) &&
         axis_test(mk3(-e0.y, e0.x, 0.0f), v0, v1, v2, h) && axis_test(mk3(-e1.y, e1.x, 0.0f), v0, v1, v2, h) &&
         axis_test(mk3(-e2.y, e2.x, 0.0f), v0, v1, v2, h);
}

// Cell range of a triangle at level L, clipped to the window (voxRange, src/core/voxel.ts:293-300). brick: cells of
// 8 voxels.
struct Range {
  i3 lo, n;
};
__device__ __forceinline__ Range vox_range(const VoxK& k, const Tri& T, int L, bool brick) {
  float cell = brick ? vox_size(k, L) * 8.0f : vox_size(k, L);
  float inv = k.inv_v0 * inv_pow2(L) * (brick ? 0.125f : 1.0f);
  bool p2 = k.pow2 != 0;
  f3 lo = region_lo(k, L);
  f3 mn = min3(min3(T.a, T.b), T.c), mx = max3(max3(T.a, T.b), T.c);
  int lim = f2i(div_cell((float)k.G * vox_size(k, L) * 8.0f, cell, inv, p2));
  i3 zero{0, 0, 0}, l3{lim, lim, lim};
  f3 a0 = mn - lo, b0 = mx - lo;
  f3 af = mk3(div_cell(a0.x, cell, inv, p2), div_cell(a0.y, cell, inv, p2), div_cell(a0.z, cell, inv, p2));
  f3 bf = mk3(div_cell(b0.x, cell, inv, p2), div_cell(b0.y, cell, inv, p2), div_cell(b0.z, cell, inv, p2));
  i3 a = clampi3(f2i3(floor3(af)), zero, l3);
  i3 b = clampi3(f2i3(floor3(bf)) + 1, zero, l3);
  return Range{a, maxi3(b - a, zero)};
}

// Cell c of level L >= 1 lies inside level L-1's window (finerHolds, src/core/voxel.ts:303-307).
__device__ __forceinline__ bool finer_holds(const VoxK& k, int L, i3 c) {
  if (L == 0) return false;
  i3 f = (origin(k, L) * 8 + c) * 2 - origin(k, L - 1) * 8;
  return inside3(f, k.E);
}
__device__ __forceinline__ bool finer_holds_range(const VoxK& k, int L, const Range& r, int s) {
  return finer_holds(k, L, r.lo * s) && finer_holds(k, L, (r.lo + r.n) * s - 1);
}
// finerBuilt / finerBuiltRange (src/core/voxel.ts:311-320): held by the finer level unless its brick was dropped.
__device__ __forceinline__ bool finer_built(const VoxK& k, int L, i3 c, uint32_t drops) {
  if (!finer_holds(k, L, c)) return false;
  if (drops == 0u) return true;
  i3 f = (origin(k, L) * 8 + c) * 2 - origin(k, L - 1) * 8;
  return k.ptr[ptr_index(k, L - 1, shr3(f, 3))] != PTR_DROP;
}

__device__ __forceinline__ i3 cell_of(const Range& r, int q) {
  return i3{r.lo.x + q % r.n.x, r.lo.y + (q / r.n.x) % r.n.y, r.lo.z + q / (r.n.x * r.n.y)};
}
// cell_of for a small range (at most 8 cells, so q < 8 and every extent <= 8): q / n as (q m) >> 12 with
// m = 4096 / n + 1, which is exact for q < 64 and n <= 8 (checked exhaustively), in place of the divisions.
struct SmallDiv {
  uint32_t mx, my;
};
__device__ __forceinline__ SmallDiv small_div(const Range& r) {
  return SmallDiv{4096u / (uint32_t)r.n.x + 1u, 4096u / (uint32_t)r.n.y + 1u};
}
__device__ __forceinline__ i3 cell_of_small(const Range& r, SmallDiv m, int q) {
  const uint32_t t = ((uint32_t)q * m.mx) >> 12, z = (t * m.my) >> 12;
  return i3{r.lo.x + q - (int)t * r.n.x, r.lo.y + (int)(t - z * (uint32_t)r.n.y), r.lo.z + (int)z};
}

__device__ __forceinline__ uint32_t lane_below(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Wavefront-aggregated append: lanes with `want` get consecutive slots of list[] past *counter; one atomic per wave.
// The counter counts every append (it may pass cap); entries at or past cap are dropped and counted in *ovf.
__device__ __forceinline__ void wave_append(bool want, uint32_t* counter, uint32_t* list, uint32_t cap, uint32_t* ovf,
                                            uint32_t value) {
  uint64_t m = __ballot(want);
  if (m == 0) return;
  int leader = __ffsll((unsigned long long)m) - 1;
  uint32_t base = 0;
  if (__lane_id() == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
  base = __shfl(base, leader);
  uint32_t slot = base + lane_below(m);
  bool keep = want && slot < cap;
  if (keep) list[slot] = value;
  uint64_t lost = __ballot(want && !keep);
  if (lost && __lane_id() == leader) atomicAdd(ovf, (uint32_t)__popcll(lost));
}

// Mark of brick pointer i: a bit of the workgroup's LDS bitmap, or (no bitmap) the pointer itself.
__device__ __forceinline__ void mark_brick(const VoxK& k, uint32_t* bits, uint32_t i) {
  if (bits)
    atomicOr(&bits[i >> 5], 1u << (i & 31u));
  else if (k.ptr[i] == PTR_EMPTY)
    k.ptr[i] = PTR_MARK;  // every writer stores the same value, and nothing else writes the table in this phase
}

// markRange (src/core/voxel.ts:348-357); SMALL: a range of at most 8 bricks (a single brick needs no division).
template <bool SMALL>
__device__ void mark_range(const VoxK& k, uint32_t* bits, const Tri& T, int L, const Range& r, int start, int stride) {
  float bs = vox_size(k, L) * 8.0f;
  f3 lo = region_lo(k, L);
  int n = r.n.x * r.n.y * r.n.z;
  f3 h = splat3(0.5f * bs * 1.0001f);
  SmallDiv m{1u, 1u};
  if (SMALL && n > 1) m = small_div(r);
  for (int j = start; j < n; j += stride) {
    i3 b = SMALL ? (n == 1 ? r.lo : cell_of_small(r, m, j)) : cell_of(r, j);
    if (finer_holds(k, L, b * 8) && finer_holds(k, L, b * 8 + 7)) continue;
    if (n == 1 || tri_box(lo + (tof3(b) + splat3(0.5f)) * bs, h, T.a, T.b, T.c)) mark_brick(k, bits, ptr_index(k, L, b));
  }
}

extern __shared__ uint32_t lds_bits[];

__device__ __forceinline__ void bits_clear(const VoxK& k) {
  uint32_t words = ((uint32_t)(k.NL * k.per) + 31u) >> 5;
  for (uint32_t i = threadIdx.x; i < words; i += blockDim.x) lds_bits[i] = 0u;
  __syncthreads();
}
// Marked pointers of the bitmap become PTR_MARK with plain stores (idempotent across workgroups).
__device__ __forceinline__ void bits_flush(const VoxK& k) {
  __syncthreads();
  uint32_t words = ((uint32_t)(k.NL * k.per) + 31u) >> 5;
  for (uint32_t i = threadIdx.x; i < words; i += blockDim.x) {
    uint32_t w = lds_bits[i];
    while (w) {
      uint32_t b = (uint32_t)__builtin_ctz(w);
      w &= w - 1u;
      k.ptr[i * 32u + b] = PTR_MARK;
    }
  }
}

__device__ __forceinline__ void clear_body(const VoxK& k, uint32_t lb, uint32_t nblk) {
  uint32_t i = lb * BLK + threadIdx.x;
  if (i < HDR_WORDS) k.hdr[i] = 0u;
  if (k.stats && i < 3u * RTGI_MAX_CASCADES) k.stats[i] = 0ull;
  if (i < (uint32_t)(k.NL * k.per)) k.ptr[i] = PTR_EMPTY;
}
// posedVertices (src/core/scenes.ts:80-102), thread per vertex of the objects' span: its object by binary search of
// the ascending firsts (vertices between objects are left alone), x = rest - pivot in double, the turn about +y and
// the offset, each value rounded to float once. The reference evaluates the same expressions in JavaScript doubles
// on the host; the same IEEE double operations in the same order (no contraction) give the same floats.
__device__ __forceinline__ void pose_body(const VoxK& k, uint32_t lb) {
  const uint32_t i = lb * BLK + threadIdx.x;
  if (i >= k.pose_span) return;
  const uint32_t v = k.pose_first + i;
  uint32_t lo = 0, hi = k.npose - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (k.pose[mid].first <= v)
      lo = mid;
    else
      hi = mid - 1;
  }
  const PoseRec& o = k.pose[lo];
  if (v < o.first || v - o.first >= o.count) return;
  const float* rp = k.rest + (size_t)i * 3;
  const float* rn = k.rest + (size_t)k.pose_span * 3 + (size_t)i * 3;
  const double x = (double)rp[0] - o.pivot[0], y = (double)rp[1] - o.pivot[1], z = (double)rp[2] - o.pivot[2];
  const double nx = rn[0], ny = rn[1], nz = rn[2];
  float* P = const_cast<float*>(k.pos) + (size_t)v * 3;
  float* N = const_cast<float*>(k.nrm) + (size_t)v * 3;
  P[0] = (float)(o.pivot[0] + o.c * x + o.s * z + o.off[0]);
  P[1] = (float)(o.pivot[1] + y + o.off[1]);
  P[2] = (float)(o.pivot[2] - o.s * x + o.c * z + o.off[2]);
  N[0] = (float)(o.c * nx + o.s * nz);
  N[1] = (float)ny;
  N[2] = (float)(-o.s * nx + o.c * nz);
}
__global__ void __launch_bounds__(BLK) k_pose(VoxK k) { pose_body(k, blockIdx.x); }
__global__ void __launch_bounds__(BLK) k_pose_b(BatchK b) {
  const uint32_t blk = batch_block(b);
  int w = batch_world(b, blk);
  pose_body(b.w[w].vx, blk - b.start[w]);
}

__global__ void __launch_bounds__(BLK) k_clear(VoxK k) { clear_body(k, blockIdx.x, gridDim.x); }
__global__ void __launch_bounds__(BLK) k_clear_b(BatchK b) {
  const uint32_t blk = batch_block(b);
  int w = batch_world(b, blk);
  clear_body(b.w[w].vx, blk - b.start[w], b.start[w + 1] - b.start[w]);
  // the batch's work queues (the build's and each trace+merge step's)
  if (b.queue && blockIdx.x == 0 && threadIdx.x < VQ_LANES * (1 + RTGI_MAX_CASCADES)) b.queue[threadIdx.x * VQ_STRIDE] = 0u;
}

// markBricks over all triangles, grid-stride (every lane of a workgroup runs the same trip count, so the wave-level
// list append sees converged waves).
__device__ __forceinline__ void mark_body(const VoxK& k, uint32_t lb, uint32_t nblk) {
  uint32_t* bits = k.lds_marks ? lds_bits : nullptr;
  if (bits) bits_clear(k);
  for (uint32_t base = lb * BLK; base < k.ntris; base += nblk * BLK) {
    uint32_t t = base + threadIdx.x;
    bool valid = t < k.ntris;
    Tri T{};
    if (valid) T = load_tri(k, t);
    for (int L = 0; L < k.NL; L++) {
      bool work = false, big = false;
      Range r{};
      if (valid) {
        r = vox_range(k, T, L, true);
        int n = r.n.x * r.n.y * r.n.z;
        if (n != 0 && !finer_holds_range(k, L, r, 8)) {
          big = n > 8;
          work = !big;
        }
      }
      wave_append(big, &k.hdr[H_BIGMARK_L + L], k.big_mark + (size_t)L * k.big_cap, k.big_cap, &k.hdr[H_BIG_OVF],
                  t);
      if (work) mark_range<true>(k, bits, T, L, r, 0, 1);
    }
  }
  if (bits) bits_flush(k);
}
__global__ void __launch_bounds__(BLK) k_mark(VoxK k) { mark_body(k, blockIdx.x, gridDim.x); }
__global__ void __launch_bounds__(BLK) k_mark_b(BatchK b) {
  const uint32_t blk = batch_block(b);
  int w = batch_world(b, blk);
  mark_body(b.w[w].vx, blk - b.start[w], b.start[w + 1] - b.start[w]);
}

// One wavefront per listed (triangle, level); entries of all levels are taken in level order.
__device__ __forceinline__ bool big_entry(const VoxK& k, uint32_t e, uint32_t hdr_base, int& L, uint32_t& idx) {
  for (L = 0; L < k.NL; L++) {
    uint32_t n = min(k.hdr[hdr_base + L], k.big_cap);
    if (e < n) {
      idx = e;
      return true;
    }
    e -= n;
  }
  return false;
}

__device__ __forceinline__ void mark_big_body(const VoxK& k, uint32_t lb, uint32_t nblk) {
  uint32_t total = 0;
  for (int L = 0; L < k.NL; L++) total += min(k.hdr[H_BIGMARK_L + L], k.big_cap);
  if (total == 0) return;  // uniform across the grid
  uint32_t* bits = k.lds_marks ? lds_bits : nullptr;
  if (bits) bits_clear(k);
  uint32_t wave = (lb * BLK + threadIdx.x) >> 6, lane = threadIdx.x & 63u;
  uint32_t waves = (nblk * BLK) >> 6;
  for (uint32_t e = wave; e < total; e += waves) {
    int L;
    uint32_t idx;
    big_entry(k, e, H_BIGMARK_L, L, idx);
    uint32_t t = k.big_mark[(size_t)L * k.big_cap + idx];
    Tri T = load_tri(k, t);
    mark_range<false>(k, bits, T, L, vox_range(k, T, L, true), (int)lane, 64);
  }
  if (bits) bits_flush(k);
}
__global__ void __launch_bounds__(BLK) k_mark_big(VoxK k) { mark_big_body(k, blockIdx.x, gridDim.x); }
__global__ void __launch_bounds__(BLK) k_mark_big_b(BatchK b) {
  const uint32_t blk = batch_block(b);
  int w = batch_world(b, blk);
  mark_big_body(b.w[w].vx, blk - b.start[w], b.start[w + 1] - b.start[w]);
}

__device__ __forceinline__ void mark_parents_body(const VoxK& k, uint32_t lb, uint32_t nblk) {
  uint32_t i = lb * BLK + threadIdx.x;
  uint32_t per = (uint32_t)k.per;
  if (i >= (uint32_t)(k.NL - 1) * per) return;
  if (k.ptr[i] != PTR_MARK) return;
  int L = (int)(i / per);
  uint32_t r = i % per, G = (uint32_t)k.G;
  i3 w = origin(k, L) + i3{(int)(r % G), (int)((r / G) % G), (int)(r / (G * G))};
  for (L++; L < k.NL; L++) {
    w = shr3(w, 1);
    i3 b = w - origin(k, L);
    if (!inside3(b, k.G)) return;
    uint32_t j = ptr_index(k, L, b);
    if (k.ptr[j] == PTR_EMPTY) k.ptr[j] = PTR_MARK;
  }
}
__global__ void __launch_bounds__(BLK) k_mark_parents(VoxK k) { mark_parents_body(k, blockIdx.x, gridDim.x); }
__global__ void __launch_bounds__(BLK) k_mark_parents_b(BatchK b) {
  const uint32_t blk = batch_block(b);
  int w = batch_world(b, blk);
  mark_parents_body(b.w[w].vx, blk - b.start[w], b.start[w + 1] - b.start[w]);
}

// Block-wide exclusive sum of one value per thread (1024 threads), thread order; *total gets the sum. Every thread
// must call it.
__device__ __forceinline__ uint32_t block_excl_sum(uint32_t x, uint32_t* wsum, uint32_t* total) {
  uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
  uint32_t v = x;
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(v, o);
    if (lane >= (uint32_t)o) v += y;
  }
  if (lane == 63u) wsum[wv] = v;
  __syncthreads();
  if (threadIdx.x < 64) {
    uint32_t nw = blockDim.x >> 6;
    uint32_t s = threadIdx.x < nw ? wsum[threadIdx.x] : 0u, u = s;
    for (int o = 1; o < 64; o <<= 1) {
      uint32_t y = __shfl_up(u, o);
      if (threadIdx.x >= (uint32_t)o) u += y;
    }
    if (threadIdx.x < nw) wsum[threadIdx.x] = u - s;
    if (threadIdx.x == nw - 1) wsum[16] = u;
  }
  __syncthreads();
  uint32_t r = wsum[wv] + v - x;
  *total = wsum[16];
  __syncthreads();
  return r;
}

// Ordered brick allocation: order o = 0.. walks level NL-1 first, pointer index ascending within a level
// (scanMarks, src/core/voxel.ts:428-448); ranks past max_bricks are dropped (allocBricks, voxel.ts:450-470).
// Each thread owns 32 consecutive orders (one level, since G^3 is a multiple of 64), read with 8 vector loads in
// flight; one block scan of the per-thread counts gives every rank. Chunks of 32768 orders.
__device__ __forceinline__ void alloc_bricks_body(const VoxK& k, uint32_t lb, uint32_t nblk) {
  __shared__ uint32_t wsum[32];
  __shared__ uint32_t lvl[RTGI_MAX_LEVELS];
  if (threadIdx.x < RTGI_MAX_LEVELS) lvl[threadIdx.x] = 0u;
  __syncthreads();
  uint32_t per = (uint32_t)k.per, NL = (uint32_t)k.NL, total = NL * per, G = (uint32_t)k.G;
  uint32_t base = 0, drops = 0;
  for (uint32_t c0 = 0; c0 < total; c0 += 32u * blockDim.x) {
    uint32_t o0 = c0 + 32u * threadIdx.x;
    uint32_t mask = 0, L = 0, r0 = 0;
    if (o0 < total) {
      L = NL - 1u - o0 / per;
      r0 = o0 % per;
      const uint4* src = reinterpret_cast<const uint4*>(k.ptr + (size_t)L * per + r0);
      uint4 q[8];
      for (int j = 0; j < 8; j++) q[j] = src[j];
      for (int j = 0; j < 8; j++) {
        mask |= (q[j].x == PTR_MARK ? 1u : 0u) << (4 * j);
        mask |= (q[j].y == PTR_MARK ? 1u : 0u) << (4 * j + 1);
        mask |= (q[j].z == PTR_MARK ? 1u : 0u) << (4 * j + 2);
        mask |= (q[j].w == PTR_MARK ? 1u : 0u) << (4 * j + 3);
      }
    }
    uint32_t cnt = (uint32_t)__popc(mask), round;
    uint32_t rank = base + block_excl_sum(cnt, wsum, &round);
    uint32_t alloc = 0;
    while (mask) {
      uint32_t j = (uint32_t)__builtin_ctz(mask);
      mask &= mask - 1u;
      uint32_t r = r0 + j, i = L * per + r;
      if (rank >= (uint32_t)k.max_bricks) {
        k.ptr[i] = PTR_DROP;
        drops++;
      } else {
        k.ptr[i] = rank;
        k.bcode[rank] = (L << 27) | ((r % G) << 18) | (((r / G) % G) << 9) | (r / (G * G));
        alloc++;
      }
      rank++;
    }
    if (alloc) atomicAdd(&lvl[L], alloc);
    base += round;
  }
  if (drops) atomicAdd(&k.hdr[H_BRICK_OVF], drops);
  __syncthreads();
  if (threadIdx.x == 0) k.hdr[H_MARKED] = base;
  if (threadIdx.x < NL) k.hdr[H_BRICKS_L + threadIdx.x] = lvl[threadIdx.x];
}
__global__ void __launch_bounds__(1024) k_alloc_bricks(VoxK k) { alloc_bricks_body(k, blockIdx.x, gridDim.x); }
__global__ void __launch_bounds__(1024) k_alloc_bricks_b(BatchK b) {
  const uint32_t blk = batch_block(b);
  int w = batch_world(b, blk);
  alloc_bricks_body(b.w[w].vx, blk - b.start[w], b.start[w + 1] - b.start[w]);
}

__device__ __forceinline__ uint32_t n_bricks(const VoxK& k) { return min(k.hdr[H_MARKED], (uint32_t)k.max_bricks); }

__device__ __forceinline__ void init_bricks_body(const VoxK& k, uint32_t lb, uint32_t nblk) {
  uint32_t nb = n_bricks(k);
  uint32_t n = nb * 512u;
  for (uint32_t i = lb * BLK + threadIdx.x; i < n; i += nblk * BLK) {
    k.rep[i] = 0u;
    if ((i & 511u) < 16u) k.occ[(i >> 9) * 16u + (i & 511u)] = 0u;
  }
}
__global__ void __launch_bounds__(BLK) k_init_bricks(VoxK k) { init_bricks_body(k, blockIdx.x, gridDim.x); }
__global__ void __launch_bounds__(BLK) k_init_bricks_b(BatchK b) {
  const uint32_t blk = batch_block(b);
  int w = batch_world(b, blk);
  init_bricks_body(b.w[w].vx, blk - b.start[w], b.start[w + 1] - b.start[w]);
}

// Closest point of a triangle, barycentrics (Ericson 5.1.5; closestBary, src/core/voxel.ts:250-267).
__device__ f3 closest_bary(f3 p, f3 a, f3 b, f3 c) {
  f3 ab = b - a, ac = c - a, ap = p - a;
  float d1 = dot3(ab, ap), d2 = dot3(ac, ap);
  if (d1 <= 0.0f && d2 <= 0.0f) return mk3(1.0f, 0.0f, 0.0f);
  f3 bp = p - b;
  float d3 = dot3(ab, bp), d4 = dot3(ac, bp);
  if (d3 >= 0.0f && d4 <= d3) return mk3(0.0f, 1.0f, 0.0f);
  float vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
    float w = d1 / (d1 - d3);
    return mk3(1.0f - w, w, 0.0f);
  }
  f3 cp = p - c;
  float d5 = dot3(ab, cp), d6 = dot3(ac, cp);
  if (d6 >= 0.0f && d5 <= d6) return mk3(0.0f, 0.0f, 1.0f);
  float vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
    float w = d2 / (d2 - d6);
    return mk3(1.0f - w, 0.0f, w);
  }
  float va = d3 * d6 - d5 * d4;
  if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
    float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    return mk3(0.0f, 1.0f - w, w);
  }
  float dn = 1.0f / (va + vb + vc);
  float v = vb * dn, w = vc * dn;
  return mk3(1.0f - v - w, v, w);
}

// surfaceAt without wantAttr (src/core/voxel.ts:325-333), the alpha test of voxelizeOne (voxel.ts:485): a voxel of
// an alpha-masked material is set only where the texture x baseColorFactor alpha, sampled at the UV of the closest
// point of triangle t to the voxel centre p, is not below the material's cutoff. Untextured worlds and unmasked
// materials pass without a sample.
__device__ __forceinline__ bool surf_pass(const VoxK& k, uint32_t t, f3 p, f3 A, f3 B, f3 Cc, int L) {
  if (!k.surf.tex) return true;
  const uint4 q = reinterpret_cast<const uint4*>(k.tri)[t];
  if (!(surf_flags(k.surf, q.w) & RTGI_SURF_ALPHA_MASK)) return true;
  float u, v;
  surf_uv(k.surf, q.x, q.y, q.z, closest_bary(p, A, B, Cc), u, v);
  return !(surf_base_color(k.surf, k.mat, q.w, u, v, L).w < surf_cutoff(k.surf, q.w));
}

// Voxel updates of a workgroup are first merged in an LDS hash table keyed by (brick slot, bit): each unique voxel
// then costs one global atomicMax of its representative and one atomicOr of its bit, however many of the
// workgroup's triangles touch it (consecutive triangles are spatially coherent; on coarse levels dozens share a
// voxel). A full table falls back to direct global atomics for the update that found no slot.
constexpr uint32_t VH_SIZE = 2048;  // entries (16 KiB of LDS)
constexpr int VH_PROBES = 16;

__device__ __forceinline__ void voxel_update(const VoxK& k, uint32_t* hkey, uint32_t* hval, uint32_t ptr, uint32_t bit,
                                             uint32_t t) {
  uint32_t key = ptr * 512u + bit + 1u;  // 0 marks an empty slot
  uint32_t h = (key * 2654435761u) >> (32 - 11);
  for (int p = 0; p < VH_PROBES; p++) {
    uint32_t cur = atomicCAS(&hkey[h], 0u, key);
    if (cur == 0u || cur == key) {
      atomicMax(&hval[h], t + 1u);
      return;
    }
    h = (h + 1u) & (VH_SIZE - 1u);
  }
  atomicOr(&k.occ[ptr * 16u + (bit >> 5)], 1u << (bit & 31u));
  atomicMax(&k.rep[(size_t)ptr * 512u + bit], t + 1u);
}

__device__ __forceinline__ void voxel_table_flush(const VoxK& k, uint32_t* hkey, uint32_t* hval) {
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < VH_SIZE; i += blockDim.x) {
    uint32_t key = hkey[i];
    if (key == 0u) continue;
    uint32_t v = key - 1u, ptr = v >> 9, bit = v & 511u;
    atomicOr(&k.occ[ptr * 16u + (bit >> 5)], 1u << (bit & 31u));
    atomicMax(&k.rep[(size_t)ptr * 512u + bit], hval[i]);
  }
}

// voxelizeOne (src/core/voxel.ts:476-490) of cell q of the small range r (n cells) of triangle t at level L.
// TX: some world of the launch has textured surfaces (a separate instantiation: the alpha test costs registers).
template <bool TX>
__device__ __forceinline__ void voxelize_cell(const VoxK& k, uint32_t* hkey, uint32_t* hval, f3 A, f3 B, f3 Cc,
                                              uint32_t t, int L, const Range& r, SmallDiv m, int q, int n,
                                              uint32_t drops) {
  i3 c = n == 1 ? r.lo : cell_of_small(r, m, q);
  if (finer_built(k, L, c, drops)) return;
  float v = vox_size(k, L);
  const f3 p = region_lo(k, L) + (tof3(c) + splat3(0.5f)) * v;
  if (n != 1 && !tri_box(p, splat3(0.5f * v * 1.0001f), A, B, Cc)) return;
  uint32_t ptr = k.ptr[ptr_index(k, L, shr3(c, 3))];
  if (ptr >= PTR_MARK) return;
  if (TX && !surf_pass(k, t, p, A, B, Cc, L)) return;
  voxel_update(k, hkey, hval, ptr, (uint32_t)((c.x & 7) + 8 * (c.y & 7) + 64 * (c.z & 7)), t);
}

// voxBits (src/core/voxel.ts:511-513) over VOX_CHUNK consecutive triangles per workgroup.
constexpr int VOX_CHUNK = 4;
// Batches of at least this many worlds fill the GPU on their own: their voxel passes pack work into fewer, fuller
// wavefronts.
constexpr int WIDE_BATCH = 8;
template <bool TX>
__device__ __forceinline__ void vox_small_body(const VoxK& k, uint32_t lb, uint32_t nblk) {
  __shared__ uint32_t hkey[VH_SIZE];
  __shared__ uint32_t hval[VH_SIZE];
  for (uint32_t i = threadIdx.x; i < VH_SIZE; i += BLK) {
    hkey[i] = 0u;
    hval[i] = 0u;
  }
  __syncthreads();
  uint32_t drops = k.hdr[H_BRICK_OVF];
  for (int sub = 0; sub < VOX_CHUNK; sub++) {
    uint32_t t = (lb * VOX_CHUNK + sub) * BLK + threadIdx.x;
    bool valid = t < k.ntris;
    Tri T{};
    if (valid) T = load_tri(k, t);
    for (int L = 0; L < k.NL; L++) {
      bool work = false, big = false;
      Range r{};
      int n = 0;
      if (valid) {
        r = vox_range(k, T, L, false);
        n = r.n.x * r.n.y * r.n.z;
        if (n != 0 && !(drops == 0u && finer_holds_range(k, L, r, 1))) {
          big = n > 8;
          work = !big;
        }
      }
      wave_append(big, &k.hdr[H_BIGVOX_L + L], k.big_vox + (size_t)L * k.big_cap, k.big_cap, &k.hdr[H_BIG_OVF],
                  t);
      // The wavefront's small ranges of this level laid end to end (a scan of the cell counts), the lanes taking the
      // cells in turn: each finds its triangle by binary search over the scan and reads it by shuffles. One range per
      // lane would run the wavefront as long as its largest range (up to 8 cells) with the other lanes idle.
      const int cnt = work ? n : 0;
      const SmallDiv md = cnt > 1 ? small_div(r) : SmallDiv{1u, 1u};
      int inc = cnt;
      for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(inc, o);
        if ((int)(threadIdx.x & 63u) >= o) inc += y;
      }
      const int start = inc - cnt, N = __shfl(inc, 63);
      for (int j0 = 0; j0 < N; j0 += 64) {
        const int j = j0 + (int)(threadIdx.x & 63u);
        int own = 0;  // the last lane whose range starts at or before cell j (empty ranges share the next start)
        for (int step = 32; step > 0; step >>= 1) {
          int cand = own + step;
          if (__shfl(start, cand) <= j) own = cand;
        }
        const int q = j - __shfl(start, own), no = __shfl(cnt, own);
        const uint32_t to = (uint32_t)__shfl((int)t, own);
        const f3 A = mk3(__shfl(T.a.x, own), __shfl(T.a.y, own), __shfl(T.a.z, own));
        const f3 B = mk3(__shfl(T.b.x, own), __shfl(T.b.y, own), __shfl(T.b.z, own));
        const f3 Cc = mk3(__shfl(T.c.x, own), __shfl(T.c.y, own), __shfl(T.c.z, own));
        Range ro;
        ro.lo = i3{__shfl(r.lo.x, own), __shfl(r.lo.y, own), __shfl(r.lo.z, own)};
        ro.n = i3{__shfl(r.n.x, own), __shfl(r.n.y, own), __shfl(r.n.z, own)};
        const SmallDiv mo{(uint32_t)__shfl((int)md.mx, own), (uint32_t)__shfl((int)md.my, own)};
        if (j < N) voxelize_cell<TX>(k, hkey, hval, A, B, Cc, to, L, ro, mo, q, no, drops);
      }
    }
  }
  voxel_table_flush(k, hkey, hval);
}
template <bool TX>
__global__ void __launch_bounds__(BLK) k_vox_small(VoxK k) { vox_small_body<TX>(k, blockIdx.x, gridDim.x); }
template <bool TX>
__global__ void __launch_bounds__(BLK) k_vox_small_b(BatchK b) {
  const uint32_t blk = batch_block(b);
  int w = batch_world(b, blk);
  vox_small_body<TX>(b.w[w].vx, blk - b.start[w], b.start[w + 1] - b.start[w]);
}

// voxelizeBig (src/core/voxel.ts:515-531) over the listed (triangle, level) entries, `chunk` (a power of two up to
// 64) to a wavefront: lane i loads entry i, a wavefront scan of the cell counts lays the ranges end to end, and the
// lanes take the cells in turn, each finding its entry by binary search over the scan and its triangle by shuffles.
// A listed range has more than 8 cells but often far fewer than 64, so one entry per wavefront left most lanes idle.
// With many worlds in a launch the chunk is 64 (fewest wavefronts, fullest lanes); one world alone takes the smallest
// chunk that its launch's wavefronts cover, keeping the GPU busy. The OR of the bit and the max of the
// representative commute, so the result does not depend on the schedule.
// Entries [base, base + chunk) of the listed (triangle, level) entries (up to 64, one per lane): a wavefront scan of
// their cell counts lays the ranges end to end and the lanes take the cells in turn, each finding its entry by binary
// search over the scan and its triangle by shuffles. Called with the same arguments by the whole wavefront.
template <bool TX>
__device__ __forceinline__ void vox_big_chunk(const VoxK& k, uint32_t base, uint32_t chunk, uint32_t total,
                                              uint32_t lane, uint32_t drops) {
  const uint32_t e = lane < chunk ? base + lane : total;
  int L = 0, n = 0;
  uint32_t t = 0u;
  Tri T{};
  Range r{};
  if (e < total) {
    uint32_t idx;
    big_entry(k, e, H_BIGVOX_L, L, idx);
    t = k.big_vox[(size_t)L * k.big_cap + idx];
    T = load_tri(k, t);
    r = vox_range(k, T, L, false);
    n = r.n.x * r.n.y * r.n.z;
  }
  int inc = n;  // inclusive scan of the cell counts; listed entries have n > 8, so the starts increase strictly
  for (int o = 1; o < 64; o <<= 1) {
    int y = __shfl_up(inc, o);
    if ((int)lane >= o) inc += y;
  }
  const int start = inc - n, N = __shfl(inc, 63);
  for (int j0 = 0; j0 < N; j0 += 64) {
    const int j = j0 + (int)lane;
    int own = 0;  // the last entry starting at or before cell j
    for (int step = 32; step > 0; step >>= 1) {
      int cand = own + step;
      if (__shfl(start, cand) <= j) own = cand;
    }
    const int q = j - __shfl(start, own);
    const int Lo = __shfl(L, own);
    const uint32_t to = (uint32_t)__shfl((int)t, own);
    const f3 A = mk3(__shfl(T.a.x, own), __shfl(T.a.y, own), __shfl(T.a.z, own));
    const f3 B = mk3(__shfl(T.b.x, own), __shfl(T.b.y, own), __shfl(T.b.z, own));
    const f3 Cc = mk3(__shfl(T.c.x, own), __shfl(T.c.y, own), __shfl(T.c.z, own));
    Range ro;
    ro.lo = i3{__shfl(r.lo.x, own), __shfl(r.lo.y, own), __shfl(r.lo.z, own)};
    ro.n = i3{__shfl(r.n.x, own), __shfl(r.n.y, own), __shfl(r.n.z, own)};
    if (j >= N) continue;
    i3 c = cell_of(ro, q);
    if (finer_built(k, Lo, c, drops)) continue;
    float v = vox_size(k, Lo);
    f3 p = region_lo(k, Lo) + (tof3(c) + splat3(0.5f)) * v;
    if (!tri_box(p, splat3(0.5f * v * 1.0001f), A, B, Cc)) continue;
    uint32_t ptr = k.ptr[ptr_index(k, Lo, shr3(c, 3))];
    if (ptr >= PTR_MARK) continue;
    if (TX && !surf_pass(k, to, p, A, B, Cc, Lo)) continue;
    uint32_t bit = (uint32_t)((c.x & 7) + 8 * (c.y & 7) + 64 * (c.z & 7));
    atomicOr(&k.occ[ptr * 16u + (bit >> 5)], 1u << (bit & 31u));
    atomicMax(&k.rep[(size_t)ptr * 512u + bit], to + 1u);
  }
}

__device__ __forceinline__ uint32_t big_total(const VoxK& k) {
  uint32_t total = 0;
  for (int L = 0; L < k.NL; L++) total += min(k.hdr[H_BIGVOX_L + L], k.big_cap);
  return total;
}

template <bool TX>
__device__ __forceinline__ void vox_big_body(const VoxK& k, uint32_t lb, uint32_t nblk, bool wide) {
  const uint32_t total = big_total(k);
  const uint32_t drops = k.hdr[H_BRICK_OVF];
  const uint32_t wave = (lb * BLK + threadIdx.x) >> 6, lane = threadIdx.x & 63u;
  const uint32_t waves = (nblk * BLK) >> 6;
  uint32_t chunk = 64u;
  if (!wide) {
    const uint32_t per = (total + waves - 1u) / waves;
    chunk = 1u;
    while (chunk < per && chunk < 64u) chunk <<= 1;
  }
  for (uint32_t base = wave * chunk; base < total; base += waves * chunk)  // the same rounds for the whole wavefront
    vox_big_chunk<TX>(k, base, chunk, total, lane, drops);
}

// voxelizeBig over every world of a batch from one work queue: a world's listed entries are only counted on the
// device (k_vox_small appends them), and their cells per entry vary by two orders of magnitude, so a split of the
// blocks among the worlds fixed on the host leaves some wavefronts with many times the average work. Here each
// wavefront takes chunks of VQ_CHUNK entries of any world until all are taken; chunk g belongs to the world whose
// prefix of chunk counts (built per workgroup in LDS from the worlds' list lengths) covers g. The chunks are dealt to
// VQ_LANES counters (chunk g to counter g mod VQ_LANES, taken by the workgroups of that residue, one device-scope
// atomic per chunk): device-scope atomics on one address serialise at about 10 ns each, which a single counter
// made the launch's limit on a batch of 64 cities. The OR of the bit and the max of the representative commute, so
// the result does not depend on which wavefront takes which chunk.
constexpr uint32_t VQ_CHUNK = 16;
constexpr int VQ_MAXW = 1024;  // worlds per batch the queue serves (larger batches use the static split)
// (the launch has at least VQ_LANES workgroups, so every class of chunks has wavefronts to take it)
template <bool TX>
__device__ __forceinline__ void vox_big_queue(const BatchK& b) {
  __shared__ uint32_t s_pre[VQ_MAXW + 1];  // chunks of worlds [0, w)
  __shared__ uint32_t s_tot[VQ_MAXW];      // listed entries of world w
  const uint32_t lane = threadIdx.x & 63u;
  if (threadIdx.x < 64) {
    uint32_t carry = 0;
    for (int w0 = 0; w0 < b.n; w0 += 64) {
      const int w = w0 + (int)lane;
      const uint32_t tot = w < b.n ? big_total(b.w[w].vx) : 0u;
      uint32_t inc = (tot + VQ_CHUNK - 1u) / VQ_CHUNK;
      for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(inc, o);
        if ((int)lane >= o) inc += y;
      }
      if (w < b.n) {
        s_tot[w] = tot;
        s_pre[w + 1] = carry + inc;
      }
      carry += __shfl(inc, 63);
    }
    if (lane == 0) s_pre[0] = 0u;
  }
  __syncthreads();
  const uint32_t chunks = s_pre[b.n], cls = blockIdx.x % VQ_LANES;
  uint32_t* const counter = b.queue + cls * VQ_STRIDE;
  for (;;) {
    uint32_t i = 0;
    if (lane == 0) i = atomicAdd(counter, 1u);
    const uint32_t g = cls + VQ_LANES * (uint32_t)__shfl((int)i, 0);
    if (g >= chunks) break;  // uniform over the wavefront: every wavefront ends here once its class is taken
    int lo = 0, hi = b.n;  // the last world whose first chunk is at or before g
    while (hi - lo > 1) {
      const int m = (lo + hi) >> 1;
      if (s_pre[m] <= g)
        lo = m;
      else
        hi = m;
    }
    const VoxK& k = b.w[lo].vx;
    vox_big_chunk<TX>(k, (g - s_pre[lo]) * VQ_CHUNK, VQ_CHUNK, s_tot[lo], lane, k.hdr[H_BRICK_OVF]);
  }
}
template <bool TX>
__global__ void __launch_bounds__(BLK) k_vox_big(VoxK k) { vox_big_body<TX>(k, blockIdx.x, gridDim.x, false); }
template <bool TX>
__global__ void __launch_bounds__(BLK) k_vox_big_b(BatchK b) {
  const uint32_t blk = batch_block(b);
  int w = batch_world(b, blk);
  vox_big_body<TX>(b.w[w].vx, blk - b.start[w], b.start[w + 1] - b.start[w], b.n >= WIDE_BATCH);
}
template <bool TX>
__global__ void __launch_bounds__(BLK) k_vox_big_q(BatchK b) { vox_big_queue<TX>(b); }

// Level L from level L-1 (reduce, src/core/voxel.ts:540-577), gathered per voxel of every level-L brick: its 8
// children's bits are ORed and their representatives maxed. One wave owns two occupancy words (64 consecutive voxels)
// and each representative has one writer, so there are no atomics. The result equals the reference's scatter form:
// a child brick is allocated only if its parent is (markParents and the coarse-first order), and a dropped child
// contributes nothing in either form.
// Only the level-L bricks are visited: the allocation hands out slots from the coarsest level down, so level L's
// bricks are one run of slots after the coarser levels' runs (their counts are the header's per-level words).
__device__ __forceinline__ void reduce_body(const VoxK& k, int L, uint32_t lb, uint32_t nblk) {
  uint32_t first = 0;
  for (int l = k.NL - 1; l > L; l--) first += k.hdr[H_BRICKS_L + l];
  const uint32_t n = (first + k.hdr[H_BRICKS_L + L]) * 512u;
  const i3 oL = origin(k, L), oF = origin(k, L - 1);
  for (uint32_t base = first * 512u + lb * BLK; base < n; base += nblk * BLK) {
    uint32_t i = base + threadIdx.x;
    uint32_t ptr = i >> 9, bit = i & 511u;
    bool any = false;
    uint32_t rv = 0;
    bool mine = false;  // this wave's brick is a level-L brick (uniform per wave: 64 voxels of one brick)
    if (i < n) {
      uint32_t code = k.bcode[ptr];
      mine = (int)(code >> 27) == L;
      if (mine) {
        i3 b{(int)((code >> 18) & 511u), (int)((code >> 9) & 511u), (int)(code & 511u)};
        i3 pc = b * 8 + i3{(int)(bit & 7u), (int)((bit >> 3) & 7u), (int)(bit >> 6)};
        i3 f0 = (oL * 8 + pc) * 2 - oF * 8;
        if (inside3(f0, k.E)) {  // the window corners are even cells, so all 8 children are inside together
          uint32_t cptr = k.ptr[ptr_index(k, L - 1, shr3(f0, 3))];  // the 8 children share one finer brick
          if (cptr < PTR_MARK) {
            const uint32_t* co = k.occ + (size_t)cptr * 16u;
            const uint32_t* cr = k.rep + (size_t)cptr * 512u;
            for (int q = 0; q < 8; q++) {
              i3 f = f0 + i3{q & 1, (q >> 1) & 1, (q >> 2) & 1};
              uint32_t cb = (uint32_t)((f.x & 7) + 8 * (f.y & 7) + 64 * (f.z & 7));
              if (co[cb >> 5] & (1u << (cb & 31u))) {
                any = true;
                rv = max(rv, cr[cb]);
              }
            }
          }
        }
      }
    }
    uint64_t m = __ballot(any);
    if (any) {
      uint32_t* rp = &k.rep[(size_t)ptr * 512u + bit];
      if (*rp < rv) *rp = rv;
    }
    uint32_t lane = threadIdx.x & 63u;
    if (mine && (lane == 0u || lane == 32u)) {
      uint32_t word = (uint32_t)(m >> lane);
      if (word) k.occ[(size_t)ptr * 16u + (bit >> 5)] |= word;
    }
  }
}
__global__ void __launch_bounds__(BLK) k_reduce(VoxK k, int L) { reduce_body(k, L, blockIdx.x, gridDim.x); }
__global__ void __launch_bounds__(BLK) k_reduce_b(BatchK b) {
  const uint32_t blk = batch_block(b);
  int w = batch_world(b, blk);
  reduce_body(b.w[w].vx, b.arg, blk - b.start[w], b.start[w + 1] - b.start[w]);
}

__device__ __forceinline__ uint32_t brick_pop(const VoxK& k, uint32_t b) {
  const uint4* o = reinterpret_cast<const uint4*>(k.occ + (size_t)b * 16u);
  uint32_t n = 0;
  for (int q = 0; q < 4; q++) {
    uint4 x = o[q];
    n += __popc(x.x) + __popc(x.y) + __popc(x.z) + __popc(x.w);
  }
  return n;
}

// Attribute slots by prefix sum over bricks in slot order (allocVoxels, src/core/voxel.ts:579-588, with the atomic
// counter's order fixed to the slot order): rounds of 4096 bricks, four consecutive bricks per thread whose 16
// occupancy loads are in flight together.
__device__ __forceinline__ void alloc_voxels_body(const VoxK& k, uint32_t lb, uint32_t nblk) {
  __shared__ uint32_t wsum[32];
  uint32_t nb = n_bricks(k);
  uint32_t first = 0, ovf = 0;
  for (uint32_t b0 = 0; b0 < nb; b0 += 4u * blockDim.x) {
    uint32_t b = b0 + 4u * threadIdx.x;
    uint32_t n[4] = {0, 0, 0, 0};
    for (int j = 0; j < 4; j++)
      if (b + j < nb) n[j] = brick_pop(k, b + j);
    uint32_t round;
    uint32_t f = first + block_excl_sum(n[0] + n[1] + n[2] + n[3], wsum, &round);
    for (int j = 0; j < 4; j++) {
      if (b + j >= nb) break;
      const bool fits = f + n[j] <= (uint32_t)k.max_voxels;
      if (!fits) {
        k.battr[b + j] = ATTR_NONE;
        ovf++;
      } else {
        k.battr[b + j] = f;
      }
      // First slot of each occupancy word (voxRank's prefix, so that a rank is one load and one popcount).
      const uint4* o = reinterpret_cast<const uint4*>(k.occ + (size_t)(b + j) * 16u);
      uint4* ws = reinterpret_cast<uint4*>(k.wslot + (size_t)(b + j) * 16u);
      uint32_t run = f;
      for (int q = 0; q < 4; q++) {
        uint4 x = o[q];
        uint4 y;
        y.x = fits ? run : ATTR_NONE;
        run += __popc(x.x);
        y.y = fits ? run : ATTR_NONE;
        run += __popc(x.y);
        y.z = fits ? run : ATTR_NONE;
        run += __popc(x.z);
        y.w = fits ? run : ATTR_NONE;
        run += __popc(x.w);
        ws[q] = y;
      }
      f += n[j];
    }
    first += round;
  }
  if (ovf) atomicAdd(&k.hdr[H_VOXEL_OVF], ovf);
  if (threadIdx.x == 0) k.hdr[H_VOXELS] = first;
}
__global__ void __launch_bounds__(1024) k_alloc_voxels(VoxK k) { alloc_voxels_body(k, blockIdx.x, gridDim.x); }
__global__ void __launch_bounds__(1024) k_alloc_voxels_b(BatchK b) {
  const uint32_t blk = batch_block(b);
  int w = batch_world(b, blk);
  alloc_voxels_body(b.w[w].vx, blk - b.start[w], b.start[w + 1] - b.start[w]);
}


// WGSL pack4x8unorm / pack2x16snorm, by their specified formulas.
__device__ __forceinline__ uint32_t unorm8(float e) {
  return (uint32_t)floorf(0.5f + 255.0f * fminf(1.0f, fmaxf(0.0f, e)));
}
__device__ __forceinline__ uint32_t snorm16(float e) {
  return (uint32_t)(int)floorf(0.5f + 32767.0f * fminf(1.0f, fmaxf(-1.0f, e))) & 0xffffu;
}

// resolveAttr (src/core/voxel.ts:596-622) of the occupied voxel (ptr, bit) with attribute slot ai.
template <bool TX>
__device__ __forceinline__ void resolve_voxel(const VoxK& k, uint32_t ptr, uint32_t bit, uint32_t ai) {
  if (k.vslot) k.vslot[ai] = (ptr << 9) | bit;
  uint32_t code = k.bcode[ptr];
  int L = (int)(code >> 27);
  i3 b{(int)((code >> 18) & 511u), (int)((code >> 9) & 511u), (int)(code & 511u)};
  i3 c = b * 8 + i3{(int)(bit & 7u), (int)((bit >> 3) & 7u), (int)(bit >> 6)};
  f3 p = region_lo(k, L) + (tof3(c) + splat3(0.5f)) * vox_size(k, L);
  uint32_t rp = k.rep[ptr * 512u + bit];
  if (rp == 0u) return;  // every occupied voxel has a representative; kept as a guard on the load below
  Tri T = load_tri(k, rp - 1u);
  // surfaceAt with wantAttr (src/core/voxel.ts:325-336): the albedo is the texture (white when untextured) times
  // baseColorFactor at the closest point's UV, on mip 3 + L; the albedo word's alpha is the toon flag.
  f3 wb = closest_bary(p, T.a, T.b, T.c);
  f3 nn = ldp(k.nrm, T.i0) * wb.x + ldp(k.nrm, T.i1) * wb.y + ldp(k.nrm, T.i2) * wb.z;
  if (!(dot3(nn, nn) > 1e-12f)) nn = cross3(T.b - T.a, T.c - T.a);
  nn = normalize3(nn);
  f3 q = T.a * wb.x + T.b * wb.y + T.c * wb.z;
  float alb[3], toon = 0.0f;
  if (TX && k.surf.tex) {
    float u, v;
    surf_uv(k.surf, T.i0, T.i1, T.i2, wb, u, v);
    const float4 tx = surf_base_color(k.surf, k.mat, T.m, u, v, L);
    alb[0] = tx.x;
    alb[1] = tx.y;
    alb[2] = tx.z;
    toon = (surf_flags(k.surf, T.m) & RTGI_SURF_TOON) ? 1.0f : 0.0f;
  } else {
    const float4 m = reinterpret_cast<const float4*>(k.mat)[T.m];
    alb[0] = 1.0f * m.x;
    alb[1] = 1.0f * m.y;
    alb[2] = 1.0f * m.z;
  }
  uint32_t pa = unorm8(alb[0]) | (unorm8(alb[1]) << 8) | (unorm8(alb[2]) << 16) | (unorm8(toon) << 24);
  // octEnc (src/core/voxel.ts:584-588).
  float s = fabsf(nn.x) + fabsf(nn.y) + fabsf(nn.z);
  float px = nn.x / s, py = nn.y / s;
  if (nn.z < 0.0f) {
    float ox = px, oy = py;
    px = (1.0f - fabsf(oy)) * (ox >= 0.0f ? 1.0f : -1.0f);
    py = (1.0f - fabsf(ox)) * (oy >= 0.0f ? 1.0f : -1.0f);
  }
  uint32_t pn = snorm16(px) | (snorm16(py) << 16);
  k.attr[ai * 2u] = pa;
  k.attr[ai * 2u + 1u] = pn;
  if (k.alb16) *reinterpret_cast<uint2*>(k.alb16 + ai * 2u) = make_uint2(f2bf(alb[0]) | (f2bf(alb[1]) << 16), f2bf(alb[2]) | (f2bf(toon) << 16));
  k.plane[ai] = dot3(nn, q - p);
}

// Position of the n-th (from 0) set bit of x, which has more than n set bits.
__device__ __forceinline__ uint32_t nth_set_bit(uint32_t x, uint32_t n) {
  uint32_t p = 0u;
  for (uint32_t s = 16u; s > 0u; s >>= 1) {
    uint32_t c = (uint32_t)__popc(x & ((1u << s) - 1u));
    if (n >= c) {
      n -= c;
      x >>= s;
      p += s;
    }
  }
  return p;
}

// The occupied voxels of R bricks (16 R occupancy words, one per lane) per round of a wavefront, laid end to end by
// a scan of the words' popcounts: the lanes take the voxels in turn, each finding its word by binary search over the
// scan and its bit by a popcount search. One lane per voxel slot of the bricks kept about a quarter of the lanes busy.
// R = 4 with many worlds in a launch (fullest lanes); 1 for one world alone, which then spreads over more wavefronts.
template <bool TX>
__device__ __forceinline__ void resolve_body(const VoxK& k, uint32_t lb, uint32_t nblk, bool wide) {
  const uint32_t nb = n_bricks(k);
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wave = (lb * BLK + threadIdx.x) >> 6, waves = (nblk * BLK) >> 6;
  const uint32_t R = wide ? 4u : 1u;
  for (uint32_t b0 = wave * R; b0 < nb; b0 += waves * R) {  // the same rounds for the whole wavefront
    uint32_t word = 0u, first = ATTR_NONE;
    if ((lane >> 4) < R && b0 + (lane >> 4) < nb) {
      word = k.occ[(size_t)b0 * 16u + lane];
      first = k.wslot[(size_t)b0 * 16u + lane];
    }
    const int cnt = first == ATTR_NONE ? 0 : __popc(word);
    int inc = cnt;
    for (int o = 1; o < 64; o <<= 1) {
      int y = __shfl_up(inc, o);
      if ((int)lane >= o) inc += y;
    }
    const int start = inc - cnt, N = __shfl(inc, 63);
    for (int j0 = 0; j0 < N; j0 += 64) {
      const int j = j0 + (int)lane;
      int own = 0;  // the last word whose voxels start at or before j
      for (int step = 32; step > 0; step >>= 1) {
        int cand = own + step;
        if (__shfl(start, cand) <= j) own = cand;
      }
      const uint32_t wd = (uint32_t)__shfl((int)word, own), fs = (uint32_t)__shfl((int)first, own);
      const uint32_t n = (uint32_t)(j - __shfl(start, own));
      if (j >= N) continue;
      const uint32_t bit = ((uint32_t)own & 15u) * 32u + nth_set_bit(wd, n);
      resolve_voxel<TX>(k, b0 + ((uint32_t)own >> 4), bit, fs + n);
    }
  }
}
// Mip level m >= 1 of every texture layer from level m - 1 (generateMips, src/core/scene_gpu.ts:38-54), one thread
// per texel; levels run in order on the stream.
__global__ void __launch_bounds__(BLK) k_surf_mip(SurfK s, uint32_t* tex, uint32_t m) {
  const uint32_t wl = s.size_log2 - m, per = 1u << (2u * wl);
  const uint64_t i = (uint64_t)blockIdx.x * BLK + threadIdx.x;
  if (i >= (uint64_t)per * s.layers) return;
  const uint32_t layer = (uint32_t)(i >> (2u * wl)), j = (uint32_t)i & (per - 1u);
  tex[(size_t)layer * s.layer_words + surf_mip_off(s.size_log2, m) + j] =
      surf_mip_texel(s, layer, m, j & ((1u << wl) - 1u), j >> wl);
}

template <bool TX>
__global__ void __launch_bounds__(BLK) k_resolve(VoxK k) { resolve_body<TX>(k, blockIdx.x, gridDim.x, false); }
template <bool TX>
__global__ void __launch_bounds__(BLK) k_resolve_b(BatchK b) {
  const uint32_t blk = batch_block(b);
  int w = batch_world(b, blk);
  resolve_body<TX>(b.w[w].vx, blk - b.start[w], b.start[w + 1] - b.start[w], b.n >= WIDE_BATCH);
}


}  // namespace

// LDS bytes of the mark bitmap (0: the table is too large and marks go to global memory).
size_t vox_mark_lds_bytes(const VoxK& k) {
  size_t bytes = (((size_t)k.NL * k.per + 31) / 32) * 4;
  return bytes <= 64 * 1024 ? bytes : 0;
}

static BatchK bk(const BatchPlan& p, int l, int arg = 0) {
  return BatchK{p.w, p.start[l], p.n, arg, p.xcd[l] ? 1 : 0, p.queue};
}

int launch_vox_build_batch(const BatchPlan& p, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (p.grid[BL_POSE]) k_pose_b<<<p.grid[BL_POSE], BLK, 0, s>>>(bk(p, BL_POSE));
  if (p.grid[BL_CLEAR]) k_clear_b<<<p.grid[BL_CLEAR], BLK, 0, s>>>(bk(p, BL_CLEAR));
  if (p.grid[BL_MARK]) k_mark_b<<<p.grid[BL_MARK], BLK, p.mark_lds, s>>>(bk(p, BL_MARK));
  if (p.grid[BL_MARK_BIG]) k_mark_big_b<<<p.grid[BL_MARK_BIG], BLK, p.mark_lds, s>>>(bk(p, BL_MARK_BIG));
  if (p.grid[BL_MARK_PARENTS]) k_mark_parents_b<<<p.grid[BL_MARK_PARENTS], BLK, 0, s>>>(bk(p, BL_MARK_PARENTS));
  if (p.grid[BL_ALLOC_BRICKS]) k_alloc_bricks_b<<<p.grid[BL_ALLOC_BRICKS], 1024, 0, s>>>(bk(p, BL_ALLOC_BRICKS));
  if (p.grid[BL_INIT]) k_init_bricks_b<<<p.grid[BL_INIT], BLK, 0, s>>>(bk(p, BL_INIT));
  const bool tx = p.surf;  // textured instantiations only where some world samples textures
  if (p.grid[BL_VOX_SMALL])
    (tx ? k_vox_small_b<true> : k_vox_small_b<false>)<<<p.grid[BL_VOX_SMALL], BLK, 0, s>>>(bk(p, BL_VOX_SMALL));
  if (p.grid[BL_VOX_BIG]) {
    if (p.queue && p.n <= VQ_MAXW)  // persistent wavefronts over the batch's work queue (zeroed by k_clear_b)
      (tx ? k_vox_big_q<true> : k_vox_big_q<false>)<<<
          std::max<uint32_t>(VQ_LANES, std::min<uint32_t>(p.grid[BL_VOX_BIG], 2048)), BLK, 0, s>>>(bk(p, BL_VOX_BIG));
    else
      (tx ? k_vox_big_b<true> : k_vox_big_b<false>)<<<p.grid[BL_VOX_BIG], BLK, 0, s>>>(bk(p, BL_VOX_BIG));
  }
  for (int L = 1; L < RTGI_MAX_LEVELS; L++)
    if (p.grid[BL_REDUCE + L - 1]) k_reduce_b<<<p.grid[BL_REDUCE + L - 1], BLK, 0, s>>>(bk(p, BL_REDUCE + L - 1, L));
  if (p.grid[BL_ALLOC_VOXELS]) k_alloc_voxels_b<<<p.grid[BL_ALLOC_VOXELS], 1024, 0, s>>>(bk(p, BL_ALLOC_VOXELS));
  if (p.grid[BL_RESOLVE])
    (tx ? k_resolve_b<true> : k_resolve_b<false>)<<<p.grid[BL_RESOLVE], BLK, 0, s>>>(bk(p, BL_RESOLVE));
  return (int)hipGetLastError();
}

static uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

void batch_blocks(const TraceK& t, uint32_t out[BL_COUNT]) {
  const VoxK& k = t.vx;
  for (int i = 0; i < BL_COUNT; i++) out[i] = 0;
  uint64_t cells = (uint64_t)k.NL * k.per, pool = (uint64_t)k.max_bricks * 512u;
  uint32_t persist = std::min<uint32_t>(2048, std::max<uint32_t>(1, cdiv(pool, BLK * 4)));
  out[BL_POSE] = k.npose ? cdiv(k.pose_span, BLK) : 0;
  out[BL_CLEAR] = cdiv(std::max<uint64_t>(cells, HDR_WORDS), BLK);
  out[BL_MARK] = k.ntris ? std::min<uint32_t>(1024, cdiv(k.ntris, BLK * 4)) : 0;
  out[BL_MARK_BIG] = std::min<uint32_t>(512, std::max<uint32_t>(1, k.ntris / 1024));
  out[BL_MARK_PARENTS] = k.NL > 1 ? cdiv((uint64_t)(k.NL - 1) * k.per, BLK) : 0;
  out[BL_ALLOC_BRICKS] = 1;
  out[BL_INIT] = persist;
  out[BL_VOX_SMALL] = k.ntris ? cdiv(k.ntris, (uint64_t)BLK * VOX_CHUNK) : 0;
  out[BL_VOX_BIG] = std::min<uint32_t>(2048, std::max<uint32_t>(16, k.ntris / 256));
  for (int L = 1; L < k.NL; L++) out[BL_REDUCE + L - 1] = persist;
  out[BL_ALLOC_VOXELS] = 1;
  out[BL_RESOLVE] = std::min<uint32_t>(2048, std::max<uint32_t>(1, cdiv((uint64_t)k.max_bricks, 16)));  // 4 bricks a wave
  if (k.dyn.hdr && k.dyn.end_tri > k.dyn.first_tri) {  // the dynamic layer (grid-stride loops)
    const uint32_t nd = k.dyn.end_tri - k.dyn.first_tri;
    const uint32_t vox = std::min<uint32_t>(4096, cdiv((uint64_t)k.dyn.cap * 512u, BLK));
    out[BL_DL_MARK] = std::min<uint32_t>(4096, cdiv(nd, BLK));
    out[BL_DL_PARENTS] = k.dyn.nw > (uint32_t)k.per / 32u ? cdiv(k.dyn.nw - (uint32_t)k.per / 32u, BLK) : 0;
    out[BL_DL_ALLOC] = 1;
    out[BL_DL_CLEAR] = vox;
    out[BL_DL_SPLAT] = std::min<uint32_t>(8192, cdiv(nd, 16));  // workgroups of 16 triangles (one wavefront)
    for (int L = 1; L < k.NL; L++) out[BL_DL_RESOLVE + L - 1] = vox;
    if (t.objects) out[BL_DL_INJECT] = vox;
  }
  // Per attribute slot: 4 lanes per voxel for the sun taps, one for the bounce injection (grid-stride loops).
  const uint64_t slots = (uint64_t)k.max_voxels;
  if (t.sunvis) out[BL_SUN] = std::min<uint32_t>(8192, std::max<uint32_t>(1, cdiv(slots * 4, BLK)));
  if (t.bounce) {
    out[BL_BNVOL] = BN_PROBES / BN_PROBES_PER_BLOCK;
    out[BL_BOUNCE] = std::min<uint32_t>(4096, std::max<uint32_t>(1, cdiv(slots, BLK)));
  }
  if (t.reloc) {
    out[BL_RELOC] = cdiv(t.total_probes, BLK);
    out[BL_LINK] = t.N > 1 ? cdiv(t.c[t.N - 1].probe_off, BLK) : 0;
  }
  out[BL_TRACE] = cdiv(t.total_rays, BLK);
  for (int st = 0; st < t.N; st++) {
    int c = t.N - 1 - st;
    out[BL_MERGE + st] = cdiv(merge_threads(t, c), BLK);
    out[BL_TM + st] = cdiv(trace_merge_threads(t, c), BLK);
  }
  out[BL_SH] = cdiv(t.c[0].np, BLK);
}

int launch_surf_mips(const SurfK& s, uint32_t* tex, void* stream) {
  for (uint32_t m = 1; m < s.mips; m++) {
    const uint64_t n = ((uint64_t)1 << (2u * (s.size_log2 - m))) * s.layers;
    if (n == 0) break;
    k_surf_mip<<<cdiv(n, BLK), BLK, 0, (hipStream_t)stream>>>(s, tex, m);
  }
  return (int)hipGetLastError();
}

int launch_pose(const VoxK& k, void* stream) {
  if (k.npose) k_pose<<<cdiv(k.pose_span, BLK), BLK, 0, (hipStream_t)stream>>>(k);
  return (int)hipGetLastError();
}

int launch_vox_build(const VoxK& k, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  TraceK t{};
  t.vx = k;
  uint32_t g[BL_COUNT];
  batch_blocks(t, g);
  size_t lds = k.lds_marks ? vox_mark_lds_bytes(k) : 0;
  k_clear<<<g[BL_CLEAR], BLK, 0, s>>>(k);
  if (g[BL_MARK]) k_mark<<<g[BL_MARK], BLK, lds, s>>>(k);
  k_mark_big<<<g[BL_MARK_BIG], BLK, lds, s>>>(k);
  if (g[BL_MARK_PARENTS]) k_mark_parents<<<g[BL_MARK_PARENTS], BLK, 0, s>>>(k);
  k_alloc_bricks<<<1, 1024, 0, s>>>(k);
  k_init_bricks<<<g[BL_INIT], BLK, 0, s>>>(k);
  const bool tx = k.surf.tex != nullptr;
  if (g[BL_VOX_SMALL]) (tx ? k_vox_small<true> : k_vox_small<false>)<<<g[BL_VOX_SMALL], BLK, 0, s>>>(k);
  (tx ? k_vox_big<true> : k_vox_big<false>)<<<g[BL_VOX_BIG], BLK, 0, s>>>(k);
  for (int L = 1; L < k.NL; L++) k_reduce<<<g[BL_REDUCE + L - 1], BLK, 0, s>>>(k, L);
  k_alloc_voxels<<<1, 1024, 0, s>>>(k);
  (tx ? k_resolve<true> : k_resolve<false>)<<<g[BL_RESOLVE], BLK, 0, s>>>(k);
  return (int)hipGetLastError();
}
