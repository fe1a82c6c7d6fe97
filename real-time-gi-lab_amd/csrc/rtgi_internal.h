// Internal declarations of librtgi: kernel argument blocks, launchers and error helpers.
#pragma once

#include <stdint.h>

#include "rtgi.h"

// Records msg as the calling thread's last error and returns code.
int rtgi_fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

// Header (counter) words at the start of the clipmap: the reference's header slots (src/core/voxel.ts:11), with
// word 0 = every marked brick (allocated = min(word 0, max_bricks)) and words 16-31 = big-list lengths.
enum : uint32_t {
  HDR_WORDS = 64,
  H_MARKED = 0,
  H_VOXELS = 1,
  H_BRICK_OVF = 2,
  H_BIG_OVF = 3,
  H_VOXEL_OVF = 4,
  H_BRICKS_L = 8,      // + level
  H_BIGVOX_L = 16,     // + level
  H_BIGMARK_L = 24,    // + level
};

constexpr uint32_t PTR_EMPTY = RTGI_PTR_EMPTY;
constexpr uint32_t PTR_DROP = RTGI_PTR_DROP;

constexpr uint32_t PTR_MARK = RTGI_PTR_MARK;
constexpr uint32_t ATTR_NONE = 0xffffffffu;

// Clipmap of one world in HBM. Structure of arrays, so that each access class is dense and aligned:
//   ptr    [levels][G^3]       brick pointer per window brick (index z-major: (bz G + by) G + bx), or a sentinel
//   occ    [max_bricks][16]    occupancy words of a brick (bit = x + 8 y + 64 z), 64 B per brick
//   battr  [max_bricks]        first attribute slot of the brick's voxels (ATTR_NONE: no slots)
//   wslot  [max_bricks][16]    first attribute slot of each occupancy word (battr + popcounts of the earlier words)
//   bcode  [max_bricks]        level << 27 | bx << 18 | by << 9 | bz (the reference's brick code word)
//   rep    [max_bricks][512]   representative triangle + 1 per voxel (build temporary, kept for readback)
//   attr   [max_voxels][2]     pack4x8unorm albedo, pack2x16snorm oct normal
//   plane  [max_voxels]        representative plane offset from the voxel centre
// Capacity of a level's big-triangle list (BIG_CAP, src/core/voxel.ts:224); entries past it are dropped and counted
// in H_BIG_OVF, as the reference does.
constexpr uint32_t BIG_CAP = 1u << 18;
inline uint32_t big_cap_for(uint32_t ntris) { return ntris < BIG_CAP ? ntris : BIG_CAP; }

// Dynamic coverage layer (src/core/dynlayer.ts): dynamic triangles [first_tri, end_tri) are splatted every update
// into bricks of 8^3 voxels, each voxel 13 words (projected areas on +x -x +y -y +z -z, area, area x radiance rgb,
// area x albedo rgb: the last three stay 0, as the reference's splat without objects receiving light), fixed point at
// level 0 and floats above after the resolve. hdr words: 0 wanted bricks, 1 dropped, 2..7 bounds (cm + 2^30),
// 8 allocated, 9 triangles outside every clipmap, 10 update counter (bitset parity), 16 + L first slot of level L,
// 24 + L bricks of level L, 32..37 and 38 bounds and outside count being marked.
constexpr uint32_t DYN_WORDS = 13;
constexpr uint32_t DYN_HDR = 64;
constexpr uint32_t DYN_EMPTY = 0xffffffffu;
constexpr uint32_t DYN_DROPPED = 0xfffffffdu;
constexpr float DYN_FIX = 65536.0f;
struct DynK {
  uint32_t* hdr;    // [DYN_HDR], or null: no layer
  uint32_t* ptr;    // [levels][G^3] slot of each brick, DYN_EMPTY or DYN_DROPPED
  uint32_t* bits;   // [2][nw] wanted bricks of this and the previous update (bit g = (levels-1-L) G^3 + brick)
  uint32_t* meta;   // [cap] pointer index (L G^3 + brick) of each slot
  uint32_t* mask;   // [cap][16] voxels with a value
  uint32_t* data;   // [cap][512][DYN_WORDS]
  int32_t cap;      // brick pool (dynBricks)
  uint32_t nw;      // levels G^3 / 32
  uint32_t first_tri, end_tri;
};

// Textured surfaces (rtgi_world_set_surface): the base-colour array with its mip chain, UVs and material words.
// tex == null: untextured (every material samples white, none is alpha-masked).
//   tex   [layers][mips] RGBA8 texels (rgb sRGB-encoded, alpha linear), level m of layer l at l * layer_words +
//         mip_off(m) with mip_off(m) = sum over j < m of (S >> j)^2
//   uv    [nverts][2]
//   mext  [nmat][4]: base-colour layer (0: white default), flags (RTGI_SURF_*), alpha cutoff (float bits), 0
//   lut   [256] sRGB-decoded value of each code (f32 of the exact value), then [255] encode thresholds: code k + 1
//         starts where the linear value reaches lut[256 + k] (the f32 of the decoded value of (k + 0.5) / 255)
constexpr uint32_t SURF_MAX_LOG2 = 12;
struct SurfK {
  const uint32_t* tex;
  const float* uv;
  const uint32_t* mext;
  const float* lut;
  uint32_t size_log2, mips, layer_words, layers;
};

struct VoxK {
  float v0;
  float inv_v0;   // 1 / v0, used only when pow2 (then x / v == x * (1/v) exactly)
  int32_t pow2;   // v0 is a power of two
  int32_t NL, G, E, per;  // per = G^3
  int32_t max_bricks, max_voxels;
  int32_t origin[RTGI_MAX_LEVELS][3];
  uint32_t* hdr;
  uint32_t* ptr;
  uint32_t* occ;
  uint32_t* battr;
  uint32_t* wslot;     // [max_bricks][16] first attribute slot of each occupancy word (ATTR_NONE: brick has none)
  uint32_t* bcode;
  uint32_t* rep;
  uint32_t* attr;
  float* plane;
  uint32_t* vslot;     // [max_voxels] brick << 9 | bit of each attribute slot (written by the resolve), or null
  uint32_t* alb16;     // [max_voxels][2] bf16 albedo of each attribute slot (precision form), or null
  uint32_t* big_vox;   // [levels][big_cap]
  uint32_t* big_mark;  // [levels][big_cap]
  const float* pos;
  const float* nrm;
  const uint32_t* tri;  // 4 words per triangle
  const float* mat;     // 4 floats per material
  uint32_t ntris;
  uint32_t big_cap;    // entries per level of each big list: min(ntris, BIG_CAP)
  int32_t lds_marks;    // brick marks accumulate in an LDS bitmap of levels * G^3 bits (0: global fallback)
  // Moving objects posed at the start of the update (rtgi_world_set_pose; npose 0: none): pose[npose] in ascending
  // vertex order; rest holds the span's rest positions then its rest normals (3 floats per vertex each).
  const struct PoseRec* pose;
  const float* rest;
  uint32_t npose, pose_first, pose_span;
  unsigned long long* stats;  // tracer statistics (3 x 8 counters) zeroed by the clear stage, or null
  DynK dyn;                   // the dynamic layer (dyn.hdr null: none)
  SurfK surf;                 // textured surfaces (surf.tex null: untextured)
};

// Bounce volume (DEFAULT_BOUNCE, src/core/bounce.ts:33): BN_DX x BN_DY x BN_DZ ambient-cube probes BN_CELL m apart
// in a window around the eye; each active probe traces BN_RAYS irradiance rays and 6 axis rays, every BN_DIV-th
// update by its world position (kept in between). A volume is [BN_PROBES][6] texels of 4 halves (rgb irradiance
// along axis +x, -x, +y, -y, +z, -z; alpha: the free reach along that axis, 0 = inactive probe).
constexpr int BN_DX = 32, BN_DY = 16, BN_DZ = 32;
constexpr uint32_t BN_PROBES = (uint32_t)(BN_DX * BN_DY * BN_DZ);
constexpr uint32_t BN_RAYS = 16;
constexpr uint32_t BN_DIV = 2;
constexpr float BN_CELL = 1.2f;
constexpr uint32_t BN_WORDS = BN_PROBES * 6u * 2u;  // u32 words of one volume
constexpr uint32_t BN_PROBES_PER_BLOCK = 8;          // 32 lanes per probe in a 256-thread block

// One posed object: vertices [first, first + count), pivot, cos / sin of the turn about +y, translation.
struct PoseRec {
  double pivot[3];
  double c, s;
  double off[3];
  uint32_t first, count;
};

struct CascK {
  int32_t n;        // probes per axis
  int32_t res;      // directions per axis
  int32_t vox_level;
  float b0, b1;     // interval in units of s0 (f32, as the Layout uniform stores them)
  uint32_t np;      // probes (n^3)
  int32_t n_log2;   // log2(n) when n is a power of two, else -1 (index arithmetic by shifts)
  int32_t res_log2; // log2(res) (res = 4 << min(c, resCap) is always a power of two)
  int32_t tile_log2; // probe storage order: 0 rows of the grid (x fastest); k > 0 (n a power of two, n >= 2^k):
                     // blocks of 2^k probes per axis in row order, x fastest inside a block
  uint32_t dir_off; // first entry of the cascade in the direction table
  uint64_t avg_off; // cascade c >= 1: first entry of its 2x2-direction average (probes x res_{c-1}^2 float4)
  uint64_t ray_off; // first ray of the cascade
  uint64_t rays;
  uint64_t probe_off; // first probe of the cascade in the relocated-probe arrays
};

// Direction table entry of an octahedral texel: the normalised direction (dirOf, src/core/rc.ts:308) and the
// reciprocal of its slab-safe components (traceVox, src/core/voxel.ts:128-130), computed once on the host with the
// same f32 operations the reference evaluates per ray.
struct DirEntry {
  float d[4];
  float inv[4];
};

struct TraceK {
  VoxK vx;
  int32_t N;
  CascK c[RTGI_MAX_CASCADES];
  int32_t G0[3];  // absolute c0 grid index of probe (0,0,0); a multiple of 2^(N-1)
  float s0;
  float vox_skip, vox_offset;
  float sun_dir[3], sun_e[3], sky[3], sun_vis;
  int32_t n_point;
  float point_pos[RTGI_MAX_POINT_LIGHTS][3];
  float point_i[RTGI_MAX_POINT_LIGHTS][3];
  float omega0[16];
  const DirEntry* dirs;  // all cascades' direction tables
  int32_t fp16;     // rad holds 4 halves per ray and sh halves (RNE at the reference's pack points), else f32
  int32_t bf16;     // bounce words hold bf16 (precision form), else halves
  int32_t no_merge; // trace only: rad keeps the traced values, no averages, no SH
  int32_t reloc;    // voxel relocation: probe positions / normals / parent visibility from the arrays below
  float lift;       // relocated probes keep lift * spacing of free space behind them
  float* ppos;      // [all cascades' probes] float4: relocated position (xyz) and s0 (w)
  float* pnrm;      // [all cascades' probes] float4: probe normal (xyz; 0: none)
  uint32_t* pvis;   // [all cascades' probes] bit k: parent k of cascade c+1 exists and is not hidden by voxels
  float* sunvis;    // [max_voxels] traced sun visibility per attribute slot, or null (frame sun_vis)
  uint32_t* bounce; // [max_voxels][4] bounce irradiance per slot: front rg, b, back rg, b (halves or bf16), or null
  uint32_t* bn_cur;        // bounce volume this update writes (BN_WORDS), or null
  const uint32_t* bn_prev; // the previous update's volume: its hits' bounce term and the kept probes
  int32_t bn_org[3];       // probe coordinates of probe 0 of bn_cur (floor(eye / BN_CELL) - dims / 2)
  int32_t bn_prev_org[3];  // the same of bn_prev
  uint32_t bn_frame;       // updates of the volume so far, this one included
  int32_t bn_level;        // levelOf(BN_CELL): the clipmap level of the near-surface test
  uint32_t bn_dir_off;     // first of the BN_RAYS irradiance directions in dirs
  int32_t objects;         // token obj with bounce and a dynamic layer: albedo splats, dynInject, bounce rays read it
  uint64_t total_probes;
  float* rad;       // 4 floats (or halves) per ray; within a cascade direction-major: ray = di * np + probe
  float* avg;       // merged cascade c >= 1 averaged over 2x2 directions at cascade c-1's resolution (upperAvg)
  uint32_t* hits;   // per ray (same order), or null
  float* sh;        // [27][probes0^3] in grid-row order (f32, or halves in the precision form)
  const float* sh_prev;  // the previous update's SH (same layout) the c0 SH blends with, or null (no history)
  int32_t hist_d[3];     // c0 grid origin of this update minus that of sh_prev: a probe's row there is at local + d
  int32_t sh_fuse;       // the c0 trace+merge launch also projects the SH (16 probes x 16 directions per block)
  uint64_t total_rays;
};

// One batched launch: world descriptors and the block prefix (n + 1 entries) in device memory; arg is the level of a
// reduce step or the cascade-step of a merge.
struct BatchK {
  const TraceK* __restrict__ w;
  const uint32_t* __restrict__ start;
  int32_t n;
  int32_t arg;
  int32_t xcd;  // remap blocks so that each XCD runs one contiguous range (worlds of about equal block counts)
  uint32_t* queue;  // the batch's work-queue counter (zeroed by k_clear_b, taken by k_vox_big_q), or null
};

// The SH projection takes two neighbouring c0 probes per thread where the radiance and SH are f16 in grid rows of
// even length: each of its reads and writes then covers both probes. The host picks the launch form from the
// descriptors of the update; the kernels test the same predicate.
#if defined(__HIPCC__)
__host__ __device__
#endif
inline bool sh_pairs(const TraceK& t) {
  return t.fp16 && t.c[0].tile_log2 == 0 && (t.c[0].n & 1) == 0 && (t.c[0].np & 1u) == 0 && (t.c[0].ray_off & 1u) == 0;
}

// Launchers (each returns a hipError_t as int).
// posedVertices of the world's objects into its vertex buffers (before the voxel build that reads them).
int launch_pose(const VoxK& k, void* stream);
// Mip levels 1.. of a world's base-colour array from its level 0 (tex: the array s.tex points to).
int launch_surf_mips(const SurfK& s, uint32_t* tex, void* stream);
size_t vox_mark_lds_bytes(const VoxK& k);
int launch_vox_build(const VoxK& k, void* stream);
int launch_trace(const TraceK& k, void* stream);
// n rays (origins o, unit directions d; device arrays of 3 floats each) traced as rays of cascade c -> rgba, hits.
int launch_trace_rays(const TraceK& k, int c, const float* o, const float* d, uint32_t n, float* rgba,
                      uint32_t* hits, void* stream);
// Gather of the irradiance field at n surface points (device arrays: 3 floats each) -> irr (3 floats each).
int launch_gather(const TraceK& k, const float* pos, const float* nrm, uint32_t n, int occl, float* irr, void* stream);
// Device address (on `device`) of the allocation-free gather's staging array: RTGI_GATHER_SCRATCH_POINTS x 9 floats
// (positions, normals, irradiance). Looked up once per device (hipGetSymbolAddress with the device current); returns
// a hipError_t as int.
constexpr int RTGI_SCRATCH_MAX_DEVICES = 64;
int gather_scratch(int device, float** out);
// Traced sun visibility of every occupied voxel (thread per brick voxel).
int launch_sun(const TraceK& k, void* stream);
// Bounce volume update (8 probes per block), then bounce injection of every occupied voxel from it (thread per
// brick voxel).
int launch_bounce(const TraceK& k, void* stream);
// Voxel relocation of every cascade's probes, then the parent visibility masks (thread per probe).
int launch_relocate(const TraceK& k, void* stream);
// Threads of cascade c's merge launch, and of its fused trace+merge launch (c = 0 with sh_fuse: 256 per 16 probes).
uint64_t merge_threads(const TraceK& k, int c);
uint64_t trace_merge_threads(const TraceK& k, int c);
// Trace and merge of cascade c in one launch (cascades N-1 .. 0 in order; c + 1's averages must be complete).
int launch_trace_merge(const TraceK& k, int c, void* stream);
int launch_merge(const TraceK& k, int c, void* stream);
int launch_sh(const TraceK& k, void* stream);


// The launches of one batched update, and for each the per-world block prefix (device memory) and grid size.
enum BatchLaunch : int {
  BL_POSE = 0,                             // posed vertices of the moving objects
  BL_CLEAR,
  BL_MARK,
  BL_MARK_BIG,
  BL_MARK_PARENTS,
  BL_ALLOC_BRICKS,
  BL_INIT,
  BL_VOX_SMALL,
  BL_VOX_BIG,
  BL_REDUCE,                               // + level - 1 (levels 1..7)
  BL_ALLOC_VOXELS = BL_REDUCE + RTGI_MAX_LEVELS - 1,
  BL_RESOLVE,
  BL_DL_MARK,                              // dynamic layer: wanted bricks of the dynamic triangles
  BL_DL_PARENTS,                           //   their ancestors
  BL_DL_ALLOC,                             //   slots in bit order (one block of 1024), counts, bounds
  BL_DL_CLEAR,                             //   the allocated bricks' words
  BL_DL_SPLAT,                             //   fixed-point splats of the dynamic triangles' samples
  BL_DL_INJECT,                            //   the bounce volume's light into the layer (token obj)
  BL_DL_RESOLVE,                           //   + level - 1 (levels 1..7): coarse values from own splats + children
  BL_SUN = BL_DL_RESOLVE + RTGI_MAX_LEVELS - 1,  // traced sun visibility of every occupied voxel
  BL_BNVOL,                                // bounce volume update (8 probes per block)
  BL_BOUNCE,                               // bounce injection from this update's volume
  BL_RELOC,                                // relocation of every cascade's probes
  BL_LINK,                                 // parent visibility of cascades 0 .. N-2
  BL_TRACE,
  BL_MERGE,                                // + step (cascade N-1-step of each world)
  BL_TM = BL_MERGE + RTGI_MAX_CASCADES,    // + step: trace and merge of cascade N-1-step in one launch
  BL_SH = BL_TM + RTGI_MAX_CASCADES,
  BL_COUNT
};
// The work-queue counters of a batch: VQ_LANES counters VQ_STRIDE words (1 KiB) apart.
constexpr uint32_t VQ_LANES = 8;
constexpr uint32_t VQ_STRIDE = 256;
// Counter sets: set 0 for the big-triangle voxelisation, set 1 + step for the trace+merge launch of each cascade step.
constexpr size_t VQ_SET = (size_t)VQ_LANES * VQ_STRIDE;
constexpr size_t RTGI_QUEUE_WORDS = VQ_SET * (1 + RTGI_MAX_CASCADES);
struct BatchPlan {
  const TraceK* w;
  int32_t n;
  const uint32_t* start[BL_COUNT];
  uint32_t grid[BL_COUNT];
  uint32_t* queue;  // RTGI_QUEUE_WORDS device words per batch for the work-queue launches (null: static only)
  size_t mark_lds;  // dynamic LDS of the mark kernels (the largest bitmap of the batch; 0: global marks)
  bool stats;       // some world of the batch counts tracer statistics
  bool fused;       // per-cascade trace+merge launches (BL_TM) instead of one trace launch and merge steps
  bool dyn;         // some world has a dynamic layer (the trace reads it)
  bool surf;        // some world has textured surfaces (the voxel passes sample them)
  bool xcd[BL_COUNT];  // the launch's worlds have about equal block counts: contiguous ranges per XCD pay
  bool sh_pairs;       // every world with an SH projection takes it in pairs (sh_pairs): the pair-only kernel
  int32_t min_casc[RTGI_MAX_CASCADES];  // trace+merge step st: the finest cascade any world of the batch traces
};
// Blocks a world needs for each launch (the single-world launches use the same counts).
void batch_blocks(const TraceK& t, uint32_t out[BL_COUNT]);
int launch_vox_build_batch(const BatchPlan& p, void* stream);  // with the pose first
// The dynamic layer's build (single world / batch), after the voxel build.
int launch_dyn_build(const VoxK& k, void* stream);
int launch_dyn_splat(const TraceK& t, void* stream);
int launch_dyn_build_batch(const BatchPlan& p, void* stream);
// dynInject (token obj): after the layer's build, before the bounce volume update that reads the layer.
int launch_dyn_inject(const TraceK& t, void* stream);
int launch_dyn_inject_batch(const BatchPlan& p, void* stream);
int launch_trace_batch(const BatchPlan& p, void* stream);
int launch_merge_batch(const BatchPlan& p, int step, void* stream);
int launch_trace_merge_batch(const BatchPlan& p, int step, void* stream);
int launch_relocate_batch(const BatchPlan& p, void* stream);
int launch_sun_batch(const BatchPlan& p, void* stream);
int launch_bounce_batch(const BatchPlan& p, void* stream);
int launch_sh_batch(const BatchPlan& p, void* stream);

// Direction i of the n irradiance rays of a bounce probe (sfDir, src/core/bounce.ts:155-160: a spherical Fibonacci
// set), evaluated in double and rounded to f32, with the slab-safe reciprocal.
void bn_dir_entry(int i, int n, DirEntry* out);
// Host-side f32 restatement of dirOf (bit-identical to the device's: same IEEE operations, no contraction).
void dir_entry(int di, int res, DirEntry* out);
