// Internals shared by the host files of librtgi (api.cpp: worlds and batches, server.cpp: the request server).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "rtgi.h"
#include "rtgi_internal.h"

#define HIPCHK(call)                                                                                 \
  do {                                                                                               \
    hipError_t e_ = (call);                                                                          \
    if (e_ != hipSuccess) return rtgi_fail(RTGI_E_HIP, "%s: %s (%s:%d)", #call, hipGetErrorString(e_), \
                                           __FILE__, __LINE__);                                      \
  } while (0)

struct rtgi_world {
  int device = 0;
  rtgi_vox_params vp{};
  rtgi_rc_params rp{};
  int NL = 0, G = 0, E = 0, per = 0;
  int origin[RTGI_MAX_LEVELS][3] = {};
  int N = 0;
  int n[RTGI_MAX_CASCADES] = {}, res[RTGI_MAX_CASCADES] = {}, vlev[RTGI_MAX_CASCADES] = {};
  float b0[RTGI_MAX_CASCADES] = {}, b1[RTGI_MAX_CASCADES] = {};
  uint64_t ray_off[RTGI_MAX_CASCADES + 1] = {};
  uint32_t dir_off[RTGI_MAX_CASCADES + 1] = {};
  uint64_t avg_off[RTGI_MAX_CASCADES + 1] = {};  // [c]: cascade c's 2x2 average; [N]: total
  uint64_t probe_off[RTGI_MAX_CASCADES + 1] = {};  // [c]: cascade c's first probe; [N]: total
  int G0[3] = {};
  float omega0[16] = {};
  rtgi_frame frame{};
  bool frame_set = false, scene_set = false;
  uint32_t ntris = 0, nverts = 0, nmat = 0;
  uint32_t cap_tris = 0, cap_verts = 0, cap_mat = 0;
  // device buffers
  uint32_t *hdr = nullptr, *ptr = nullptr, *occ = nullptr, *battr = nullptr, *bcode = nullptr, *rep = nullptr,
           *attr = nullptr, *wslot = nullptr;
  float* plane = nullptr;
  uint32_t *big_vox = nullptr, *big_mark = nullptr;
  float *pos = nullptr, *nrm = nullptr, *mat = nullptr;
  uint32_t* tri = nullptr;
  float* rad = nullptr;
  uint32_t* hits = nullptr;
  // c0 SH in grid-row order, [27][P0] f32 or halves. Two buffers with the history on (the last update's SH is the
  // next one's history), one with RTGI_RC_NO_HISTORY.
  float* shb[2] = {};
  int sh_par = 0;            // the buffer of the last SH written
  bool sh_written = false;   // an SH has been written (a history exists)
  bool hist_prev = false;    // the SH being written blends with shb[1 - sh_par]
  int sh_G0[3] = {};         // c0 grid origin of the SH in shb[sh_par]
  int hist_G0[3] = {};       // c0 grid origin of the history it blends with
  int32_t data_tile = 1;     // probe tile the radiance, averages and probe arrays were last written in
  DirEntry* dirs = nullptr;
  float* avg = nullptr;
  unsigned long long* stats = nullptr;  // 3 x 8 tracer counters (RTGI_RC_STATS)
  float *ppos = nullptr, *pnrm = nullptr;  // relocated probes (RTGI_RC_RELOCATE): float4 each
  uint32_t* pvis = nullptr;               // parent visibility masks (RTGI_RC_RELOCATE)
  float* sunvis = nullptr;                // traced sun visibility per attribute slot (RTGI_RC_SUN_TRACE)
  uint32_t* vslot = nullptr;              // brick << 9 | bit per attribute slot (RTGI_RC_SUN_TRACE or _BOUNCE)
  uint32_t* alb16 = nullptr;              // bf16 albedo per attribute slot: r | g << 16, b | toon << 16 (RTGI_RC_BF16)
  // Dynamic coverage layer (rtgi_world_set_dynamic): triangles [dyn_first, ntris) go to it instead of the bricks.
  uint32_t dyn_first = 0xffffffffu;
  int32_t dyn_cap = 0;
  uint32_t *dyn_hdr = nullptr, *dyn_ptr = nullptr, *dyn_bits = nullptr, *dyn_meta = nullptr, *dyn_mask = nullptr,
           *dyn_data = nullptr;
  // Textured surfaces (rtgi_world_set_surface): base-colour array with its mips, UVs, material words, sRGB tables.
  bool surf_on = false;
  uint32_t* tex = nullptr;
  size_t tex_words = 0;
  float* uv = nullptr;
  uint32_t uv_verts = 0;
  uint32_t* mext = nullptr;
  uint32_t mext_n = 0;
  float* lut = nullptr;
  uint32_t tex_log2 = 0, tex_mips = 0, tex_layer_words = 0, tex_layers = 0;
  uint32_t* bounce = nullptr;             // bounce irradiance per attribute slot, 4 words (RTGI_RC_BOUNCE)
  uint32_t* bnvol[2] = {};                // the two bounce volumes (RTGI_RC_BOUNCE), BN_WORDS each
  int bn_par = 1;                         // the volume the last update wrote
  int bn_org[2][3] = {};                  // origins of the two volumes
  uint32_t bn_frame = 0;                  // volume updates so far
  uint32_t bn_dir_off = 0;                // first bounce ray direction in dirs
  // Moving objects (rtgi_world_set_objects, rtgi_world_set_pose): the rest mesh of their vertex span (positions,
  // then normals), their pose records on the device and two pinned staging copies (the one being filled was last
  // copied two poses ago).
  std::vector<rtgi_object> objs;
  uint32_t pose_first = 0, pose_span = 0;
  float* rest = nullptr;
  PoseRec* d_pose = nullptr;
  PoseRec* h_pose[2] = {};
  hipEvent_t pose_copied[2] = {};
  int pose_cur = 0;
  bool pose_set = false;  // every update poses the objects first
  uint64_t bytes = 0;
  char* arena = nullptr;     // one device allocation holding every fixed buffer above (scene and layer excluded)
  size_t arena_bytes = 0;
  hipStream_t stream = nullptr;
  bool timing = false;
  hipEvent_t ev[5] = {};
  uint32_t last_stages = 0;
  hipEvent_t ext_done = nullptr;  // completion of the last batched update that included this world
  hipEvent_t order_ev = nullptr;  // orders work on a caller's stream after the world's own (asynchronous gather)
  rtgi_tuning tuning{1, -1};
};

// log2 of the probe block edge cascade c is stored in (0: rows), and the device index of probe (x, y, z).
int probe_tile_log2(const rtgi_world* w, int c);
uint32_t probe_index_host(const rtgi_world* w, int c, uint32_t x, uint32_t y, uint32_t z);

// Launch arguments of a world (its current frame, buffers and layout; probes in the order of data_tile).
VoxK vox_args(const rtgi_world* w);
TraceK trace_args(const rtgi_world* w);
// Marks the start of an update's trace: its radiance and probe arrays are written in the tuning's probe order.
inline void trace_begin(rtgi_world* w) { w->data_tile = w->tuning.probe_tile; }
// Advances the bounce volume by one update (the next volume written, its origin from the frame's eye, the update
// counter); called once for each enqueued update that includes the trace stage, before its launch arguments.
void bounce_advance(rtgi_world* w);
// Advances the c0 SH history by one update (the next SH buffer written, the one it blends with and their grid
// origins); called once for each enqueued update that projects SH, before its launch arguments.
void sh_advance(rtgi_world* w);
// Buffers of the c0 SH (1, or 2 with the history on), and floats per buffer.
int sh_buffers(const rtgi_rc_params* rp);
size_t sh_floats_of(const rtgi_world* w);
// Waits for the world's own stream and for the last batched update that included it.
hipError_t world_wait(rtgi_world* w);
// Device bytes a world of these parameters holds once a scene of this size is uploaded (rtgi_world_estimate_bytes).
uint64_t world_bytes(const rtgi_vox_params* vp, const rtgi_rc_params* rp, uint32_t nverts, uint32_t ntris,
                     uint32_t nmat);

// Window origins and c0 grid origin of a frame for these parameters (writeUniform, src/core/voxel.ts:872-875, and the
// dense probe grid), or RTGI_E_INVALID for a frame no world of them accepts (bad light count, eye not finite or too
// far from the origin). rtgi_world_set_frame and the server's submit share it.
int frame_origins(const rtgi_vox_params* vp, const rtgi_rc_params* rp, const rtgi_frame* f,
                  int origin[RTGI_MAX_LEVELS][3], int G0[3]);
// Starts world w's part of a batched update with plan p: advances its per-update state (trace order, bounce volume,
// SH history) and writes its launch descriptor.
void batch_world_begin(rtgi_world* w, const BatchPlan& p, TraceK* out);
// Whether every world with an SH projection takes it in pairs under its update descriptor (desc[i]).
bool plan_sh_pairs(rtgi_world* const* worlds, const TraceK* desc, int n);
// Temporal state a world carries from one update to the next (the SH history and the bounce volume); the server
// saves it when it evicts a world and restores it when the world is admitted again.
struct WorldState {
  bool valid = false;
  std::vector<uint8_t> sh;  // the last SH (device bytes), empty if none was written
  int sh_G0[3] = {};
  std::vector<uint32_t> bnvol[2];
  int bn_par = 1, bn_org[2][3] = {};
  uint32_t bn_frame = 0;
};
int world_save_state(rtgi_world* w, WorldState* st);
int world_restore_state(rtgi_world* w, const WorldState& st);
// Block prefix arrays of a batched launch over worlds[0..n): start holds BL_COUNT * (n + 1) entries (launch-major);
// fills p.n, p.grid, p.mark_lds and p.stats (the caller sets p.w and p.start to device copies).
void plan_blocks(rtgi_world* const* worlds, int n, uint32_t* start, BatchPlan& p);
// The stage launches of one batched update on stream s; ev (5 events or null) brackets the four stages.
int plan_launch(const BatchPlan& p, hipStream_t s, hipEvent_t* ev);
