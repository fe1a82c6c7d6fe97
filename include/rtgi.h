/*
 * rtgi.h -- C ABI of the MI355X (gfx950) per-frame GI update: clipmap voxel rebuild, radiance-cascade probe-ray DDA
 * trace, cascade merge and L2 SH projection.
 *
 * The entry points replace the reference renderer's pass interface (TypeScript factories recording WGSL compute passes
 * into a Dawn encoder). Each declaration cites the reference interface it replaces; paths are relative to the
 * reference repository (owoDra/real-time-gi-lab).
 *
 * Conventions (replacing the reference's synchronous JS throws, src/core/voxel.ts:749, src/core/rc.ts:1022):
 *   - every entry point returns an int status: RTGI_OK (0) or a negative RTGI_E_* code;
 *   - rtgi_last_error() returns a thread-local message for the last failing call of the calling thread;
 *   - GPU-side pool overflow is not an error: it is reported through rtgi_counters (src/core/voxel.ts:908-918);
 *   - host arrays passed in are copied during the call; readback buffers are caller-allocated;
 *   - a world handle is not re-entrant: calls on one handle must not overlap; different handles may be driven from
 *     different threads.
 * Only plain C types cross this boundary (no HIP or torch types): streams are passed as void* (a hipStream_t).
 */
#ifndef RTGI_H
#define RTGI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RTGI_OK 0
#define RTGI_E_INVALID (-1)  /* bad argument or configuration */
#define RTGI_E_HIP (-2)      /* a HIP runtime call failed */
#define RTGI_E_NOMEM (-3)    /* device or host allocation failed, or the memory budget refused it */
#define RTGI_E_STATE (-4)    /* call out of order (e.g. enqueue before a scene was uploaded) */
#define RTGI_E_NODEVICE (-5) /* no gfx950 device visible */

#define RTGI_MAX_LEVELS 8
#define RTGI_MAX_CASCADES 8
#define RTGI_MAX_POINT_LIGHTS 8

/* Pointer-table sentinels, identical to the reference's VOX_EMPTY / VOX_DROP / VOX_MARK (src/core/voxel.ts:64-66). */
#define RTGI_PTR_EMPTY 0xffffffffu
#define RTGI_PTR_DROP 0xfffffffeu
#define RTGI_PTR_MARK 0xfffffffdu
/* Miss value of a per-ray hit id. */
#define RTGI_NO_HIT 0xffffffffu

/* Voxel clipmap parameters; mirrors VoxParams (src/core/voxel.ts:31-38) without the dynamic-layer pool. */
typedef struct rtgi_vox_params {
  double v0;          /* finest voxel size (m); level L has voxels of v0 * 2^L */
  int32_t levels;     /* 1..8 */
  int32_t G;          /* bricks per axis per level (a brick is 8^3 voxels); G^3 must be a multiple of 64 */
  int32_t max_bricks; /* brick pool, all levels */
  int32_t max_voxels; /* voxel attribute pool, all levels */
} rtgi_vox_params;

/* Flags of rtgi_rc_params.flags. */
#define RTGI_RC_RECORD_HITS 1u /* also write the first-hit voxel id of every ray (rtgi_world_read_cascade) */
#define RTGI_RC_NO_MERGE 2u    /* stop after the trace (the cascade buffers hold the traced intervals) */
#define RTGI_RC_RELOCATE 16u   /* voxel relocation 'vrel' (relocateVox, rc.ts:413-472): probes are moved off occupied
                                * voxels along a normal; merge parents and gather probes hidden by voxels are skipped
                                * (linkParents rc.ts:559-566, segOccluded rc.ts:845-869) */
#define RTGI_RC_SUN_TRACE 32u  /* per-voxel sun visibility (voxSun, voxel.ts:631-673) with the shadow-map taps replaced
                                * by voxel shadow rays toward the sun through the clipmap levels; hit shading reads
                                * it instead of the frame's sun_vis */
#define RTGI_RC_BOUNCE 64u     /* multi-bounce (createBounce with DEFAULT_BOUNCE, bounce.ts:33, 312-379): each update
                                * first rebuilds the bounce volume, 32 x 16 x 32 ambient-cube probes 1.2 m apart
                                * around the eye (bnClassify, bnUpdate: 16 rays per probe whose hits read the previous
                                * update's volume, every 2nd update per probe), then stores its irradiance on both
                                * sides of every voxel (bnInject); cascade hits add albedo / pi times it (voxBounce,
                                * voxSide, shadeHitV, rc.ts:107-126, 198-202) */
#define RTGI_RC_OBJECTS 128u /* token obj (rc.ts:79-81, not in the reference's defaults): with RTGI_RC_BOUNCE and a dynamic
                              * layer (rtgi_world_set_dynamic), the layer also splats albedo, its voxels receive the
                              * previous bounce volume's light (dynInject, dynlayer.ts:166-190, dynE of bounce.ts:267-298)
                              * and the volume's irradiance rays read it (bnTraceDyn, bounce.ts:173-178) */
#define RTGI_RC_STATS 8u       /* count traces, DDA steps and step-cap hits per cascade (setStats(1), rc.ts:232-235) */
#define RTGI_RC_FP16 4u        /* precision form: cascade radiance and SH stored as f16 (round to nearest even) at the
                                  reference's own pack points (packRad, src/core/rc.ts:309, 621, 707; projectSH
                                  pack2x16float, rc.ts:770-771); all arithmetic stays f32 */
#define RTGI_RC_BF16 512u      /* precision form (with RTGI_RC_FP16: FP16 radiance): the voxel albedo is kept as bf16 rgb
                                * (round to nearest even from the material's f32 value) beside the reference's RGBA8
                                * word, and hit shading reads it; with RTGI_RC_BOUNCE the injected bounce light of every
                                * voxel (its emission toward the cascades) is stored as bf16 instead of f16. All
                                * arithmetic stays f32 */
#define RTGI_RC_NO_HISTORY 256u /* turns the c0 SH history off. By default (the reference's world probes: HIST_ON =
                                 * !screen, rc.ts:1071, 1112) each update's c0 SH is blended with the SH the probe at
                                 * the same grid point had in the previous update: sh = prev * 0.5 + new * 0.5
                                 * (histBlend, mix with HIST_W = 0.5, rc.ts:193-194, 784-814); probes whose grid point
                                 * the previous update did not hold (the first update, the grid moved with the eye)
                                 * keep their own value. The blended SH is what the next update reads, stored as the SH
                                 * is (f16 in the precision form, as the reference's pack2x16float history). */

/* Radiance-cascade parameters; the subset of RcParams (src/core/rc.ts:42-91) that the world-space voxel tracer uses,
 * with probes on a dense camera-centred grid instead of screen-registered sparse probes (relocation 'none'). */
typedef struct rtgi_rc_params {
  int32_t cascades;   /* N, 1..8 */
  int32_t probes0;    /* c0 probes per axis; cascade c has probes0 >> c per axis; must be a multiple of 2^(N-1) */
  double s0;          /* c0 probe spacing (m); cascade c is spaced s0 * 2^c */
  double k_interval;  /* kInterval: interval [b_c, b_{c+1}] * s0, b_k = k (4^k - 1) / 3 */
  double level_map;   /* levelMap */
  double vox_scale;   /* voxScale */
  double vox_offset;  /* voxOffset: traces start voxOffset voxels along the probe normal (0 without relocation) */
  double vox_skip;    /* voxSkip */
  int32_t res_cap;    /* resCap: res_c = 4 << min(c, resCap) */
  uint32_t flags;     /* RTGI_RC_* */
  double lift;        /* lift (rc.ts:53): relocated probes keep lift * spacing of free space behind them (0.25) */
} rtgi_rc_params;

/* Per-frame inputs; the part of the Frame uniform (src/core/frame.ts:4-23) the path reads, plus the eye that
 * VoxWorld.update takes (src/core/voxel.ts:868-892). */
typedef struct rtgi_frame {
  double eye[3];      /* clipmap and probe-grid centre (world m) */
  float sun_dir[3];   /* towards the sun (normalised by the caller, as packFrame does) */
  float sun_e[3];     /* sun irradiance */
  float sky[3];       /* sky radiance returned by rays that pass the last interval */
  float sun_vis;      /* per-voxel sun visibility used at every hit (1: unshadowed; the shadow map is not in this path) */
  int32_t n_point;    /* point lights (0..8); 0 keeps the reference's sun-only shading */
  float point_pos[RTGI_MAX_POINT_LIGHTS][3];
  float point_i[RTGI_MAX_POINT_LIGHTS][3]; /* intensity (W/sr-like units): irradiance = I * max(n.l, 0) / r^2 */
} rtgi_frame;

/* Counters; the same semantics as VoxWorld.readCounters (src/core/voxel.ts:908-918). big_overflow is always 0 here:
 * the big-triangle lists hold every triangle. */
typedef struct rtgi_counters {
  uint32_t bricks;          /* allocated bricks (min(marked, max_bricks)) */
  uint32_t marked;          /* marked bricks before allocation */
  uint32_t voxels;          /* occupied voxels requested (may exceed max_voxels) */
  uint32_t brick_overflow;  /* bricks dropped by a full pool */
  uint32_t big_overflow;
  uint32_t voxel_overflow;  /* bricks whose voxels found no attribute slots */
  uint32_t bricks_per_level[RTGI_MAX_LEVELS];
  uint32_t big_per_level[RTGI_MAX_LEVELS];  /* triangles handed to the big (one workgroup each) voxel list */
} rtgi_counters;

/* Sizes of a world's buffers, for readback allocation. */
typedef struct rtgi_layout {
  int32_t levels, G, E;                       /* E = 8 G voxels per axis per level */
  int32_t window_origin[RTGI_MAX_LEVELS][3];  /* first brick of each level's window (world brick coordinates) */
  int32_t cascades;
  int32_t probes[RTGI_MAX_CASCADES];          /* probes per axis of each cascade */
  int32_t res[RTGI_MAX_CASCADES];             /* octahedral directions per axis of each cascade */
  int32_t vox_level[RTGI_MAX_CASCADES];       /* voxLevel(s0, c) (src/core/rc.ts:210-213) */
  int32_t grid_origin[3];                     /* absolute c0 grid index of probe (0,0,0) */
  uint64_t rays[RTGI_MAX_CASCADES];           /* probes^3 * res^2 */
  uint64_t total_rays;
  float b0[RTGI_MAX_CASCADES], b1[RTGI_MAX_CASCADES]; /* interval bounds in units of s0 */
  float omega0[16];                           /* solid angle of each c0 texel (octSolidAngles(4), src/core/frame.ts:105-114) */
  uint64_t device_bytes;                      /* device memory held by the world */
} rtgi_layout;

/* Stages for rtgi_world_enqueue_stages (bit mask) and rtgi_world_stage_ms. */
#define RTGI_STAGE_VOX 1u    /* clear .. resolveAttr (src/core/voxel.ts:708, VOX_STAGES without 'sun') */
#define RTGI_STAGE_TRACE 2u  /* giTrace (src/core/rc.ts:1201-1210) */
#define RTGI_STAGE_MERGE 4u  /* giMerge: mergeWorld N-1..0 (src/core/rc.ts:1212-1214) */
#define RTGI_STAGE_SH 8u     /* projectSH (src/core/rc.ts:1216) */
#define RTGI_STAGE_ALL 15u

typedef struct rtgi_world rtgi_world;

/* Message of the calling thread's last failing call ("" if none). */
const char* rtgi_last_error(void);

/* Number of visible HIP devices (0 without a GPU; never fails for that reason). */
int rtgi_device_count(int32_t* n);

/* Library version string. */
const char* rtgi_version(void);

/* octSolidAngles(res, samples) (src/core/frame.ts:105-114): solid angle of each texel of a res x res octahedral map,
 * from a Fibonacci-sphere count in double precision. out: res*res floats. */
int rtgi_oct_solid_angles(int32_t res, int32_t samples, float* out);

/* Creates a world on the current HIP device. Replaces both factories createVoxWorld (src/core/voxel.ts:746-747) and
 * createRc (src/core/rc.ts:1019), called by createRunContext (src/core/rcrun.ts:62-63) and runConfig
 * (src/core/rcrun.ts:269-271). Throws become RTGI_E_INVALID: levels > 8 (voxel.ts:749), cascades > 8 (rc.ts:1022).
 * Also refused (sizes this library addresses in 32 bits): a cascade of 2^32 rays or more, and a cascade c > 0 whose
 * 2x2 averages (n_c^3 res_{c-1}^2 entries of 16 B) reach 4 GiB. */
int rtgi_world_create(const rtgi_vox_params* vox, const rtgi_rc_params* rc, rtgi_world** out);

/* Scene upload: positions and normals (3 floats per vertex), the triangle table (4 u32 per triangle: absolute vertex
 * indices and material, triangleTable src/core/voxel.ts:730-739) and materials (baseColorFactor rgba per material;
 * untextured: the white layer 0 of src/core/scene_gpu.ts:64-68). Replaces uploadScene (src/core/scene_gpu.ts:56) and the
 * tri-table upload (src/core/rcrun.ts:53-59). Every triangle is voxelised (firstDynamicTri = triangleCount). */
int rtgi_world_upload_scene(rtgi_world* w, const float* pos, const float* nrm, uint32_t nverts, const uint32_t* tri,
                            uint32_t ntris, const float* materials, uint32_t nmat);

/* Replaces the posed-vertex writeBuffer (src/core/rcrun.ts:213-214): vertices [first, first+count). */
int rtgi_world_update_vertices(rtgi_world* w, uint32_t first, uint32_t count, const float* pos, const float* nrm);

/* ---- Textured and alpha-tested surfaces (surfaceAt, src/core/voxel.ts:325-336; shadeSample, dynlayer.ts:451-452) ---- */

/* The surface words of one material (struct Material of uploadScene, src/core/scene_gpu.ts:95-103). */
typedef struct rtgi_surface_material {
  uint32_t base_color_layer;  /* Material.layers.x: 0 the white default layer, k >= 1 texel layer k - 1 below */
  uint32_t flags;             /* Material.layers.w: RTGI_SURF_ALPHA_MASK, RTGI_SURF_TOON (other bits ignored) */
  float alpha_cutoff;         /* Material.params.z */
} rtgi_surface_material;
#define RTGI_SURF_ALPHA_MASK 1u /* voxels (and dynamic-layer samples) whose sampled alpha is below the cutoff are empty */
#define RTGI_SURF_TOON 4u       /* the toon flag, stored in the albedo word's alpha (pack4x8unorm(albedo, toon)) */

/* Base-colour textures, per-vertex UVs and per-material surface words of the uploaded scene, the inputs surfaceAt
 * samples (uploadScene's baseColorImage array, uvs buffer and material words, src/core/scene_gpu.ts:56-106; the
 * voxel passes' sampler, voxel.ts:793: linear min/mag/mip filter, repeat addressing). texels: `layers` image layers
 * of size x size RGBA8 texels (rgb sRGB-encoded, alpha linear; row-major, 4 bytes per texel; the array's layers 1..layers,
 * its layer 0 being the white default); size a power of two <= 4096. Mip levels are built on the device as the
 * reference's generateMips builds them (each texel of level m the bilinear sample of level m-1 at its centre: the
 * mean of 2 x 2 texels in linear space, stored sRGB-encoded, scene_gpu.ts:38-54). surfaceAt then samples level
 * min(3 + L, last) bilinearly at the UV interpolated at the voxel centre's closest point: resolveAttr stores
 * texture x baseColorFactor as the albedo and the toon flag as its alpha; with RTGI_SURF_ALPHA_MASK, voxBits leaves
 * a voxel empty whose sampled alpha x factor alpha is below the cutoff; the dynamic layer's samples take the same
 * albedo and alpha test. uvs: nverts x 2 floats (nverts and nmat those of the uploaded scene). uvs NULL: untextured
 * again (every material samples white and none is masked). Allocates; call between updates. */
int rtgi_world_set_surface(rtgi_world* w, const float* uvs, uint32_t nverts, const rtgi_surface_material* mats,
                           uint32_t nmat, const uint8_t* texels, uint32_t size, uint32_t layers);

/* ---- Moving objects, posed on the device (DynamicSpec and posedVertices, src/core/scenes.ts:41-53, 80-102) ---- */

/* A rigid object: vertices [first_vertex, first_vertex + vertex_count) turning about the vertical axis through pivot
 * (DynamicObject, scenes.ts:41; its height serves only the wave and bend deformations, which are not built). */
typedef struct rtgi_object {
  uint32_t first_vertex, vertex_count;
  double pivot[3];
} rtgi_object;

/* Pose of one object (DynamicPose.poses[i], scenes.ts:47): c = cos(rotY) and s = sin(rotY) as the host evaluates
 * them (posedVertices, scenes.ts:89), then the translation offset. */
typedef struct rtgi_pose {
  double c, s;
  double offset[3];
} rtgi_pose;

/* Registers n moving objects (n = 0: none) with their rest mesh, which stays resident on the device: rest_pos and
 * rest_nrm hold 3 floats per vertex of [objects[0].first_vertex, objects[n-1].first_vertex +
 * objects[n-1].vertex_count), the scene.positions / scene.normals that posedVertices reads. Objects come in ascending
 * vertex order without overlap, inside the uploaded scene; vertices between two objects are left alone. No pose
 * applies until rtgi_world_set_pose; a new scene upload removes the objects. Synchronises; allocates. */
int rtgi_world_set_objects(rtgi_world* w, const rtgi_object* objects, uint32_t n, const float* rest_pos,
                           const float* rest_nrm);

/* Poses of the n registered objects for the following updates: from then on every update begins by writing the
 * objects' positions and normals from the rest mesh (posedVertices, scenes.ts:80-102: x = rest - pivot in double,
 * pivot + (c x + s z, y, -s x + c z) + offset and the normal turned the same way, each value rounded to float once).
 * Replaces posedVertices and the writeBuffer of positions and normals per frame (src/core/rcrun.ts:194-217); vertices
 * written by rtgi_world_update_vertices inside an object are posed over. Asynchronous on the world's stream (the
 * poses are copied at once); waits for a batched update of the world still in flight. */
int rtgi_world_set_pose(rtgi_world* w, const rtgi_pose* poses);

/* Replaces VoxWorld.update(eye) (src/core/voxel.ts:896, writeUniform 868-892) and the Frame uniform fields the path reads
 * (packFrame src/core/frame.ts:76-94: sunDir, sunE, sky). Also places the dense probe grid around the eye. */
int rtgi_world_set_frame(rtgi_world* w, const rtgi_frame* f);

/* Enqueues one GI update (voxel rebuild, trace, merge, SH) on `stream` (a hipStream_t; NULL: the world's own stream).
 * Replaces VoxWorld.encode (src/core/voxel.ts:897-899) followed by Rc.encode's giTrace and giMerge passes
 * (src/core/rc.ts:1169-1216) and the queue submit (src/core/rcrun.ts:333). Asynchronous. */
int rtgi_world_enqueue(rtgi_world* w, void* stream);

/* The same restricted to a mask of RTGI_STAGE_* (run in pipeline order). Replaces encodeStaged
 * (src/core/voxel.ts:900-907). With timing on, per-stage HIP events are recorded around each stage. */
int rtgi_world_enqueue_stages(rtgi_world* w, uint32_t stages, void* stream);

/* Per-stage timing with HIP events (replaces timestampWrites, src/core/rcrun.ts:325-326). on: 0/1. */
int rtgi_world_set_timing(rtgi_world* w, int32_t on);
/* Milliseconds of each stage of the last enqueue (after rtgi_world_sync): ms[0] vox, [1] trace, [2] merge, [3] sh. */
int rtgi_world_stage_ms(rtgi_world* w, float ms[4]);

/* Waits for the world's work (replaces the mapAsync waits, src/core/rcrun.ts:349-358). */
int rtgi_world_sync(rtgi_world* w);

/* Replaces VoxWorld.readCounters (src/core/voxel.ts:908-918). Synchronises. */
int rtgi_world_read_counters(rtgi_world* w, rtgi_counters* out);

int rtgi_world_get_layout(const rtgi_world* w, rtgi_layout* out);

/* Canonical (layout-independent) readback of clipmap level L; replaces reading the vox buffer (layout
 * src/core/voxel.ts:11,52-53,753-760). Arrays are indexed by the window cell x + E (y + E z), E = 8 G; any may be NULL:
 *   occ_bits     E^3/32 words: bit (i & 31) of word (i >> 5) set if cell i is occupied;
 *   rep          E^3: representative triangle + 1 (0: empty), the rep rule of src/core/voxel.ts:489 and :571-574;
 *   albedo       E^3: pack4x8unorm(albedo, toon) of occupied cells with an attribute slot, else 0;
 *   normal       E^3: pack2x16snorm(octEnc(n)), else 0;
 *   plane        E^3: offset of the representative's plane from the cell centre (src/core/voxel.ts:619), else 0;
 *   brick_state  G^3: 0 empty, 1 allocated with attributes, 2 dropped by a full pool, 3 allocated without attribute
 *                slots (voxel pool full). Synchronises. */
int rtgi_world_read_clipmap(rtgi_world* w, int32_t level, uint32_t* occ_bits, uint32_t* rep, uint32_t* albedo,
                            uint32_t* normal, float* plane, uint8_t* brick_state);

/* Cascade c after the last update (f16 storage is widened to f32): rgba (4 floats per ray: radiance, transmittance;
 * after the merge a = 1), in ray order
 * probe-major (probe x + n (y + n z)), direction di = u + res v within a probe; hits (if recorded): first-hit voxel id
 * level << 29 | (x + E (y + E z)) of the window cell, RTGI_NO_HIT for a miss. Either pointer may be NULL. Replaces reading
 * the rad buffer (src/core/rc.ts:1064). Synchronises. */
int rtgi_world_read_cascade(rtgi_world* w, int32_t c, float* rgba, uint32_t* hits);

/* Ray query over the clipmap of the last update: n rays from origins o along unit directions d (3 floats each)
 * traced as rays of cascade c (its interval [b_c, b_c+1] s0, level voxLevelFor and skip), shaded as the cascade trace
 * (traceWorld -> traceVox -> voxShade, src/core/rc.ts:216-222, src/core/voxel.ts:121-202) -> rgba (4 floats: radiance,
 * transmittance; f16-rounded in the precision form) and hits (first-hit voxel ids, may be NULL). Synchronous. The
 * reference has no ray query entry point; this one serves error measurements on arbitrary (non-texel) directions. */
int rtgi_world_trace_rays(rtgi_world* w, int32_t c, const float* o, const float* d, uint32_t n, float* rgba,
                          uint32_t* hits);

/* Relocated probes of cascade c after the last update (RTGI_RC_RELOCATE), in grid order x + n (y + n z): positions
 * and normals (3 floats each; normal 0 where no voxel is near, relocateVox rc.ts:413-472) and the merge parents in use
 * (bit k: parent (g >> 1) + (k & 1, k >> 1 & 1, k >> 2 & 1) exists and is not hidden by voxels, linkParents
 * rc.ts:559-566; 0 for the last cascade). Any pointer may be NULL. Replaces reading the probes / parents buffers
 * (rc.ts:1066-1068). Synchronises. */
int rtgi_world_read_probes(rtgi_world* w, int32_t c, float* pos, float* nrm, uint32_t* parents);

/* Dynamic coverage layer (createVoxWorld's firstDynamicTri, src/core/voxel.ts:746-747, 822, and VoxParams.dynBricks;
 * createDynLayer, src/core/dynlayer.ts:650-761). Triangles [first_dynamic_tri, ntris) of the uploaded scene are then
 * not voxelised into the bricks: every update splats them into a layer of dyn_bricks bricks of coverage and mean
 * radiance (mark, parents, alloc, clear, splat, resolve), and the cascade trace reads it in front of each ray's
 * static hit (traceDyn, dynlayer.ts:118-167; traceWorld, rc.ts:224-231). Samples are shaded with the frame's
 * sun_vis, untextured. first_dynamic_tri >= ntris, the default: no layer (every triangle in the bricks). Allocates;
 * call between updates. */
int rtgi_world_set_dynamic(rtgi_world* w, uint32_t first_dynamic_tri, int32_t dyn_bricks);
/* The layer's counters after the last update (DynLayer.readCounters, dynlayer.ts:743-754): allocated bricks, bricks
 * dropped for want of the pool, triangles outside every clipmap. Synchronises. */
int rtgi_world_read_dyn_counters(rtgi_world* w, uint32_t* bricks, uint32_t* brick_overflow, uint32_t* outside);
/* The raw layer after the last update: hdr [64] header words, ptr [levels][G^3] slots, meta [dyn_bricks] pointer
 * index of each slot, mask [dyn_bricks][16] voxels with values, data [dyn_bricks][512][13] words (fixed point 2^-16 at
 * level 0, f32 bits above). Any pointer may be NULL. Synchronises. */
int rtgi_world_read_dynamic(rtgi_world* w, uint32_t* hdr, uint32_t* ptr, uint32_t* meta, uint32_t* mask,
                            uint32_t* data);

/* The bounce volume the last update wrote (RTGI_RC_BOUNCE): texels [32 * 16 * 32 probes, x fastest][6 axes +x, -x,
 * +y, -y, +z, -z][4 halves: rgb irradiance, reach along the axis (0: inactive probe)] and the probe coordinates of
 * probe 0 (origin: probe i sits at (origin + (x, y, z)) * 1.2 m). Replaces reading the bounce texture (bounce.ts:
 * 303-311: six axes stacked along z, the origin in an extra texel). Synchronises. */
int rtgi_world_read_bounce_volume(rtgi_world* w, uint16_t* texels, int32_t origin[3]);

/* Gather (gatherWorld, src/core/rc.ts:871-927; evalSH :819-829) of the last update's irradiance field at n surface
 * points, position pos and unit normal nrm (3 floats each) -> irradiance irr (3 floats each). The c0 probes of the grid
 * cell around a point are weighted by their trilinear weight + 1e-4 times ((side + 1)^2 / 4 + 0.02), side the cosine
 * between the normal and the direction to the probe; E = sum w evalSH(probe, n) / sum w (0 with no probe). With
 * RTGI_GATHER_OCCLUSION, a probe whose segment from the point offset one voxel along the normal crosses an occupied
 * voxel (segOccluded, rc.ts:845-869) is left out, and when every probe is hidden they are retried from two voxels out
 * (the reference's vrel gather). The points stand in for the reference's G-buffer pixels (depth and normal textures;
 * the toon-style receiver operator is out of scope). Synchronous. */
#define RTGI_GATHER_OCCLUSION 1u
int rtgi_world_gather(rtgi_world* w, const float* pos, const float* nrm, uint32_t n, uint32_t flags, float* irr);

/* The gather at frame rate, on device arrays: the caller's G-buffer positions and normals (3 floats per point each)
 * in, 3 floats of irradiance per point out, all in device memory. Same computation, results and errors as
 * rtgi_world_gather, enqueued on `stream` (NULL: the world's own) after the world's last update and before its next
 * one, without allocation or synchronisation. Replaces gatherWorld's dispatch per frame (rc.ts:1218-1220) where the
 * G-buffer already lives on the GPU. The arrays must stay valid until the stream has run the gather. */
int rtgi_world_gather_async(rtgi_world* w, const float* d_pos, const float* d_nrm, uint32_t n, uint32_t flags,
                            float* d_irr, void* stream);

/* rtgi_world_gather for callers that cannot allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). No counterpart in the reference. It allocates nothing: the points and the irradiance are
 * staged in a fixed device array of RTGI_GATHER_SCRATCH_POINTS points (9 floats each, 9 MiB) that each GPU the library
 * runs on holds until the process ends; n above that is refused (RTGI_E_INVALID). Every call on one GPU shares that
 * array, so calls that overlap on one GPU (from several threads) race. Same arguments, results and errors as
 * rtgi_world_gather. Synchronous. */
#define RTGI_GATHER_SCRATCH_POINTS 262144u
int rtgi_world_gather_noalloc(rtgi_world* w, const float* pos, const float* nrm, uint32_t n, uint32_t flags,
                              float* irr);

/* Tracer statistics of the last update (RTGI_RC_STATS): per cascade the world traces, the DDA loop iterations and
 * the traces that reached the 1024-iteration cap (the reference's counts[24..32), [32..40), [64..72), rc.ts:232-235). */
typedef struct rtgi_trace_stats {
  uint64_t traces[RTGI_MAX_CASCADES];
  uint64_t steps[RTGI_MAX_CASCADES];
  uint64_t capped[RTGI_MAX_CASCADES];
} rtgi_trace_stats;
int rtgi_world_read_trace_stats(rtgi_world* w, rtgi_trace_stats* out);

/* SH irradiance of the c0 probes in grid order x + n (y + n z): 27 floats per probe (9 coefficients x rgb, scaled by
 * {pi, 2pi/3, pi/4}), blended with the previous update's SH unless RTGI_RC_NO_HISTORY (histBlend, rc.ts:784-814);
 * replaces reading the sh buffer written by projectSH (src/core/rc.ts:752-774; f16 storage widened to f32).
 * Synchronises. */
int rtgi_world_read_sh(rtgi_world* w, float* sh);

/* Frees the world's device memory (replaces VoxWorld.destroy src/core/voxel.ts:923 and Rc.destroy rc.ts:1234). */
int rtgi_world_destroy(rtgi_world* w);

/* Launch tuning of a world: choices that change speed and device storage order, never results (the readbacks are
 * in the same order whatever the tuning). No reference counterpart (the reference's workgroup sizes are constants).
 *   probe_tile  probes per axis of a trace block: 1 stores probes in grid rows (a wavefront traces 64 probes of a
 *               row), 2 or 4 in blocks of 2^3 / 4^3 (a wavefront traces a compact block of parallel rays); cascades
 *               whose probes per axis are not a power of two at least probe_tile keep rows. Default 1.
 *   fused       1: each cascade is traced and merged in one launch, cascades N-1 .. 0 in sequence (the traced values
 *               are never stored unmerged; the trace stage then includes the merge and the merge stage is empty);
 *               0: one trace launch over all cascades, then the merge steps; -1 (default): fused in batches, not in
 *               single-world updates (a single world's top cascades alone do not fill the GPU). A batch runs fused
 *               when none of its worlds says 0. */
typedef struct rtgi_tuning {
  int32_t probe_tile;
  int32_t fused;
} rtgi_tuning;
int rtgi_world_set_tuning(rtgi_world* w, const rtgi_tuning* t);
int rtgi_world_get_tuning(const rtgi_world* w, rtgi_tuning* t);

/* Device bytes a world of these parameters holds once a scene of nverts / ntris / nmat is uploaded (what
 * rtgi_layout.device_bytes then reads), without allocating: the admission test of a memory budget. No reference
 * counterpart (the reference sizes its GPUBuffers inside createVoxWorld / createRc, voxel.ts:753-767, rc.ts:1061-1072).
 * Errors as rtgi_world_create. */
int rtgi_world_estimate_bytes(const rtgi_vox_params* vox, const rtgi_rc_params* rc, uint32_t nverts, uint32_t ntris,
                              uint32_t nmat, uint64_t* bytes);

/* ---- Batched update of many worlds ----
 * No reference counterpart: the reference updates one world per frame. A batch runs one GI update of each of its
 * worlds with one launch per stage (clipmap stages, trace, each merge step, SH) whatever their sizes; the sequence
 * can be captured once as a hipGraph and replayed. Worlds keep their own buffers and parameters; their frames and
 * scenes are read at each enqueue. A batch holds pointers to its worlds: destroy it before them. */
typedef struct rtgi_batch rtgi_batch;

int rtgi_batch_create(rtgi_world* const* worlds, int32_t n, rtgi_batch** out);
/* Capture the launch sequence as a hipGraph on the next enqueue and replay it (on: 1) or launch directly (0). */
int rtgi_batch_set_graph(rtgi_batch* b, int32_t on);
/* One update of every world of the batch on `stream` (NULL: the batch's own stream). Asynchronous; the worlds' read
 * functions wait for it. */
int rtgi_batch_enqueue(rtgi_batch* b, void* stream);
int rtgi_batch_sync(rtgi_batch* b);
/* Per-stage HIP-event timing of enqueues, direct or graph replays (with timing on, a graph batch replays one graph
 * per stage with the events recorded between them); ms as for rtgi_world_stage_ms, of the last enqueue. */
int rtgi_batch_set_timing(rtgi_batch* b, int32_t on);
int rtgi_batch_stage_ms(rtgi_batch* b, float ms[4]);
int rtgi_batch_destroy(rtgi_batch* b);

/* ---- Request server: continuous batching of GI updates over resident worlds ----
 * Replaces the reference's frame loop (src/core/rcrun.ts:300-340: one world, one command buffer per frame, two frames
 * in flight) for many worlds at once. Worlds are registered with their scene (copied to host memory) and become
 * resident on the device when a request needs them and their bytes (rtgi_world_estimate_bytes) fit in the memory
 * budget; idle resident worlds are evicted least recently used first to make room (an evicted world is rebuilt from
 * its host copy on its next request; every update is a full rebuild, and the state one update hands the next, the c0
 * SH history and the bounce volume, is copied to host memory at eviction and back at admission, so results do not
 * depend on residency). A request's frame is checked at submission as rtgi_world_set_frame checks it.
 * Requests carry a frame (viewpoint and lights). A scheduler thread gathers the queued requests whose worlds are idle
 * and resident or admissible, at most one per world and max_batch per launch, into one batched update (one launch per
 * stage) and starts it on one of `slots` streams as soon as a slot is free, so requests that arrive while others run
 * are admitted as slots and memory free up and their work overlaps on the device. Requests of one world run in
 * submission order. A completion thread waits for each launch, copies the optional per-request outputs, and reports
 * completions through rtgi_server_poll and the optional callback (called on that thread). */
typedef struct rtgi_server rtgi_server;

typedef struct rtgi_server_params {
  uint64_t hbm_budget; /* device bytes for resident worlds (0: 90% of the free device memory at creation) */
  int32_t max_batch;   /* worlds per batched launch (0: 64) */
  int32_t slots;       /* batched launches in flight at once (0: 2) */
} rtgi_server_params;

typedef struct rtgi_request {
  int32_t world;               /* id from rtgi_server_add_world */
  uint64_t tag;                /* the caller's id, returned in the completion */
  rtgi_frame frame;            /* eye and lights of this update */
  float* sh_out;               /* optional (NULL): receives this update's SH (rtgi_world_read_sh layout) */
  rtgi_counters* counters_out; /* optional (NULL): receives this update's clipmap counters */
} rtgi_request;

typedef struct rtgi_completion {
  uint64_t tag;
  int32_t world;
  int32_t status;                /* RTGI_OK, or the RTGI_E_* code of the failed update */
  int32_t batch_size;            /* worlds in the batched launch that ran it */
  double t_submit, t_launch, t_done; /* seconds since the server's creation (steady clock) */
} rtgi_completion;

typedef struct rtgi_server_stats {
  uint64_t submitted, completed, batches, admissions, evictions;
  uint64_t resident_bytes, peak_resident_bytes, budget;
  int32_t worlds, resident_worlds, max_batch_seen, queued;
} rtgi_server_stats;

typedef void (*rtgi_completion_fn)(void* user, const rtgi_completion* c);

int rtgi_server_create(const rtgi_server_params* p, rtgi_server** out);
/* Registers a world (parameters and scene, copied). RTGI_E_NOMEM if its bytes exceed the whole budget. */
int rtgi_server_add_world(rtgi_server* s, const rtgi_vox_params* vox, const rtgi_rc_params* rc, const float* pos,
                          const float* nrm, uint32_t nverts, const uint32_t* tri, uint32_t ntris, const float* materials,
                          uint32_t nmat, int32_t* id);
/* Queues one update. The output pointers must stay valid until the request's completion. */
int rtgi_server_submit(rtgi_server* s, const rtgi_request* r);
/* Up to max completions (oldest first) into out, waiting up to wait_ms (0: no wait, < 0: until one is there). */
int rtgi_server_poll(rtgi_server* s, rtgi_completion* out, int32_t max, int32_t wait_ms, int32_t* n);
/* Completion callback (NULL: none); completions are still delivered to rtgi_server_poll as well. */
int rtgi_server_set_callback(rtgi_server* s, rtgi_completion_fn fn, void* user);
/* Waits until every submitted request has completed. */
int rtgi_server_drain(rtgi_server* s);
int rtgi_server_get_stats(rtgi_server* s, rtgi_server_stats* out);
/* The resident world of id (RTGI_E_STATE if it is not resident); valid until its next eviction. Read it only while
 * no request for it is queued or running. */
int rtgi_server_world(rtgi_server* s, int32_t id, rtgi_world** out);
/* Drains, stops the threads and frees every world. */
int rtgi_server_destroy(rtgi_server* s);

/* ---- Seeded procedural scenes (inputs of the benchmark configurations; no reference counterpart beyond the Builder
 * idiom of src/core/scenes.ts:176-211 and the rigid posing of src/core/scenes.ts:80-102). ---- */
#define RTGI_SCENE_ATRIUM 1 /* 20 x 20 x 10 m atrium: walls, 24 pillars, tessellated arches, 16 moving boxes */
#define RTGI_SCENE_CITY 2   /* city block: seeded buildings that translate and rotate per frame, street furniture */

typedef struct rtgi_scene_params {
  int32_t kind;
  uint64_t seed;
  int32_t frame;       /* pose of the moving parts */
  double extent;       /* city: side of the square block (m); atrium: ignored (0: default) */
  int32_t buildings;   /* city: building count (0: default 64) */
  double detail;       /* tessellation step scale (1: default triangle count) */
} rtgi_scene_params;

/* Motion schedule of a generated object (its DynamicSpec.at, scenes.ts:52): at frame f, rotY = rot0 + w f and
 * offset = amp sin(wt f + ph). */
typedef struct rtgi_motion {
  double rot0, w, wt, ph;
  double amp[3];
} rtgi_motion;

typedef struct rtgi_scene {
  uint32_t nverts, ntris, nmat;
  float* pos;        /* nverts * 3 */
  float* nrm;        /* nverts * 3 */
  uint32_t* tri;     /* ntris * 4 */
  float* materials;  /* nmat * 4 */
  double bounds_lo[3], bounds_hi[3];
  double eye[3];     /* a default eye for the scene */
  float sun_dir[3];
  uint32_t first_dynamic_tri;  /* the moving parts are triangles [first_dynamic_tri, ntris) (atrium: the 16 boxes;
                                * city: ntris, every building moving, rebuilt with the rest) */
  uint32_t nobj;               /* moving objects (atrium: the 16 boxes; city: the buildings), ascending, contiguous */
  rtgi_object* objects;        /* nobj */
  rtgi_motion* motion;         /* nobj: each object's schedule */
  float* rest_pos;             /* the objects' vertices at rest, [objects[0].first_vertex, end of the last) x 3; pos and */
  float* rest_nrm;             /* nrm hold them posed at the frame (rtgi_motion_pose_at, then posedVertices) */
} rtgi_scene;

int rtgi_scene_generate(const rtgi_scene_params* p, rtgi_scene** out);
int rtgi_scene_free(rtgi_scene* s);
/* Poses of n scheduled objects at a frame (c and s by the C library's cos and sin). */
int rtgi_motion_pose_at(const rtgi_motion* m, uint32_t n, double frame, rtgi_pose* out);

#ifdef __cplusplus
}
#endif
#endif /* RTGI_H */
