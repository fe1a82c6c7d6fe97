/*
 * rtgi_napi.c -- Node-API addon over librtgi's C ABI (include/rtgi.h): the thin binding the reference's TypeScript
 * host calls in place of its Dawn compute passes (createVoxWorld, src/core/voxel.ts:746; createRc, src/core/rc.ts:1019;
 * encode/submit, src/core/rcrun.ts:283-298, 333; readCounters, src/core/voxel.ts:908-918).
 *
 * Every function validates its arguments, calls the C ABI and turns a non-zero status into a thrown JS Error carrying
 * rtgi_last_error(), like the reference's synchronous throws. Typed arrays are passed by pointer (copied by the
 * library during the call).
 */
#include <node_api.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "rtgi.h"

#define NAPI_CALL(env, call)                                            \
  do {                                                                  \
    if ((call) != napi_ok) {                                            \
      napi_throw_error((env), NULL, "rtgi addon: N-API call failed");   \
      return NULL;                                                      \
    }                                                                   \
  } while (0)

static napi_value throw_status(napi_env env, int st) {
  char msg[600];
  snprintf(msg, sizeof msg, "rtgi error %d: %s", st, rtgi_last_error());
  napi_throw_error(env, NULL, msg);
  return NULL;
}

static int get_f64(napi_env env, napi_value obj, const char* key, double* out) {
  bool has = false;
  napi_value v;
  if (napi_has_named_property(env, obj, key, &has) != napi_ok || !has) return 0;
  if (napi_get_named_property(env, obj, key, &v) != napi_ok) return 0;
  bool b;
  if (napi_get_value_bool(env, v, &b) == napi_ok) {  /* the reference's switches are booleans (RcParams, rc.ts:42-91) */
    *out = b ? 1.0 : 0.0;
    return 1;
  }
  return napi_get_value_double(env, v, out) == napi_ok;
}

static int get_vec(napi_env env, napi_value obj, const char* key, double* out, int n) {
  bool has = false;
  napi_value arr, e;
  if (napi_has_named_property(env, obj, key, &has) != napi_ok || !has) return 0;
  if (napi_get_named_property(env, obj, key, &arr) != napi_ok) return 0;
  for (int i = 0; i < n; i++) {
    if (napi_get_element(env, arr, (uint32_t)i, &e) != napi_ok) return 0;
    if (napi_get_value_double(env, e, &out[i]) != napi_ok) return 0;
  }
  return 1;
}

static void* typed_data(napi_env env, napi_value v, napi_typedarray_type want, size_t* len) {
  bool is = false;
  napi_typedarray_type t;
  void* data = NULL;
  napi_value ab;
  size_t off;
  if (napi_is_typedarray(env, v, &is) != napi_ok || !is) return NULL;
  if (napi_get_typedarray_info(env, v, &t, len, &data, &ab, &off) != napi_ok || t != want) return NULL;
  return data;
}

static rtgi_world* world_of(napi_env env, napi_value v) {
  void* p = NULL;
  if (napi_get_value_external(env, v, &p) != napi_ok) return NULL;
  return (rtgi_world*)p;
}

static napi_value make_u32(napi_env env, uint32_t x) {
  napi_value v;
  napi_create_uint32(env, x, &v);
  return v;
}

/* version() -> string */
static napi_value js_version(napi_env env, napi_callback_info info) {
  napi_value s;
  (void)info;
  NAPI_CALL(env, napi_create_string_utf8(env, rtgi_version(), NAPI_AUTO_LENGTH, &s));
  return s;
}

/* deviceCount() -> number */
static napi_value js_device_count(napi_env env, napi_callback_info info) {
  int32_t n = 0;
  (void)info;
  int st = rtgi_device_count(&n);
  if (st) return throw_status(env, st);
  return make_u32(env, (uint32_t)n);
}

/* createWorld(vox: {v0, levels, G, maxBricks, maxVoxels}, rc: {cascades, probes0, s0, kInterval, levelMap, voxScale,
 * voxOffset, voxSkip, resCap, recordHits, relocate, lift, sunTrace, bounce, objects, fp16, bf16, history})
 * -> handle */
static napi_value js_create_world(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 2) {
    napi_throw_type_error(env, NULL, "createWorld(vox, rc)");
    return NULL;
  }
  double x;
  rtgi_vox_params vp = {0.2, 5, 32, 4608, 524288};
  if (get_f64(env, argv[0], "v0", &x)) vp.v0 = x;
  if (get_f64(env, argv[0], "levels", &x)) vp.levels = (int32_t)x;
  if (get_f64(env, argv[0], "G", &x)) vp.G = (int32_t)x;
  if (get_f64(env, argv[0], "maxBricks", &x)) vp.max_bricks = (int32_t)x;
  if (get_f64(env, argv[0], "maxVoxels", &x)) vp.max_voxels = (int32_t)x;
  rtgi_rc_params rp = {5, 64, 0.2, 1.0, 1.0, 0.25, 1.0, 0.5, 8, 0, 0.25};
  if (get_f64(env, argv[1], "cascades", &x)) rp.cascades = (int32_t)x;
  if (get_f64(env, argv[1], "probes0", &x)) rp.probes0 = (int32_t)x;
  if (get_f64(env, argv[1], "s0", &x)) rp.s0 = x;
  if (get_f64(env, argv[1], "kInterval", &x)) rp.k_interval = x;
  if (get_f64(env, argv[1], "levelMap", &x)) rp.level_map = x;
  if (get_f64(env, argv[1], "voxScale", &x)) rp.vox_scale = x;
  if (get_f64(env, argv[1], "voxOffset", &x)) rp.vox_offset = x;
  if (get_f64(env, argv[1], "voxSkip", &x)) rp.vox_skip = x;
  if (get_f64(env, argv[1], "resCap", &x)) rp.res_cap = (int32_t)x;
  if (get_f64(env, argv[1], "recordHits", &x) && x != 0) rp.flags |= RTGI_RC_RECORD_HITS;
  if (get_f64(env, argv[1], "lift", &x)) rp.lift = x;
  if (get_f64(env, argv[1], "relocate", &x) && x != 0) rp.flags |= RTGI_RC_RELOCATE;
  if (get_f64(env, argv[1], "sunTrace", &x) && x != 0) rp.flags |= RTGI_RC_SUN_TRACE;
  if (get_f64(env, argv[1], "bounce", &x) && x != 0) rp.flags |= RTGI_RC_BOUNCE;
  if (get_f64(env, argv[1], "objects", &x) && x != 0) rp.flags |= RTGI_RC_OBJECTS;
  if (get_f64(env, argv[1], "fp16", &x) && x != 0) rp.flags |= RTGI_RC_FP16;
  if (get_f64(env, argv[1], "bf16", &x) && x != 0) rp.flags |= RTGI_RC_BF16;
  /* history: false (the reference's screen-probe setting, HIST_ON = !screen, rc.ts:1071) turns the c0 history off */
  if (get_f64(env, argv[1], "history", &x) && x == 0) rp.flags |= RTGI_RC_NO_HISTORY;
  rtgi_world* w = NULL;
  int st = rtgi_world_create(&vp, &rp, &w);
  if (st) return throw_status(env, st);
  napi_value h;
  NAPI_CALL(env, napi_create_external(env, w, NULL, NULL, &h));
  return h;
}

/* uploadScene(h, pos: Float32Array, nrm: Float32Array, tri: Uint32Array, materials: Float32Array) */
static napi_value js_upload_scene(napi_env env, napi_callback_info info) {
  size_t argc = 5;
  napi_value argv[5];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  rtgi_world* w = argc == 5 ? world_of(env, argv[0]) : NULL;
  size_t np = 0, nn = 0, nt = 0, nm = 0;
  float* pos = typed_data(env, argv[1], napi_float32_array, &np);
  float* nrm = typed_data(env, argv[2], napi_float32_array, &nn);
  uint32_t* tri = typed_data(env, argv[3], napi_uint32_array, &nt);
  float* mat = typed_data(env, argv[4], napi_float32_array, &nm);
  if (!w || !pos || !nrm || !tri || !mat || np != nn || np % 3 || nt % 4 || nm % 4) {
    napi_throw_type_error(env, NULL, "uploadScene(h, Float32Array pos, Float32Array nrm, Uint32Array tri, Float32Array materials)");
    return NULL;
  }
  int st = rtgi_world_upload_scene(w, pos, nrm, (uint32_t)(np / 3), tri, (uint32_t)(nt / 4), mat, (uint32_t)(nm / 4));
  if (st) return throw_status(env, st);
  return NULL;
}

/* setFrame(h, {eye: [3], sunDir: [3], sunE: [3], sky: [3], sunVis}) */
static napi_value js_set_frame(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  rtgi_world* w = argc == 2 ? world_of(env, argv[0]) : NULL;
  if (!w) {
    napi_throw_type_error(env, NULL, "setFrame(h, frame)");
    return NULL;
  }
  rtgi_frame f;
  memset(&f, 0, sizeof f);
  double v[3] = {0, 0, 0};
  if (!get_vec(env, argv[1], "eye", v, 3)) {
    napi_throw_type_error(env, NULL, "frame.eye: [x, y, z] required");
    return NULL;
  }
  for (int a = 0; a < 3; a++) f.eye[a] = v[a];
  double sd[3] = {-0.15, 1.0, 0.1}, se[3] = {9, 8.5, 7.8}, sk[3] = {0.45, 0.6, 0.9}, vis = 1.0;
  get_vec(env, argv[1], "sunDir", sd, 3);
  get_vec(env, argv[1], "sunE", se, 3);
  get_vec(env, argv[1], "sky", sk, 3);
  get_f64(env, argv[1], "sunVis", &vis);
  double l = sd[0] * sd[0] + sd[1] * sd[1] + sd[2] * sd[2];
  l = l > 0 ? 1.0 / __builtin_sqrt(l) : 0;  /* packFrame normalises sunDir (src/core/frame.ts:84) */
  for (int a = 0; a < 3; a++) {
    f.sun_dir[a] = (float)(sd[a] * l);
    f.sun_e[a] = (float)se[a];
    f.sky[a] = (float)sk[a];
  }
  f.sun_vis = (float)vis;
  int st = rtgi_world_set_frame(w, &f);
  if (st) return throw_status(env, st);
  return NULL;
}

static napi_value one_world_call(napi_env env, napi_callback_info info, int which) {
  size_t argc = 1;
  napi_value argv[1];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  rtgi_world* w = argc == 1 ? world_of(env, argv[0]) : NULL;
  if (!w) {
    napi_throw_type_error(env, NULL, "expected a world handle");
    return NULL;
  }
  int st = which == 0 ? rtgi_world_enqueue(w, NULL) : which == 1 ? rtgi_world_sync(w) : rtgi_world_destroy(w);
  if (st) return throw_status(env, st);
  return NULL;
}
static napi_value js_encode(napi_env env, napi_callback_info info) { return one_world_call(env, info, 0); }
static napi_value js_sync(napi_env env, napi_callback_info info) { return one_world_call(env, info, 1); }
static napi_value js_destroy(napi_env env, napi_callback_info info) { return one_world_call(env, info, 2); }

/* readCounters(h) -> {bricks, voxels, brickOverflow, bigOverflow, voxelOverflow, bricksPerLevel, bigPerLevel} */
static napi_value js_read_counters(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  rtgi_world* w = argc == 1 ? world_of(env, argv[0]) : NULL;
  if (!w) {
    napi_throw_type_error(env, NULL, "readCounters(h)");
    return NULL;
  }
  rtgi_counters c;
  rtgi_layout lay;
  int st = rtgi_world_read_counters(w, &c);
  if (!st) st = rtgi_world_get_layout(w, &lay);
  if (st) return throw_status(env, st);
  napi_value o, a1, a2;
  NAPI_CALL(env, napi_create_object(env, &o));
  napi_set_named_property(env, o, "bricks", make_u32(env, c.bricks));
  napi_set_named_property(env, o, "voxels", make_u32(env, c.voxels));
  napi_set_named_property(env, o, "brickOverflow", make_u32(env, c.brick_overflow));
  napi_set_named_property(env, o, "bigOverflow", make_u32(env, c.big_overflow));
  napi_set_named_property(env, o, "voxelOverflow", make_u32(env, c.voxel_overflow));
  NAPI_CALL(env, napi_create_array_with_length(env, (size_t)lay.levels, &a1));
  NAPI_CALL(env, napi_create_array_with_length(env, (size_t)lay.levels, &a2));
  for (int L = 0; L < lay.levels; L++) {
    napi_set_element(env, a1, (uint32_t)L, make_u32(env, c.bricks_per_level[L]));
    napi_set_element(env, a2, (uint32_t)L, make_u32(env, c.big_per_level[L]));
  }
  napi_set_named_property(env, o, "bricksPerLevel", a1);
  napi_set_named_property(env, o, "bigPerLevel", a2);
  return o;
}

/* readSh(h) -> Float32Array (27 per c0 probe) ; readCascade(h, c) -> Float32Array (4 per ray) */
static napi_value js_read_sh(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  rtgi_world* w = argc == 1 ? world_of(env, argv[0]) : NULL;
  rtgi_layout lay;
  if (!w || rtgi_world_get_layout(w, &lay)) {
    napi_throw_type_error(env, NULL, "readSh(h)");
    return NULL;
  }
  size_t np = (size_t)lay.probes[0] * lay.probes[0] * lay.probes[0];
  void* data;
  napi_value ab, ta;
  NAPI_CALL(env, napi_create_arraybuffer(env, np * 27 * 4, &data, &ab));
  int st = rtgi_world_read_sh(w, (float*)data);
  if (st) return throw_status(env, st);
  NAPI_CALL(env, napi_create_typedarray(env, napi_float32_array, np * 27, ab, 0, &ta));
  return ta;
}

static napi_value js_read_cascade(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  rtgi_world* w = argc == 2 ? world_of(env, argv[0]) : NULL;
  int32_t c = -1;
  rtgi_layout lay;
  if (!w || napi_get_value_int32(env, argv[1], &c) != napi_ok || rtgi_world_get_layout(w, &lay) || c < 0 ||
      c >= lay.cascades) {
    napi_throw_type_error(env, NULL, "readCascade(h, c)");
    return NULL;
  }
  size_t n = (size_t)lay.rays[c] * 4;
  void* data;
  napi_value ab, ta;
  NAPI_CALL(env, napi_create_arraybuffer(env, n * 4, &data, &ab));
  int st = rtgi_world_read_cascade(w, c, (float*)data, NULL);
  if (st) return throw_status(env, st);
  NAPI_CALL(env, napi_create_typedarray(env, napi_float32_array, n, ab, 0, &ta));
  return ta;
}

static napi_value typed(napi_env env, napi_typedarray_type t, size_t n, size_t esize, void** data) {
  napi_value ab, ta;
  if (napi_create_arraybuffer(env, n * esize, data, &ab) != napi_ok) return NULL;
  if (napi_create_typedarray(env, t, n, ab, 0, &ta) != napi_ok) return NULL;
  return ta;
}

/* gather(h, pos: Float32Array, nrm: Float32Array, occlusion) -> Float32Array irradiance (3 per point); replaces
 * gatherWorld (src/core/rc.ts:871-927) over G-buffer pixels */
static napi_value js_gather(napi_env env, napi_callback_info info) {
  size_t argc = 4;
  napi_value argv[4];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  rtgi_world* w = argc >= 3 ? world_of(env, argv[0]) : NULL;
  size_t np = 0, nn = 0;
  float* pos = w ? typed_data(env, argv[1], napi_float32_array, &np) : NULL;
  float* nrm = w ? typed_data(env, argv[2], napi_float32_array, &nn) : NULL;
  bool occl = false;
  if (argc == 4) napi_get_value_bool(env, argv[3], &occl);
  if (!w || !pos || !nrm || np != nn || np % 3) {
    napi_throw_type_error(env, NULL, "gather(h, Float32Array pos, Float32Array nrm, occlusion)");
    return NULL;
  }
  void* data;
  napi_value ta = typed(env, napi_float32_array, np, 4, &data);
  if (!ta) return NULL;
  int st = rtgi_world_gather(w, pos, nrm, (uint32_t)(np / 3), occl ? RTGI_GATHER_OCCLUSION : 0u, (float*)data);
  if (st) return throw_status(env, st);
  return ta;
}

/* readProbes(h, c) -> {pos: Float32Array, nrm: Float32Array, parents: Uint32Array} (relocated probes of cascade c) */
static napi_value js_read_probes(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  rtgi_world* w = argc == 2 ? world_of(env, argv[0]) : NULL;
  int32_t c = -1;
  rtgi_layout lay;
  if (!w || napi_get_value_int32(env, argv[1], &c) != napi_ok || rtgi_world_get_layout(w, &lay) || c < 0 ||
      c >= lay.cascades) {
    napi_throw_type_error(env, NULL, "readProbes(h, c)");
    return NULL;
  }
  size_t n = (size_t)lay.probes[c] * lay.probes[c] * lay.probes[c];
  void *p, *q, *r;
  napi_value tp = typed(env, napi_float32_array, n * 3, 4, &p), tq = typed(env, napi_float32_array, n * 3, 4, &q),
             tr = typed(env, napi_uint32_array, n, 4, &r), o;
  if (!tp || !tq || !tr) return NULL;
  int st = rtgi_world_read_probes(w, c, (float*)p, (float*)q, (uint32_t*)r);
  if (st) return throw_status(env, st);
  NAPI_CALL(env, napi_create_object(env, &o));
  napi_set_named_property(env, o, "pos", tp);
  napi_set_named_property(env, o, "nrm", tq);
  napi_set_named_property(env, o, "parents", tr);
  return o;
}

/* readBounceVolume(h) -> {texels: Uint16Array (halves: probes x 6 axes x rgba), origin: Int32Array(3)} */
static napi_value js_read_bounce_volume(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  rtgi_world* w = argc == 1 ? world_of(env, argv[0]) : NULL;
  if (!w) {
    napi_throw_type_error(env, NULL, "readBounceVolume(h)");
    return NULL;
  }
  size_t n = (size_t)32 * 16 * 32 * 6 * 4;
  void *t, *g;
  napi_value tt = typed(env, napi_uint16_array, n, 2, &t), tg = typed(env, napi_int32_array, 3, 4, &g), o;
  if (!tt || !tg) return NULL;
  int st = rtgi_world_read_bounce_volume(w, (uint16_t*)t, (int32_t*)g);
  if (st) return throw_status(env, st);
  NAPI_CALL(env, napi_create_object(env, &o));
  napi_set_named_property(env, o, "texels", tt);
  napi_set_named_property(env, o, "origin", tg);
  return o;
}

/* setDynamic(h, firstDynamicTri, dynBricks): triangles from firstDynamicTri on go to the dynamic coverage layer */
static napi_value js_set_dynamic(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  rtgi_world* w = argc == 3 ? world_of(env, argv[0]) : NULL;
  double first = 0, bricks = 0;
  if (!w || napi_get_value_double(env, argv[1], &first) != napi_ok ||
      napi_get_value_double(env, argv[2], &bricks) != napi_ok) {
    napi_throw_type_error(env, NULL, "setDynamic(h, firstDynamicTri, dynBricks)");
    return NULL;
  }
  int st = rtgi_world_set_dynamic(w, (uint32_t)first, (int32_t)bricks);
  if (st) return throw_status(env, st);
  return NULL;
}

/* setObjects(h, objects: [{firstVertex, vertexCount, pivot: [3]}], restPos: Float32Array, restNrm: Float32Array):
 * the DynamicSpec's objects (src/core/scenes.ts:41) and the rest mesh of their vertex span */
static napi_value js_set_objects(napi_env env, napi_callback_info info) {
  size_t argc = 4;
  napi_value argv[4], e;
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  rtgi_world* w = argc == 4 ? world_of(env, argv[0]) : NULL;
  uint32_t n = 0;
  size_t np = 0, nn = 0;
  float* rp = argc == 4 ? typed_data(env, argv[2], napi_float32_array, &np) : NULL;
  float* rn = argc == 4 ? typed_data(env, argv[3], napi_float32_array, &nn) : NULL;
  if (!w || napi_get_array_length(env, argv[1], &n) != napi_ok || (n && (!rp || !rn || np != nn || np % 3))) {
    napi_throw_type_error(env, NULL, "setObjects(h, [{firstVertex, vertexCount, pivot}], Float32Array restPos, "
                                     "Float32Array restNrm)");
    return NULL;
  }
  rtgi_object* objs = (rtgi_object*)calloc(n ? n : 1, sizeof(rtgi_object));
  if (!objs) {
    napi_throw_error(env, NULL, "rtgi addon: out of memory");
    return NULL;
  }
  int ok = 1;
  for (uint32_t i = 0; i < n && ok; i++) {
    double f = 0, c = 0;
    ok = napi_get_element(env, argv[1], i, &e) == napi_ok && get_f64(env, e, "firstVertex", &f) &&
         get_f64(env, e, "vertexCount", &c) && get_vec(env, e, "pivot", objs[i].pivot, 3);
    objs[i].first_vertex = (uint32_t)f;
    objs[i].vertex_count = (uint32_t)c;
  }
  /* the rest arrays must cover the span the objects name */
  if (ok && n && (uint64_t)objs[n - 1].first_vertex + objs[n - 1].vertex_count - objs[0].first_vertex > np / 3) ok = 0;
  int st = ok ? rtgi_world_set_objects(w, objs, n, rp, rn) : 0;
  free(objs);
  if (!ok) {
    napi_throw_type_error(env, NULL, "setObjects: bad object record or rest arrays shorter than the objects' span");
    return NULL;
  }
  if (st) return throw_status(env, st);
  return NULL;
}

/* setSurface(h, uvs: Float32Array | null, materials: [{baseColorLayer, flags, alphaCutoff}], texels: Uint8Array, size,
 * layers): what surfaceAt samples (uploadScene's uvs buffer, baseColor array and material words,
 * src/core/scene_gpu.ts:56-106); uvs null: untextured again. */
static napi_value js_set_surface(napi_env env, napi_callback_info info) {
  size_t argc = 6;
  napi_value argv[6], e;
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  rtgi_world* w = argc >= 2 ? world_of(env, argv[0]) : NULL;
  napi_valuetype t = napi_undefined;
  if (w) napi_typeof(env, argv[1], &t);
  if (w && (t == napi_null || t == napi_undefined)) {
    int st = rtgi_world_set_surface(w, NULL, 0, NULL, 0, NULL, 1, 0);
    if (st) return throw_status(env, st);
    return NULL;
  }
  size_t nuv = 0, ntex = 0;
  uint32_t n = 0;
  double size = 0, layers = 0;
  float* uv = argc == 6 ? typed_data(env, argv[1], napi_float32_array, &nuv) : NULL;
  uint8_t* tex = argc == 6 ? typed_data(env, argv[3], napi_uint8_array, &ntex) : NULL;
  int ok = w && uv && nuv % 2 == 0 && napi_get_array_length(env, argv[2], &n) == napi_ok &&
           napi_get_value_double(env, argv[4], &size) == napi_ok && napi_get_value_double(env, argv[5], &layers) == napi_ok &&
           size >= 1 && size <= 4096 && layers >= 0 && layers < 65536 && (tex || layers == 0) &&
           (double)ntex >= size * size * 4 * layers;
  if (!ok) {
    napi_throw_type_error(env, NULL, "setSurface(h, Float32Array uvs, [{baseColorLayer, flags, alphaCutoff}], "
                                     "Uint8Array texels (layers x size x size x 4), size, layers)");
    return NULL;
  }
  rtgi_surface_material* m = (rtgi_surface_material*)calloc(n ? n : 1, sizeof(rtgi_surface_material));
  if (!m) {
    napi_throw_error(env, NULL, "rtgi addon: out of memory");
    return NULL;
  }
  for (uint32_t i = 0; i < n && ok; i++) {
    double l = 0, f = 0, c = 0;
    ok = napi_get_element(env, argv[2], i, &e) == napi_ok && get_f64(env, e, "baseColorLayer", &l) &&
         get_f64(env, e, "flags", &f) && get_f64(env, e, "alphaCutoff", &c);
    m[i].base_color_layer = (uint32_t)l;
    m[i].flags = (uint32_t)f;
    m[i].alpha_cutoff = (float)c;
  }
  int st = ok ? rtgi_world_set_surface(w, uv, (uint32_t)(nuv / 2), m, n, tex, (uint32_t)size, (uint32_t)layers) : 0;
  free(m);
  if (!ok) {
    napi_throw_type_error(env, NULL, "setSurface: bad material record");
    return NULL;
  }
  if (st) return throw_status(env, st);
  return NULL;
}

/* setPose(h, poses: [{c, s, offset: [3]}]): each object's cos(rotY), sin(rotY) and offset, as posedVertices forms
 * them (src/core/scenes.ts:87-89) */
static napi_value js_set_pose(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2], e;
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  rtgi_world* w = argc == 2 ? world_of(env, argv[0]) : NULL;
  uint32_t n = 0;
  if (!w || napi_get_array_length(env, argv[1], &n) != napi_ok) {
    napi_throw_type_error(env, NULL, "setPose(h, [{c, s, offset}])");
    return NULL;
  }
  rtgi_pose* ps = (rtgi_pose*)calloc(n ? n : 1, sizeof(rtgi_pose));
  if (!ps) {
    napi_throw_error(env, NULL, "rtgi addon: out of memory");
    return NULL;
  }
  int ok = 1;
  for (uint32_t i = 0; i < n && ok; i++)
    ok = napi_get_element(env, argv[1], i, &e) == napi_ok && get_f64(env, e, "c", &ps[i].c) &&
         get_f64(env, e, "s", &ps[i].s) && get_vec(env, e, "offset", ps[i].offset, 3);
  int st = ok ? rtgi_world_set_pose(w, ps) : 0;
  free(ps);
  if (!ok) {
    napi_throw_type_error(env, NULL, "setPose: each pose needs c, s and offset [3]");
    return NULL;
  }
  if (st) return throw_status(env, st);
  return NULL;
}

/* readDynCounters(h) -> {bricks, brickOverflow, outside} (DynLayer.readCounters) */
static napi_value js_read_dyn_counters(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1], o;
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  rtgi_world* w = argc == 1 ? world_of(env, argv[0]) : NULL;
  if (!w) {
    napi_throw_type_error(env, NULL, "readDynCounters(h)");
    return NULL;
  }
  uint32_t b = 0, ov = 0, out = 0;
  int st = rtgi_world_read_dyn_counters(w, &b, &ov, &out);
  if (st) return throw_status(env, st);
  NAPI_CALL(env, napi_create_object(env, &o));
  napi_set_named_property(env, o, "bricks", make_u32(env, b));
  napi_set_named_property(env, o, "brickOverflow", make_u32(env, ov));
  napi_set_named_property(env, o, "outside", make_u32(env, out));
  return o;
}

/* generateScene(kind: 'atrium'|'city', seed, frame) -> {pos, nrm, tri, materials, eye, sunDir, firstDynamicTri,
 * objects: [{firstVertex, vertexCount, pivot}], motion: [{rot0, w, wt, ph, amp}], restPos, restNrm} */
static napi_value js_generate_scene(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  char kind[16] = "";
  size_t kl = 0;
  double seed = 1, frame = 0;
  if (argc < 1 || napi_get_value_string_utf8(env, argv[0], kind, sizeof kind, &kl) != napi_ok) {
    napi_throw_type_error(env, NULL, "generateScene(kind, seed, frame)");
    return NULL;
  }
  if (argc > 1) napi_get_value_double(env, argv[1], &seed);
  if (argc > 2) napi_get_value_double(env, argv[2], &frame);
  rtgi_scene_params p = {strcmp(kind, "city") == 0 ? RTGI_SCENE_CITY : RTGI_SCENE_ATRIUM, (uint64_t)seed, (int32_t)frame,
                         0.0, 0, 1.0};
  rtgi_scene* s = NULL;
  int st = rtgi_scene_generate(&p, &s);
  if (st) return throw_status(env, st);
  napi_value o, ab, ta, arr, e;
  void* data;
  NAPI_CALL(env, napi_create_object(env, &o));
  struct { const char* name; void* src; size_t n; napi_typedarray_type t; } f[4] = {
      {"pos", s->pos, (size_t)s->nverts * 3, napi_float32_array},
      {"nrm", s->nrm, (size_t)s->nverts * 3, napi_float32_array},
      {"tri", s->tri, (size_t)s->ntris * 4, napi_uint32_array},
      {"materials", s->materials, (size_t)s->nmat * 4, napi_float32_array}};
  for (int i = 0; i < 4; i++) {
    if (napi_create_arraybuffer(env, f[i].n * 4, &data, &ab) != napi_ok) break;
    memcpy(data, f[i].src, f[i].n * 4);
    napi_create_typedarray(env, f[i].t, f[i].n, ab, 0, &ta);
    napi_set_named_property(env, o, f[i].name, ta);
  }
  napi_create_array_with_length(env, 3, &arr);
  for (int a = 0; a < 3; a++) {
    napi_create_double(env, s->eye[a], &e);
    napi_set_element(env, arr, (uint32_t)a, e);
  }
  napi_set_named_property(env, o, "eye", arr);
  napi_create_array_with_length(env, 3, &arr);
  for (int a = 0; a < 3; a++) {
    napi_create_double(env, s->sun_dir[a], &e);
    napi_set_element(env, arr, (uint32_t)a, e);
  }
  napi_set_named_property(env, o, "sunDir", arr);
  napi_set_named_property(env, o, "firstDynamicTri", make_u32(env, s->first_dynamic_tri));
  napi_value objs, mot, rec, pv;
  napi_create_array_with_length(env, s->nobj, &objs);
  napi_create_array_with_length(env, s->nobj, &mot);
  for (uint32_t i = 0; i < s->nobj; i++) {
    napi_create_object(env, &rec);
    napi_set_named_property(env, rec, "firstVertex", make_u32(env, s->objects[i].first_vertex));
    napi_set_named_property(env, rec, "vertexCount", make_u32(env, s->objects[i].vertex_count));
    napi_create_array_with_length(env, 3, &pv);
    for (int a = 0; a < 3; a++) {
      napi_create_double(env, s->objects[i].pivot[a], &e);
      napi_set_element(env, pv, (uint32_t)a, e);
    }
    napi_set_named_property(env, rec, "pivot", pv);
    napi_set_element(env, objs, i, rec);
    const rtgi_motion* m = &s->motion[i];
    napi_create_object(env, &rec);
    const char* keys[4] = {"rot0", "w", "wt", "ph"};
    const double vals[4] = {m->rot0, m->w, m->wt, m->ph};
    for (int k = 0; k < 4; k++) {
      napi_create_double(env, vals[k], &e);
      napi_set_named_property(env, rec, keys[k], e);
    }
    napi_create_array_with_length(env, 3, &pv);
    for (int a = 0; a < 3; a++) {
      napi_create_double(env, m->amp[a], &e);
      napi_set_element(env, pv, (uint32_t)a, e);
    }
    napi_set_named_property(env, rec, "amp", pv);
    napi_set_element(env, mot, i, rec);
  }
  napi_set_named_property(env, o, "objects", objs);
  napi_set_named_property(env, o, "motion", mot);
  const size_t span = s->nobj ? (size_t)s->objects[s->nobj - 1].first_vertex + s->objects[s->nobj - 1].vertex_count -
                                    s->objects[0].first_vertex
                              : 0;
  const char* rk[2] = {"restPos", "restNrm"};
  float* rs[2] = {s->rest_pos, s->rest_nrm};
  for (int i = 0; i < 2; i++) {
    if (napi_create_arraybuffer(env, span * 12, &data, &ab) != napi_ok) break;
    if (span) memcpy(data, rs[i], span * 12);
    napi_create_typedarray(env, napi_float32_array, span * 3, ab, 0, &ta);
    napi_set_named_property(env, o, rk[i], ta);
  }
  rtgi_scene_free(s);
  return o;
}

static napi_value init(napi_env env, napi_value exports) {
  napi_property_descriptor d[] = {
      {"version", 0, js_version, 0, 0, 0, napi_default, 0},
      {"deviceCount", 0, js_device_count, 0, 0, 0, napi_default, 0},
      {"createWorld", 0, js_create_world, 0, 0, 0, napi_default, 0},
      {"uploadScene", 0, js_upload_scene, 0, 0, 0, napi_default, 0},
      {"setFrame", 0, js_set_frame, 0, 0, 0, napi_default, 0},
      {"encode", 0, js_encode, 0, 0, 0, napi_default, 0},
      {"sync", 0, js_sync, 0, 0, 0, napi_default, 0},
      {"readCounters", 0, js_read_counters, 0, 0, 0, napi_default, 0},
      {"readCascade", 0, js_read_cascade, 0, 0, 0, napi_default, 0},
      {"readSh", 0, js_read_sh, 0, 0, 0, napi_default, 0},
      {"gather", 0, js_gather, 0, 0, 0, napi_default, 0},
      {"readProbes", 0, js_read_probes, 0, 0, 0, napi_default, 0},
      {"readBounceVolume", 0, js_read_bounce_volume, 0, 0, 0, napi_default, 0},
      {"setDynamic", 0, js_set_dynamic, 0, 0, 0, napi_default, 0},
      {"setObjects", 0, js_set_objects, 0, 0, 0, napi_default, 0},
      {"setPose", 0, js_set_pose, 0, 0, 0, napi_default, 0},
      {"setSurface", 0, js_set_surface, 0, 0, 0, napi_default, 0},
      {"readDynCounters", 0, js_read_dyn_counters, 0, 0, 0, napi_default, 0},
      {"destroy", 0, js_destroy, 0, 0, 0, napi_default, 0},
      {"generateScene", 0, js_generate_scene, 0, 0, 0, napi_default, 0},
  };
  if (napi_define_properties(env, exports, sizeof d / sizeof d[0], d) != napi_ok) return NULL;
  return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)
