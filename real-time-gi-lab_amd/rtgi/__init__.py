"""Host-side mirror of the reference renderer's pass interface for the per-frame GI update, over librtgi's C ABI.

Reference interface (owoDra/real-time-gi-lab) and what stands in for it here:
  VoxParams / DEFAULT_VOX / voxParams      src/core/voxel.ts:31-50     -> VoxParams, DEFAULT_VOX, vox_params
  RcParams / DEFAULT_PARAMS                src/core/rc.ts:42-91        -> RcParams, DEFAULT_PARAMS (world-tracer subset)
  createVoxWorld(...) -> VoxWorld          src/core/voxel.ts:746-925   -> World (voxel half: update, encode_staged,
                                                                          read_counters, used_bytes)
  createRc(...) -> Rc                      src/core/rc.ts:1019-1236    -> World (cascade half: read_cascade, read_sh)
  SUN_E, SKY                               src/core/rcrun.ts:26-27     -> SUN_E, SKY
  uploadScene's uvs, baseColor array and   src/core/scene_gpu.ts:56-106 -> Surface, World.set_surface
  material words (what surfaceAt samples)
  parseConfig / named configs              src/runner/configs.ts:47-73 -> CONFIGS, config()

Errors follow the reference's: bad configuration raises synchronously (RtgiError, like its JS throws); brick pool
overflow is not an error and shows in read_counters().
"""
import ctypes as C
import dataclasses
import math

import numpy as np

from . import _lib as _L
from ._lib import RtgiError, check, lib, ptr

__all__ = ["pack_irradiance", "write_irradiance", "read_irradiance", "VoxParams", "RcParams", "DEFAULT_VOX", "DEFAULT_PARAMS", "vox_params", "Frame", "Scene", "Surface", "SURF_ALPHA_MASK", "SURF_TOON", "World", "Batch",
           "Server", "estimate_bytes", "generate_scene", "CONFIGS", "config", "RtgiError", "SUN_E", "SKY",
           "device_count", "oct_solid_angles", "lib"]

SUN_E = (9.0, 8.5, 7.8)
SKY = (0.45, 0.6, 0.9)


@dataclasses.dataclass
class VoxParams:
    """VoxParams (src/core/voxel.ts:31-38) without the dynamic-layer pool."""
    v0: float = 0.2
    levels: int = 5
    G: int = 32
    maxBricks: int = 4608
    maxVoxels: int = 524288


DEFAULT_VOX = VoxParams()


def vox_params(**o):
    """voxParams (src/core/voxel.ts:46-50): DEFAULT_VOX with overrides; unknown keys raise."""
    for k in o:
        if k not in VoxParams.__dataclass_fields__:
            raise ValueError(f"unknown voxel parameter {k}")
    return dataclasses.replace(DEFAULT_VOX, **o)


@dataclasses.dataclass
class RcParams:
    """The RcParams fields (src/core/rc.ts:42-91) the world-space voxel tracer uses, plus the dense probe grid."""
    cascades: int = 5
    probes0: int = 64        # c0 probes per axis of the dense grid (replaces screen registration)
    s0: float = 0.2          # c0 spacing (m); the reference's sMin raised to v0 (src/core/rcrun.ts:242)
    kInterval: float = 1.0
    levelMap: float = 1.0
    voxScale: float = 0.25
    voxOffset: float = 1.0
    voxSkip: float = 0.5
    resCap: int = 8
    recordHits: bool = False
    noMerge: bool = False
    fp16: bool = False       # precision form: radiance and SH stored as f16 at the reference's pack points
    stats: int = 0           # setStats (src/core/rc.ts:1233): 1 counts traces, DDA steps and step-cap hits
    relocate: bool = False   # relocation 'vrel' (relocateVox, src/core/rc.ts:413-472) with occlusion-aware parents
    lift: float = 0.25       # relocated probes keep lift * spacing free behind them (src/core/rc.ts:53, 88)
    sunTrace: bool = False   # per-voxel sun visibility by voxel shadow rays (voxSun, src/core/voxel.ts:631-673)
    bounce: bool = False     # multi-bounce from the previous update's field (voxBounce, src/core/rc.ts:198-202)
    objects: bool = False    # token obj (src/core/rc.ts:79-81): with bounce and a dynamic layer, the moving objects
                             # receive the bounce volume's light (dynInject) and block its irradiance rays (BN_DYN)
    history: bool = True     # c0 SH history of world probes (HIST_ON = !screen, src/core/rc.ts:1071; histBlend :784)
    bf16: bool = False       # precision form: voxel albedo and injected bounce light stored as bf16


DEFAULT_PARAMS = RcParams()


@dataclasses.dataclass
class Frame:
    eye: tuple
    sunDir: tuple = (-0.3, 1.0, 0.25)
    sunE: tuple = SUN_E
    sky: tuple = SKY
    sunVis: float = 1.0
    points: tuple = ()       # ((pos3, intensity3), ...)

    def to_c(self):
        f = _L.FrameC()
        sd = np.asarray(self.sunDir, np.float64)
        sd = sd / np.linalg.norm(sd)  # packFrame normalises sunDir (src/core/frame.ts:84)
        for a in range(3):
            f.eye[a] = float(self.eye[a])
            f.sun_dir[a] = float(sd[a])
            f.sun_e[a] = float(self.sunE[a])
            f.sky[a] = float(self.sky[a])
        f.sun_vis = float(self.sunVis)
        if len(self.points) > _L.MAX_POINT:
            raise ValueError("at most 8 point lights")
        f.n_point = len(self.points)
        for k, (p, i) in enumerate(self.points):
            for a in range(3):
                f.point_pos[k][a] = float(p[a])
                f.point_i[k][a] = float(i[a])
        return f


SURF_ALPHA_MASK = 1  # Material.layers.w bits (src/core/scene_gpu.ts:95): alpha-tested
SURF_TOON = 4        # toon style (stored as the albedo word's alpha)
SURFACE_MATERIAL_DTYPE = np.dtype([("base_color_layer", np.uint32), ("flags", np.uint32),
                                   ("alpha_cutoff", np.float32)])


@dataclasses.dataclass
class Surface:
    """Textured surfaces of a scene (what surfaceAt samples, src/core/voxel.ts:325-336): per-vertex UVs, the
    base-colour image layers (layers x size x size x 4 uint8: rgb sRGB-encoded, alpha linear; the reference array's
    layers 1.., its layer 0 being white) and per-material texture layer (0: white), flags (SURF_ALPHA_MASK,
    SURF_TOON) and alpha cutoff (Material.layers.x, layers.w, params.z of uploadScene, src/core/scene_gpu.ts:95-103)."""
    uvs: np.ndarray
    images: np.ndarray
    layer: np.ndarray
    flags: np.ndarray
    cutoff: np.ndarray

    def materials(self):
        m = np.zeros(len(self.layer), SURFACE_MATERIAL_DTYPE)
        m["base_color_layer"] = self.layer
        m["flags"] = self.flags
        m["alpha_cutoff"] = self.cutoff
        return m


@dataclasses.dataclass
class Scene:
    """SceneData subset: positions, normals, triangle table (v0, v1, v2, material) and materials (rgba); surface:
    UVs, base-colour textures and material surface words (None: untextured, every material white and opaque)."""
    pos: np.ndarray
    nrm: np.ndarray
    tri: np.ndarray
    materials: np.ndarray
    eye: tuple = (0.0, 0.0, 0.0)
    sunDir: tuple = (-0.3, 1.0, 0.25)
    bounds: tuple = None
    firstDynamicTri: int = None  # the moving parts are triangles [firstDynamicTri, ntris) (None: none)
    # Moving objects (DynamicSpec, src/core/scenes.ts:41-53): rtgi_object records, each one's motion schedule
    # (rot0, w, wt, ph, amp xyz) and the rest mesh of their vertex span; pos / nrm hold it posed at the frame.
    objects: np.ndarray = None
    motion: np.ndarray = None
    rest_pos: np.ndarray = None
    rest_nrm: np.ndarray = None
    surface: Surface = None

    @property
    def ntris(self):
        return int(self.tri.shape[0])

    def poses_at(self, frame):
        """Poses (c, s, offset xyz per object; DynamicSpec.at) of the scene's objects at a frame."""
        return motion_poses(self.motion, frame)


def motion_poses(motion, frame):
    """rtgi_motion_pose_at: (n, 5) doubles (c = cos(rotY), s = sin(rotY), offset) of n motion schedules at a frame."""
    m = np.ascontiguousarray(motion, np.float64).reshape(-1, 7)
    out = np.zeros((m.shape[0], 5), np.float64)
    check(lib().rtgi_motion_pose_at(ptr(m), m.shape[0], float(frame), ptr(out)))
    return out


def generate_scene(kind, seed=1, frame=0, extent=0.0, buildings=0, detail=1.0):
    """Seeded procedural scene from the library ('atrium' or 'city')."""
    kinds = {"atrium": _L.SCENE_ATRIUM, "city": _L.SCENE_CITY}
    p = _L.SceneParamsC(kinds[kind], seed, frame, extent, buildings, detail)
    out = C.POINTER(_L.SceneC)()
    check(lib().rtgi_scene_generate(C.byref(p), C.byref(out)))
    s = out.contents
    try:
        objs = np.frombuffer(C.string_at(s.objects, s.nobj * _L.OBJECT_DTYPE.itemsize),
                             _L.OBJECT_DTYPE).copy() if s.nobj else np.zeros(0, _L.OBJECT_DTYPE)
        span = int(objs["first_vertex"][-1] + objs["vertex_count"][-1] - objs["first_vertex"][0]) if s.nobj else 0
        sc = Scene(pos=_L.as_np(s.pos, s.nverts * 3, np.float32).reshape(-1, 3),
                   nrm=_L.as_np(s.nrm, s.nverts * 3, np.float32).reshape(-1, 3),
                   tri=_L.as_np(s.tri, s.ntris * 4, np.uint32).reshape(-1, 4),
                   materials=_L.as_np(s.materials, s.nmat * 4, np.float32).reshape(-1, 4),
                   eye=tuple(s.eye), sunDir=tuple(s.sun_dir),
                   bounds=(tuple(s.bounds_lo), tuple(s.bounds_hi)), firstDynamicTri=int(s.first_dynamic_tri),
                   objects=objs, motion=_L.as_np(s.motion, s.nobj * 7, np.float64).reshape(-1, 7),
                   rest_pos=_L.as_np(s.rest_pos, span * 3, np.float32).reshape(-1, 3),
                   rest_nrm=_L.as_np(s.rest_nrm, span * 3, np.float32).reshape(-1, 3))
    finally:
        lib().rtgi_scene_free(out)
    return sc


def device_count():
    n = C.c_int32()
    check(lib().rtgi_device_count(C.byref(n)))
    return n.value


def oct_solid_angles(res, samples=1 << 20):
    """octSolidAngles (src/core/frame.ts:105-114)."""
    out = np.zeros(res * res, np.float32)
    check(lib().rtgi_oct_solid_angles(res, samples, ptr(out)))
    return out


def pack_irradiance(rgb, tag):
    """packIrradiance (src/runner/report.ts:54-59): per pixel r, g, b and the tag (0 invalid, 1 photoreal, >= 2 toon)
    as one float32 array [n, 4], rows top to bottom: the layout of the reference's .irr.f32 dumps (--dump-irr), which
    its comparators read back (src/native/compare_runs.ts:36-45). rgb: [n, 3] (a gather's output), tag: [n]."""
    rgb = np.asarray(rgb, np.float32).reshape(-1, 3)
    tag = np.asarray(tag).reshape(-1)
    if tag.shape[0] != rgb.shape[0]:
        raise ValueError(f"pack_irradiance: {rgb.shape[0]} rgb pixels but {tag.shape[0]} tags")
    out = np.empty((rgb.shape[0], 4), np.float32)
    out[:, :3] = rgb
    out[:, 3] = tag
    return out


def write_irradiance(path, rgb, tag):
    """pack_irradiance written as a .irr.f32 file: raw little-endian float32, no header."""
    pack_irradiance(rgb, tag).astype("<f4").tofile(path)


def read_irradiance(path):
    """A .irr.f32 file back as (rgb [n, 3] float32, tag [n] uint8, valid [n] bool: tag > 0), as the reference's
    comparators unpack it (compare_runs.ts:38-42)."""
    v = np.fromfile(path, "<f4")
    if v.size % 4:
        raise ValueError(f"read_irradiance: {path} holds {v.size} floats, not a multiple of 4")
    v = v.reshape(-1, 4)
    tag = v[:, 3].astype(np.uint8)
    return v[:, :3].astype(np.float32), tag, tag > 0


def _vox_c(vox):
    return _L.VoxParamsC(vox.v0, vox.levels, vox.G, vox.maxBricks, vox.maxVoxels)


def _rc_c(rc):
    flags = ((_L.RC_RECORD_HITS if rc.recordHits else 0) | (_L.RC_NO_MERGE if rc.noMerge else 0)
             | (_L.RC_FP16 if rc.fp16 else 0) | (_L.RC_STATS if rc.stats else 0)
             | (_L.RC_RELOCATE if rc.relocate else 0) | (_L.RC_SUN_TRACE if rc.sunTrace else 0)
             | (_L.RC_BOUNCE if rc.bounce else 0) | (_L.RC_OBJECTS if rc.objects else 0)
             | (0 if rc.history else _L.RC_NO_HISTORY) | (_L.RC_BF16 if rc.bf16 else 0))
    return _L.RcParamsC(rc.cascades, rc.probes0, rc.s0, rc.kInterval, rc.levelMap, rc.voxScale, rc.voxOffset,
                        rc.voxSkip, rc.resCap, flags, rc.lift)


def _scene_arrays(scene):
    return (np.ascontiguousarray(scene.pos, np.float32), np.ascontiguousarray(scene.nrm, np.float32),
            np.ascontiguousarray(scene.tri, np.uint32), np.ascontiguousarray(scene.materials, np.float32))


def estimate_bytes(vox: VoxParams, rc: RcParams, nverts, ntris, nmat):
    """Device bytes of a world with a scene of this size (rtgi_world_estimate_bytes), without allocating."""
    out = C.c_uint64()
    check(lib().rtgi_world_estimate_bytes(C.byref(_vox_c(vox)), C.byref(_rc_c(rc)), nverts, ntris, nmat,
                                          C.byref(out)))
    return out.value


def _counters_dict(c, NL):
    return dict(bricks=c.bricks, marked=c.marked, voxels=c.voxels, brickOverflow=c.brick_overflow,
                bigOverflow=c.big_overflow, voxelOverflow=c.voxel_overflow,
                bricksPerLevel=list(c.bricks_per_level[:NL]), bigPerLevel=list(c.big_per_level[:NL]))


class World:
    """One GI world on the current GPU: createVoxWorld + createRc of the reference, backed by an rtgi_world."""

    def __init__(self, vox: VoxParams, rc: RcParams):
        self.vox, self.rc = vox, rc
        vp, rp = _vox_c(vox), _rc_c(rc)
        h = C.c_void_p()
        check(lib().rtgi_world_create(C.byref(vp), C.byref(rp), C.byref(h)))
        self._h = h
        self.scene = None

    # -- lifetime
    def destroy(self):
        if self._h:
            lib().rtgi_world_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.destroy()

    # -- inputs
    def upload_scene(self, scene: Scene):
        pos = np.ascontiguousarray(scene.pos, np.float32)
        nrm = np.ascontiguousarray(scene.nrm, np.float32)
        tri = np.ascontiguousarray(scene.tri, np.uint32)
        mat = np.ascontiguousarray(scene.materials, np.float32)
        check(lib().rtgi_world_upload_scene(self._h, ptr(pos), ptr(nrm), pos.shape[0], ptr(tri), tri.shape[0],
                                            ptr(mat), mat.shape[0]))
        self.scene = scene
        self._nobj = 0  # a new scene has no objects until set_objects
        if scene.surface is not None:
            self.set_surface(scene.surface)

    def update_vertices(self, first, pos, nrm):
        pos = np.ascontiguousarray(pos, np.float32)
        nrm = np.ascontiguousarray(nrm, np.float32)
        check(lib().rtgi_world_update_vertices(self._h, first, pos.shape[0], ptr(pos), ptr(nrm)))

    def set_surface(self, surface):
        """rtgi_world_set_surface: the scene's UVs, base-colour textures and material surface words (None: untextured
        again). Mips are built on the device."""
        if surface is None:
            check(lib().rtgi_world_set_surface(self._h, None, 0, None, 0, None, 1, 0))
            return
        uv = np.ascontiguousarray(surface.uvs, np.float32).reshape(-1, 2)
        img = np.ascontiguousarray(surface.images, np.uint8)
        if img.ndim != 4 or img.shape[1] != img.shape[2] or img.shape[3] != 4:
            raise ValueError("images must be layers x size x size x 4 uint8")
        mats = np.ascontiguousarray(surface.materials())
        check(lib().rtgi_world_set_surface(self._h, ptr(uv), uv.shape[0], mats.ctypes.data if len(mats) else None,
                                           len(mats), img.ctypes.data if img.shape[0] else None, img.shape[1],
                                           img.shape[0]))

    def set_objects(self, objects, rest_pos, rest_nrm):
        """rtgi_world_set_objects: moving objects (OBJECT_DTYPE records) and the rest mesh of their vertex span, posed
        on the device by every update once set_pose has been called. An empty array removes them."""
        objs = np.ascontiguousarray(objects, _L.OBJECT_DTYPE)
        rp = np.ascontiguousarray(rest_pos, np.float32)
        rn = np.ascontiguousarray(rest_nrm, np.float32)
        check(lib().rtgi_world_set_objects(self._h, objs.ctypes.data if len(objs) else None, len(objs),
                                           ptr(rp), ptr(rn)))
        self._nobj = len(objs)

    def set_scene_objects(self, scene):
        """set_objects with a generated scene's objects and rest mesh."""
        self.set_objects(scene.objects, scene.rest_pos, scene.rest_nrm)

    def set_pose(self, poses):
        """rtgi_world_set_pose: (n, 5) doubles (c, s, offset xyz) for the n registered objects, from the next update."""
        q = np.ascontiguousarray(poses, np.float64).reshape(-1, 5)
        n = getattr(self, "_nobj", 0)
        if n and q.shape[0] != n:  # (without objects the library refuses the call)
            raise ValueError(f"{q.shape[0]} poses for {n} objects")
        check(lib().rtgi_world_set_pose(self._h, ptr(q)))

    def set_frame(self, frame: Frame):
        self.frame = frame
        f = frame.to_c()
        check(lib().rtgi_world_set_frame(self._h, C.byref(f)))

    def update(self, eye):
        """VoxWorld.update(eye) (src/core/voxel.ts:896): re-centres the clipmap (and the probe grid) on eye."""
        f = dataclasses.replace(self.frame, eye=tuple(eye)) if getattr(self, "frame", None) else Frame(eye=tuple(eye))
        self.set_frame(f)

    # -- execution
    def encode(self, stream=None):
        check(lib().rtgi_world_enqueue(self._h, stream))

    def encode_staged(self, stages=_L.STAGE_ALL, stream=None):
        check(lib().rtgi_world_enqueue_stages(self._h, stages, stream))

    def set_timing(self, on=True):
        check(lib().rtgi_world_set_timing(self._h, 1 if on else 0))

    def stage_ms(self):
        ms = np.zeros(4, np.float32)
        check(lib().rtgi_world_stage_ms(self._h, ptr(ms)))
        return dict(zip(("vox", "trace", "merge", "sh"), ms.tolist()))

    def sync(self):
        check(lib().rtgi_world_sync(self._h))

    # -- outputs
    def layout(self):
        lo = _L.LayoutC()
        check(lib().rtgi_world_get_layout(self._h, C.byref(lo)))
        N = lo.cascades
        return dict(levels=lo.levels, G=lo.G, E=lo.E,
                    window_origin=[list(lo.window_origin[L]) for L in range(lo.levels)],
                    cascades=N, probes=list(lo.probes[:N]), res=list(lo.res[:N]), vox_level=list(lo.vox_level[:N]),
                    grid_origin=list(lo.grid_origin), rays=list(lo.rays[:N]), total_rays=lo.total_rays,
                    b0=list(lo.b0[:N]), b1=list(lo.b1[:N]), omega0=np.array(lo.omega0[:], np.float32),
                    device_bytes=lo.device_bytes)

    def read_counters(self):
        """VoxWorld.readCounters (src/core/voxel.ts:908-918)."""
        c = _L.CountersC()
        check(lib().rtgi_world_read_counters(self._h, C.byref(c)))
        return _counters_dict(c, self.vox.levels)

    def used_bytes(self, c=None):
        """VoxWorld.usedBytes (src/core/voxel.ts:919-922) for this layout: persistent and build-temporary bytes."""
        c = c or self.read_counters()
        NL, G = self.vox.levels, self.vox.G
        vox = min(c["voxels"], self.vox.maxVoxels)
        return dict(persistent=(64 + NL * G ** 3) * 4 + c["bricks"] * (64 + 8) + vox * 12,
                    buildTemp=c["bricks"] * 512 * 4)

    def read_clipmap(self, level, rep=True, attrs=True):
        E = self.vox.G * 8
        n = E ** 3
        occ = np.zeros(n // 32, np.uint32)
        rp = np.zeros(n, np.uint32) if rep else None
        alb = np.zeros(n, np.uint32) if attrs else None
        nrm = np.zeros(n, np.uint32) if attrs else None
        pl = np.zeros(n, np.float32) if attrs else None
        st = np.zeros(self.vox.G ** 3, np.uint8)
        check(lib().rtgi_world_read_clipmap(self._h, level, ptr(occ), ptr(rp), ptr(alb), ptr(nrm), ptr(pl), ptr(st)))
        return dict(occ=occ, rep=rp, albedo=alb, normal=nrm, plane=pl, brick_state=st)

    def read_cascade(self, c, hits=False):
        lo = self.layout()
        n = lo["rays"][c]
        rgba = np.zeros((n, 4), np.float32)
        h = np.zeros(n, np.uint32) if hits else None
        check(lib().rtgi_world_read_cascade(self._h, c, ptr(rgba), ptr(h)))
        return (rgba, h) if hits else rgba

    def set_tuning(self, probe_tile=None, fused=None):
        """Launch tuning (rtgi_world_set_tuning): speed and storage order only, never results. None keeps a
        setting."""
        t = self.tuning()
        t = _L.TuningC(t["probe_tile"] if probe_tile is None else probe_tile,
                       t["fused"] if fused is None else int(fused))
        check(lib().rtgi_world_set_tuning(self._h, C.byref(t)))

    def tuning(self):
        t = _L.TuningC()
        check(lib().rtgi_world_get_tuning(self._h, C.byref(t)))
        return dict(probe_tile=t.probe_tile, fused=t.fused)

    def trace_rays(self, c, origins, dirs):
        """Rays from origins along unit dirs (n x 3) traced as cascade c's rays over the last update's clipmap."""
        o = np.ascontiguousarray(origins, np.float32)
        d = np.ascontiguousarray(dirs, np.float32)
        n = o.shape[0]
        rgba = np.zeros((n, 4), np.float32)
        hits = np.zeros(n, np.uint32)
        check(lib().rtgi_world_trace_rays(self._h, c, ptr(o), ptr(d), n, ptr(rgba), ptr(hits)))
        return rgba, hits

    def read_probes(self, c):
        """Relocated probes of cascade c (grid order): positions, normals (n x 3) and parent masks."""
        n = (self.rc.probes0 >> c) ** 3
        pos = np.zeros((n, 3), np.float32)
        nrm = np.zeros((n, 3), np.float32)
        par = np.zeros(n, np.uint32)
        check(lib().rtgi_world_read_probes(self._h, c, ptr(pos), ptr(nrm), ptr(par)))
        return pos, nrm, par

    def set_dynamic(self, first_dynamic_tri, dyn_bricks=512):
        """Triangles from first_dynamic_tri on go to the dynamic coverage layer (createVoxWorld's firstDynamicTri,
        VoxParams.dynBricks; src/core/dynlayer.ts) instead of the bricks."""
        check(lib().rtgi_world_set_dynamic(self._h, int(first_dynamic_tri), int(dyn_bricks)))
        self._dyn_bricks = int(dyn_bricks)

    def read_dyn_counters(self):
        """Allocated bricks, dropped bricks and triangles outside every clipmap of the last update."""
        v = (C.c_uint32 * 3)()
        check(lib().rtgi_world_read_dyn_counters(self._h, C.byref(v, 0), C.byref(v, 4), C.byref(v, 8)))
        return dict(bricks=int(v[0]), brickOverflow=int(v[1]), outside=int(v[2]))

    def read_dynamic(self):
        """The raw layer: hdr [64], ptr [levels, G^3], meta [cap], mask [cap, 16], data [cap, 512, 13] (uint32)."""
        cap = self._dyn_bricks
        per = self.vox.G ** 3
        out = dict(hdr=np.zeros(64, np.uint32), ptr=np.zeros((self.vox.levels, per), np.uint32),
                   meta=np.zeros(cap, np.uint32), mask=np.zeros((cap, 16), np.uint32),
                   data=np.zeros((cap, 512, 13), np.uint32))
        check(lib().rtgi_world_read_dynamic(self._h, ptr(out["hdr"]), ptr(out["ptr"]), ptr(out["meta"]),
                                            ptr(out["mask"]), ptr(out["data"])))
        return out

    def read_bounce_volume(self):
        """The bounce volume of the last update: texels (z, y, x, axis, rgba) as float16 (axes +x -x +y -y +z -z;
        alpha: reach, 0 = inactive probe) and the probe coordinates of probe 0."""
        dx, dy, dz = _L.BN_DIMS
        t = np.zeros((dz, dy, dx, 6, 4), np.uint16)
        org = np.zeros(3, np.int32)
        check(lib().rtgi_world_read_bounce_volume(self._h, ptr(t), ptr(org)))
        return t.view(np.float16), org

    def gather(self, pos, nrm, occlusion=False):
        """gatherWorld (src/core/rc.ts:871-927): irradiance (n x 3) of the last update's field at surface points."""
        p = np.ascontiguousarray(pos, np.float32)
        n = np.ascontiguousarray(nrm, np.float32)
        irr = np.zeros((p.shape[0], 3), np.float32)
        check(lib().rtgi_world_gather(self._h, ptr(p), ptr(n), p.shape[0], _L.GATHER_OCCLUSION if occlusion else 0,
                                      ptr(irr)))
        return irr

    def gather_async(self, d_pos, d_nrm, n, d_irr, occlusion=False, stream=None):
        """rtgi_world_gather_async: the gather on device arrays (addresses of n x 3 float32 positions, normals and
        irradiance, e.g. torch tensors' data_ptr()), enqueued on a HIP stream (None: the world's own), no allocation
        and no wait."""
        check(lib().rtgi_world_gather_async(self._h, C.c_void_p(int(d_pos)), C.c_void_p(int(d_nrm)), int(n),
                                            _L.GATHER_OCCLUSION if occlusion else 0, C.c_void_p(int(d_irr)),
                                            C.c_void_p(int(stream)) if stream else None))

    def read_trace_stats(self):
        """Per cascade: world traces, DDA steps, traces at the 1024-step cap (rc.ts:232-235)."""
        st = _L.TraceStatsC()
        check(lib().rtgi_world_read_trace_stats(self._h, C.byref(st)))
        N = self.rc.cascades
        return dict(traces=list(st.traces[:N]), steps=list(st.steps[:N]), capped=list(st.capped[:N]))

    def read_sh(self):
        n = self.rc.probes0 ** 3
        sh = np.zeros((n, 27), np.float32)
        check(lib().rtgi_world_read_sh(self._h, ptr(sh)))
        return sh


class Batch:
    """One GI update of many worlds with one launch per stage (rtgi_batch_*); optionally replayed as a hipGraph."""

    def __init__(self, worlds, graph=False):
        self.worlds = list(worlds)
        arr = (C.c_void_p * len(self.worlds))(*[w._h.value for w in self.worlds])
        h = C.c_void_p()
        check(lib().rtgi_batch_create(arr, len(self.worlds), C.byref(h)))
        self._h = h
        if graph:
            check(lib().rtgi_batch_set_graph(self._h, 1))

    def encode(self, stream=None):
        check(lib().rtgi_batch_enqueue(self._h, stream))

    def sync(self):
        check(lib().rtgi_batch_sync(self._h))

    def set_timing(self, on=True):
        check(lib().rtgi_batch_set_timing(self._h, 1 if on else 0))

    def stage_ms(self):
        ms = np.zeros(4, np.float32)
        check(lib().rtgi_batch_stage_ms(self._h, ptr(ms)))
        return dict(zip(("vox", "trace", "merge", "sh"), ms.tolist()))

    def destroy(self):
        if self._h:
            lib().rtgi_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.destroy()


class Server:
    """Request server (rtgi_server_*): GI-update requests over registered worlds, continuously batched, with the
    worlds resident on the device under a memory budget (0: 90% of the free device memory)."""

    def __init__(self, budget=0, max_batch=64, slots=2):
        p = _L.ServerParamsC(budget, max_batch, slots)
        h = C.c_void_p()
        check(lib().rtgi_server_create(C.byref(p), C.byref(h)))
        self._h = h
        self._worlds = {}   # id -> (VoxParams, RcParams)
        self._out = {}      # tag -> (sh array, counters struct, world id)
        self._next = 1

    def add_world(self, scene, vox: VoxParams, rc: RcParams):
        pos, nrm, tri, mat = _scene_arrays(scene)
        i = C.c_int32()
        check(lib().rtgi_server_add_world(self._h, C.byref(_vox_c(vox)), C.byref(_rc_c(rc)), ptr(pos), ptr(nrm),
                                          pos.shape[0], ptr(tri), tri.shape[0], ptr(mat), mat.shape[0], C.byref(i)))
        self._worlds[i.value] = (vox, rc)
        return i.value

    def submit(self, world, frame, sh=False, counters=False, tag=None):
        """Queues one update of `world` at `frame`; returns its tag. sh / counters: capture those outputs."""
        if tag is None:
            tag = self._next
            self._next += 1
        vox, rc = self._worlds[world]
        sh_arr = np.zeros((rc.probes0 ** 3, 27), np.float32) if sh else None
        cnt = _L.CountersC() if counters else None
        r = _L.RequestC(world, tag, frame.to_c(), ptr(sh_arr) if sh else None,
                        C.cast(C.pointer(cnt), C.c_void_p) if counters else None)
        self._out[tag] = (sh_arr, cnt, world)
        try:
            check(lib().rtgi_server_submit(self._h, C.byref(r)))
        except Exception:
            del self._out[tag]
            raise
        return tag

    def poll(self, max_n=1024, wait_ms=0):
        """Completed requests (oldest first): dicts with tag, world, status, batch, t_submit/t_launch/t_done (s),
        latency_ms and the captured sh / counters."""
        buf = (_L.CompletionC * max_n)()
        n = C.c_int32()
        check(lib().rtgi_server_poll(self._h, buf, max_n, wait_ms, C.byref(n)))
        out = []
        for c in buf[:n.value]:
            sh_arr, cnt, world = self._out.pop(c.tag, (None, None, c.world))
            d = dict(tag=c.tag, world=c.world, status=c.status, batch=c.batch_size, t_submit=c.t_submit,
                     t_launch=c.t_launch, t_done=c.t_done, latency_ms=(c.t_done - c.t_submit) * 1e3, sh=sh_arr,
                     counters=_counters_dict(cnt, self._worlds[world][0].levels) if cnt is not None else None)
            out.append(d)
        return out

    def drain(self):
        check(lib().rtgi_server_drain(self._h))

    def stats(self):
        st = _L.ServerStatsC()
        check(lib().rtgi_server_get_stats(self._h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_}

    def destroy(self):
        if self._h:
            lib().rtgi_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.destroy()


def _atrium_points(seed=1):
    rng = np.random.default_rng(seed)
    pts = []
    for _ in range(4):
        p = (float(rng.uniform(-7, 7)), float(rng.uniform(1.5, 6.0)), float(rng.uniform(-7, 7)))
        i = tuple(float(x) for x in rng.uniform(2.0, 6.0, 3))
        pts.append((p, i))
    return tuple(pts)


# Named configurations (the role of src/runner/configs.ts). 'atrium' and 'city' are the plain FP32 forms of the
# 50k-triangle atrium (4 x 64^3 clipmap at 0.125 m, 32^3 c0 probes, 4 cascades) and the 500k-triangle city block
# (4 x 128^3 clipmap at 0.25 m, 64^3 c0 probes, 5 cascades).
CONFIGS = {
    "atrium": dict(scene=dict(kind="atrium", seed=1), vox=VoxParams(v0=0.125, levels=4, G=8, maxBricks=4 * 8 ** 3,
                                                                    maxVoxels=1 << 20),
                   rc=RcParams(cascades=4, probes0=32, s0=0.625), points=True),
    "city": dict(scene=dict(kind="city", seed=1), vox=VoxParams(v0=0.25, levels=4, G=16, maxBricks=4 * 16 ** 3,
                                                                maxVoxels=1 << 22),
                 rc=RcParams(cascades=5, probes0=64, s0=1.0), points=False),
}


def config(name, **rc_over):
    """(scene, VoxParams, RcParams, Frame) of a named configuration."""
    cf = CONFIGS[name]
    scene = generate_scene(**cf["scene"])
    rc = dataclasses.replace(cf["rc"], **rc_over)
    frame = Frame(eye=scene.eye, sunDir=scene.sunDir, points=_atrium_points() if cf["points"] else ())
    return scene, cf["vox"], rc, frame
