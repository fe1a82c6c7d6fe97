"""TEST INFRASTRUCTURE ONLY: ctypes wrapper of the CPU oracle (liboracle.so, rtgi_oracle.c).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import this module, and only as the checker.

What pins each oracle function (beyond being an operation-by-operation restatement of the reference WGSL, whose
citations are in rtgi_oracle.c); the reference ships no golden vectors or tests for this path:
  triBox (voxel.ts:231-248)          exact rational SAT over the 13 axes on random inputs (tests/test_oracle.py)
  closestBary (voxel.ts:250-267)     brute-force minimisation of the distance over a barycentric grid
  pack4x8unorm / pack2x16snorm       WGSL specification formulas and known values (tests/golden/pack_vectors.json)
  octEnc / octDecode                 round trip; closed-form values on the axes and diagonals
  build (voxel.ts:338-620)           brute force: occupancy = {cells whose box the triangle overlaps, by exact SAT}
                                     at level 0, coarse level = OR of its 8 children where the finer window holds
                                     it, rep = max overlapping triangle + 1; ordered overflow drops the finest bricks
  traceVox (voxel.ts:121-202)        FP64 voxel walk (first hit equal away from ties); axis-aligned rays closed form;
                                     empty world -> T = 1; single occupied voxel
  mergeWorld (rc.ts:681-708)         equal-weight parent mean = trilinear interpolation on the dense grid
  projectSH (rc.ts:745-774)          constant radiance -> only L0, = pi * 0.282095 * 4 pi * L; sum(omega) = 4 pi
  octSolidAngles (frame.ts:105-114)  sum = 4 pi, 8-fold symmetry of the 4x4 texels, library value agrees
  posedVertices (scenes.ts:80-102)   exact rational evaluation on dyadic inputs (every double operation exact);
                                     quarter turns and the identity in closed form; distances to the pivot kept
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


class OrcVox(C.Structure):
    _fields_ = [("v0", C.c_float), ("levels", C.c_int32), ("G", C.c_int32), ("max_bricks", C.c_int32),
                ("max_voxels", C.c_int32), ("origin", (C.c_int32 * 3) * 8)]


class OrcRc(C.Structure):
    _fields_ = [("cascades", C.c_int32), ("probes0", C.c_int32), ("res_cap", C.c_int32), ("s0", C.c_float),
                ("vox_skip", C.c_float), ("vox_level", C.c_int32 * 8), ("b0", C.c_float * 8), ("b1", C.c_float * 8),
                ("G0", C.c_int32 * 3), ("sun_dir", C.c_float * 3), ("sun_e", C.c_float * 3), ("sky", C.c_float * 3),
                ("sun_vis", C.c_float), ("n_point", C.c_int32), ("point_pos", (C.c_float * 3) * 8),
                ("point_i", (C.c_float * 3) * 8), ("omega0", C.c_float * 16), ("fp16", C.c_int32),
                ("vox_offset", C.c_float), ("lift", C.c_float), ("reloc", C.c_int32), ("ppos", C.c_void_p * 8),
                ("pnrm", C.c_void_p * 8), ("pvis", C.c_void_p * 8), ("sunvis", C.c_void_p),
                ("bounce", C.c_void_p), ("bn_prev", C.c_void_p), ("bn_prev_org", C.c_int32 * 3), ("dyn", C.c_void_p),
                ("objects", C.c_int32), ("bf16", C.c_int32)]


class OrcHist(C.Structure):
    _fields_ = [("cap", C.c_int32), ("H", C.c_uint32), ("limit", C.c_int32), ("par", C.c_int32), ("frames", C.c_uint32),
                ("misses", C.c_uint64), ("hash", C.c_void_p), ("grid", C.c_void_p), ("sh", C.c_void_p),
                ("G0", (C.c_int32 * 3) * 2), ("n", C.c_int32)]


class OrcSurface(C.Structure):
    _fields_ = [("tex", C.c_void_p), ("uv", C.c_void_p), ("mext", C.c_void_p), ("lut", C.c_void_p),
                ("size_log2", C.c_uint32), ("mips", C.c_uint32), ("layer_words", C.c_uint32), ("layers", C.c_uint32)]


class OrcDyn(C.Structure):
    _fields_ = [("cap", C.c_int32), ("first_tri", C.c_uint32), ("end_tri", C.c_uint32), ("hdr", C.c_void_p),
                ("ptr", C.c_void_p), ("bits", C.c_void_p), ("meta", C.c_void_p), ("mask", C.c_void_p),
                ("data", C.c_void_p)]


_P = C.c_void_p
_SIG = {
    "orc_vox_words": (C.c_size_t, [C.POINTER(OrcVox)]),
    "orc_origins": (None, [C.c_double, C.c_int, C.c_int, _P, _P]),
    "orc_cascade_setup": (None, [C.POINTER(OrcRc), C.c_double, C.c_int, C.c_double, C.c_double, C.c_double,
                                 C.c_double, _P]),
    "orc_oct_solid_angles": (None, [C.c_int, C.c_int, _P]),
    "orc_build": (C.c_int, [C.POINTER(OrcVox), _P, _P, _P, C.c_uint32, _P, _P, _P]),
    "orc_canonical": (C.c_int, [C.POINTER(OrcVox), _P, _P, C.c_int, _P, _P, _P, _P, _P, _P]),
    "orc_rays": (C.c_uint64, [C.POINTER(OrcRc), C.c_int]),
    "orc_trace": (C.c_int, [C.POINTER(OrcVox), _P, C.POINTER(OrcRc), C.c_uint64, C.c_uint64, _P, _P, _P, C.c_int]),
    "orc_sun": (C.c_int, [C.POINTER(OrcVox), _P, C.POINTER(OrcRc), _P]),
    "orc_relocate": (C.c_int, [C.POINTER(OrcVox), _P, C.POINTER(OrcRc), C.c_int, _P, _P]),
    "orc_link": (C.c_int, [C.POINTER(OrcVox), _P, C.POINTER(OrcRc), C.c_int, _P, _P, _P]),
    "orc_bn_origin": (None, [_P, _P]),
    "orc_dyn_init": (C.c_int, [C.POINTER(OrcVox), C.POINTER(OrcDyn)]),
    "orc_dyn_inject": (C.c_int, [C.POINTER(OrcVox), C.POINTER(OrcDyn), _P, _P]),
    "orc_dyn_build": (C.c_int, [C.POINTER(OrcVox), C.POINTER(OrcRc), _P, _P, _P, _P, C.POINTER(OrcDyn)]),
    "orc_bounce_volume": (C.c_int, [C.POINTER(OrcVox), _P, C.POINTER(OrcRc), _P, _P, C.c_uint32, _P, _P]),
    "orc_bounce_inject": (C.c_int, [C.POINTER(OrcVox), _P, _P, _P, _P, C.c_int]),
    "orc_gather": (C.c_int, [C.POINTER(OrcVox), _P, C.POINTER(OrcRc), _P, _P, _P, C.c_uint64, C.c_int, _P]),
    "orc_trace_rays": (C.c_int, [C.POINTER(OrcVox), _P, C.POINTER(OrcRc), C.c_int, _P, _P, C.c_uint64, _P, _P,
                                 C.c_int]),
    "orc_trace_rays_f64": (C.c_int, [C.POINTER(OrcVox), _P, C.POINTER(OrcRc), C.c_int, _P, _P, C.c_uint64, _P, _P,
                                     _P, C.c_double, C.c_int]),
    "orc_merge_sh": (C.c_int, [C.POINTER(OrcRc), _P, _P]),
    "orc_merge_sh_hist": (C.c_int, [C.POINTER(OrcRc), _P, _P, C.POINTER(OrcHist)]),
    "orc_hist_init": (C.c_int, [C.POINTER(OrcHist), C.c_int32]),
    "orc_trace_f64": (C.c_int, [C.POINTER(OrcVox), _P, C.POINTER(OrcRc), C.c_uint64, C.c_uint64, _P, _P, _P,
                                C.c_double, C.c_int]),
    "orc_merge_sh_f64": (C.c_int, [C.POINTER(OrcRc), _P, _P]),
    "orc_trace_f64_take": (C.c_int, [C.POINTER(OrcVox), _P, C.POINTER(OrcRc), C.c_uint64, C.c_uint64, _P, _P, _P,
                                     C.c_double, C.c_int]),
    "orc_tri_box": (C.c_int, [_P, _P, _P, _P, _P]),
    "orc_set_surface": (None, [_P]),
    "orc_srgb_tables": (None, [_P]),
    "orc_surf_mips": (None, [_P, C.c_uint32, C.c_uint32, _P, _P]),
    "orc_surf_sample": (None, [_P, C.c_uint32, C.c_float, C.c_float, C.c_uint32, _P]),
    "orc_closest_bary": (None, [_P, _P, _P, _P, _P]),
    "orc_pack4x8unorm": (C.c_uint32, [_P]),
    "orc_pack2x16snorm": (C.c_uint32, [C.c_float, C.c_float]),
    "orc_oct_enc": (None, [_P, _P]),
    "orc_f2h": (C.c_uint16, [C.c_float]),
    "orc_h2f": (C.c_float, [C.c_uint16]),
    "orc_log2": (C.c_float, [C.c_float]),
    "orc_f2bf": (C.c_uint16, [C.c_float]),
    "orc_bf2f": (C.c_float, [C.c_uint16]),
    "orc_pose": (None, [C.c_uint32, _P, _P, _P, _P, _P, _P, _P, _P]),
}
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is not built (make -C oracle)")
        L = C.CDLL(LIB_PATH)
        for n, (r, a) in _SIG.items():
            f = getattr(L, n)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def f32(x):
    return np.ascontiguousarray(x, np.float32)


class DynState:
    """The dynamic coverage layer's state, kept across updates (as the device keeps it): triangles from first_tri on
    of each scene built with it are dynamic."""

    def __init__(self, vox, cap, first_tri):
        per = vox.G ** 3
        self.cap, self.first_tri = int(cap), int(first_tri)
        self.hdr = np.zeros(64, np.uint32)
        self.ptr = np.zeros((vox.levels, per), np.uint32)
        self.bits = np.zeros(2 * vox.levels * per // 32, np.uint32)
        self.meta = np.zeros(self.cap, np.uint32)
        self.mask = np.zeros((self.cap, 16), np.uint32)
        self.data = np.zeros((self.cap, 512, 13), np.uint32)
        self._vox = vox
        vp = OrcVox()
        vp.levels, vp.G = vox.levels, vox.G
        assert lib().orc_dyn_init(C.byref(vp), C.byref(self.struct(0))) == 0

    def struct(self, end_tri):
        d = OrcDyn()
        d.cap, d.first_tri, d.end_tri = self.cap, self.first_tri, int(end_tri)
        for k in ("hdr", "ptr", "bits", "meta", "mask", "data"):
            setattr(d, k, getattr(self, k).ctypes.data)
        return d


class History:
    """The c0 SH history of one world's world probes (histBlend / clearHist, rc.ts:775-818), kept across updates as
    the device keeps it: each merge_sh(..., hist=) of an update blends the c0 SH with the previous update's SH of the
    probe at the same grid point (mix(prev, new, 0.5)) and records this update's SH.

    limit: slots a hash search visits. The reference's 64 (limit=64) misses a few probes of dense grids of 64^3 and
    more (a long linear-probing cluster; which ones depends on the GPU's insertion order); limit=0 searches the whole
    table, so every probe whose grid point the previous update held finds its history. The device path keys the
    history by grid point directly, which is the limit=0 rule."""

    def __init__(self, probes0, limit=0):
        cap = int(probes0) ** 3
        self.s = OrcHist()
        assert lib().orc_hist_init(C.byref(self.s), cap) == 0
        self.s.limit = int(limit)
        self.hash = np.zeros(2 * self.s.H, np.uint32)
        self.grid = np.zeros((2, cap, 4), np.int32)
        self.sh = np.zeros((2, cap, 32), np.float32)
        self.s.hash, self.s.grid, self.s.sh = self.hash.ctypes.data, self.grid.ctypes.data, self.sh.ctypes.data

    @property
    def misses(self):
        """Probes whose grid point the previous update held but the reference's 64-slot hash search missed."""
        return int(self.s.misses)


BN_DIMS = (32, 16, 32)   # bounce volume probes per axis (x, y, z)
BN_CELL = np.float32(1.2)


def bn_origin(eye):
    """Probe coordinates of probe 0 of the bounce volume around eye (bnOriginNow)."""
    e = np.ascontiguousarray(eye, np.float32)
    o = np.zeros(3, np.int32)
    lib().orc_bn_origin(_p(e), _p(o))
    return o


def tri_box(c, h, a, b, cc):
    return bool(lib().orc_tri_box(_p(f32(c)), _p(f32(h)), _p(f32(a)), _p(f32(b)), _p(f32(cc))))


def closest_bary(p, a, b, c):
    out = np.zeros(3, np.float32)
    lib().orc_closest_bary(_p(f32(p)), _p(f32(a)), _p(f32(b)), _p(f32(c)), _p(out))
    return out


def pack4x8unorm(e):
    return int(lib().orc_pack4x8unorm(_p(f32(e))))


def pack2x16snorm(x, y):
    return int(lib().orc_pack2x16snorm(float(np.float32(x)), float(np.float32(y))))


def oct_enc(n):
    out = np.zeros(2, np.float32)
    lib().orc_oct_enc(_p(f32(n)), _p(out))
    return out


def f2h(x):
    return int(lib().orc_f2h(float(np.float32(x))))


def pose(first, count, pivot, poses, rest_pos, rest_nrm, pos, nrm):
    """posedVertices of objects (first, count, pivot (n, 3)) at poses (n, 5: c, s, offset) from the rest mesh of their
    span; returns posed copies of the whole-scene arrays pos / nrm."""
    first = np.ascontiguousarray(first, np.uint32)
    count = np.ascontiguousarray(count, np.uint32)
    pivot = np.ascontiguousarray(pivot, np.float64).reshape(-1, 3)
    poses = np.ascontiguousarray(poses, np.float64).reshape(-1, 5)
    rp = np.ascontiguousarray(rest_pos, np.float32)
    rn = np.ascontiguousarray(rest_nrm, np.float32)
    out_p = np.array(pos, np.float32, order="C")
    out_n = np.array(nrm, np.float32, order="C")
    lib().orc_pose(len(first), _p(first), _p(count), _p(pivot), _p(poses), _p(rp), _p(rn), _p(out_p), _p(out_n))
    return out_p, out_n


def pose_scene(scene, poses):
    """pose() with a generated scene's objects and rest mesh, over its vertex arrays."""
    o = scene.objects
    return pose(o["first_vertex"], o["vertex_count"], o["pivot"], poses, scene.rest_pos, scene.rest_nrm, scene.pos,
                scene.nrm)


def oct_solid_angles(res, samples=1 << 20):
    out = np.zeros(res * res, np.float32)
    lib().orc_oct_solid_angles(res, samples, _p(out))
    return out


def origins(v0, levels, G, eye):
    e = np.ascontiguousarray(eye, np.float64)
    out = np.zeros((8, 3), np.int32)
    lib().orc_origins(v0, levels, G, _p(e), _p(out))
    return out[:levels]


def srgb_tables():
    lut = np.zeros(511, np.float32)
    lib().orc_srgb_tables(_p(lut))
    return lut


class SurfaceData:
    """A scene's textured surface in the oracle's form (orc_surface): the mip chain of its images, its UVs, material
    words and the sRGB tables. `struct` is what orc_set_surface takes (valid while this object lives)."""

    def __init__(self, surface):
        img = np.ascontiguousarray(surface.images, np.uint8)
        layers, size = int(img.shape[0]), int(img.shape[1])
        sl = int(size).bit_length() - 1
        self.per = sum(1 << (2 * (sl - m)) for m in range(sl + 1))
        self.lut = srgb_tables()
        self.tex = np.zeros(max(1, layers * self.per), np.uint32)
        lib().orc_surf_mips(_p(img) if layers else None, size, layers, _p(self.lut), _p(self.tex))
        self.uv = np.ascontiguousarray(surface.uvs, np.float32).reshape(-1, 2)
        n = len(surface.layer)
        self.mext = np.zeros((max(1, n), 4), np.uint32)
        self.mext[:n, 0] = np.asarray(surface.layer, np.uint32)
        self.mext[:n, 1] = np.asarray(surface.flags, np.uint32)
        self.mext[:n, 2] = np.asarray(surface.cutoff, np.float32).view(np.uint32)
        s = OrcSurface()
        s.tex, s.uv, s.mext, s.lut = (self.tex.ctypes.data, self.uv.ctypes.data, self.mext.ctypes.data,
                                      self.lut.ctypes.data)
        s.size_log2, s.mips, s.layer_words, s.layers = sl, sl + 1, self.per, layers
        self.struct = s

    def sample(self, layer, u, v, lod):
        out = np.zeros(4, np.float32)
        lib().orc_surf_sample(C.byref(self.struct), layer, u, v, lod, _p(out))
        return out


class Oracle:
    """CPU restatement of one world: build -> canonical clipmap; trace -> merge -> SH (f32), and the FP64 tracer."""

    def __init__(self, vox, rc, frame, omega0=None):
        self.vp = OrcVox()
        self.vp.v0 = vox.v0
        self.vp.levels = vox.levels
        self.vp.G = vox.G
        self.vp.max_bricks = vox.maxBricks
        self.vp.max_voxels = vox.maxVoxels
        org = origins(vox.v0, vox.levels, vox.G, frame.eye)
        for L in range(vox.levels):
            for a in range(3):
                self.vp.origin[L][a] = int(org[L][a])
        r = OrcRc()
        r.cascades = rc.cascades
        r.probes0 = rc.probes0
        r.res_cap = rc.resCap
        r.vox_skip = rc.voxSkip
        r.fp16 = 1 if getattr(rc, "fp16", False) else 0
        r.vox_offset = rc.voxOffset
        r.lift = getattr(rc, "lift", 0.25)
        r.reloc = 1 if getattr(rc, "relocate", False) else 0
        r.objects = 1 if getattr(rc, "objects", False) else 0
        r.bf16 = 1 if getattr(rc, "bf16", False) else 0
        self._sun_trace = bool(getattr(rc, "sunTrace", False))
        self._bounce = bool(getattr(rc, "bounce", False))
        self._eye = np.asarray(frame.eye, np.float32)
        eye = np.ascontiguousarray(frame.eye, np.float64)
        lib().orc_cascade_setup(C.byref(r), vox.v0, vox.levels, rc.s0, rc.kInterval, rc.levelMap, rc.voxScale,
                                _p(eye))
        sd = np.asarray(frame.sunDir, np.float64)
        sd = sd / np.linalg.norm(sd)
        for a in range(3):
            r.sun_dir[a] = float(sd[a])
            r.sun_e[a] = float(frame.sunE[a])
            r.sky[a] = float(frame.sky[a])
        r.sun_vis = frame.sunVis
        r.n_point = len(frame.points)
        for k, (p, i) in enumerate(frame.points):
            for a in range(3):
                r.point_pos[k][a] = float(p[a])
                r.point_i[k][a] = float(i[a])
        om = omega0 if omega0 is not None else oct_solid_angles(4)
        for i in range(16):
            r.omega0[i] = float(om[i])
        self.rc = r
        self.vox = vox
        self.E = vox.G * 8
        self.words = lib().orc_vox_words(C.byref(self.vp))
        self.buf = None

    def cascade_rays(self):
        return [int(lib().orc_rays(C.byref(self.rc), c)) for c in range(self.rc.cascades)]

    def build(self, scene, dyn=None):
        """The clipmap of the scene; with dyn (a DynState), triangles from dyn.first_tri on are dynamic: they are
        splatted into the layer (this update of it) instead of the bricks, and the trace reads it."""
        self.buf = np.zeros(self.words, np.uint32)
        self.rep = np.zeros(self.vox.maxBricks * 512, np.uint32)
        self._scene = (f32(scene.pos), f32(scene.nrm), np.ascontiguousarray(scene.tri, np.uint32),
                       f32(scene.materials))
        pos, nrm, tri, mat = self._scene
        nstatic = tri.shape[0] if dyn is None else min(tri.shape[0], dyn.first_tri)
        surf = getattr(scene, "surface", None)
        self.surface = SurfaceData(surf) if surf is not None else None
        lib().orc_set_surface(C.byref(self.surface.struct) if self.surface is not None else None)
        try:
            rc = lib().orc_build(C.byref(self.vp), _p(pos), _p(nrm), _p(tri), nstatic, _p(mat), _p(self.buf),
                                 _p(self.rep))
            assert rc == 0
            if dyn is not None:
                self._dyn_struct = dyn.struct(tri.shape[0])
                assert lib().orc_dyn_build(C.byref(self.vp), C.byref(self.rc), _p(pos), _p(nrm), _p(tri), _p(mat),
                                           C.byref(self._dyn_struct)) == 0
                self.rc.dyn = C.addressof(self._dyn_struct)
                self.dyn = dyn
        finally:
            lib().orc_set_surface(None)
        if self._sun_trace:
            self.sunvis = np.zeros(self.vox.maxVoxels, np.float32)
            assert lib().orc_sun(C.byref(self.vp), _p(self.buf), C.byref(self.rc), _p(self.sunvis)) == 0
            self.rc.sunvis = self.sunvis.ctypes.data
        if self.rc.reloc:
            self.relocate()
        if self._bounce:
            self.bounce = np.zeros((self.vox.maxVoxels, 4), np.uint32)
            self.rc.bounce = self.bounce.ctypes.data

    def bounce_update(self, prev=None):
        """The bounce volume of this frame (bnClassify + bnUpdate, hits reading the previous frame's volume) and its
        injection into this frame's voxels (bnInject), as an update with multi-bounce runs them before the trace.
        prev: the previous frame's (volume, origin, update number) as returned here, or None (first update).
        Returns this frame's (volume [probes][6][4] halves as uint16, origin, update number)."""
        assert self._bounce and self.buf is not None
        org = bn_origin(self._eye)
        n = 1 if prev is None else prev[2] + 1
        vol = np.zeros((BN_DIMS[0] * BN_DIMS[1] * BN_DIMS[2], 6, 4), np.uint16)
        pv, po = (None, None) if prev is None else (np.ascontiguousarray(prev[0], np.uint16),
                                                     np.ascontiguousarray(prev[1], np.int32))
        if self.rc.objects and self.rc.dyn and pv is not None:
            # dynInject: the dynamic layer receives the previous volume's light before this volume's rays read it
            assert lib().orc_dyn_inject(C.byref(self.vp), C.byref(self._dyn_struct), _p(pv), _p(po)) == 0
        assert lib().orc_bounce_volume(C.byref(self.vp), _p(self.buf), C.byref(self.rc), _p(pv), _p(po), n, _p(org),
                                       _p(vol)) == 0
        self.inject_volume(vol, org)
        return vol, org, n

    def inject_volume(self, vol, org):
        """bnInject from a given volume ([probes][6][4] halves as uint16; origin in probe coordinates)."""
        assert self._bounce and self.buf is not None
        vol = np.ascontiguousarray(vol, np.uint16)
        org = np.ascontiguousarray(org, np.int32)
        assert lib().orc_bounce_inject(C.byref(self.vp), _p(self.buf), _p(vol), _p(org), _p(self.bounce),
                                       self.rc.bf16) == 0

    def relocate(self):
        """relocateVox for every cascade, then the parent visibility masks (linkParents with voxel relocation);
        the trace, merge and gather then use them."""
        N = self.rc.cascades
        self._pos, self._nrm, self._vis = [], [], []
        for c in range(N):
            n = (self.rc.probes0 >> c) ** 3
            pos = np.zeros((n, 3), np.float32)
            nrm = np.zeros((n, 3), np.float32)
            assert lib().orc_relocate(C.byref(self.vp), _p(self.buf), C.byref(self.rc), c, _p(pos), _p(nrm)) == 0
            self._pos.append(pos)
            self._nrm.append(nrm)
        for c in range(N):
            vis = np.zeros((self.rc.probes0 >> c) ** 3, np.uint32)
            if c + 1 < N:
                assert lib().orc_link(C.byref(self.vp), _p(self.buf), C.byref(self.rc), c, _p(self._pos[c]),
                                      _p(self._pos[c + 1]), _p(vis)) == 0
            self._vis.append(vis)
            self.rc.ppos[c] = self._pos[c].ctypes.data
            self.rc.pnrm[c] = self._nrm[c].ctypes.data
            self.rc.pvis[c] = vis.ctypes.data

    def probes(self, c):
        """Relocated positions, normals and parent masks of cascade c (grid order)."""
        return self._pos[c], self._nrm[c], self._vis[c]

    def counters(self):
        b = self.buf
        NL = self.vox.levels
        return dict(bricks=int(min(b[0], self.vox.maxBricks)), marked=int(b[0]), voxels=int(b[1]),
                    brickOverflow=int(b[2]), bigOverflow=int(b[3]), voxelOverflow=int(b[4]),
                    bricksPerLevel=[int(x) for x in b[8:8 + NL]], bigPerLevel=[int(x) for x in b[16:16 + NL]])

    def clipmap(self, level):
        n = self.E ** 3
        occ = np.zeros(n // 32, np.uint32)
        rp = np.zeros(n, np.uint32)
        alb = np.zeros(n, np.uint32)
        nrm = np.zeros(n, np.uint32)
        pl = np.zeros(n, np.float32)
        st = np.zeros(self.vox.G ** 3, np.uint8)
        rc = lib().orc_canonical(C.byref(self.vp), _p(self.buf), _p(self.rep), level, _p(occ), _p(rp), _p(alb),
                                 _p(nrm), _p(pl), _p(st))
        assert rc == 0
        return dict(occ=occ, rep=rp, albedo=alb, normal=nrm, plane=pl, brick_state=st)

    def trace(self, r0=0, r1=None, threads=0, steps=False):
        total = sum(self.cascade_rays())
        r1 = total if r1 is None else r1
        rgba = np.zeros((r1 - r0, 4), np.float32)
        hits = np.zeros(r1 - r0, np.uint32)
        st = np.zeros(r1 - r0, np.uint32) if steps else None
        lib().orc_trace(C.byref(self.vp), _p(self.buf), C.byref(self.rc), r0, r1, _p(rgba), _p(hits), _p(st), threads)
        return (rgba, hits, st) if steps else (rgba, hits)

    def merge_sh(self, rgba, hist=None):
        """mergeWorld and projectSH; with hist (a History), the c0 SH history blend of this update, which advances
        hist. Without it, the SH of a first update (no history yet)."""
        rgba = np.ascontiguousarray(rgba, np.float32).copy()
        n = self.rc.probes0 ** 3
        sh = np.zeros((n, 27), np.float32)
        if hist is None:
            lib().orc_merge_sh(C.byref(self.rc), _p(rgba), _p(sh))
        else:
            assert lib().orc_merge_sh_hist(C.byref(self.rc), _p(rgba), _p(sh), C.byref(hist.s)) == 0
        return rgba, sh

    def trace_f64(self, r0=0, r1=None, eps=1e-6, threads=0):
        total = sum(self.cascade_rays())
        r1 = total if r1 is None else r1
        rgbt = np.zeros((r1 - r0, 4), np.float64)
        hits = np.zeros(r1 - r0, np.uint32)
        tie = np.zeros(r1 - r0, np.uint8)
        lib().orc_trace_f64(C.byref(self.vp), _p(self.buf), C.byref(self.rc), r0, r1, _p(rgbt), _p(hits), _p(tie),
                            eps, threads)
        return rgbt, hits, tie

    def trace_f64_take(self, hits32, eps=1e-6, threads=0):
        """The FP64 trace of every ray, taking the f32 walk's first-hit voxel (hits32, as trace() returns them) on tie
        rays and on rays whose FP64 first hit differs: rgbt (double) and the rays taken so."""
        n = sum(self.cascade_rays())
        hits32 = np.ascontiguousarray(hits32, np.uint32)
        assert hits32.shape[0] == n
        rgbt = np.zeros((n, 4), np.float64)
        taken = np.zeros(n, np.uint8)
        assert lib().orc_trace_f64_take(C.byref(self.vp), _p(self.buf), C.byref(self.rc), 0, n, _p(hits32), _p(rgbt),
                                        _p(taken), eps, threads) == 0
        return rgbt, taken

    def trace_rays(self, c, o, d, threads=0):
        """Rays from origins o along unit directions d (n x 3 float32) traced as cascade c's rays (f32)."""
        o = np.ascontiguousarray(o, np.float32)
        d = np.ascontiguousarray(d, np.float32)
        n = o.shape[0]
        rgba = np.zeros((n, 4), np.float32)
        hits = np.zeros(n, np.uint32)
        assert lib().orc_trace_rays(C.byref(self.vp), _p(self.buf), C.byref(self.rc), c, _p(o), _p(d), n, _p(rgba),
                                    _p(hits), threads) == 0
        return rgba, hits

    def trace_rays_f64(self, c, o, d, eps=1e-6, threads=0):
        """The same rays in double: rgbt, hit ids, tie flags."""
        o = np.ascontiguousarray(o, np.float64)
        d = np.ascontiguousarray(d, np.float64)
        n = o.shape[0]
        rgbt = np.zeros((n, 4), np.float64)
        hits = np.zeros(n, np.uint32)
        tie = np.zeros(n, np.uint8)
        assert lib().orc_trace_rays_f64(C.byref(self.vp), _p(self.buf), C.byref(self.rc), c, _p(o), _p(d), n,
                                        _p(rgbt), _p(hits), _p(tie), eps, threads) == 0
        return rgbt, hits, tie

    def gather(self, sh, pos, nrm, occlusion=False):
        """gatherWorld at surface points (n x 3 pos, unit nrm) from sh ([P0][27]) -> irradiance (n x 3)."""
        sh = np.ascontiguousarray(sh, np.float32)
        pos = np.ascontiguousarray(pos, np.float32)
        nrm = np.ascontiguousarray(nrm, np.float32)
        irr = np.zeros((pos.shape[0], 3), np.float32)
        lib().orc_gather(C.byref(self.vp), _p(self.buf), C.byref(self.rc), _p(sh), _p(pos), _p(nrm), pos.shape[0],
                         1 if occlusion else 0, _p(irr))
        return irr

    def merge_sh_f64(self, rgbt):
        rgbt = np.ascontiguousarray(rgbt, np.float64).copy()
        n = self.rc.probes0 ** 3
        sh = np.zeros((n, 27), np.float64)
        lib().orc_merge_sh_f64(C.byref(self.rc), _p(rgbt), _p(sh))
        return rgbt, sh

    def split(self, arr):
        """Splits a full ray array into per-cascade arrays."""
        out, o = [], 0
        for n in self.cascade_rays():
            out.append(arr[o:o + n])
            o += n
        return out
