"""Pins the CPU oracle (oracle/rtgi_oracle.c) against facts that do not come from its own formulas.

The reference ships no golden vectors or tests for this path, so each oracle function is checked against an
independent statement of what it computes: LP feasibility for triangle/box overlap, brute-force minimisation for the
closest point, the WGSL specification's packing values, a double-precision brute-force voxeliser, trilinear
interpolation for the merge, scipy's real spherical harmonics for the SH basis, and the FP64 tracer for first hits.
"""
import dataclasses
import json
import os

import numpy as np
import pytest
from scipy.optimize import linprog

import rtgi
from conftest import random_scene
from oracle import oracle as orc
from oracle.oracle import DynState, Oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ---------------------------------------------------------------------------------------------------------------
# triBox (voxel.ts:231-248) against LP feasibility: max s such that a point of the triangle lies in the box shrunk by s.

def overlap_margin(c, h, a, b, cc):
    V = np.stack([a, b, cc]).astype(np.float64)  # 3 x 3
    # variables: l0 l1 l2 s ; maximise s  <=>  minimise -s
    A_ub, b_ub = [], []
    for ax in range(3):
        row = np.zeros(4)
        row[:3] = V[:, ax]
        row[3] = 1.0
        A_ub.append(row.copy())  # p_ax + s <= c + h
        b_ub.append(c[ax] + h[ax])
        row[:3] = -V[:, ax]
        A_ub.append(row)  # -p_ax + s <= -(c - h)
        b_ub.append(-(c[ax] - h[ax]))
    r = linprog([0, 0, 0, -1], A_ub=np.array(A_ub), b_ub=np.array(b_ub), A_eq=[[1, 1, 1, 0]], b_eq=[1],
                bounds=[(0, None)] * 3 + [(None, None)], method="highs")
    assert r.status == 0
    return -r.fun


def test_tri_box_against_lp():
    rng = np.random.default_rng(5)
    agree, checked = 0, 0
    for _ in range(3000):
        c = rng.uniform(-1, 1, 3)
        h = rng.uniform(0.05, 0.5, 3)
        a, b, cc = (c + rng.normal(0, 0.8, 3) for _ in range(3))
        c, h, a, b, cc = (np.float32(x) for x in (c, h, a, b, cc))
        m = overlap_margin(c, h, a, b, cc)
        if abs(m) < 1e-4:
            continue  # within rounding of touching
        checked += 1
        agree += orc.tri_box(c, h, a, b, cc) == (m > 0)
    assert checked > 2500 and agree == checked


def test_tri_box_exact_dyadic_cases():
    # Dyadic coordinates make every f32 operation exact: touching faces count as overlap (closed boxes).
    h = [0.5, 0.5, 0.5]
    assert orc.tri_box([0, 0, 0], h, [0.5, 0, 0], [1, 1, 0], [1, 0, 1])
    assert not orc.tri_box([0, 0, 0], h, [0.5 + 2 ** -10, 0, 0], [1, 1, 0], [1, 0, 1])
    # The triangle plane passes the box corner exactly (plane test boundary).
    assert orc.tri_box([0, 0, 0], h, [1.5, 0, 0], [0, 1.5, 0], [0, 0, 1.5])
    assert not orc.tri_box([0, 0, 0], h, [1.5 + 2 ** -8, 0, 0], [0, 1.5 + 2 ** -8, 0], [0, 0, 1.5 + 2 ** -8])
    # An edge axis separates: a sliver crossing near a box edge but outside it (only the 9 cross axes catch this).
    assert not orc.tri_box([0, 0, 0], h, [1.0, 0.75, -2], [0.75, 1.0, -2], [0.875, 0.875, 2])


# ---------------------------------------------------------------------------------------------------------------
# closestBary (voxel.ts:250-267) against brute-force minimisation of the distance.

def test_closest_bary_brute_force():
    rng = np.random.default_rng(7)
    g = np.linspace(0, 1, 401)
    L0, L1 = np.meshgrid(g, g)
    keep = L0 + L1 <= 1
    bary = np.stack([L0[keep], L1[keep], 1 - L0[keep] - L1[keep]], 1)
    for _ in range(200):
        a, b, c = (rng.normal(0, 1, 3).astype(np.float32) for _ in range(3))
        p = rng.normal(0, 1.5, 3).astype(np.float32)
        w = orc.closest_bary(p, a, b, c)
        assert np.all(w >= -1e-5) and abs(w.sum() - 1) < 1e-5
        q = w[0] * a.astype(np.float64) + w[1] * b + w[2] * c
        d_orc = np.linalg.norm(q - p)
        pts = bary @ np.stack([a, b, c]).astype(np.float64)
        d_bf = np.min(np.linalg.norm(pts - p, axis=1))
        assert d_orc <= d_bf + 1e-4
        assert d_orc >= d_bf - 1e-2  # the grid is 1/400 fine


# ---------------------------------------------------------------------------------------------------------------
# pack4x8unorm / pack2x16snorm (WGSL spec) and octEnc / octDecode.

def test_pack_golden_vectors():
    vec = json.load(open(os.path.join(GOLDEN, "pack_vectors.json")))
    for e, expect in vec["pack4x8unorm"]:
        assert orc.pack4x8unorm(e) == int(expect, 16), e
    for (x, y), expect in vec["pack2x16snorm"]:
        assert orc.pack2x16snorm(x, y) == int(expect, 16), (x, y)


def oct_decode(u, v):
    """octDecode (rc.ts:291-298) in double, for round trips."""
    f = np.array([u * 2 - 1, v * 2 - 1])
    n = np.array([f[0], f[1], 1 - abs(f[0]) - abs(f[1])])
    t = max(-n[2], 0)
    n[0] += -t if n[0] >= 0 else t
    n[1] += -t if n[1] >= 0 else t
    return n / np.linalg.norm(n)


def test_oct_encode_axes_and_round_trip():
    assert np.allclose(orc.oct_enc([0, 0, 1]), [0, 0])
    assert np.allclose(orc.oct_enc([1, 0, 0]), [1, 0])
    assert np.allclose(orc.oct_enc([0, -1, 0]), [0, -1])
    assert np.allclose(orc.oct_enc([0, 0, -1]), [1, 1])
    rng = np.random.default_rng(3)
    for _ in range(500):
        n = rng.normal(0, 1, 3)
        n /= np.linalg.norm(n)
        e = orc.oct_enc(n.astype(np.float32))
        assert np.all(np.abs(e) <= 1 + 1e-6)
        m = oct_decode(e[0] * 0.5 + 0.5, e[1] * 0.5 + 0.5)
        assert np.allclose(m, n, atol=1e-5)


def test_f16_rounding_matches_ieee():
    """pack2x16float's f32 -> f16 conversion (round to nearest even) against numpy's IEEE binary16 cast."""
    rng = np.random.default_rng(0)
    vals = np.concatenate([rng.normal(0, 1, 20000) * np.exp(rng.uniform(-30, 12, 20000)),
                           [0.0, -0.0, 65504.0, 65519.99, 6e-8, 5.96e-8, 2.98e-8, 2 ** -25, 2 ** -24, 1e-40,
                            1.0 + 2 ** -11, 1.0 + 3 * 2 ** -11]]).astype(np.float32)
    with np.errstate(over="ignore"):
        ref = vals.astype(np.float16).view(np.uint16)
    got = np.array([orc.f2h(v) for v in vals], np.uint16)
    assert np.array_equal(got, ref)


def test_oct_solid_angles():
    w = orc.oct_solid_angles(4)
    assert abs(w.astype(np.float64).sum() - 4 * np.pi) < 1e-5
    W = w.reshape(4, 4)
    # Mirror symmetries of the octahedral map (u -> 1-u, v -> 1-v, transpose) hold up to the Fibonacci count noise.
    for M in (W[:, ::-1], W[::-1, :], W.T):
        assert np.allclose(M, W, rtol=2e-3)
    assert np.array_equal(w, rtgi.oct_solid_angles(4))  # the library computes the same table


# ---------------------------------------------------------------------------------------------------------------
# Voxel build against a brute-force voxeliser in double precision.

def sat_margin_np(P, h, a, b, c):
    """Vectorised SAT in double for boxes (centres P: n x 3, half h) against one triangle: the minimum over the 13
    axes of the overlap slack (> 0 overlap, < 0 separated), normalised by the axis length."""
    v0, v1, v2 = a - P, b - P, c - P
    e = [v1 - v0, v2 - v1, v0 - v2]
    axes = [np.eye(3)[i] for i in range(3)]
    n = np.cross(b - a, c - a)
    axes.append(n)
    for ei in (b - a, c - b, a - c):
        for k in range(3):
            axes.append(np.cross(np.eye(3)[k], ei))
    slack = np.full(P.shape[0], np.inf)
    for ax in axes:
        ln = np.linalg.norm(ax)
        if ln < 1e-12:
            continue
        p0, p1, p2 = v0 @ ax, v1 @ ax, v2 @ ax
        r = np.abs(ax) @ h
        s = np.minimum(r - np.minimum(np.minimum(p0, p1), p2), np.maximum(np.maximum(p0, p1), p2) + r) / ln
        slack = np.minimum(slack, s)
    del e
    return slack


def brute_force_clipmap(scene, vox, origins):
    """Occupancy and representative per level by direct overlap tests (double), coarse cells inside the finer
    window as the OR / max of their 8 children, clipped single-cell ranges occupied without a test. Returns
    (occ, rep, ambiguous) per level as E^3 arrays."""
    G, E, NL = vox.G, vox.G * 8, vox.levels
    idx = np.arange(E)
    Z, Y, X = np.meshgrid(idx, idx, idx, indexing="ij")
    cells = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    out = []
    for L in range(NL):
        v = vox.v0 * 2 ** L
        lo = np.array(origins[L], np.float64) * v * 8
        rep = np.zeros(E ** 3, np.int64)
        amb = np.zeros(E ** 3, bool)
        if L > 0:
            f = (np.array(origins[L]) * 8 + cells) * 2 - np.array(origins[L - 1]) * 8
            held = np.all((f >= 0) & (f < E), 1)
        else:
            held = np.zeros(E ** 3, bool)
        P = lo + (cells + 0.5) * v
        h = np.full(3, 0.5 * v * 1.0001)
        for t in range(scene.ntris):
            i0, i1, i2, _ = scene.tri[t]
            a, b, c = (scene.pos[i].astype(np.float64) for i in (i0, i1, i2))
            mn, mx = np.minimum(np.minimum(a, b), c), np.maximum(np.maximum(a, b), c)
            ra = np.clip(np.floor((mn - lo) / v), 0, E).astype(int)
            rb = np.clip(np.floor((mx - lo) / v) + 1, 0, E).astype(int)
            n = np.maximum(rb - ra, 0)
            if n.prod() == 0:
                continue
            inr = np.all((cells >= ra) & (cells < rb), 1) & ~held
            if n.prod() == 1:
                hit = inr
            else:
                s = sat_margin_np(P[inr], h, a, b, c)
                hit = np.zeros(E ** 3, bool)
                hit[np.nonzero(inr)[0][s >= 0]] = True
                amb[np.nonzero(inr)[0][np.abs(s) < 1e-5]] = True
            rep[hit] = np.maximum(rep[hit], t + 1)
        if L > 0:
            prev_rep, prev_amb = out[L - 1][1], out[L - 1][2]
            fine = np.stack(np.unravel_index(np.arange(E ** 3), (E, E, E))[::-1], 1)  # x, y, z of finer cells
            parent = ((np.array(origins[L - 1]) * 8 + fine) >> 1) - np.array(origins[L]) * 8
            pi = parent[:, 0] + E * (parent[:, 1] + E * parent[:, 2])
            np.maximum.at(rep, pi, prev_rep)
            np.logical_or.at(amb, pi, prev_amb)
        out.append((rep > 0, rep, amb))
    return out


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_build_against_brute_force(seed):
    rng = np.random.default_rng(seed)
    scene = random_scene(rng, 25, -3.0, 3.0, 0.5)
    vox = rtgi.VoxParams(v0=0.25, levels=3, G=2, maxBricks=3 * 8, maxVoxels=1 << 14)
    rc = rtgi.RcParams(cascades=2, probes0=4, s0=0.5)
    frame = rtgi.Frame(eye=(0.1, 0.05, -0.2))
    o = Oracle(vox, rc, frame)
    o.build(scene)
    org = orc.origins(vox.v0, vox.levels, vox.G, frame.eye)
    bf = brute_force_clipmap(scene, vox, org)
    E = vox.G * 8
    for L in range(vox.levels):
        cm = o.clipmap(L)
        occ = np.unpackbits(cm["occ"].view(np.uint8), bitorder="little").astype(bool)[:E ** 3]
        occ_bf, rep_bf, amb = bf[L]
        ok = ~amb
        assert np.array_equal(occ[ok], occ_bf[ok]), f"level {L}"
        assert np.array_equal(cm["rep"][ok & occ].astype(np.int64), rep_bf[ok & occ]), f"level {L}"
        assert occ_bf.sum() > 0


def test_build_invariants_atrium():
    scene, vox, rc, frame = rtgi.config("atrium")
    o = Oracle(vox, rc, frame)
    o.build(scene)
    E = vox.G * 8
    org = orc.origins(vox.v0, vox.levels, vox.G, frame.eye)
    prev = None
    for L in range(vox.levels):
        cm = o.clipmap(L)
        occ = np.unpackbits(cm["occ"].view(np.uint8), bitorder="little").astype(bool)[:E ** 3].reshape(E, E, E)
        rep = cm["rep"].reshape(E, E, E)
        assert np.array_equal(occ, rep > 0)
        assert rep.max() <= scene.ntris
        # Every occupied voxel carries attributes (no voxel overflow here) and a unit-length oct normal code.
        assert np.all(cm["albedo"].reshape(E, E, E)[occ] != 0)
        if prev is not None:
            # coarse = OR of children wherever the finer window holds the cell (voxel.ts:16-19).
            pocc, prep = prev
            d = (np.array(org[L - 1]) * 8 - np.array(org[L]) * 16) // 2  # finer window origin in coarse cells
            sub = occ[d[2]:d[2] + E // 2, d[1]:d[1] + E // 2, d[0]:d[0] + E // 2]
            kids = pocc.reshape(E // 2, 2, E // 2, 2, E // 2, 2).any(axis=(1, 3, 5))
            assert np.array_equal(sub, kids)
            kr = prep.reshape(E // 2, 2, E // 2, 2, E // 2, 2).max(axis=(1, 3, 5))
            assert np.array_equal(rep[d[2]:d[2] + E // 2, d[1]:d[1] + E // 2, d[0]:d[0] + E // 2], kr)
        prev = (occ, rep)


def test_pool_overflow_drops_finest_first():
    rng = np.random.default_rng(9)
    scene = random_scene(rng, 300, -4.0, 4.0, 0.6)
    base = rtgi.VoxParams(v0=0.25, levels=3, G=4, maxBricks=3 * 64, maxVoxels=1 << 16)
    rc = rtgi.RcParams(cascades=2, probes0=4, s0=0.5)
    frame = rtgi.Frame(eye=(0.1, 0.2, 0.3))
    full = Oracle(base, rc, frame)
    full.build(scene)
    nb = full.counters()["bricks"]
    small = dataclasses.replace(base, maxBricks=nb - 5)
    o = Oracle(small, rc, frame)
    o.build(scene)
    c = o.counters()
    assert c["brickOverflow"] == 5 and c["bricks"] == nb - 5
    # Coarse levels keep every brick; only level-0 bricks are dropped (src/core/voxel.ts:405-410).
    assert c["bricksPerLevel"][1:] == full.counters()["bricksPerLevel"][1:]
    st0 = o.clipmap(0)["brick_state"]
    assert np.count_nonzero(st0 == 2) == 5
    # Dropped level-0 cells are voxelised at level 1 from the triangles, so level 1 still covers every occupied
    # level-0 cell of the full build.
    E = base.G * 8
    occ1 = np.unpackbits(o.clipmap(1)["occ"].view(np.uint8), bitorder="little")[:E ** 3].reshape(E, E, E)
    occ0f = np.unpackbits(full.clipmap(0)["occ"].view(np.uint8), bitorder="little")[:E ** 3].reshape(E, E, E)
    org = orc.origins(base.v0, base.levels, base.G, frame.eye)
    d = (np.array(org[0]) * 8 - np.array(org[1]) * 16) // 2
    kids = occ0f.reshape(E // 2, 2, E // 2, 2, E // 2, 2).any(axis=(1, 3, 5))
    assert np.all(occ1[d[2]:d[2] + E // 2, d[1]:d[1] + E // 2, d[0]:d[0] + E // 2][kids])


# ---------------------------------------------------------------------------------------------------------------
# Trace against the FP64 oracle, special cases, merge and SH.

@pytest.mark.parametrize("s0,min_clean,sh_median", [(0.625, 0.35, 0.05), (0.61, 0.8, 1e-3)])
def test_trace_matches_fp64_away_from_ties(s0, min_clean, sh_median):
    """First hits equal the FP64 voxel walk on every ray that passes no voxel edge, corner, interval end or skip
    distance within 1e-5 (relative); with probes on voxel planes (s0 = 0.625 = 5 v0) many rays pass exact edges."""
    scene, vox, rc, frame = rtgi.config("atrium", s0=s0)
    o = Oracle(vox, rc, frame)
    o.build(scene)
    rgba, hits = o.trace()
    rgbt, hits64, tie = o.trace_f64(eps=1e-5)
    clean = tie == 0
    assert clean.mean() > min_clean
    assert np.array_equal(hits[clean], hits64[clean])
    assert (hits != hits64).mean() < 0.02
    hit = clean & (hits != 0xFFFFFFFF)
    assert hit.sum() > 1000
    err = np.abs(rgba[hit, :3] - rgbt[hit, :3]) / np.maximum(np.abs(rgbt[hit, :3]), 1e-3)
    assert err.max() < 1e-4 and np.median(err) < 1e-6
    # Merged SH against the FP64 pipeline: dominated by the rays whose first hit differs at exact edges (each probe
    # averages many upper-cascade rays), so the bound is on the median and depends on how many rays pass edges.
    m32, sh32 = o.merge_sh(rgba)
    m64, sh64 = o.merge_sh_f64(rgbt)
    rel = np.abs(sh32 - sh64).max(axis=1) / np.maximum(np.abs(sh64).max(axis=1), 1e-6)
    assert np.median(rel) < sh_median


def sphere_rays(o, c, n, seed):
    """n rays of cascade c: origins at uniformly drawn probe positions, directions uniform on the sphere."""
    rng = np.random.default_rng(seed)
    N = o.rc.probes0 >> c
    g = rng.integers(0, N, (n, 3)) + np.array(o.rc.G0[:3]) // (1 << c)
    org = (g.astype(np.float32) * np.float32(o.rc.s0 * 2 ** c)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    return org, d


def test_ray_queries_match_cascade_trace_and_fp64():
    """orc_trace_rays on a cascade's own probe origins and table directions reproduces orc_trace bit for bit; on
    random sphere directions its first hits equal the FP64 walk's away from ties, and radiance agrees to 1e-4."""
    scene, vox, rc, frame = rtgi.config("atrium", s0=0.61)
    o = Oracle(vox, rc, frame)
    o.build(scene)
    rgba, hits = o.trace(r0=0, r1=4096)
    res = 4
    pi, di = np.arange(4096) // (res * res), np.arange(4096) % (res * res)
    n = rc.probes0
    g = np.stack([pi % n, (pi // n) % n, pi // (n * n)], 1) + np.array(o.rc.G0[:3])
    org = (g.astype(np.float32) * np.float32(rc.s0)).astype(np.float32)
    dirs = np.array([orc_dir(int(x), res) for x in di], np.float32)
    r2, h2 = o.trace_rays(0, org, dirs)
    assert np.array_equal(h2, hits) and np.array_equal(r2.view(np.uint32), rgba.view(np.uint32))
    for c in range(rc.cascades):
        org, d = sphere_rays(o, c, 4000, seed=c)
        r32, h32 = o.trace_rays(c, org, d)
        r64, h64, tie = o.trace_rays_f64(c, org.astype(np.float64), d.astype(np.float64), eps=1e-5)
        clean = tie == 0
        # Tie flags are conservative: a start on the window's face (rays entering from outside, most of the top
        # cascade's long intervals) counts as a tie.
        assert clean.sum() > 500 and np.array_equal(h32[clean], h64[clean])
        hit = clean & (h32 != 0xFFFFFFFF)
        err = np.abs(r32[hit, :3] - r64[hit, :3]) / np.maximum(np.abs(r64[hit, :3]), 1e-3)
        assert hit.sum() > 50 and err.max() < 1e-4


def orc_dir(di, res):
    """dirOf (src/core/rc.ts:291-298) in f32 with the operation order of the device."""
    f = np.float32
    u = (f(di % res) + f(0.5)) / f(res)
    v = (f(di // res) + f(0.5)) / f(res)
    fx, fy = u * f(2) - f(1), v * f(2) - f(1)
    nz = f(1) - abs(fx) - abs(fy)
    t = max(-nz, f(0))
    fx = fx - t if fx >= 0 else fx + t
    fy = fy - t if fy >= 0 else fy + t
    ln = np.sqrt(fx * fx + fy * fy + nz * nz, dtype=np.float32)
    return [fx / ln, fy / ln, nz / ln]


def wall_world(x=0.35):
    """A wall of two triangles in the plane x (0.35: occupying the voxels 0.25 <= x < 0.375 at v0 = 0.125)."""
    pos = np.array([[x, -6, -6], [x, 6, -6], [x, 6, 6], [x, -6, 6]], np.float32)
    nrm = np.tile(np.array([[1.0, 0, 0]], np.float32), (4, 1))
    tri = np.array([[0, 1, 2, 0], [0, 2, 3, 0]], np.uint32)
    scene = rtgi.Scene(pos=pos, nrm=nrm, tri=tri, materials=np.array([[0.5, 0.5, 0.5, 1]], np.float32))
    vox = rtgi.VoxParams(v0=0.125, levels=3, G=8, maxBricks=3 * 512, maxVoxels=1 << 16)
    rc = rtgi.RcParams(cascades=2, probes0=16, s0=0.5)
    return scene, vox, rc, rtgi.Frame(eye=(0.0, 0.0, 0.0))


def test_gather_closed_forms():
    """gatherWorld on hand-made SH fields. (1) Every probe holds the same SH: any weighting gives that SH's value,
    so the result is evalSH of it (L0 only: c * 0.282095 for every normal). (2) A point on a probe: the trilinear
    weight puts 1 + 1e-4 on that probe and 1e-4 on the others, so the result is within 1e-3 of that probe's value,
    which pins the corner order and the probe index. (3) Occlusion: probes behind a wall are left out, so a point
    whose far-side probes are all hidden gets exactly the near side's value."""
    scene, vox, rc, frame = wall_world()
    o = Oracle(vox, rc, frame)
    o.build(scene)
    N = rc.probes0
    rng = np.random.default_rng(3)
    pos = rng.uniform(-3, 3, (500, 3)).astype(np.float32)
    nrm = rng.normal(size=(500, 3))
    nrm = (nrm / np.linalg.norm(nrm, axis=1, keepdims=True)).astype(np.float32)
    sh = np.zeros((N ** 3, 27), np.float32)
    sh[:, 0:3] = [2.0, 3.0, 5.0]
    irr = o.gather(sh, pos, nrm)
    inside = np.all((pos / rc.s0 >= np.array(o.rc.G0[:3])) & (pos / rc.s0 < np.array(o.rc.G0[:3]) + N - 1), axis=1)
    np.testing.assert_allclose(irr[inside], np.tile(np.float32(0.282095) * np.array([2, 3, 5], np.float32),
                                                    (inside.sum(), 1)), rtol=2e-6)
    # (2) distinct L0 per probe, points on probes.
    sh2 = np.zeros_like(sh)
    sh2[:, 0] = np.arange(N ** 3, dtype=np.float32) + 1
    g = rng.integers(1, N - 1, (200, 3))
    p2 = ((g + np.array(o.rc.G0[:3])) * rc.s0).astype(np.float32)
    out = o.gather(sh2, p2, np.tile(np.array([[0, 1, 0]], np.float32), (200, 1)))
    idx = g[:, 0] + N * (g[:, 1] + N * g[:, 2])
    np.testing.assert_allclose(out[:, 0], np.float32(0.282095) * (idx + 1), rtol=1e-3)
    # (3) the wall at x = 0.35: near-side probes (x <= 0) hold 1, far-side ones (x >= 0.5) hold 100.
    G0 = np.array(o.rc.G0[:3])
    gx = np.arange(N) + G0[0]
    sh3 = np.zeros_like(sh)
    L0 = np.where(gx * rc.s0 <= 0.0, 1.0, 100.0).astype(np.float32)
    sh3[:, 0] = np.tile(L0, N * N)  # probe index x + N (y + N z): x fastest
    m = 100
    p3 = np.stack([np.full(m, 0.1), rng.uniform(-2, 2, m), rng.uniform(-2, 2, m)], 1).astype(np.float32)
    n3 = np.tile(np.array([[-1.0, 0, 0]], np.float32), (m, 1))
    occ = o.gather(sh3, p3, n3, occlusion=True)
    np.testing.assert_allclose(occ[:, 0], np.float32(0.282095), rtol=1e-6)
    free = o.gather(sh3, p3, n3, occlusion=False)
    assert np.all(free[:, 0] > 1.5 * 0.282095)


def test_relocation_closed_forms():
    """relocateVox (rc.ts:413-472) worked by hand for a wall in the plane x = 0.52 (in the level-0 voxels
    0.5 <= x < 0.625 and the level-1 voxels 0.5 <= x < 0.75), c0 spacing 0.5, lift 0.25:
    - a c0 probe on the grid plane x = 0.5 sits in an occupied voxel. Its 4^3 cube at level 1 has occupied cells
      only at +0.125, so the normal is -x. Marching by max(v/2, 1.5 s/16) = 0.0625 with 0.125 of free space required
      behind it, the first free step is the third: x = 0.3125.
    - a c1 probe at x = 1 (spacing 1, cube at level 2 occupied at -0.25) gets normal +x and stays: its voxel and the
      one 0.25 behind it are free.
    - probes with no occupied voxel in their cube keep their grid point and get normal 0.
    - linkParents: from the relocated x = 0.3125 probes, the parents at x = 1 (behind the wall) are hidden and those
      at x = 0 are used."""
    scene, vox, rc, frame = wall_world(x=0.52)
    rc = dataclasses.replace(rc, relocate=True)
    o = Oracle(vox, rc, frame)
    o.build(scene)
    G0 = np.array(o.rc.G0[:3])
    n0 = rc.probes0
    pos0, nrm0, vis0 = o.probes(0)
    idx = np.arange(n0 ** 3)
    g = np.stack([idx % n0, (idx // n0) % n0, idx // (n0 * n0)], 1) + G0
    grid = (g * np.float32(rc.s0)).astype(np.float32)
    wallcol = (g[:, 0] * rc.s0 == 0.5) & (np.abs(grid[:, 1]) < 3) & (np.abs(grid[:, 2]) < 3)
    assert wallcol.sum() > 50
    np.testing.assert_array_equal(pos0[wallcol, 0], np.float32(0.3125))
    np.testing.assert_array_equal(pos0[wallcol, 1:], grid[wallcol, 1:])
    np.testing.assert_array_equal(nrm0[wallcol], np.tile(np.array([[-1, 0, 0]], np.float32), (wallcol.sum(), 1)))
    far = np.abs(grid[:, 0] - 0.52) > 1.6
    np.testing.assert_array_equal(pos0[far], grid[far])
    assert not np.any(nrm0[far])
    pos1, nrm1, _ = o.probes(1)
    n1 = n0 // 2
    i1 = np.arange(n1 ** 3)
    g1 = np.stack([i1 % n1, (i1 // n1) % n1, i1 // (n1 * n1)], 1) + G0 // 2
    at1 = (g1[:, 0] == 1) & (np.abs(g1[:, 1]) < 3) & (np.abs(g1[:, 2]) < 3)
    np.testing.assert_array_equal(pos1[at1, 0], np.float32(1.0))
    np.testing.assert_array_equal(nrm1[at1], np.tile(np.array([[1, 0, 0]], np.float32), (at1.sum(), 1)))
    # Parents of the relocated wall-column probes: bit k has o.x = k & 1.
    ox1 = sum(1 << k for k in range(8) if k & 1)
    inner = wallcol & (np.abs(grid[:, 1]) < 2) & (np.abs(grid[:, 2]) < 2)
    assert np.all(vis0[inner] & ox1 == 0)
    assert np.all(vis0[inner] & 1)  # the o = 0 parent (x = 0) is in use


def roof_world():
    """A floor quad at y = 0 (x, z in [-6, 6]) under a roof quad at y = 2 over x in [-6, 0]; sun straight up."""
    q = lambda a, b, c, d: [a, b, c, d]
    pos = np.array(q([-6, 0, -6], [6, 0, -6], [6, 0, 6], [-6, 0, 6]) + q([-6, 2, -6], [0, 2, -6], [0, 2, 6], [-6, 2, 6]),
                   np.float32)
    nrm = np.tile(np.array([[0.0, 1, 0]], np.float32), (8, 1))
    tri = np.array([[0, 1, 2, 0], [0, 2, 3, 0], [4, 5, 6, 0], [4, 6, 7, 0]], np.uint32)
    scene = rtgi.Scene(pos=pos, nrm=nrm, tri=tri, materials=np.array([[0.5, 0.6, 0.7, 1]], np.float32))
    vox = rtgi.VoxParams(v0=0.125, levels=3, G=8, maxBricks=3 * 512, maxVoxels=1 << 17)
    rc = rtgi.RcParams(cascades=2, probes0=16, s0=0.5)
    return scene, vox, rc, rtgi.Frame(eye=(0.0, 1.0, 0.0), sunDir=(0.0, 1.0, 0.0))


def test_traced_sun_visibility_closed_form():
    """voxSun with voxel shadow rays: floor voxels under the roof see no sun, the others all of it, so a ray that
    hits the floor under the roof returns 0 and one outside returns exactly the plain shading (visibility 1)."""
    scene, vox, rc, frame = roof_world()
    plain = Oracle(vox, rc, frame)
    plain.build(scene)
    traced = Oracle(vox, dataclasses.replace(rc, sunTrace=True), frame)
    traced.build(scene)
    rng = np.random.default_rng(4)
    m = 200
    xs = np.concatenate([rng.uniform(-3.5, -1.5, m), rng.uniform(1.5, 3.5, m)])
    org = np.stack([xs, np.full(2 * m, 0.3), rng.uniform(-3, 3, 2 * m)], 1).astype(np.float32)
    d = np.tile(np.array([[0.0, -1.0, 0.0]], np.float32), (2 * m, 1))
    rp, hp = plain.trace_rays(0, org, d)
    rt, ht = traced.trace_rays(0, org, d)
    assert np.array_equal(hp, ht) and np.all(hp != 0xFFFFFFFF)
    assert np.all(rp[:, :3] > 0)
    assert np.all(rt[:m, :3] == 0)
    assert np.array_equal(rt[m:], rp[m:])
    vis = traced.sunvis
    assert set(np.unique(vis[vis != 0]).tolist()) <= {0.25, 0.5, 0.75, 1.0}


def _volume(org, rgb_of_x, reach=1.8):
    """A bounce volume (probes x fastest, 6 axes, 4 halves) whose every probe is active with the given reach on all
    axes and, on all axes, the rgb rgb_of_x(world x of the probe)."""
    dx, dy, dz = orc.BN_DIMS
    px = (org[0] + np.arange(dx)) * np.float32(1.2)
    vol = np.zeros((dz, dy, dx, 6, 4), np.float16)
    vol[..., :3] = np.asarray([rgb_of_x(x) for x in px], np.float16)[None, None, :, None, :]
    vol[..., 3] = np.float16(reach)
    return vol.reshape(dx * dy * dz, 6, 4).view(np.uint16)


def test_bounce_inject_closed_form():
    """bnInject + voxBounce on a hand-made volume: probes at x <= 0 hold 1 on every axis, those beyond the wall at
    x = 0.35 hold 100, all active with reach 1.8 (> the 1.2 m spacing, so the reach test keeps every probe). bnRead
    weights a probe by smoothstep(-0.05, 0.25, (probe - p) . n / 1.2): the wall voxels' front (+x) side reads only
    the far probes and the back side only the near ones, and an axis-aligned normal picks that axis alone, so every
    voxel stores 100 in front and 1 behind (f16). A cascade ray from +x hits the front face, which the sun (x < 0) does
    not light: exactly albedo / pi * 100; a ray from -x returns the plain shading plus albedo / pi * 1."""
    scene, vox, rc, frame = wall_world()
    plain = Oracle(vox, rc, frame)
    plain.build(scene)
    o = Oracle(vox, dataclasses.replace(rc, bounce=True), frame)
    o.build(scene)
    rng = np.random.default_rng(6)
    m = 200
    yz = rng.uniform(-1.5, 1.5, (m, 2))
    dx = np.tile(np.array([[1.0, 0, 0]], np.float32), (m, 1))
    alb = np.float32(128 / 255) / np.float32(3.14159265)
    # Cascade 0 rays hit the wall at level 0 (read at the voxel centre), cascade 1 rays at level 1 (half a voxel out).
    cases = ((0, 0.6, -0.1), (1, 0.9, -0.4))
    for k, (c, xf, xb) in enumerate(cases):
        front = np.concatenate([np.full((m, 1), xf), yz], 1).astype(np.float32)
        back = np.concatenate([np.full((m, 1), xb), yz], 1).astype(np.float32)
        pb, ph = plain.trace_rays(c, back, dx)
        assert np.all(ph != 0xFFFFFFFF) and np.all(ph >> 29 == c) and np.all(pb[:, :3] > 0)
        if k == 0:
            # Before injection the bounce term is zero: identical to the plain shading.
            assert np.array_equal(o.trace_rays(c, back, dx)[0], pb)
            org = orc.bn_origin(frame.eye)
            o.inject_volume(_volume(org, lambda x: (1.0,) * 3 if x <= 0 else (100.0,) * 3), org)
        rf, hf = o.trace_rays(c, front, -dx)
        assert np.all(hf != 0xFFFFFFFF) and np.all(hf >> 29 == c)
        np.testing.assert_array_equal(rf[:, :3], np.full((m, 3), alb * np.float32(100.0), np.float32))
        rb, hb = o.trace_rays(c, back, dx)
        assert np.array_equal(hb, ph)
        np.testing.assert_allclose(rb[:, :3], pb[:, :3] + alb, rtol=1e-6)
    h = np.frombuffer(o.bounce[:o.counters()["voxels"]].tobytes(), np.float16).reshape(-1, 8)
    assert np.all(h[:, [0, 1, 2]] == 100) and np.all(h[:, [4, 5, 6]] == 1) and np.all(h[:, [3, 7]] == 0)


def floor_world(y=0.1, extent=60.0):
    pos = np.array([[-extent, y, -extent], [extent, y, -extent], [extent, y, extent], [-extent, y, extent]], np.float32)
    nrm = np.tile(np.array([[0.0, 1, 0]], np.float32), (4, 1))
    tri = np.array([[0, 2, 1, 0], [0, 3, 2, 0]], np.uint32)
    scene = rtgi.Scene(pos=pos, nrm=nrm, tri=tri, materials=np.array([[0.6, 0.5, 0.4, 1]], np.float32))
    vox = rtgi.VoxParams(v0=0.25, levels=4, G=8, maxBricks=4 * 512, maxVoxels=1 << 18)
    rc = rtgi.RcParams(cascades=2, probes0=8, s0=0.5, bounce=True)
    return scene, vox, rc, rtgi.Frame(eye=(0.3, 1.0, -0.2), sunDir=(0.2, 1.0, 0.1))


def test_bounce_volume_floor_closed_form():
    """bnClassify + bnUpdate over a floor at y = 0.1 (60 m around): the probes 1.2 m up are the only active ones (a
    level-2 voxel of the floor lies half a cell below them; those at y = 0 sit in a floor voxel, the others have no
    surface near). Every downward ray hits the floor, whose hits all shade alike (same albedo and normal at every
    level, sun visibility 1, no bounce on the first update), and every upward ray sees the sky, so the ambient cube
    holds pi * floor radiance along -y and pi * sky along +y; the -y reach is the distance down to the first floor
    voxel (0.95 m to the level-0 voxel top at 0.25 near the eye), the others the 1.8 m cap. The second update keeps
    the probes whose turn it is not (every 2nd update by x + y + z) from the previous volume, texel for texel, and
    traces the rest, whose floor hits now add albedo / pi times the first volume's irradiance above the floor."""
    scene, vox, rc, frame = floor_world()
    o = Oracle(vox, rc, frame)
    o.build(scene)
    vol, org, n = o.bounce_update(None)
    assert n == 1 and list(org) == [-16, -8, -17]
    dx, dy, dz = orc.BN_DIMS
    V = vol.view(np.float16).reshape(dz, dy, dx, 6, 4).astype(np.float32)
    iy = 1 - org[1]  # the probe layer at y = 1.2
    assert np.all(np.delete(V, [iy - 1, iy], axis=1) == 0)
    # The layer at y = 0 lies in the floor's level-0 voxels inside the level-0 window (inactive); outside it the
    # level-0 test reads nothing (occAt is false out of the window), so those probes count as free and active.
    lo0 = np.array([o.vp.origin[0][a] for a in range(3)]) * 8 * vox.v0
    xw0 = (org[0] + np.arange(dx)) * np.float32(1.2)
    zw0 = (org[2] + np.arange(dz)) * np.float32(1.2)
    in0 = ((zw0[:, None] >= lo0[2]) & (zw0[:, None] < lo0[2] + 16) & (xw0[None, :] >= lo0[0]) & (xw0[None, :] < lo0[0] + 16))
    assert np.array_equal(V[:, iy - 1, :, 0, 3] > 0, ~in0) and in0.sum() > 100
    A = V[:, iy]                        # (z, x, axis, rgba)
    assert np.all(A[..., 0, 3] > 0)     # every probe of the layer is active
    alb = np.array([np.float32(np.round(255 * c) / 255) for c in (0.6, 0.5, 0.4)], np.float32)
    sd = np.array(frame.sunDir, np.float64)
    sd = (sd / np.linalg.norm(sd)).astype(np.float32)
    floor_l = alb / np.float32(3.14159265) * np.asarray(frame.sunE, np.float32) * sd[1]
    np.testing.assert_allclose(A[..., 3, :3], np.broadcast_to(np.float32(3.14159265) * floor_l, A[..., 3, :3].shape),
                               rtol=1e-3)
    np.testing.assert_allclose(A[..., 2, :3], np.broadcast_to(np.float32(3.14159265) * np.asarray(frame.sky, np.float32),
                                                               A[..., 2, :3].shape), rtol=1e-3)
    cap = np.float32(np.float16(np.float32(1.5) * np.float32(1.2)))
    assert np.all(A[..., [0, 1, 2, 4, 5], 3] == cap)
    xw = (org[0] + np.arange(dx)) * 1.2
    zw = (org[2] + np.arange(dz)) * 1.2
    near = (np.abs(zw[:, None] - 0.3) < 5) & (np.abs(xw[None, :] - 0.3) < 5)
    np.testing.assert_allclose(A[near][:, 3, 3], 0.95, rtol=1e-3)
    # Second update, same eye: kept probes copy the previous volume; traced ones gain the bounce.
    vol2, org2, n2 = o.bounce_update((vol, org, n))
    assert n2 == 2 and np.array_equal(org2, org)
    V2 = vol2.view(np.float16).reshape(dz, dy, dx, 6, 4).astype(np.float32)
    A2 = V2[:, iy]
    wsum = (org[0] + np.arange(dx))[None, :] + (org[1] + iy) + (org[2] + np.arange(dz))[:, None]
    kept = (wsum + 1048576 + 2) % 2 != 0
    assert kept.any() and (~kept).any()
    assert np.array_equal(A2[kept], A[kept])
    assert np.all(A2[~kept][:, 3, :3] > A[~kept][:, 3, :3])   # floor hits add the bounce
    assert np.array_equal(A2[~kept][:, :, 3], A[~kept][:, :, 3])  # reach does not depend on light
    assert np.array_equal(A2[~kept][:, 2, :3], A[~kept][:, 2, :3])  # the sky side is unchanged


def _quad(c, u, v):
    c, u, v = (np.asarray(x, np.float64) for x in (c, u, v))
    return [c - u - v, c + u - v, c + u + v, c - u + v]


def dyn_world(xw=0.3125, zw=0.0, xs=-0.6):
    """A static wall (6 m square in the plane x = xs, facing +x) and, as the scene's dynamic part, a 1 m square in the
    plane x = xw (0.3125: the centre plane of a column of v0 = 0.125 voxels) facing +x, albedo (0.8, 0.4, 0.2)."""
    pos = np.array(_quad([xs, 0, 0], [0, 3, 0], [0, 0, 3]) + _quad([xw, 0, zw], [0, 0.5, 0], [0, 0, 0.5]), np.float32)
    nrm = np.tile(np.array([[1.0, 0, 0]], np.float32), (8, 1))
    tri = np.array([[0, 1, 2, 0], [0, 2, 3, 0], [4, 5, 6, 1], [4, 6, 7, 1]], np.uint32)
    mat = np.array([[0.5, 0.5, 0.5, 1], [0.8, 0.4, 0.2, 1]], np.float32)
    scene = rtgi.Scene(pos=pos, nrm=nrm, tri=tri, materials=mat, firstDynamicTri=2)
    vox = rtgi.VoxParams(v0=0.125, levels=3, G=8, maxBricks=3 * 512, maxVoxels=1 << 16)
    rc = rtgi.RcParams(cascades=2, probes0=16, s0=0.5)
    return scene, vox, rc, rtgi.Frame(eye=(0.0, 0.0, 0.0), sunDir=(1.0, 0.0, 0.0))


def _level_sums(d, L, per=512):
    """Sums over the allocated voxels of level L of each of the 13 words (level 0: fixed point, 1/65536)."""
    slots = np.nonzero(d.meta[:d.hdr[8]] // per == L)[0]
    w = d.data[slots]
    if L == 0:
        return w.astype(np.float64).sum(axis=(0, 1)) / 65536.0
    return w.view(np.float32).astype(np.float64).sum(axis=(0, 1))


def test_dynamic_layer_closed_forms():
    """The dynamic coverage layer (dynlayer.ts) for one 1 m square facing +x, lit head-on by the sun.
    (1) Area: the splat spreads each sample's area over 8 voxels with weights summing to 1, so the level-0 coverage
        words sum to A / v0^2 = 64 (up to fixed-point rounding), all of it on the +x side word; each coarser level
        holds a quarter of the level below (own 0 + children / 4), i.e. A / v_L^2.
    (2) Bricks: the square with a v0 margin covers 1 x 2 x 2 bricks at every level (brick edges at multiples of
        1, 2, 4 m), allocated coarse level first; its bounds are floor/ceil of (min - 0.5, max + 0.5) * 100.
    (3) Radiance: every voxel holds albedo / pi * sunE (n . l = 1).
    (4) A ray hitting the square's front sees colour = coverage * that radiance; from behind it sees nothing.
    (5) Composition with a static hit behind: colour = c R + (1 - c) S with one c for all three channels.
    (6) A square behind the static wall changes nothing; the static bricks are those of the static triangles alone."""
    scene, vox, rc, frame = dyn_world()
    o = Oracle(vox, rc, frame)
    d = DynState(vox, 64, scene.firstDynamicTri)
    o.build(scene, dyn=d)
    h = d.hdr
    # fixed-point rounding: at most 1/2^17 per addition, 2 x 12^2 samples x 8 voxels -> 0.0176 on the level-0 sum
    sums = [_level_sums(d, L) for L in range(3)]
    np.testing.assert_allclose([sums[0][6], sums[0][0]], [64.0, 64.0], atol=0.0176)
    for L in range(3):
        assert np.all(sums[L][1:6] == 0) and np.all(sums[L][10:] == 0)
        if L > 0:
            np.testing.assert_allclose(sums[L][[0, 6]], sums[L - 1][[0, 6]] / 4, rtol=1e-6)
    assert h[0] == 12 and h[8] == 12 and h[1] == 0 and h[9] == 0 and h[10] == 1
    assert list(h[16:19]) == [8, 4, 0] and list(h[24:27]) == [4, 4, 4]
    B = 1 << 30
    assert list(h[2:8]) == [B - 19, B - 100, B - 100, B + 82, B + 100, B + 100]
    per = vox.G ** 3
    alloc = d.meta[:h[8]]
    assert np.all(d.ptr.reshape(-1)[alloc] == np.arange(h[8]))
    assert np.sum(d.ptr != 0xFFFFFFFF) == 12
    lev0 = np.nonzero(alloc // per == 0)[0]
    w = d.data[lev0].astype(np.float64)
    m = w[..., 6] > 0.01 * 65536
    R = np.array([0.8, 0.4, 0.2]) / np.pi * np.array(frame.sunE)
    np.testing.assert_allclose(w[..., 7:10][m] / w[..., 6][m, None], np.tile(R, (m.sum(), 1)), rtol=2e-3)

    rng = np.random.default_rng(5)
    n = 64
    yz = rng.uniform(-0.35, 0.35, (n, 2))
    mx = np.tile(np.array([[-1.0, 0, 0]], np.float32), (n, 1))
    front = np.stack([np.full(n, 0.6125), yz[:, 0], yz[:, 1]], 1).astype(np.float32)
    r, hid = o.trace_rays(0, front, mx)
    cov = 1.0 - r[:, 3]
    assert np.all(cov > 0.5) and np.all(hid == 0xFFFFFFFF)
    np.testing.assert_allclose(r[:, :3] / cov[:, None], np.tile(R, (n, 1)), rtol=1e-3)
    back = front.copy()
    back[:, 0] = 0.0125
    rb, _ = o.trace_rays(0, back, -mx)
    assert np.all(rb == np.array([0, 0, 0, 1], np.float32))

    static_only = dataclasses.replace(scene, tri=scene.tri[:2], firstDynamicTri=None)
    plain = Oracle(vox, rc, frame)
    plain.build(static_only)
    assert np.array_equal(plain.buf, o.buf)
    far = front.copy()
    far[:, 0] = 1.0
    rs, hs = plain.trace_rays(1, far, mx)
    rd, hd = o.trace_rays(1, far, mx)
    assert np.array_equal(hs, hd) and np.all(hs != 0xFFFFFFFF)
    c = (rd[:, 0] - rs[:, 0]) / (R[0] - rs[:, 0])
    assert np.all((c > 0.5) & (c <= 1.0 + 1e-5))
    np.testing.assert_allclose(rd[:, :3], c[:, None] * R + (1 - c[:, None]) * rs[:, :3], atol=2e-4)
    hidden_scene = dyn_world(xw=-1.1875)[0]
    hidden = Oracle(vox, rc, frame)
    hidden.build(hidden_scene, dyn=DynState(vox, 64, 2))
    rh, hh = hidden.trace_rays(1, far, mx)
    assert np.array_equal(rh, rs) and np.array_equal(hh, hs)


def test_objects_inject_closed_form():
    """Token obj (dynInject with the bounce volume's dynE): on a hand-made previous volume whose every probe is active
    (reach 1.8) with the same rgb E0 on every axis, the dynamic square (facing +x) reads its +x face only, and only the
    probes in front of it (smoothstep of their offset along +x) count: E = E0. Every voxel of the layer, at every
    level, gains area x albedo / pi x E0 on its radiance words, so its mean radiance rises by albedo / pi x E0; the
    albedo words hold area x albedo. Without dynamic triangles the token changes nothing: the same bounce volume bit
    for bit and an empty layer."""
    scene, vox, rc, frame = dyn_world()
    rcb = dataclasses.replace(rc, bounce=True, objects=True)
    E0 = np.array([10.0, 20.0, 30.0], np.float32)
    alb = np.array([0.8, 0.4, 0.2], np.float32)
    o = Oracle(vox, rcb, frame)
    d = DynState(vox, 64, scene.firstDynamicTri)
    o.build(scene, dyn=d)
    before = d.data.copy()
    n, per = int(d.hdr[8]), vox.G ** 3
    org = orc.bn_origin(frame.eye)
    o.bounce_update((_volume(org, lambda x: tuple(E0.tolist())), org, 1))
    for s in range(n):
        L = int(d.meta[s]) // per
        b, a = before[s], d.data[s]
        if L == 0:
            w6 = b[:, 6].astype(np.float64)
            m = w6 > 0.1 * 65536
            albw = b[:, 10:13][m].astype(np.float64) / w6[m][:, None]
            tol = 0.5 / w6[m][:, None]  # the injected add is rounded to 1/65536
            gain = (a[m][:, 7:10].astype(np.float64) - b[m][:, 7:10]) / w6[m][:, None]
        else:
            fb, fa = b.view(np.float32).astype(np.float64), a.view(np.float32).astype(np.float64)
            w6 = fb[:, 6]
            m = w6 > 0.1
            albw = fb[m][:, 10:13] / w6[m][:, None]
            tol = 0.0
            gain = (fa[m][:, 7:10] - fb[m][:, 7:10]) / w6[m][:, None]
        assert m.any()
        # the albedo words: area x albedo (fixed-point rounding of the level-0 sums shows in the 4th digit)
        np.testing.assert_allclose(albw, np.tile(alb, (m.sum(), 1)), rtol=1e-3, atol=0)
        # the gain: albedo-area / pi x E0, per voxel
        want = albw / np.float32(3.14159265) * E0
        assert np.all(np.abs(gain - want) <= tol + 1e-5 * want), L
    # No dynamic triangles: identical to the plain bounce volume.
    plain = Oracle(vox, dataclasses.replace(rc, bounce=True), frame)
    plain.build(scene)
    obj = Oracle(vox, rcb, frame)
    d0 = DynState(vox, 64, scene.ntris)
    obj.build(scene, dyn=d0)
    assert d0.hdr[8] == 0
    v1 = plain.bounce_update(None)
    v2 = obj.bounce_update(None)
    assert np.array_equal(v1[0], v2[0])


def test_log2_series_against_float64():
    """The level blend's log2 (a fixed sequence of float operations, so that the device gives the same bits) against
    numpy's float64 log2: within 2^-22 absolute plus one float step of the result over [1, 2^10] (the blend's range:
    cone width over v0), exact at powers of two, and increasing."""
    lib = orc.lib()
    x = np.concatenate([np.geomspace(1.0, 1024.0, 20001), 2.0 ** np.arange(11), np.sqrt(2.0) ** np.arange(21)])
    x = x.astype(np.float32)
    got = np.array([lib.orc_log2(float(v)) for v in x], np.float64)
    np.testing.assert_allclose(got, np.log2(x.astype(np.float64)), rtol=2.0 ** -23, atol=2.0 ** -22)
    assert np.array_equal(got[20001:20012], np.arange(11.0))
    s = np.sort(x)
    g = np.array([lib.orc_log2(float(v)) for v in s])
    assert np.all(np.diff(g) >= 0)


def test_dynamic_layer_moves_and_overflows():
    """Across updates the layer keeps only this update's bricks: moving the square 2 m along z frees the old bricks
    (their pointers return to empty) and allocates the new ones, with the same area. With 6 slots for the
    12 bricks, allocation is coarse level first: level 2 gets its 4, level 1 2 of 4, level 0 none (marked dropped).
    The splats land on level 0 only, so they are lost with its bricks; the coarser levels, which only gather
    children, stay empty, and a ray reads nothing (dynlayer.ts: splats into dropped bricks are skipped)."""
    scene, vox, rc, frame = dyn_world()
    o = Oracle(vox, rc, frame)
    d = DynState(vox, 64, 2)
    o.build(scene, dyn=d)
    before = set(np.nonzero(d.ptr.reshape(-1) != 0xFFFFFFFF)[0].tolist())
    o.build(dyn_world(zw=2.0)[0], dyn=d)
    after = set(np.nonzero(d.ptr.reshape(-1) != 0xFFFFFFFF)[0].tolist())
    # at z in [1.375, 2.625]: 2 level-0 bricks along z, 2 of level 1, 1 of level 2 (edge at z = 0 and 4)
    assert len(after) == 10 and d.hdr[0] == 10 and list(d.hdr[24:27]) == [4, 4, 2] and d.hdr[10] == 2
    assert before != after and set(d.meta[:10].tolist()) == after
    np.testing.assert_allclose(_level_sums(d, 0)[6], 64.0, atol=0.0176)

    small = Oracle(vox, rc, frame)
    ds = DynState(vox, 6, 2)
    small.build(scene, dyn=ds)
    h = ds.hdr
    assert h[0] == 12 and h[8] == 6 and h[1] == 6
    assert list(h[16:19]) == [6, 4, 0] and list(h[24:27]) == [0, 2, 4]
    per = vox.G ** 3
    ptr = ds.ptr.reshape(-1)
    assert np.sum(ptr[:per] == 0xFFFFFFFD) == 4 and np.sum(ptr[per:2 * per] == 0xFFFFFFFD) == 2
    assert np.sum(ptr[2 * per:] < 6) == 4
    n = 32
    yz = np.random.default_rng(6).uniform(-0.35, 0.35, (n, 2))
    front = np.stack([np.full(n, 0.6125), yz[:, 0], yz[:, 1]], 1).astype(np.float32)
    r, _ = small.trace_rays(0, front, np.tile(np.array([[-1.0, 0, 0]], np.float32), (n, 1)))
    assert np.all(r == np.array([0, 0, 0, 1], np.float32))
    assert np.all(ds.data[:6] == 0) and np.all(ds.mask[:6] == 0)


def test_trace_empty_world_and_single_voxel():
    vox = rtgi.VoxParams(v0=0.25, levels=2, G=2, maxBricks=16, maxVoxels=1024)
    rc = rtgi.RcParams(cascades=2, probes0=2, s0=0.5)
    frame = rtgi.Frame(eye=(0.0, 0.0, 0.0))
    empty = rtgi.Scene(pos=np.zeros((3, 3), np.float32), nrm=np.zeros((3, 3), np.float32),
                       tri=np.zeros((0, 4), np.uint32), materials=np.ones((1, 4), np.float32))
    o = Oracle(vox, rc, frame)
    o.build(empty)
    rgba, hits = o.trace()
    assert np.all(hits == 0xFFFFFFFF) and np.all(rgba[:, 3] == 1) and np.all(rgba[:, :3] == 0)
    m, sh = o.merge_sh(rgba)
    assert np.allclose(m[:, :3], np.float32(frame.sky))
    # One tiny triangle inside the level-0 cell (x, y, z) = (9, 8, 8) (centre (0.375, 0.125, 0.125) m): every hit
    # is that voxel, and the ray from the probe at the origin along +x (within its direction bin) hits it.
    p = np.array([[0.30, 0.10, 0.10], [0.45, 0.10, 0.12], [0.35, 0.20, 0.15]], np.float32)
    one = rtgi.Scene(pos=p, nrm=np.tile([[0, 0, 1]], (3, 1)).astype(np.float32),
                     tri=np.array([[0, 1, 2, 0]], np.uint32), materials=np.array([[0.5, 0.5, 0.5, 1]], np.float32))
    o.build(one)
    rgba, hits = o.trace()
    E = vox.G * 8
    got = set(int(h) for h in hits if h != 0xFFFFFFFF)
    org = orc.origins(vox.v0, vox.levels, vox.G, frame.eye)
    cell = (np.array([0.375, 0.125, 0.125]) / 0.25 - np.array(org[0]) * 8).astype(int)
    expect0 = int(cell[0] + E * (cell[1] + E * cell[2]))
    assert expect0 in got
    assert all((h >> 29) in (0, 1) for h in got)
    assert all((h & ((1 << 29) - 1)) == expect0 for h in got if (h >> 29) == 0)


def trilinear_merge_reference(rgba_per_c, rc, sky):
    """Merge by trilinear interpolation of the upper cascade at the probe (dense grid: equal-weight parents = trilinear
    weights, rc.ts:700-705), with the upper 2x2 direction average."""
    from scipy.ndimage import map_coordinates
    N = rc.cascades
    out = [None] * N
    for c in range(N - 1, -1, -1):
        n, res = rc.probes0 >> c, 4 << min(c, rc.resCap)
        v = rgba_per_c[c].astype(np.float64).reshape(n, n, n, res, res, 4)  # z y x | v u
        if c == N - 1:
            up = np.broadcast_to(np.asarray(sky, np.float64), v[..., :3].shape)
        else:
            un, ur = n // 2, 4 << min(c + 1, rc.resCap)
            U = out[c + 1].reshape(un, un, un, ur, ur, 4)[..., :3]
            if ur != res:
                U = U.reshape(un, un, un, res, 2, res, 2, 3).mean(axis=(4, 6))
            g = np.arange(n) / 2.0
            Zc, Yc, Xc = np.meshgrid(g, g, g, indexing="ij")
            coords = np.stack([Zc.ravel(), Yc.ravel(), Xc.ravel()])
            up = np.empty(v[..., :3].shape)
            for dv in range(res):
                for du in range(res):
                    for ch in range(3):
                        up[:, :, :, dv, du, ch] = map_coordinates(U[:, :, :, dv, du, ch], coords, order=1,
                                                                  mode="nearest").reshape(n, n, n)
        m = v.copy()
        m[..., :3] = v[..., :3] + v[..., 3:4] * up
        m[..., 3] = 1
        out[c] = m.reshape(-1, 4)
    return out


def test_merge_is_trilinear_on_the_dense_grid():
    rng = np.random.default_rng(4)
    vox = rtgi.VoxParams(v0=0.25, levels=2, G=2, maxBricks=16, maxVoxels=1024)
    for rc in (rtgi.RcParams(cascades=3, probes0=8, s0=0.5), rtgi.RcParams(cascades=3, probes0=8, s0=0.5, resCap=1)):
        frame = rtgi.Frame(eye=(0.0, 0.0, 0.0), sky=(0.3, 0.5, 0.7))
        o = Oracle(vox, rc, frame)
        total = sum(o.cascade_rays())
        rgba = np.zeros((total, 4), np.float32)
        rgba[:, :3] = rng.uniform(0, 2, (total, 3))
        rgba[:, 3] = rng.integers(0, 2, total)
        m, _ = o.merge_sh(rgba)
        ref = trilinear_merge_reference(o.split(rgba), rc, frame.sky)
        for c, (got, exp) in enumerate(zip(o.split(m), ref)):
            np.testing.assert_allclose(got, exp, rtol=1e-5, atol=1e-5, err_msg=f"cascade {c}")


def real_sh(n):
    """Real SH of degree <= 2 in the basis order of rc.ts:745-749 (0; y, z, x; xy, yz, 3z^2-1, xz, x^2-y^2), from
    scipy's complex harmonics (Condon-Shortley phase removed)."""
    from scipy.special import sph_harm_y
    x, y, z = n
    theta = np.arccos(np.clip(z, -1, 1))
    phi = np.arctan2(y, x)
    Y = lambda l, m: sph_harm_y(l, abs(m), theta, phi)  # noqa: E731
    def r(l, m):
        if m == 0:
            return Y(l, 0).real
        s = np.sqrt(2) * (-1) ** m
        return s * (Y(l, m).imag if m < 0 else Y(l, m).real)
    return np.array([r(0, 0), r(1, -1), r(1, 0), r(1, 1), r(2, -2), r(2, -1), r(2, 0), r(2, 1), r(2, 2)])


def test_sh_projection_basis_and_constant_field():
    vox = rtgi.VoxParams(v0=0.25, levels=2, G=2, maxBricks=16, maxVoxels=1024)
    rc = rtgi.RcParams(cascades=1, probes0=1, s0=0.5)
    frame = rtgi.Frame(eye=(0.0, 0.0, 0.0), sky=(0.0, 0.0, 0.0))
    o = Oracle(vox, rc, frame)
    om = orc.oct_solid_angles(4)
    a = np.array([np.pi] + [2 * np.pi / 3] * 3 + [np.pi / 4] * 5)
    for d in range(16):
        rgba = np.zeros((16, 4), np.float32)
        rgba[d, :3] = (1.0, 2.0, 0.5)
        _, sh = o.merge_sh(rgba)
        n = oct_decode((d % 4 + 0.5) / 4, (d // 4 + 0.5) / 4)
        exp = np.outer(om[d] * real_sh(n) * a, [1.0, 2.0, 0.5]).ravel()
        np.testing.assert_allclose(sh[0], exp, rtol=2e-5, atol=1e-6)
    rgba = np.zeros((16, 4), np.float32)
    rgba[:, :3] = 1.5
    _, sh = o.merge_sh(rgba)
    assert abs(sh[0, 0] - np.pi * 0.282095 * 4 * np.pi * 1.5) < 1e-4


def test_golden_world_regression():
    """The oracle still reproduces the committed fixture (tests/golden/make_golden.py)."""
    g = np.load(os.path.join(GOLDEN, "tiny_world.npz"))
    scene = rtgi.Scene(pos=g["pos"], nrm=g["nrm"], tri=g["tri"], materials=g["mat"])
    vox = rtgi.VoxParams(**json.loads(str(g["vox"])))
    rc = rtgi.RcParams(**json.loads(str(g["rc"])))
    frame = rtgi.Frame(**json.loads(str(g["frame"])))
    o = Oracle(vox, rc, frame)
    o.build(scene)
    for L in range(vox.levels):
        cm = o.clipmap(L)
        assert np.array_equal(cm["occ"], g[f"occ{L}"]) and np.array_equal(cm["rep"], g[f"rep{L}"])
    rgba, hits = o.trace()
    assert np.array_equal(hits, g["hits"])
    m, sh = o.merge_sh(rgba)
    np.testing.assert_array_equal(m, g["merged"])
    np.testing.assert_array_equal(sh, g["sh"])


def _hist_rgba(o, frame_no):
    """Per-ray radiance of a one-cascade dense grid that depends on the probe's world grid point, the direction and
    the frame: a wrong history row shows as a wrong value."""
    n = o.rc.probes0
    g0 = np.array(o.rc.G0[:3])
    pi = np.arange(n ** 3)
    g = np.stack([pi % n, (pi // n) % n, pi // (n * n)], 1) + g0
    d = np.arange(16)
    L = (1.0 + 0.05 * g[:, 0] - 0.03 * g[:, 1] + 0.02 * g[:, 2])[:, None] + 0.1 * d[None, :] + frame_no
    rgba = np.zeros((n ** 3 * 16, 4), np.float32)
    rgba[:, 0] = L.ravel()
    rgba[:, 1] = 0.5 * L.ravel()
    rgba[:, 2] = 2.0 - 0.01 * L.ravel()
    return rgba, g


def test_history_blend_closed_form():
    """histBlend (rc.ts:784-814): each probe's SH is prev * 0.5 + new * 0.5 of the previous update's stored SH at the
    same world grid point, the frame's own SH where the previous grid did not hold that point; the parity flips per
    update, so the third update reads the second's blended values. The expectation indexes the plain per-update SH by
    world grid point in numpy (not through the oracle's hash)."""
    vox = rtgi.VoxParams(v0=0.25, levels=2, G=2, maxBricks=16, maxVoxels=1024)
    rc = rtgi.RcParams(cascades=1, probes0=8, s0=0.5)
    eyes = [(0.1, 0.2, 0.3), (1.1, 0.2, -0.2), (1.1, 0.2, -0.2), (-0.9, 0.7, -0.2)]  # shifts (2, 0, -1), 0, (-4, 1, 0)
    for limit in (0, 64):
        hist = orc.History(rc.probes0, limit=limit)
        prev = None  # {world grid point: stored SH}
        shifted = kept = 0
        for f, eye in enumerate(eyes):
            o = Oracle(vox, rc, rtgi.Frame(eye=eye, sky=(0.0, 0.0, 0.0)))
            rgba, g = _hist_rgba(o, f)
            _, plain = o.merge_sh(rgba)
            _, got = o.merge_sh(rgba, hist=hist)
            exp = plain.copy()
            if prev is not None:
                for p in range(g.shape[0]):
                    h = prev.get(tuple(g[p]))
                    if h is not None:
                        exp[p] = h * np.float32(0.5) + plain[p] * np.float32(0.5)
                        kept += 1
                    else:
                        shifted += 1
            np.testing.assert_array_equal(got, exp, err_msg=f"update {f + 1}, limit {limit}")
            prev = {tuple(g[p]): got[p] for p in range(g.shape[0])}
        assert kept > 0 and shifted > 0 and hist.misses == 0
    # A constant field keeps its SH through the blend; the history of a moved grid is exactly the overlap's.
    hist = orc.History(rc.probes0)
    for eye in eyes:
        o = Oracle(vox, rc, rtgi.Frame(eye=eye, sky=(0.0, 0.0, 0.0)))
        rgba = np.zeros((8 ** 3 * 16, 4), np.float32)
        rgba[:, :3] = 1.5
        _, got = o.merge_sh(rgba, hist=hist)
        assert np.all(np.abs(got[:, 0] - np.pi * 0.282095 * 4 * np.pi * 1.5) < 1e-4)


def test_history_hash_limit_on_dense_grids():
    """The reference's 64-slot linear-probing search (rc.ts:791-800, hashSize = nextPow2(2 cap)) finds every probe of
    dense grids up to 32^3 but misses a few at 64^3 (long clusters at load 1/2); the whole-table search (the device's
    grid-point index) misses none. Both blend every probe they find identically."""
    vox = rtgi.VoxParams(v0=0.25, levels=2, G=2, maxBricks=16, maxVoxels=1024)
    for n, ref_misses in ((16, 0), (64, None)):
        rc = rtgi.RcParams(cascades=1, probes0=n, s0=0.5)
        h64, h0 = orc.History(n, limit=64), orc.History(n, limit=0)
        for f, eye in enumerate([(0.0, 0.0, 0.0), (0.0, 0.0, 0.0), (2.0, 0.5, -1.0)]):
            o = Oracle(vox, rc, rtgi.Frame(eye=eye, sky=(0.0, 0.0, 0.0)))
            rgba, _ = _hist_rgba(o, f)
            _, a = o.merge_sh(rgba, hist=h64)
            _, b = o.merge_sh(rgba, hist=h0)
            differ = np.count_nonzero(np.any(a != b, axis=1))
            assert differ <= h64.misses  # only the probes the limited search missed differ
        assert h0.misses == 0
        if ref_misses is not None:
            assert h64.misses == ref_misses
        else:
            assert 0 < h64.misses < 64  # a handful of 262,144 probes per update


def test_fp64_trace_taking_the_f32_voxel():
    """orc_trace_f64_take: rays whose walks agree keep the FP64 walk's result; the others are shaded in double at the
    f32 walk's voxel. The irradiance it gives then differs from the f32 path by the arithmetic alone (well below
    1e-4 relative), while the plain FP64 walk differs on tie rays by whole voxels."""
    scene, vox, rc, frame = rtgi.config("atrium")
    rc = dataclasses.replace(rc, probes0=8, cascades=2)
    frame = dataclasses.replace(frame, eye=(0.3, 4.0, -0.2))
    o = Oracle(vox, rc, frame)
    o.build(scene)
    rgba, hits = o.trace()
    rgbt, hit64, tie = o.trace_f64(eps=1e-6)
    take, taken = o.trace_f64_take(hits, eps=1e-6)
    keep = taken == 0
    assert np.array_equal(take[keep], rgbt[keep])
    assert np.all(taken[(tie != 0) | (hit64 != hits)] == 1)
    # where the two walks chose the same voxel the recomputed entry changes only the point-light terms' last bits
    same = (taken == 1) & (hit64 == hits)
    np.testing.assert_allclose(take[same], rgbt[same], rtol=1e-9, atol=1e-12)
    # where they differ, the taken ray follows the f32 walk: a miss there is (0, 0, 0, 1)
    miss = (taken == 1) & (hits == 0xFFFFFFFF)
    assert np.all(take[miss] == np.array([0, 0, 0, 1.0]))
    _, sh32 = o.merge_sh(rgba)
    _, sh64 = o.merge_sh_f64(take)
    rel = np.abs(sh32 - sh64).max(axis=1) / np.maximum(np.abs(sh64).max(axis=1), 1e-6)
    assert rel.max() < 1e-5, rel.max()
    assert (hit64 != hits).sum() > 0  # the case is not vacuous: some rays do pick another voxel in double


def tilted_triangle_world():
    """One large tilted triangle with three different vertex normals over a one-level 32^3 window."""
    a = np.array([-9.0, -3.0, -8.0])
    b = np.array([10.0, -1.0, -6.0])
    c = np.array([-2.0, 5.0, 9.0])
    na, nb, nc = (np.array(v) / np.linalg.norm(v) for v in ([0.2, 0.9, -0.3], [-0.4, 0.8, 0.1], [0.1, 0.7, 0.6]))
    scene = rtgi.Scene(pos=np.array([a, b, c], np.float32), nrm=np.array([na, nb, nc], np.float32),
                       tri=np.array([[0, 1, 2, 0]], np.uint32), materials=np.array([[0.3, 0.6, 0.9, 1.0]], np.float32))
    vox = rtgi.VoxParams(v0=0.25, levels=1, G=4, maxBricks=64, maxVoxels=1 << 15)
    return scene, vox, rtgi.RcParams(cascades=1, probes0=4, s0=1.0), rtgi.Frame(eye=(0.0, 0.0, 0.0))


def test_resolve_attr_tilted_triangle_closed_form():
    """resolveAttr (voxel.ts:590-620) on one large tilted triangle with three different vertex normals, worked in
    double: for every level-0 voxel whose centre projects inside the triangle, the closest point is the centre's
    projection onto the plane, the stored normal is octEnc(normalize(sum of the barycentric-weighted vertex
    normals)) and the plane offset is that normal's dot with (projection - centre). The voxel's albedo is the
    material's, packed as pack4x8unorm."""
    scene, vox, rc, frame = tilted_triangle_world()
    a, b, c = scene.pos.astype(np.float64)
    na, nb, nc = scene.nrm.astype(np.float64)
    o = Oracle(vox, rc, frame)
    o.build(scene)
    cm = o.clipmap(0)
    E = vox.G * 8
    org = orc.origins(vox.v0, 1, vox.G, frame.eye)[0] * 8
    A, B, Cc = (np.asarray(x, np.float32).astype(np.float64) for x in (a, b, c))
    NA, NB, NC = (np.asarray(x, np.float32).astype(np.float64) for x in (na, nb, nc))
    nf = np.cross(B - A, Cc - A)
    nf /= np.linalg.norm(nf)
    T = np.stack([B - A, Cc - A], 1)
    idx = np.nonzero(np.unpackbits(cm["occ"].view(np.uint8), bitorder="little")[:E ** 3])[0]
    checked = 0
    for i in idx:
        cell = np.array([i % E, (i // E) % E, i // (E * E)])
        p = (org + cell + 0.5) * vox.v0
        q = p - np.dot(p - A, nf) * nf
        st = np.linalg.lstsq(T, q - A, rcond=None)[0]
        w = np.array([1 - st.sum(), st[0], st[1]])
        if w.min() < 0.02:
            continue  # the closest point is on an edge: not the plane projection
        n = w[0] * NA + w[1] * NB + w[2] * NC
        n /= np.linalg.norm(n)
        code = int(cm["normal"][i])
        sx, sy = [max(np.int16(np.uint16(v)) / 32767.0, -1.0) for v in (code & 0xFFFF, code >> 16)]
        got = oct_decode(sx * 0.5 + 0.5, sy * 0.5 + 0.5)
        assert np.degrees(np.arccos(min(1.0, float(np.dot(got, n))))) < 0.01, (i, got, n)
        assert abs(float(cm["plane"][i]) - np.dot(n, q - p)) < 2e-5, (i, cm["plane"][i], np.dot(n, q - p))
        assert int(cm["albedo"][i]) == orc.pack4x8unorm([0.3, 0.6, 0.9, 0.0])
        checked += 1
    assert checked > 500


# ---------------------------------------------------------------------------------------------------------------
# posedVertices (scenes.ts:80-102).

def _dyadic(rng, n, scale, bits):
    return rng.integers(-scale * 2 ** bits, scale * 2 ** bits + 1, size=n) / 2.0 ** bits


def test_pose_exact_on_dyadic_inputs():
    """Dyadic rest positions, pivots, cos/sin stand-ins and offsets with few bits make every double operation of
    posedVertices exact, so each output is the float32 rounding of the exact rational value of
    (pivot + c x + s z + off, pivot + y + off, pivot - s x + c z + off) and (c nx + s nz, ny, -s nx + c nz): computed
    here with fractions, independent of operation order. A wrong sign, a swapped c and s or a wrong axis fails."""
    from fractions import Fraction as Fr
    rng = np.random.default_rng(17)
    n_obj, per = 5, 7
    first = np.arange(n_obj, dtype=np.uint32) * per + 3  # objects start at vertex 3
    count = np.full(n_obj, per, np.uint32)
    pivot = _dyadic(rng, n_obj * 3, 40, 2).reshape(-1, 3)
    poses = np.stack([_dyadic(rng, n_obj, 1, 5), _dyadic(rng, n_obj, 1, 5)] +
                     [_dyadic(rng, n_obj, 3, 3) for _ in range(3)], axis=1)
    nv = 3 + n_obj * per + 2
    rest_p = _dyadic(rng, n_obj * per * 3, 50, 6).reshape(-1, 3).astype(np.float32)
    rest_n = _dyadic(rng, n_obj * per * 3, 1, 6).reshape(-1, 3).astype(np.float32)
    base_p = rng.normal(size=(nv, 3)).astype(np.float32)
    base_n = rng.normal(size=(nv, 3)).astype(np.float32)
    pos, nrm = orc.pose(first, count, pivot, poses, rest_p, rest_n, base_p, base_n)
    assert np.array_equal(pos[:3], base_p[:3]) and np.array_equal(pos[-2:], base_p[-2:])  # outside the objects
    assert np.array_equal(nrm[:3], base_n[:3]) and np.array_equal(nrm[-2:], base_n[-2:])
    for i in range(n_obj):
        c, s = Fr(poses[i, 0]), Fr(poses[i, 1])
        off = [Fr(v) for v in poses[i, 2:]]
        pv = [Fr(v) for v in pivot[i]]
        for v in range(int(first[i]), int(first[i] + count[i])):
            r = [Fr(float(t)) for t in rest_p[v - 3]]
            q = [Fr(float(t)) for t in rest_n[v - 3]]
            x, y, z = r[0] - pv[0], r[1] - pv[1], r[2] - pv[2]
            want_p = [pv[0] + c * x + s * z + off[0], pv[1] + y + off[1], pv[2] - s * x + c * z + off[2]]
            want_n = [c * q[0] + s * q[2], q[1], -s * q[0] + c * q[2]]
            for a in range(3):
                assert pos[v, a] == np.float32(float(want_p[a])), (i, v, a)
                assert nrm[v, a] == np.float32(float(want_n[a])), (i, v, a)


def test_pose_quarter_turns_identity_and_rigidity():
    """Closed forms: c = 1, s = 0 with zero offset and a dyadic pivot gives the rest mesh back; c = 0, s = 1 maps
    (x, y, z) - pivot to (z, y, -x); any true rotation keeps every vertex's distance to the pivot (plus the offset)
    and unit normals unit, to float rounding."""
    rng = np.random.default_rng(3)
    n = 40
    rest_p = (rng.normal(size=(n, 3)) * 8).astype(np.float32)
    rest_n = rng.normal(size=(n, 3))
    rest_n = (rest_n / np.linalg.norm(rest_n, axis=1, keepdims=True)).astype(np.float32)
    first, count, pivot = [0], [n], np.array([[1.5, -2.25, 4.0]])
    z = np.zeros((n, 3), np.float32)
    p, q = orc.pose(first, count, pivot, [[1.0, 0.0, 0.0, 0.0, 0.0]], rest_p, rest_n, z, z)
    assert np.array_equal(p, rest_p) and np.array_equal(q, rest_n)
    p, q = orc.pose(first, count, pivot, [[0.0, 1.0, 0.0, 0.0, 0.0]], rest_p, rest_n, z, z)
    rel = rest_p.astype(np.float64) - pivot[0]
    want = np.stack([pivot[0, 0] + rel[:, 2], pivot[0, 1] + rel[:, 1], pivot[0, 2] - rel[:, 0]], axis=1)
    np.testing.assert_allclose(p, want, rtol=0, atol=1e-5)
    np.testing.assert_array_equal(q, np.stack([rest_n[:, 2], rest_n[:, 1], -rest_n[:, 0]], axis=1))
    for ang in rng.uniform(-np.pi, np.pi, size=6):
        off = rng.normal(size=3)
        p, q = orc.pose(first, count, pivot, [[np.cos(ang), np.sin(ang), *off]], rest_p, rest_n, z, z)
        d0 = np.linalg.norm(rel, axis=1)
        d1 = np.linalg.norm(p.astype(np.float64) - pivot[0] - off, axis=1)
        np.testing.assert_allclose(d1, d0, rtol=1e-5, atol=1e-5)
        np.testing.assert_allclose(np.linalg.norm(q.astype(np.float64), axis=1), 1.0, atol=1e-6)
        assert np.array_equal(p[:, 1], (np.float64(pivot[0, 1]) + rel[:, 1] + off[1]).astype(np.float32))


@pytest.mark.parametrize("kind", ["atrium", "city"])
def test_generated_scene_is_its_rest_mesh_posed(kind):
    """A generated scene at frame f is its rest mesh posed by the oracle at the scene's own schedule for f; the rest
    mesh does not depend on f, and only the objects' vertices move."""
    s0 = rtgi.generate_scene(kind, seed=4, frame=0)
    s7 = rtgi.generate_scene(kind, seed=4, frame=7)
    assert len(s0.objects) == (16 if kind == "atrium" else 64)
    assert np.array_equal(s0.rest_pos, s7.rest_pos) and np.array_equal(s0.objects, s7.objects)
    o = s0.objects
    assert np.all(o["first_vertex"][1:] == o["first_vertex"][:-1] + o["vertex_count"][:-1])  # contiguous
    for s, f in ((s0, 0), (s7, 7)):
        p, n = orc.pose_scene(s, s.poses_at(f))
        assert np.array_equal(p, s.pos) and np.array_equal(n, s.nrm)
    moved = np.any(s0.pos != s7.pos, axis=1)
    lo, hi = int(o["first_vertex"][0]), int(o["first_vertex"][-1] + o["vertex_count"][-1])
    assert moved[lo:hi].mean() > 0.9 and not moved[:lo].any() and not moved[hi:].any()


def test_bf16_rounding_is_nearest_even():
    """The precision form's f32 -> bf16 (orc_f2bf): on random floats of every exponent, edge patterns and exact ties,
    the result is the nearer of the two bf16 values bracketing the input (found by truncation, independent of the
    rounding formula), the even one on a tie; infinities stay, NaN stays NaN; bf16 -> f32 widens exactly."""
    rng = np.random.default_rng(23)
    bits = rng.integers(0, 2 ** 32, 20000, dtype=np.uint64).astype(np.uint32)
    ties = (rng.integers(0, 2 ** 16, 2000, dtype=np.uint64).astype(np.uint32) << 16) | np.uint32(0x8000)
    edge = np.array([0, 0x80000000, 1, 0x7F7FFFFF, 0xFF7FFFFF, 0x7F800000, 0xFF800000, 0x00807FFF, 0x3F808000,
                     0x3F818000, 0x7F7F8000], np.uint32)
    for u in np.concatenate([bits, ties, edge]).tolist():
        f = np.array([u], np.uint32).view(np.float32)[0]
        h = orc.lib().orc_f2bf(float(f))
        back = np.float32(orc.lib().orc_bf2f(h))
        assert np.array_equal(np.array([back], np.float32).view(np.uint32), np.array([h << 16], np.uint32))
        if np.isnan(f):
            assert np.isnan(back)
            continue
        lo = np.array([u & 0xFFFF0000], np.uint32).view(np.float32)[0]  # truncation toward zero
        hi_bits = (u & 0xFFFF0000) + 0x10000
        if (u & 0xFFFF) == 0:
            assert h << 16 == u
            continue
        # the next bf16 value away from zero, as an unbounded binary float (past the largest finite it is 2^128,
        # which IEEE rounding treats as the overflow threshold)
        E = (u >> 23) & 0xFF
        ulp = 2.0 ** ((E if E else 1) - 127 - 7)
        hi_val = abs(float(lo)) + ulp
        dl, dh = abs(abs(float(f)) - abs(float(lo))), abs(hi_val - abs(float(f)))
        if (u & 0xFFFF) == 0x8000:  # exact tie in the low half: the even one of the two
            want = u & 0xFFFF0000 if ((u >> 16) & 1) == 0 else hi_bits & 0xFFFFFFFF
        else:
            want = u & 0xFFFF0000 if dl < dh else hi_bits & 0xFFFFFFFF
        assert h == want >> 16, hex(u)
