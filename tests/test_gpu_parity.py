"""GPU parity: librtgi (HIP, gfx950) against the CPU oracle on the same seeded inputs.

Bars: occupancy, representative triangle ids, packed attributes, brick states, counters and first-hit voxel ids
bit-exact; FP32 radiance (traced, merged) and SH within 1e-4 relative (atol 1e-6).
"""
import dataclasses

import numpy as np
import pytest

import rtgi
from conftest import random_scene
from oracle.oracle import History, Oracle

pytestmark = pytest.mark.gpu

RTOL, ATOL = 1e-4, 1e-6


def say(msg):
    print(msg, flush=True)


def run_world(scene, vox, rc, frame, hits=True):
    w = rtgi.World(vox, dataclasses.replace(rc, recordHits=hits))
    w.upload_scene(scene)
    w.set_frame(frame)
    w.encode()
    w.sync()
    return w


def check_clipmap(w, o, vox):
    cg, co = w.read_counters(), o.counters()
    for k in ("bricks", "marked", "voxels", "brickOverflow", "voxelOverflow", "bricksPerLevel", "bigPerLevel"):
        assert cg[k] == co[k], (k, cg[k], co[k])
    for L in range(vox.levels):
        g, c = w.read_clipmap(L), o.clipmap(L)
        for k in ("brick_state", "occ", "rep", "albedo", "normal"):
            assert np.array_equal(g[k], c[k]), f"level {L} {k}: {np.count_nonzero(g[k] != c[k])} differ"
        assert np.array_equal(g["plane"].view(np.uint32), c["plane"].view(np.uint32)), f"level {L} plane"


def check_cascades(w, o, rc, merged=True, hist=None):
    """Cascades, hit ids and SH against the oracle; hist: the oracle's c0 SH history of the world's earlier updates
    (advanced here), None for a world's first update."""
    rgba, hits = o.trace()
    per_rgba, per_hits = o.split(rgba), o.split(hits)
    if not merged:
        for c in range(rc.cascades):
            g, h = w.read_cascade(c, hits=True)
            assert np.array_equal(h, per_hits[c]), f"cascade {c}: {np.count_nonzero(h != per_hits[c])} hit ids differ"
            np.testing.assert_allclose(g, per_rgba[c], rtol=RTOL, atol=ATOL)
        return
    m, sh = o.merge_sh(rgba, hist=hist)
    per_m = o.split(m)
    for c in range(rc.cascades):
        g, h = w.read_cascade(c, hits=True)
        assert np.array_equal(h, per_hits[c]), f"cascade {c}: {np.count_nonzero(h != per_hits[c])} hit ids differ"
        np.testing.assert_allclose(g, per_m[c], rtol=RTOL, atol=ATOL)
    np.testing.assert_allclose(w.read_sh(), sh, rtol=RTOL, atol=ATOL)


def tiny_case(seed=1, ntris=300, **vover):
    rng = np.random.default_rng(seed)
    scene = random_scene(rng, ntris, -5.0, 5.0, 0.6)
    vox = rtgi.VoxParams(v0=0.25, levels=3, G=4, maxBricks=3 * 64, maxVoxels=1 << 16)
    vox = dataclasses.replace(vox, **vover)
    rc = rtgi.RcParams(cascades=3, probes0=8, s0=0.5)
    frame = rtgi.Frame(eye=(0.3, 0.2, -0.1), sunDir=(0.3, 1.0, -0.2))
    return scene, vox, rc, frame


def test_tiny_random_build_trace_merge(require_gpu):
    scene, vox, rc, frame = tiny_case()
    say("case: tiny random scene, 300 triangles, G=4, 3 levels, 3 cascades")
    w = run_world(scene, vox, rc, frame)
    o = Oracle(vox, rc, frame)
    o.build(scene)
    check_clipmap(w, o, vox)
    check_cascades(w, o, rc)
    w.destroy()


def test_unmerged_trace(require_gpu):
    scene, vox, rc, frame = tiny_case(seed=2)
    rc = dataclasses.replace(rc, noMerge=True)
    say("case: tiny random scene, trace only")
    w = run_world(scene, vox, rc, frame)
    o = Oracle(vox, rc, frame)
    o.build(scene)
    check_cascades(w, o, rc, merged=False)
    w.destroy()


@pytest.mark.parametrize("max_bricks", [40, 7, 1])
def test_brick_pool_overflow(require_gpu, max_bricks):
    """A full pool drops the finest bricks and reads them one level coarser (voxel.ts:39-45, 146-167, 450-470)."""
    scene, vox, rc, frame = tiny_case(seed=3, ntris=500, maxBricks=max_bricks)
    say(f"case: brick overflow, maxBricks={max_bricks}")
    w = run_world(scene, vox, rc, frame)
    o = Oracle(vox, rc, frame)
    o.build(scene)
    assert o.counters()["brickOverflow"] > 0
    check_clipmap(w, o, vox)
    check_cascades(w, o, rc)
    w.destroy()


def test_voxel_pool_overflow(require_gpu):
    scene, vox, rc, frame = tiny_case(seed=4, ntris=500, maxVoxels=300)
    say("case: voxel attribute overflow, maxVoxels=300")
    w = run_world(scene, vox, rc, frame)
    o = Oracle(vox, rc, frame)
    o.build(scene)
    assert o.counters()["voxelOverflow"] > 0
    check_clipmap(w, o, vox)
    check_cascades(w, o, rc)
    w.destroy()


def test_degenerate_and_far_triangles(require_gpu):
    """Ragged geometry: a third of the triangles degenerate (a repeated vertex, or three collinear vertices: zero
    area and a zero face normal), some far outside every window, some with zero vertex normals (the resolve falls
    back to the face normal), plus axis-aligned slivers exactly on voxel faces. Clipmap and cascades at the bars."""
    rng = np.random.default_rng(9)
    scene, vox, rc, frame = tiny_case(seed=9, ntris=600)
    pos = scene.pos.reshape(-1, 3, 3).copy()
    nrm = scene.nrm.reshape(-1, 3, 3).copy()
    n = pos.shape[0]
    pos[0:100, 2] = pos[0:100, 1]                                              # a repeated vertex
    pos[100:200, 2] = 0.25 * pos[100:200, 0] + 0.75 * pos[100:200, 1]          # collinear (approximately in f32)
    pos[200:250] += np.float32(1e4)                                            # outside every window
    nrm[250:300] = 0.0                                                         # no vertex normals
    g = rng.integers(-16, 16, (100, 3)).astype(np.float32) * np.float32(0.25)  # slivers on voxel faces
    pos[300:400, 0] = g
    pos[300:400, 1] = g + np.array([0.25, 0.0, 0.0], np.float32)
    pos[300:400, 2] = g + np.array([0.0, 0.25, 0.0], np.float32)
    scene = dataclasses.replace(scene, pos=pos.reshape(-1, 3), nrm=nrm.reshape(-1, 3))
    say("case: degenerate, far and axis-aligned triangles, 600 triangles")
    w = run_world(scene, vox, rc, frame)
    o = Oracle(vox, rc, frame)
    o.build(scene)
    check_clipmap(w, o, vox)
    check_cascades(w, o, rc)
    w.destroy()


def test_empty_scene(require_gpu):
    scene, vox, rc, frame = tiny_case(seed=5, ntris=1)
    scene = dataclasses.replace(scene, tri=scene.tri[:0])
    say("case: empty scene")
    w = run_world(scene, vox, rc, frame)
    assert w.read_counters()["bricks"] == 0
    for c in range(rc.cascades):
        g, h = w.read_cascade(c, hits=True)
        assert np.all(h == 0xFFFFFFFF)
        # Every ray misses: merged radiance is the sky through every cascade (T = 1 all the way up).
        np.testing.assert_array_equal(g[:, 3], 1.0)
        np.testing.assert_allclose(g[:, :3], np.broadcast_to(np.float32(frame.sky), g[:, :3].shape), rtol=1e-6)
    w.destroy()


def test_point_lights_and_shading(require_gpu):
    scene, vox, rc, frame = tiny_case(seed=6, ntris=400)
    frame = dataclasses.replace(frame, sunVis=0.7, points=(((1.0, 2.0, 0.5), (3.0, 2.0, 1.0)),
                                                           ((-2.0, -1.0, 1.5), (1.0, 4.0, 2.0))))
    say("case: point lights, sun visibility 0.7")
    w = run_world(scene, vox, rc, frame)
    o = Oracle(vox, rc, frame)
    o.build(scene)
    check_cascades(w, o, rc)
    w.destroy()


def test_res_cap_and_interval(require_gpu):
    """resCap stops the direction growth (upperAvg reads the same texel, rc.ts:682); another kInterval, levelMap 2."""
    scene, vox, rc, frame = tiny_case(seed=7, ntris=400)
    rc = dataclasses.replace(rc, resCap=1, kInterval=1.5, levelMap=2.0, voxSkip=0.0)
    say("case: resCap 1, kInterval 1.5, levelMap 2, voxSkip 0")
    w = run_world(scene, vox, rc, frame)
    o = Oracle(vox, rc, frame)
    o.build(scene)
    lay = w.layout()
    assert lay["res"] == [4, 8, 8]
    check_cascades(w, o, rc)
    w.destroy()


def test_atrium_parity(require_gpu):
    scene, vox, rc, frame = rtgi.config("atrium")
    say(f"case: atrium {scene.ntris} triangles, 4 x 64^3, 32^3 c0 probes, 4 cascades")
    w = run_world(scene, vox, rc, frame)
    o = Oracle(vox, rc, frame)
    o.build(scene)
    lay = w.layout()
    assert lay["vox_level"] == list(o.rc.vox_level)[:rc.cascades]
    assert lay["grid_origin"] == list(o.rc.G0)
    check_clipmap(w, o, vox)
    check_cascades(w, o, rc)
    w.destroy()


def test_city_parity_and_determinism(require_gpu):
    scene, vox, rc, frame = rtgi.config("city")
    say(f"case: city {scene.ntris} triangles, 4 x 128^3, 64^3 c0 probes, 5 cascades")
    w = run_world(scene, vox, rc, frame)
    o = Oracle(vox, rc, frame)
    o.build(scene)
    check_clipmap(w, o, vox)
    check_cascades(w, o, rc)
    # A second update of the same inputs is bit-identical (no atomic-order dependence in any output).
    first = [w.read_cascade(c) for c in range(rc.cascades)]
    sh0 = w.read_sh()
    w.encode()
    w.sync()
    for c in range(rc.cascades):
        assert np.array_equal(w.read_cascade(c), first[c])
    assert np.array_equal(w.read_sh(), sh0)
    w.destroy()


def test_maximum_levels_and_cascades(require_gpu):
    """The largest clipmap and cascade counts the reference accepts (8 levels, voxel.ts:749; 8 cascades,
    rc.ts:1022): 128^3 c0 probes (the smallest grid that halves 7 times), directions capped at 16^2 from c2 on,
    59.9M rays; a random scene spread over all 8 levels. Clipmap and cascades at the parity bars."""
    rng = np.random.default_rng(8)
    scene = random_scene(rng, 400, -5.0, 5.0, 0.6)
    vox = rtgi.VoxParams(v0=0.05, levels=8, G=4, maxBricks=8 * 64, maxVoxels=1 << 16)
    rc = rtgi.RcParams(cascades=8, probes0=128, s0=0.05, resCap=2)
    frame = rtgi.Frame(eye=(0.3, 0.2, -0.1), sunDir=(0.3, 1.0, -0.2))
    say("case: 8 levels x 32^3, 8 cascades, 128^3 c0 probes, 59.9M rays")
    w = run_world(scene, vox, rc, frame)
    o = Oracle(vox, rc, frame)
    o.build(scene)
    assert all(x > 0 for x in o.counters()["bricksPerLevel"])
    check_clipmap(w, o, vox)
    check_cascades(w, o, rc)
    w.destroy()


@pytest.mark.parametrize("case", ["atrium", "city"])
def test_reference_default_voxel_params(require_gpu, case):
    """The reference's DEFAULT_VOX (v0 0.2, 5 levels of 256^3, 4,608 bricks, 524,288 voxels; voxel.ts:46): on the
    atrium every level fits; on the city 10,488 bricks are wanted, so the finest 5,880 are dropped and their rays
    read the coarser levels (voxAt's fallback). Clipmap and cascades at the parity bars."""
    scene, _, _, frame = rtgi.config(case)
    vox = rtgi.DEFAULT_VOX
    rc = rtgi.RcParams(cascades=4, probes0=32, s0=0.625)
    say(f"case: {case} {scene.ntris} triangles at the reference's default voxel parameters")
    w = run_world(scene, vox, rc, frame)
    o = Oracle(vox, rc, frame)
    o.build(scene)
    check_clipmap(w, o, vox)
    if case == "city":
        assert o.counters()["brickOverflow"] > 0
    check_cascades(w, o, rc)
    w.destroy()


def test_moving_frames(require_gpu):
    """Full rebuild per frame: posed vertices of frame 3 uploaded into a world built at frame 0 match a fresh oracle."""
    scene0 = rtgi.generate_scene("atrium", seed=2, frame=0)
    scene3 = rtgi.generate_scene("atrium", seed=2, frame=3)
    _, vox, rc, frame = rtgi.config("atrium")
    say("case: atrium seed 2, frame 0 then vertices of frame 3")
    w = run_world(scene0, vox, rc, frame)
    hist = History(rc.probes0)  # the second update's SH blends with the first's
    o0 = Oracle(vox, rc, frame)
    o0.build(scene0)
    o0.merge_sh(o0.trace()[0], hist=hist)
    w.update_vertices(0, scene3.pos, scene3.nrm)
    w.update((0.5, 5.0, -0.5))
    w.encode()
    w.sync()
    frame3 = dataclasses.replace(frame, eye=(0.5, 5.0, -0.5))
    o = Oracle(vox, rc, frame3)
    o.build(scene3)
    check_clipmap(w, o, vox)
    check_cascades(w, o, rc, hist=hist)
    w.destroy()


def test_golden_fixture(require_gpu):
    """The GPU path reproduces the committed fixture (tests/golden/tiny_world.npz) without running the oracle."""
    import json
    import os
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "tiny_world.npz"))
    scene = rtgi.Scene(pos=g["pos"], nrm=g["nrm"], tri=g["tri"], materials=g["mat"])
    vox = rtgi.VoxParams(**json.loads(str(g["vox"])))
    rc = rtgi.RcParams(**json.loads(str(g["rc"])))
    frame = rtgi.Frame(**json.loads(str(g["frame"])))
    say("case: golden fixture tiny_world.npz")
    w = run_world(scene, vox, rc, frame)
    for L in range(vox.levels):
        cm = w.read_clipmap(L)
        assert np.array_equal(cm["occ"], g[f"occ{L}"]) and np.array_equal(cm["rep"], g[f"rep{L}"])
    lay = w.layout()
    off = 0
    for c in range(rc.cascades):
        rgba, h = w.read_cascade(c, hits=True)
        n = lay["rays"][c]
        assert np.array_equal(h, g["hits"][off:off + n])
        np.testing.assert_allclose(rgba, g["merged"][off:off + n], rtol=RTOL, atol=ATOL)
        off += n
    np.testing.assert_allclose(w.read_sh(), g["sh"], rtol=RTOL, atol=ATOL)
    w.destroy()


@pytest.mark.parametrize("case", ["tiny", "atrium", "city"])
def test_fp16_precision_form(require_gpu, case):
    """Radiance and SH stored as f16 at the reference's pack points: identical to the oracle's quantised values."""
    if case == "tiny":
        scene, vox, rc, frame = tiny_case(seed=8, ntris=500)
    else:
        scene, vox, rc, frame = rtgi.config(case)
    rc = dataclasses.replace(rc, fp16=True)
    say(f"case: f16 storage, {case}")
    w = run_world(scene, vox, rc, frame)
    o = Oracle(vox, rc, frame)
    o.build(scene)
    rgba, hits = o.trace()
    m, sh = o.merge_sh(rgba)
    per_m, per_h = o.split(m), o.split(hits)
    for c in range(rc.cascades):
        g, h = w.read_cascade(c, hits=True)
        assert np.array_equal(h, per_h[c])
        assert np.array_equal(g, per_m[c]), f"cascade {c}: {np.count_nonzero(g != per_m[c])} values differ"
    assert np.array_equal(w.read_sh(), sh)
    w.destroy()


@pytest.mark.parametrize("case", ["tiny", "overflow", "atrium"])
def test_trace_stats(require_gpu, case):
    """setStats counters (rc.ts:232-235): traces, DDA loop iterations and step-cap hits per cascade equal the
    oracle's per-ray voxSteps summed per cascade; a second update starts them from zero again."""
    if case == "atrium":
        scene, vox, rc, frame = rtgi.config("atrium")
    else:
        scene, vox, rc, frame = tiny_case(seed=4, ntris=500, **({"maxBricks": 7} if case == "overflow" else {}))
    rc = dataclasses.replace(rc, stats=1)
    say(f"case: tracer statistics, {case}")
    w = run_world(scene, vox, rc, frame)
    o = Oracle(vox, rc, frame)
    o.build(scene)
    _, _, steps = o.trace(steps=True)
    per = o.split(steps)
    for rep in range(2):
        st = w.read_trace_stats()
        for c in range(rc.cascades):
            s = per[c].astype(np.uint64)
            assert st["traces"][c] == s.size, (rep, c)
            assert st["steps"][c] == int(s.sum()), (rep, c, st["steps"][c], int(s.sum()))
            assert st["capped"][c] == int(np.count_nonzero(s >= 1024)), (rep, c)
        w.encode()
        w.sync()
    w.destroy()


@pytest.mark.parametrize("case", ["tiny", "tiny-nomerge", "atrium", "city"])
def test_tuning_keeps_results(require_gpu, case):
    """Launch tuning changes no result: probe storage order (probe_tile 1 / 2 / 4) and per-cascade fused trace+merge
    launches read back cascades, hit ids, SH and tracer statistics bit-identical, plain and FP16 forms."""
    if case.startswith("tiny"):
        scene, vox, rc, frame = tiny_case(seed=9, ntris=400)
        rc = dataclasses.replace(rc, noMerge=case.endswith("nomerge"))
    else:
        scene, vox, rc, frame = rtgi.config(case)
    for fp16 in (False, True):
        rcf = dataclasses.replace(rc, fp16=fp16, recordHits=True, stats=1)
        say(f"case: tuning, {case}, fp16={fp16}")
        w = rtgi.World(vox, rcf)
        w.upload_scene(scene)
        w.set_frame(frame)
        ref = None
        for tile, fused in ((1, 0), (2, 0), (4, 0), (1, 1), (4, 1)):
            w.set_tuning(probe_tile=tile, fused=fused)
            assert w.tuning() == {"probe_tile": tile, "fused": fused}
            w.encode()
            w.sync()
            out = [w.read_cascade(c, hits=True) for c in range(rc.cascades)]
            if not rc.noMerge:
                out.append((w.read_sh(), None))
            stats = w.read_trace_stats()
            if ref is None:
                ref = (out, stats)
                continue
            assert stats == ref[1], (tile, fused)
            for (a, ha), (b, hb) in zip(out, ref[0]):
                assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (tile, fused, case)
                assert ha is None or np.array_equal(ha, hb)
        w.destroy()


def test_tilted_triangle_attributes(require_gpu):
    """The tilted triangle of the resolveAttr closed form (tests/test_oracle.py): occupancy, normals, plane offsets
    and albedo bit for bit against the oracle, whose values that test pins."""
    from test_oracle import tilted_triangle_world
    scene, vox, rc, frame = tilted_triangle_world()
    say("case: one tilted triangle with three vertex normals")
    w = run_world(scene, vox, rc, frame)
    o = Oracle(vox, rc, frame)
    o.build(scene)
    check_clipmap(w, o, vox)
    w.destroy()


def test_readback_after_set_tuning(require_gpu):
    """The probe order can change between an update and its readbacks: cascades, hit ids, SH, relocated probes and the
    gather keep reading that update's values (the buffers remember the order they were written in)."""
    from conftest import surface_points
    scene, vox, rc, frame = tiny_case(seed=3)
    rcf = dataclasses.replace(rc, recordHits=True, relocate=True)
    say("case: readbacks after set_tuning(probe_tile=2), tiny scene with relocation")
    w = run_world(scene, vox, rcf, frame)
    pos, nrm = surface_points(scene, 500, seed=4)

    def reads():
        out = [w.read_cascade(c, hits=True) for c in range(rc.cascades)]
        out += [(w.read_sh(), None), (w.gather(pos, nrm), None)]
        out += [(np.concatenate(w.read_probes(c)[:2], 1), w.read_probes(c)[2]) for c in range(rc.cascades)]
        return out

    before = reads()
    w.set_tuning(probe_tile=2)
    for (a, ha), (b, hb) in zip(reads(), before):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
        assert ha is None or np.array_equal(ha, hb)
    w.destroy()
