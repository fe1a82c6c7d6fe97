"""Batched updates (one launch per stage for many worlds of different sizes): every world's result is identical to its
solo update, in direct launches and in hipGraph replays, and after frames change between updates."""
import dataclasses

import numpy as np
import pytest

import rtgi
from conftest import random_scene

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def make_specs(seed=0):
    """Uneven worlds: different clipmap sizes, levels, cascade counts, direction caps, pools and scenes."""
    rng = np.random.default_rng(seed)
    specs = []
    specs.append((random_scene(rng, 400, -4, 4, 0.5), rtgi.VoxParams(0.25, 3, 4, 192, 1 << 16),
                  rtgi.RcParams(cascades=3, probes0=8, s0=0.5), (0.2, 0.1, -0.3)))
    specs.append((random_scene(rng, 900, -6, 6, 0.7), rtgi.VoxParams(0.25, 2, 8, 40, 1 << 16),  # brick overflow
                  rtgi.RcParams(cascades=2, probes0=16, s0=0.5, resCap=0), (0.0, 0.3, 0.1)))
    specs.append((random_scene(rng, 300, -3, 3, 0.4), rtgi.VoxParams(0.125, 4, 12, 4 * 1728, 2000),  # voxel overflow
                  rtgi.RcParams(cascades=4, probes0=16, s0=0.375), (0.1, 0.2, 0.3)))
    at = rtgi.generate_scene("atrium", seed=3)
    specs.append((at, rtgi.VoxParams(0.125, 4, 8, 2048, 1 << 20), rtgi.RcParams(cascades=4, probes0=32, s0=0.625),
                  at.eye))
    specs.append((random_scene(rng, 50, -2, 2, 0.3), rtgi.VoxParams(0.2, 2, 4, 128, 1 << 14),  # non power-of-two v0
                  rtgi.RcParams(cascades=1, probes0=4, s0=0.4), (0.05, 0.0, 0.0)))
    specs.append((random_scene(rng, 200, -4, 4, 0.5), rtgi.VoxParams(0.25, 3, 4, 192, 1 << 16),
                  rtgi.RcParams(cascades=3, probes0=8, s0=0.5, noMerge=True), (0.0, 0.0, 0.0)))
    specs.append((random_scene(rng, 600, -5, 5, 0.6), rtgi.VoxParams(0.25, 3, 8, 3 * 512, 1 << 16),  # f16 storage
                  rtgi.RcParams(cascades=4, probes0=16, s0=0.5, fp16=True), (0.3, -0.2, 0.1)))
    return specs


def make_worlds(seed=0):
    worlds = []
    for i, (sc, vox, rc, eye) in enumerate(make_specs(seed)):
        w = rtgi.World(vox, dataclasses.replace(rc, recordHits=True))
        w.upload_scene(sc)
        w.set_frame(rtgi.Frame(eye=eye, sunDir=(0.2 + 0.1 * i, 1.0, -0.3),
                               points=(((1.0, 2.0, 0.5), (2.0, 3.0, 1.0)),) if i % 2 else ()))
        worlds.append(w)
    return worlds


def snapshot(w):
    out = {"counters": w.read_counters()}
    for L in range(w.vox.levels):
        cm = w.read_clipmap(L)
        for k, v in cm.items():
            out[f"{k}{L}"] = v.copy()
    for c in range(w.rc.cascades):
        rgba, hits = w.read_cascade(c, hits=True)
        out[f"rad{c}"] = rgba
        out[f"hits{c}"] = hits
    if not w.rc.noMerge:
        out["sh"] = w.read_sh()
    return out


def assert_same(a, b, what):
    assert a["counters"] == b["counters"], what
    for k in a:
        if k == "counters":
            continue
        x, y = a[k], b[k]
        if x.dtype.kind == "f":
            x, y = x.view(np.uint32), y.view(np.uint32)
        assert np.array_equal(x, y), f"{what}: {k} differs in {np.count_nonzero(x != y)} places"


def solo_results(worlds):
    res = []
    for w in worlds:
        w.encode()
        w.sync()
        res.append(snapshot(w))
    return res


@pytest.mark.parametrize("fused", [0, 1])
def test_batch_matches_solo_direct(require_gpu, fused):
    worlds = make_worlds()
    say(f"case: batch of {len(worlds)} uneven worlds, direct launches, fused={fused}")
    solo = solo_results(worlds)
    for w in worlds:
        w.set_tuning(fused=fused, probe_tile=4 if fused else 1)
    with rtgi.Batch(worlds) as b:
        b.encode()
        b.sync()
        for i, w in enumerate(worlds):
            assert_same(snapshot(w), solo[i], f"world {i}")
    for w in worlds:
        w.destroy()


def test_batch_graph_replays(require_gpu):
    worlds = make_worlds(seed=1)
    say(f"case: batch of {len(worlds)} uneven worlds, hipGraph replays with eager work between")
    solo = solo_results(worlds)
    with rtgi.Batch(worlds, graph=True) as b:
        for rep in range(3):
            b.encode()
            b.sync()
            for i, w in enumerate(worlds):
                assert_same(snapshot(w), solo[i], f"replay {rep} world {i}")
            # Eager work between replays: a solo update of one world (same inputs, same results).
            worlds[rep].encode()
            worlds[rep].sync()
        # A frame change between replays reaches the graph through the descriptors. Repeated updates at one frame
        # keep the SH (the c0 history blends equal values), so a twin world updated once at the first eye and once
        # at the moved one holds the moved world's result, history included.
        worlds[0].update((0.9, -0.4, 0.6))
        b.encode()
        b.sync()
        moved = snapshot(worlds[0])
    sc, vox, rc, eye = make_specs(seed=1)[0]
    twin = rtgi.World(vox, dataclasses.replace(rc, recordHits=True))
    twin.upload_scene(sc)
    for e in (eye, (0.9, -0.4, 0.6)):
        twin.set_frame(rtgi.Frame(eye=e, sunDir=(0.2, 1.0, -0.3)))
        twin.encode()
        twin.sync()
    assert_same(snapshot(twin), moved, "moved eye")
    twin.destroy()
    for w in worlds:
        w.destroy()


def test_graph_stage_timing_keeps_results(require_gpu):
    """Per-stage HIP events captured into the graph: every replay reports positive stage times, and the results stay
    the solo ones; switching the timing off re-captures a graph without the records."""
    worlds = make_worlds(seed=2)
    say(f"case: batch of {len(worlds)} uneven worlds, hipGraph replays with captured stage timing")
    solo = solo_results(worlds)
    with rtgi.Batch(worlds, graph=True) as b:
        b.set_timing(True)
        for rep in range(3):
            b.encode()
            b.sync()
            ms = b.stage_ms()
            assert ms["vox"] > 0 and ms["trace"] > 0 and ms["sh"] > 0, ms
            for i, w in enumerate(worlds):
                assert_same(snapshot(w), solo[i], f"timed replay {rep} world {i}")
        b.set_timing(False)
        b.encode()
        b.sync()
        for i, w in enumerate(worlds):
            assert_same(snapshot(w), solo[i], f"untimed replay world {i}")
    for w in worlds:
        w.destroy()


def test_batch_of_cities_matches_solo(require_gpu):
    say("case: batch of 3 city worlds (500k triangles each) vs solo, fused hipGraph")
    worlds = []
    for seed in (1, 2, 3):
        sc, vox, rc, frame = rtgi.config("city")
        sc = rtgi.generate_scene("city", seed=seed)
        w = rtgi.World(vox, dataclasses.replace(rc, recordHits=True))
        w.upload_scene(sc)
        w.set_frame(dataclasses.replace(frame, eye=sc.eye))
        worlds.append(w)
    solo = solo_results(worlds)
    for w in worlds:
        w.set_tuning(fused=1)
    with rtgi.Batch(worlds, graph=True) as b:
        for rep in range(2):
            b.encode()
            b.sync()
            for i, w in enumerate(worlds):
                assert_same(snapshot(w), solo[i], f"city {i}")
    for w in worlds:
        w.destroy()


def test_autotune_applies_a_candidate_and_keeps_results(require_gpu):
    from rtgi.autotune import autotune

    worlds = make_worlds(seed=2)[:4]
    say(f"case: autotune over {len(worlds)} uneven worlds")
    solo = solo_results(worlds)
    res = autotune(worlds, widths=(2, 4), repeats=1)
    assert len(res["table"]) == 3 * 2 * 2 and res["best"] in res["table"]
    for w in worlds:
        assert w.tuning() == {"probe_tile": res["best"]["probe_tile"], "fused": res["best"]["fused"]}
    with rtgi.Batch(worlds, graph=True) as b:
        b.encode()
        b.sync()
        for i, w in enumerate(worlds):
            assert_same(snapshot(w), solo[i], f"tuned world {i}")
    for w in worlds:
        w.destroy()
