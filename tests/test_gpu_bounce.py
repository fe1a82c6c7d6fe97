"""Multi-bounce (createBounce with DEFAULT_BOUNCE, src/core/bounce.ts) on the GPU against the oracle, over three
updates with a moving eye (the last one at the second's eye, so that every second probe keeps its texels): the bounce
volume of each update (bnClassify + bnUpdate: hits read the previous volume) bit for bit with its origin, then the
cascades, first-hit ids and SH of the update, whose hits add the injected volume irradiance (bnInject, voxBounce),
under the plain bars (ids exact, radiance within 1e-4; FP16 exact). The batched, graph-replayed path gives the
single-world path's volumes and SH bit for bit."""
import dataclasses

import numpy as np
import pytest

import rtgi
from conftest import random_scene
from oracle.oracle import History, Oracle
from test_oracle import floor_world

pytestmark = pytest.mark.gpu
RTOL, ATOL = 1e-4, 1e-6


def say(msg):
    print(msg, flush=True)


def tiny():
    rng = np.random.default_rng(5)
    scene = random_scene(rng, 600, -4, 4, 0.6)
    vox = rtgi.VoxParams(0.25, 3, 4, 192, 1 << 16)
    rc = rtgi.RcParams(cascades=3, probes0=8, s0=0.5)
    return scene, vox, rc, rtgi.Frame(eye=(0.2, 0.1, -0.3))


def moved(frame, d):
    return dataclasses.replace(frame, eye=tuple(float(a + b) for a, b in zip(frame.eye, d)))


def check_update(w, o, rc, fp16, hist):
    rgba, hits = o.trace()
    m, sh = o.merge_sh(rgba, hist=hist)
    per_m, per_h = o.split(m), o.split(hits)
    for c in range(rc.cascades):
        g, h = w.read_cascade(c, hits=True)
        assert np.array_equal(h, per_h[c])
        if fp16:
            assert np.array_equal(g, per_m[c])
        else:
            np.testing.assert_allclose(g, per_m[c], rtol=RTOL, atol=ATOL)
    if fp16:
        assert np.array_equal(w.read_sh(), sh)
    else:
        np.testing.assert_allclose(w.read_sh(), sh, rtol=RTOL, atol=ATOL)


def check_volume(w, state):
    vol, org = w.read_bounce_volume()
    assert np.array_equal(org, state[1])
    assert np.array_equal(vol.view(np.uint16).reshape(-1, 6, 4), state[0])
    return int(np.count_nonzero(vol[..., 0, 3] > 0))


@pytest.mark.parametrize("case", ["tiny", "floor", "atrium", "city"])
def test_bounce_matches_oracle(require_gpu, case):
    if case == "tiny":
        scene, vox, rc, f1 = tiny()
    elif case == "floor":
        scene, vox, rc, f1 = floor_world()
    else:
        scene, vox, rc, f1 = rtgi.config(case)
    f2 = moved(f1, (2.5, 0.0, -1.7))
    frames = [f1, f2, f2]
    # the last form keeps the albedo and the injected bounce light as bf16 (the FP16 radiance + BF16 albedo precision form)
    for fp16, reloc, sun, bf16 in ((False, False, False, False), (True, True, True, False), (True, False, False, True)):
        rcf = dataclasses.replace(rc, bounce=True, relocate=reloc, sunTrace=sun, fp16=fp16, recordHits=True,
                                  bf16=bf16)
        w = rtgi.World(vox, rcf)
        w.upload_scene(scene)
        state = None
        hist = History(rc.probes0)
        for k, f in enumerate(frames):
            say(f"case: multi-bounce, {case}, fp16={fp16}, relocate={reloc}, sunTrace={sun}, bf16={bf16}, "
                f"update {k + 1}")
            w.set_frame(f)
            w.encode()
            w.sync()
            o = Oracle(vox, rcf, f)
            o.build(scene)
            state = o.bounce_update(state)
            active = check_volume(w, state)
            assert active > 0
            check_update(w, o, rc, fp16, hist)
        w.destroy()


def test_bounce_batch_matches_single(require_gpu):
    scene, vox, rc, f1 = rtgi.config("atrium")
    rcf = dataclasses.replace(rc, bounce=True, relocate=True, fp16=True)
    frames = [f1, moved(f1, (2.0, 0.0, 1.0)), moved(f1, (2.0, 0.0, 1.0)), moved(f1, (-1.5, 0.0, 0.5))]
    say("case: multi-bounce, atrium, single world vs batch of 2 (graph), 4 updates")
    single = rtgi.World(vox, rcf)
    pair = [rtgi.World(vox, rcf) for _ in range(2)]
    for w in [single] + pair:
        w.upload_scene(scene)
    b = rtgi.Batch(pair, graph=True)
    try:
        for f in frames:
            for w in [single] + pair:
                w.set_frame(f)
            single.encode()
            single.sync()
            b.encode()
            b.sync()
            ref = single.read_sh()
            rv, ro = single.read_bounce_volume()
            for w in pair:
                assert np.array_equal(w.read_sh(), ref)
                v, og = w.read_bounce_volume()
                assert np.array_equal(og, ro) and np.array_equal(v.view(np.uint16), rv.view(np.uint16))
    finally:
        b.destroy()
        for w in [single] + pair:
            w.destroy()
