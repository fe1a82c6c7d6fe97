"""Multi-bounce (bnInject, src/core/bounce.ts:224-262, read back by voxBounce, src/core/rc.ts:198-202) on the GPU
against the oracle: the second of two updates with a moved eye injects the first update's field (its c0 grid and,
with relocation, its probe positions) into the second clipmap's voxels and shades hits with it. Cascades, first-hit
ids and SH under the plain bars (ids exact, radiance within 1e-4; FP16 exact); and the batched, graph-replayed path
gives the single-world path's results bit for bit."""
import dataclasses

import numpy as np
import pytest

import rtgi
from conftest import random_scene
from oracle.oracle import Oracle

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


def check_frame(w, o, rc, fp16):
    rgba, hits = o.trace()
    m, sh = o.merge_sh(rgba)
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


@pytest.mark.parametrize("case", ["tiny", "atrium", "city"])
def test_bounce_matches_oracle(require_gpu, case):
    scene, vox, rc, f1 = tiny() if case == "tiny" else rtgi.config(case)
    f2 = moved(f1, (3.0 * rc.s0, 0.0, -2.0 * rc.s0))
    for fp16, reloc in ((False, False), (True, True)):
        rcf = dataclasses.replace(rc, bounce=True, relocate=reloc, fp16=fp16, recordHits=True)
        say(f"case: multi-bounce, {case}, fp16={fp16}, relocate={reloc}, update 1")
        w = rtgi.World(vox, rcf)
        w.upload_scene(scene)
        w.set_frame(f1)
        w.encode()
        w.sync()
        sh1 = w.read_sh()
        o1 = Oracle(vox, rcf, f1)
        o1.build(scene)
        if case == "tiny":
            check_frame(w, o1, rc, fp16)  # no previous field: no bounce
        say(f"case: multi-bounce, {case}, fp16={fp16}, relocate={reloc}, update 2")
        w.set_frame(f2)
        w.encode()
        w.sync()
        o2 = Oracle(vox, rcf, f2)
        o2.build(scene)
        o2.inject_bounce(o1, sh1)
        assert np.count_nonzero(o2.bounce) > 0
        check_frame(w, o2, rc, fp16)
        w.destroy()


def test_bounce_batch_matches_single(require_gpu):
    scene, vox, rc, f1 = rtgi.config("atrium")
    rcf = dataclasses.replace(rc, bounce=True, relocate=True, fp16=True)
    frames = [f1, moved(f1, (2.0, 0.0, 1.0)), moved(f1, (2.5, 0.0, -1.0))]
    say("case: multi-bounce, atrium, single world vs batch of 2 (graph), 3 updates")
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
            for w in pair:
                assert np.array_equal(w.read_sh(), ref)
    finally:
        b.destroy()
        for w in [single] + pair:
            w.destroy()
