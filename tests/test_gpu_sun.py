"""Traced per-voxel sun visibility (voxSun, src/core/voxel.ts:631-673, with voxel shadow rays for the shadow-map
taps) on the GPU against the oracle: cascades, first-hit ids and SH under the plain bars (ids exact, radiance within
1e-4; FP16 exact), with and without relocation; and the roof case's closed form through the ray query."""
import dataclasses

import numpy as np
import pytest

import rtgi
from oracle.oracle import Oracle
from test_oracle import roof_world

pytestmark = pytest.mark.gpu
RTOL, ATOL = 1e-4, 1e-6


def say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("case", ["roof", "atrium", "city"])
def test_sun_trace_matches_oracle(require_gpu, case):
    scene, vox, rc, frame = roof_world() if case == "roof" else rtgi.config(case)
    for fp16, reloc in ((False, False), (True, True)):
        rcf = dataclasses.replace(rc, sunTrace=True, relocate=reloc, fp16=fp16, recordHits=True)
        say(f"case: traced sun visibility, {case}, fp16={fp16}, relocate={reloc}")
        w = rtgi.World(vox, rcf)
        w.upload_scene(scene)
        w.set_frame(frame)
        w.encode()
        w.sync()
        o = Oracle(vox, rcf, frame)
        o.build(scene)
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
        w.destroy()


def test_sun_trace_roof_closed_form(require_gpu):
    scene, vox, rc, frame = roof_world()
    say("case: roof, rays onto shadowed and sunlit floor")
    out = {}
    for st in (False, True):
        w = rtgi.World(vox, dataclasses.replace(rc, sunTrace=st))
        w.upload_scene(scene)
        w.set_frame(frame)
        w.encode()
        w.sync()
        rng = np.random.default_rng(4)
        m = 200
        xs = np.concatenate([rng.uniform(-3.5, -1.5, m), rng.uniform(1.5, 3.5, m)])
        org = np.stack([xs, np.full(2 * m, 0.3), rng.uniform(-3, 3, 2 * m)], 1).astype(np.float32)
        d = np.tile(np.array([[0.0, -1.0, 0.0]], np.float32), (2 * m, 1))
        out[st] = w.trace_rays(0, org, d)
        w.destroy()
    (rp, hp), (rt, ht) = out[False], out[True]
    assert np.array_equal(hp, ht)
    assert np.all(rp[:, :3] > 0) and np.all(rt[:200, :3] == 0) and np.array_equal(rt[200:], rp[200:])
