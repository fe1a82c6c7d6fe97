"""Voxel relocation ('vrel': relocateVox, src/core/rc.ts:413-472; occlusion-aware parents, linkParents :559-566) on
the GPU against the oracle: relocated positions, normals and parent masks bit-exact; traced and merged cascades,
first-hit ids and SH as in the plain path's bars (ids exact, radiance within 1e-4; the FP16 form exact); gather with
occlusion from the relocated probes; batched and fused launches identical to solo ones."""
import dataclasses

import numpy as np
import pytest

import rtgi
from conftest import random_scene, surface_points
from oracle.oracle import Oracle
from test_oracle import wall_world

pytestmark = pytest.mark.gpu
RTOL, ATOL = 1e-4, 1e-6


def say(msg):
    print(msg, flush=True)


def case_params(case):
    if case == "tiny":
        rng = np.random.default_rng(11)
        scene = random_scene(rng, 800, -4, 4, 0.7)
        return (scene, rtgi.VoxParams(0.25, 3, 4, 192, 1 << 16), rtgi.RcParams(cascades=3, probes0=8, s0=0.5),
                rtgi.Frame(eye=(0.3, -0.2, 0.1)))
    if case == "wall":
        return wall_world(x=0.52)
    return rtgi.config(case)


@pytest.mark.parametrize("case", ["wall", "tiny", "atrium", "city"])
def test_relocation_matches_oracle(require_gpu, case):
    scene, vox, rc, frame = case_params(case)
    for fp16 in (False, True):
        rcf = dataclasses.replace(rc, relocate=True, recordHits=True, fp16=fp16)
        say(f"case: relocation, {case}, fp16={fp16}")
        w = rtgi.World(vox, rcf)
        w.upload_scene(scene)
        w.set_frame(frame)
        w.encode()
        w.sync()
        o = Oracle(vox, rcf, frame)
        o.build(scene)
        moved = 0
        for c in range(rc.cascades):
            gp, gn, gv = w.read_probes(c)
            op, on, ov = o.probes(c)
            assert np.array_equal(gp.view(np.uint32), op.view(np.uint32)), f"cascade {c} positions"
            assert np.array_equal(gn.view(np.uint32), on.view(np.uint32)), f"cascade {c} normals"
            assert np.array_equal(gv, ov), f"cascade {c} parents"
            moved += int(np.count_nonzero(np.any(gn != 0, axis=1)))
        assert moved > 0
        rgba, hits = o.trace()
        m, sh = o.merge_sh(rgba)
        per_m, per_h = o.split(m), o.split(hits)
        for c in range(rc.cascades):
            g, h = w.read_cascade(c, hits=True)
            assert np.array_equal(h, per_h[c]), f"cascade {c}: {np.count_nonzero(h != per_h[c])} hit ids differ"
            if fp16:
                assert np.array_equal(g, per_m[c])
            else:
                np.testing.assert_allclose(g, per_m[c], rtol=RTOL, atol=ATOL)
        gsh = w.read_sh()
        if fp16:
            assert np.array_equal(gsh, sh)
        else:
            np.testing.assert_allclose(gsh, sh, rtol=RTOL, atol=ATOL)
        pos, nrm = surface_points(scene, 5000, seed=3)
        np.testing.assert_allclose(w.gather(pos, nrm, occlusion=True), o.gather(gsh, pos, nrm, occlusion=True),
                                   rtol=RTOL, atol=ATOL)
        w.destroy()


def test_relocated_batch_and_fused_equal_solo(require_gpu):
    say("case: relocated worlds, solo vs fused vs batched")
    worlds, solo = [], []
    for case in ("tiny", "atrium"):
        scene, vox, rc, frame = case_params(case)
        w = rtgi.World(vox, dataclasses.replace(rc, relocate=True, recordHits=True))
        w.upload_scene(scene)
        w.set_frame(frame)
        w.encode()
        w.sync()
        solo.append([w.read_cascade(c, hits=True) for c in range(rc.cascades)] + [(w.read_sh(), None)])
        worlds.append(w)
    for w in worlds:
        w.set_tuning(fused=1, probe_tile=4)
    with rtgi.Batch(worlds, graph=True) as b:
        for rep in range(2):
            b.encode()
            b.sync()
            for w, ref in zip(worlds, solo):
                out = [w.read_cascade(c, hits=True) for c in range(w.rc.cascades)] + [(w.read_sh(), None)]
                for (x, hx), (y, hy) in zip(out, ref):
                    assert np.array_equal(x.view(np.uint32), y.view(np.uint32))
                    assert hx is None or np.array_equal(hx, hy)
    for w in worlds:
        w.destroy()
