"""Gather (gatherWorld, src/core/rc.ts:871-927) on the GPU against the oracle's restatement at the same surface
points, from the same SH (the GPU update's own): irradiance within 1e-4 relative (in practice identical), with and
without the occlusion tests, plain and FP16 forms; and the wall case's closed form."""
import dataclasses

import numpy as np
import pytest

import rtgi
from conftest import random_scene, surface_points
from oracle.oracle import Oracle
from test_oracle import wall_world

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("case", ["tiny", "atrium", "city"])
def test_gather_matches_oracle(require_gpu, case):
    if case == "tiny":
        rng = np.random.default_rng(5)
        scene = random_scene(rng, 600, -4, 4, 0.6)
        vox = rtgi.VoxParams(0.25, 3, 4, 192, 1 << 16)
        rc = rtgi.RcParams(cascades=3, probes0=8, s0=0.5)
        frame = rtgi.Frame(eye=(0.2, 0.1, -0.3))
    else:
        scene, vox, rc, frame = rtgi.config(case)
    pos, nrm = surface_points(scene, 20000, seed=1)
    for fp16 in (False, True):
        say(f"case: gather, {case}, fp16={fp16}")
        rcf = dataclasses.replace(rc, fp16=fp16)
        w = rtgi.World(vox, rcf)
        w.upload_scene(scene)
        w.set_frame(frame)
        w.encode()
        w.sync()
        sh = w.read_sh()
        o = Oracle(vox, rcf, frame)
        o.build(scene)
        for occl in (False, True):
            g = w.gather(pos, nrm, occlusion=occl)
            r = o.gather(sh, pos, nrm, occlusion=occl)
            np.testing.assert_allclose(g, r, rtol=1e-4, atol=1e-6)
            assert np.count_nonzero(g) > 0
        w.destroy()


def test_gather_wall_occlusion(require_gpu):
    """Probes behind a wall: with occlusion the point sees only its side's irradiance (the field below is the
    merged field of the wall world itself, so the check is against the oracle and the near-side-only average)."""
    scene, vox, rc, frame = wall_world()
    say("case: gather behind a wall")
    w = rtgi.World(vox, rc)
    w.upload_scene(scene)
    w.set_frame(frame)
    w.encode()
    w.sync()
    sh = w.read_sh()
    o = Oracle(vox, rc, frame)
    o.build(scene)
    rng = np.random.default_rng(2)
    m = 200
    p = np.stack([np.full(m, 0.1), rng.uniform(-2, 2, m), rng.uniform(-2, 2, m)], 1).astype(np.float32)
    n = np.tile(np.array([[-1.0, 0, 0]], np.float32), (m, 1))
    g = w.gather(p, n, occlusion=True)
    np.testing.assert_allclose(g, o.gather(sh, p, n, occlusion=True), rtol=1e-4, atol=1e-6)
    # Near-side probes only (x = 0): replace every far-side probe's SH by a huge value; occluded gather unchanged.
    N = rc.probes0
    gx = np.arange(N) + np.array(o.rc.G0[:3])[0]
    far = np.tile(gx * rc.s0 > 0.0, N * N)
    sh_far = sh.copy()
    sh_far[far] = 1e6
    np.testing.assert_allclose(o.gather(sh_far, p, n, occlusion=True), g, rtol=1e-6)
    w.destroy()


def test_gather_async_on_device_arrays(require_gpu):
    """rtgi_world_gather_async on device arrays and on a stream of the caller's equals the synchronous gather bit for
    bit, with and without occlusion; a following update waits for it (the field it read is the old one)."""
    from rtgi.hip import DeviceArray, Stream

    scene, vox, rc, frame = rtgi.config("atrium")
    pos, nrm = surface_points(scene, 50000, seed=3)
    say("case: asynchronous gather on device arrays, atrium, 50,000 points")
    w = rtgi.World(vox, rc)
    try:
        w.upload_scene(scene)
        w.set_frame(frame)
        w.encode()
        w.sync()
        dp, dn = DeviceArray.from_numpy(pos), DeviceArray.from_numpy(nrm)
        s = Stream()
        for occl in (False, True):
            want = w.gather(pos, nrm, occlusion=occl)
            di = DeviceArray.from_numpy(np.full((pos.shape[0], 3), -1.0, np.float32))
            w.gather_async(dp.ptr, dn.ptr, pos.shape[0], di.ptr, occlusion=occl, stream=s.handle)
            # the next update is ordered after the gather, so the gather read this update's field
            w.update((frame.eye[0] + 3.0, frame.eye[1], frame.eye[2]))
            w.encode()
            s.synchronize()
            got = di.numpy()
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), occl
            w.update(frame.eye)  # the next round gathers from a field of the original eye again
            w.encode()
            w.sync()
        di = DeviceArray.from_numpy(np.zeros((pos.shape[0], 3), np.float32))
        w.gather_async(dp.ptr, dn.ptr, pos.shape[0], di.ptr)  # the world's own stream
        w.sync()
        assert np.count_nonzero(di.numpy()) > 0
        s.destroy()
    finally:
        w.destroy()
