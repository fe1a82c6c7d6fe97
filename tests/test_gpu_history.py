"""c0 SH history (histBlend, rc.ts:784-814; on for world probes, HIST_ON = !screen, rc.ts:1071) on the GPU against the
oracle, over four updates with the eye moving: the probe grid shifts, so probes whose grid point the previous update
held keep their history and the others start from their own value; a repeated frame blends equal values. Three paths
run in lockstep on the same frames: the single-world update (trace, merge steps, SH launch), the fused single-world
update (the c0 trace+merge launch projects the SH) and a batched hipGraph of two worlds (fused and not). Each SH meets
the oracle's (FP16 bit for bit, FP32 within 1e-4 relative) and the paths meet each other bit for bit."""
import dataclasses

import numpy as np
import pytest

import rtgi
from oracle.oracle import History, Oracle

pytestmark = pytest.mark.gpu
RTOL, ATOL = 1e-4, 1e-6


def say(msg):
    print(msg, flush=True)


def moved(frame, d):
    return dataclasses.replace(frame, eye=tuple(float(a + b) for a, b in zip(frame.eye, d)))


@pytest.mark.parametrize("case,fp16", [("atrium", False), ("atrium", True), ("city", False), ("city", True)])
def test_history_matches_oracle(require_gpu, case, fp16):
    scene, vox, rc, f1 = rtgi.config(case)
    rcf = dataclasses.replace(rc, fp16=fp16)
    top = rc.s0 * 2 ** (rc.cascades - 1)  # the grid moves in steps of the top cascade's spacing
    f2 = moved(f1, (2.2 * top, 0.0, -1.3 * top))
    f4 = moved(f2, (-3.1 * top, 1.1 * top, 0.6 * top))
    frames = [f1, f2, f2, f4]
    plain, fused, b_fused, b_split = (rtgi.World(vox, rcf) for _ in range(4))
    fused.set_tuning(fused=1)
    b_split.set_tuning(fused=0)
    worlds = [plain, fused, b_fused, b_split]
    for w in worlds:
        w.upload_scene(scene)
    batches = [rtgi.Batch([b_fused], graph=True), rtgi.Batch([b_split], graph=True)]
    hist = History(rc.probes0)
    grids = []
    try:
        for k, f in enumerate(frames):
            say(f"case: c0 history, {case}, fp16={fp16}, update {k + 1}")
            for w in worlds:
                w.set_frame(f)
            for w in (plain, fused):
                w.encode()
                w.sync()
            for b in batches:
                b.encode()
                b.sync()
            grids.append(tuple(plain.layout()["grid_origin"]))
            o = Oracle(vox, rcf, f)
            o.build(scene)
            rgba, _ = o.trace()
            _, own = o.merge_sh(rgba)
            _, sh = o.merge_sh(rgba, hist=hist)
            got = plain.read_sh()
            if fp16:
                assert np.array_equal(got, sh), f"update {k + 1}: {np.count_nonzero(got != sh)} SH values differ"
            else:
                np.testing.assert_allclose(got, sh, rtol=RTOL, atol=ATOL, err_msg=f"update {k + 1}")
            for w, name in ((fused, "fused"), (b_fused, "batch fused"), (b_split, "batch")):
                assert np.array_equal(w.read_sh().view(np.uint32), got.view(np.uint32)), f"{name}, update {k + 1}"
            blended = np.any(sh != own, axis=1)
            if k in (1, 3):  # the grid moved: some probes kept a history that differs from their own value
                assert grids[k] != grids[k - 1]
                assert 0 < np.count_nonzero(blended) < sh.shape[0]
        assert hist.misses == 0
    finally:
        for b in batches:
            b.destroy()
        for w in worlds:
            w.destroy()


def test_history_off_and_repeated_frames(require_gpu):
    """history=False (RTGI_RC_NO_HISTORY) gives every update's own SH; with the history on, updates at one frame keep
    the first update's SH exactly (prev * 0.5 + new * 0.5 of equal values)."""
    scene, vox, rc, f1 = rtgi.config("atrium")
    f2 = moved(f1, (3.0, 0.0, 2.0))
    on, off = rtgi.World(vox, rc), rtgi.World(vox, dataclasses.replace(rc, history=False))
    try:
        for w in (on, off):
            w.upload_scene(scene)
        say("case: c0 history on and off, atrium, frames f1 f1 f2 f2")
        seq = []
        for f in (f1, f1, f2, f2):
            for w in (on, off):
                w.set_frame(f)
                w.encode()
                w.sync()
            seq.append((on.read_sh(), off.read_sh()))
        assert np.array_equal(seq[0][0], seq[0][1]) and np.array_equal(seq[1][0], seq[0][0])
        assert np.array_equal(seq[3][1], seq[2][1])  # no history: the frame's own SH
        assert not np.array_equal(seq[2][0], seq[2][1])  # the moved frame blends with f1's SH
        assert not np.array_equal(seq[3][0], seq[2][0])  # and moves towards its own
    finally:
        on.destroy()
        off.destroy()
