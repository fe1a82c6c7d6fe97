"""The dynamic coverage layer (src/core/dynlayer.ts) on the GPU against the oracle. The atrium's 16 moving boxes (the
tail of its triangle table) go to the layer instead of the bricks; over updates at successive scene frames (the
boxes move, the eye too) the layer's words are compared bit for bit: header (counts, per-level slot ranges, bounds,
frame counter), brick pointers, slot owners, masks and the 13 words of every allocated brick (level 0 fixed-point
sums, coarser levels float prefilters). The cascades, first-hit ids and SH that read the layer then meet the plain
bars (ids exact, radiance within 1e-4; FP16 exact). A small hand-made world with one dynamic square and a pool too
small for it covers the dropped-brick path; the batched, graph-replayed path gives the single-world path's layer and
SH bit for bit."""
import dataclasses

import numpy as np
import pytest

import rtgi
from oracle.oracle import DynState, History, Oracle
from test_oracle import dyn_world

pytestmark = pytest.mark.gpu
RTOL, ATOL = 1e-4, 1e-6


def say(msg):
    print(msg, flush=True)


def check_layer(w, d):
    got = w.read_dynamic()
    n = int(d.hdr[8])
    assert np.array_equal(got["hdr"], d.hdr)
    assert np.array_equal(got["ptr"], d.ptr)
    assert np.array_equal(got["meta"][:n], d.meta[:n])
    assert np.array_equal(got["mask"][:n], d.mask[:n])
    assert np.array_equal(got["data"][:n], d.data[:n])
    c = w.read_dyn_counters()
    assert c == dict(bricks=n, brickOverflow=int(d.hdr[1]), outside=int(d.hdr[9]))
    return n


def check_update(w, o, rc, fp16, hist=None):
    """The update's cascades, hit ids and SH against the oracle's; hist: the oracle's c0 SH history of the world's
    earlier updates (advanced here)."""
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


def atrium_frames(n=3, step=7):
    """The atrium at scene frames 0, step, 2 step, ... with the eye moving by 0.6 m a frame along x."""
    _, vox, rc, f0 = rtgi.config("atrium")
    out = []
    for k in range(n):
        scene = rtgi.generate_scene("atrium", seed=1, frame=k * step)
        out.append((scene, dataclasses.replace(f0, eye=(f0.eye[0] + 0.6 * k, f0.eye[1], f0.eye[2]))))
    return vox, rc, out


@pytest.mark.parametrize("fp16,reloc,sun", [(False, False, False), (True, True, True)])
def test_dynamic_layer_atrium_matches_oracle(require_gpu, fp16, reloc, sun):
    vox, rc, frames = atrium_frames()
    first = frames[0][0].firstDynamicTri
    assert 0 < first < frames[0][0].ntris
    rcf = dataclasses.replace(rc, fp16=fp16, relocate=reloc, sunTrace=sun, recordHits=True)
    w = rtgi.World(vox, rcf)
    try:
        w.upload_scene(frames[0][0])
        w.set_dynamic(first, 512)
        d = DynState(vox, 512, first)
        hist = History(rc.probes0)
        moved = []
        for k, (scene, f) in enumerate(frames):
            say(f"case: dynamic layer, atrium, fp16={fp16}, relocate={reloc}, sunTrace={sun}, update {k + 1}")
            if k:
                w.update_vertices(0, scene.pos, scene.nrm)
            w.set_frame(f)
            w.encode()
            w.sync()
            o = Oracle(vox, rcf, f)
            o.build(scene, dyn=d)
            n = check_layer(w, d)
            assert n > 0 and d.hdr[1] == 0
            moved.append(d.data[:n].copy())
            check_update(w, o, rc, fp16, hist)
        assert not np.array_equal(moved[0], moved[-1])
    finally:
        w.destroy()


def test_dynamic_layer_closed_form_world_and_overflow(require_gpu):
    """The hand-made world of one dynamic square (area 1 m^2, 12 bricks over 3 levels): the layer, the rays that see
    it and the cascades match the oracle, with a pool of 64 slots and with 6 (coarse-first allocation, level 0
    dropped), and after the square moves 2 m."""
    scene, vox, rc, frame = dyn_world()
    for cap in (64, 6):
        rcf = dataclasses.replace(rc, recordHits=True)
        w = rtgi.World(vox, rcf)
        try:
            w.upload_scene(scene)
            w.set_dynamic(scene.firstDynamicTri, cap)
            d = DynState(vox, cap, scene.firstDynamicTri)
            hist = History(rc.probes0)
            for k, zw in enumerate((0.0, 2.0)):
                say(f"case: dynamic layer, one square, pool {cap}, z = {zw}")
                sc = dyn_world(zw=zw)[0]
                if k:
                    w.update_vertices(0, sc.pos, sc.nrm)
                w.set_frame(frame)
                w.encode()
                w.sync()
                o = Oracle(vox, rcf, frame)
                o.build(sc, dyn=d)
                n = check_layer(w, d)
                assert n == min(cap, 12 if zw == 0 else 10)
                check_update(w, o, rc, False, hist)
                org = np.array([[0.6125, 0.05 * j - 0.3, zw + 0.02 * j - 0.2] for j in range(16)], np.float32)
                dirs = np.tile(np.array([[-1.0, 0.0, 0.0]], np.float32), (16, 1))
                gr, gh = w.trace_rays(0, org, dirs)
                orr, oh = o.trace_rays(0, org, dirs)
                assert np.array_equal(gh, oh)
                np.testing.assert_allclose(gr, orr, rtol=RTOL, atol=ATOL)
                if cap == 64:
                    assert np.all(gr[:, 3] < 0.5)
        finally:
            w.destroy()


def test_dynamic_batch_matches_single(require_gpu):
    vox, rc, frames = atrium_frames(n=4, step=5)
    first = frames[0][0].firstDynamicTri
    rcf = dataclasses.replace(rc, fp16=True, relocate=True, sunTrace=True, bounce=True)
    say("case: dynamic layer, atrium, single world vs batch of 2 (graph), 4 updates")
    single = rtgi.World(vox, rcf)
    pair = [rtgi.World(vox, rcf) for _ in range(2)]
    for w in [single] + pair:
        w.upload_scene(frames[0][0])
        w.set_dynamic(first, 512)
    b = rtgi.Batch(pair, graph=True)
    try:
        for k, (scene, f) in enumerate(frames):
            for w in [single] + pair:
                if k:
                    w.update_vertices(0, scene.pos, scene.nrm)
                w.set_frame(f)
            single.encode()
            single.sync()
            b.encode()
            b.sync()
            ref, lay = single.read_sh(), single.read_dynamic()
            n = int(lay["hdr"][8])
            for w in pair:
                assert np.array_equal(w.read_sh(), ref)
                got = w.read_dynamic()
                assert np.array_equal(got["hdr"], lay["hdr"]) and np.array_equal(got["ptr"], lay["ptr"])
                assert np.array_equal(got["data"][:n], lay["data"][:n])
    finally:
        b.destroy()
        for w in [single] + pair:
            w.destroy()


def test_dynamic_and_plain_worlds_in_one_batch(require_gpu):
    """A batch holding a world with the layer and one without runs the layer-reading trace for both: the plain
    world's update equals its solo update bit for bit, and so does the layer world's."""
    vox, rc, frames = atrium_frames(n=2, step=9)
    first = frames[0][0].firstDynamicTri
    say("case: dynamic layer, atrium, batch of a layer world and a plain world vs solo, 2 updates")
    solo_dyn, solo_plain = rtgi.World(vox, rc), rtgi.World(vox, rc)
    b_dyn, b_plain = rtgi.World(vox, rc), rtgi.World(vox, rc)
    for w in (solo_dyn, solo_plain, b_dyn, b_plain):
        w.upload_scene(frames[0][0])
    solo_dyn.set_dynamic(first, 512)
    b_dyn.set_dynamic(first, 512)
    b = rtgi.Batch([b_plain, b_dyn], graph=True)
    try:
        for k, (scene, f) in enumerate(frames):
            for w in (solo_dyn, solo_plain, b_dyn, b_plain):
                if k:
                    w.update_vertices(0, scene.pos, scene.nrm)
                w.set_frame(f)
            for w in (solo_dyn, solo_plain):
                w.encode()
                w.sync()
            b.encode()
            b.sync()
            assert np.array_equal(b_plain.read_sh(), solo_plain.read_sh())
            assert np.array_equal(b_dyn.read_sh(), solo_dyn.read_sh())
            assert not np.array_equal(solo_dyn.read_sh(), solo_plain.read_sh())
            assert b_plain.read_counters() == solo_plain.read_counters()
    finally:
        b.destroy()
        for w in (solo_dyn, solo_plain, b_dyn, b_plain):
            w.destroy()


def test_dynamic_layer_reference_default_voxel_params(require_gpu):
    """The layer at the reference's DEFAULT_VOX (5 levels of 256^3): its wanted-brick bitset (5 x 32^3 bits) is too
    large for the mark kernel's LDS copy, so the marks go straight to the layer's bitset. Words and cascades as above
    over two updates of the moving atrium."""
    _, _, frames = atrium_frames(n=2, step=11)
    vox = rtgi.DEFAULT_VOX
    rc = rtgi.RcParams(cascades=4, probes0=32, s0=0.625, recordHits=True)
    first = frames[0][0].firstDynamicTri
    w = rtgi.World(vox, rc)
    try:
        w.upload_scene(frames[0][0])
        w.set_dynamic(first, 512)
        d = DynState(vox, 512, first)
        hist = History(rc.probes0)
        for k, (scene, f) in enumerate(frames):
            say(f"case: dynamic layer, atrium at the default voxel parameters, update {k + 1}")
            if k:
                w.update_vertices(0, scene.pos, scene.nrm)
            w.set_frame(f)
            w.encode()
            w.sync()
            o = Oracle(vox, rc, f)
            o.build(scene, dyn=d)
            assert check_layer(w, d) > 0
            check_update(w, o, rc, False, hist)
    finally:
        w.destroy()


def test_set_dynamic_and_tuning_after_batch_creation(require_gpu):
    """A batch re-plans when a world's dynamic layer or tuning changes after the batch was made: the layer is then
    built and read in the batched update too, equal to the single-world update."""
    vox, rc, frames = atrium_frames(n=1)
    scene, f = frames[0]
    first = scene.firstDynamicTri
    say("case: dynamic layer and tuning set after the batch was created (graph)")
    single, bw = rtgi.World(vox, rc), rtgi.World(vox, rc)
    for w in (single, bw):
        w.upload_scene(scene)
        w.set_frame(f)
    b = rtgi.Batch([bw], graph=True)
    try:
        # Both worlds run the same sequence of updates (the second blends with the first's SH history).
        single.encode()
        single.sync()
        b.encode()
        b.sync()
        plain = bw.read_sh()
        assert np.array_equal(plain, single.read_sh())
        for w in (single, bw):
            w.set_dynamic(first, 512)
        bw.set_tuning(fused=0)
        single.encode()
        single.sync()
        b.encode()
        b.sync()
        assert np.array_equal(bw.read_sh(), single.read_sh())
        assert not np.array_equal(bw.read_sh(), plain)
        assert bw.read_dyn_counters() == single.read_dyn_counters()
    finally:
        b.destroy()
        for w in (single, bw):
            w.destroy()


@pytest.mark.parametrize("fp16", [False, True])
def test_objects_token_matches_oracle(require_gpu, fp16):
    """Token obj (rc.ts:79-81) with the bounce volume: the layer also splats albedo (words 10-12), receives the
    previous volume's light (dynInject, dynlayer.ts:166-190, with dynE, bounce.ts:267-298) and the volume's
    irradiance rays read it (bnTraceDyn, bounce.ts:173-178). Over three updates of the moving atrium: the layer's 13
    words and the bounce volume bit for bit, the cascades and SH to the plain bars. The FP16 case also runs a batch
    of two worlds through a graph and meets the single world bit for bit."""
    vox, rc, frames = atrium_frames()
    first = frames[0][0].firstDynamicTri
    rcf = dataclasses.replace(rc, fp16=fp16, bounce=True, objects=True, recordHits=True)
    worlds = [rtgi.World(vox, rcf) for _ in range(3 if fp16 else 1)]
    for w in worlds:
        w.upload_scene(frames[0][0])
        w.set_dynamic(first, 512)
    w = worlds[0]
    b = rtgi.Batch(worlds[1:], graph=True) if fp16 else None
    d = DynState(vox, 512, first)
    hist = History(rc.probes0)
    state = None
    try:
        for k, (scene, f) in enumerate(frames):
            say(f"case: objects token, atrium, fp16={fp16}, update {k + 1}")
            for x in worlds:
                if k:
                    x.update_vertices(0, scene.pos, scene.nrm)
                x.set_frame(f)
            w.encode()
            w.sync()
            o = Oracle(vox, rcf, f)
            o.build(scene, dyn=d)
            before = d.data[:int(d.hdr[8]), :, 7:10].copy()
            state = o.bounce_update(state)
            n = check_layer(w, d)
            assert n > 0
            # from the second update on the previous volume lights the boxes: dynInject raised their radiance words
            assert k == 0 or (np.any(d.data[:n, :, 7:10] != before) and np.all(d.data[:n, :, 7:10] >= before))
            vol, org = w.read_bounce_volume()
            assert np.array_equal(org, state[1])
            assert np.array_equal(vol.view(np.uint16).reshape(-1, 6, 4), state[0])
            check_update(w, o, rc, fp16, hist)
            if b is not None:
                b.encode()
                b.sync()
                lay = w.read_dynamic()
                for x in worlds[1:]:
                    got = x.read_dynamic()
                    assert np.array_equal(got["hdr"], lay["hdr"])
                    assert np.array_equal(got["data"][:n], lay["data"][:n])
                    assert np.array_equal(x.read_bounce_volume()[0], vol)
                    assert np.array_equal(x.read_sh(), w.read_sh())
    finally:
        if b is not None:
            b.destroy()
        for x in worlds:
            x.destroy()


def test_objects_and_plain_bounce_worlds_in_one_batch(require_gpu):
    """A batch of a token-obj world, a bounce world without the token and a token-obj world without a layer (the
    token then does nothing): the inject launch has blocks for the first world only, and every world's update equals
    its solo update bit for bit (SH, bounce volume, layer)."""
    vox, rc, frames = atrium_frames(n=3, step=6)
    first = frames[0][0].firstDynamicTri
    rb = dataclasses.replace(rc, fp16=True, bounce=True)
    ro = dataclasses.replace(rb, objects=True)
    say("case: objects token, batch of [obj + layer, bounce + layer, obj without layer] vs solo, 3 updates")
    kinds = [(ro, True), (rb, True), (ro, False)]
    solo = [rtgi.World(vox, r) for r, _ in kinds]
    inb = [rtgi.World(vox, r) for r, _ in kinds]
    for ws in (solo, inb):
        for w, (_, layer) in zip(ws, kinds):
            w.upload_scene(frames[0][0])
            if layer:
                w.set_dynamic(first, 512)
    b = rtgi.Batch(inb, graph=True)
    try:
        for k, (scene, f) in enumerate(frames):
            for w in solo + inb:
                if k:
                    w.update_vertices(0, scene.pos, scene.nrm)
                w.set_frame(f)
            for w in solo:
                w.encode()
                w.sync()
            b.encode()
            b.sync()
            for s, w, (_, layer) in zip(solo, inb, kinds):
                assert np.array_equal(w.read_sh(), s.read_sh())
                assert np.array_equal(w.read_bounce_volume()[0], s.read_bounce_volume()[0])
                if layer:
                    got, ref = w.read_dynamic(), s.read_dynamic()
                    n = int(ref["hdr"][8])
                    assert np.array_equal(got["hdr"], ref["hdr"]) and np.array_equal(got["data"][:n], ref["data"][:n])
            if k:
                # the token changes the lit layer world's update, not the one without a layer
                assert not np.array_equal(solo[0].read_sh(), solo[1].read_sh())
    finally:
        b.destroy()
        for w in solo + inb:
            w.destroy()
