"""GPU parity of textured and alpha-tested surfaces (surfaceAt, src/core/voxel.ts:325-336; shadeSample,
src/core/dynlayer.ts:451-452) through rtgi_world_set_surface: the HIP path against the oracle on the same textured
scenes. Bars as for the untextured path: occupancy, representatives, packed attributes (albedo sampled from the mips
built on the device), counters, hit ids and dynamic-layer words bit-exact; FP32 radiance and SH within 1e-4
relative (atol 1e-6)."""
import dataclasses

import numpy as np
import pytest

import rtgi
from conftest import textured
from oracle.oracle import DynState, History, Oracle
from test_gpu_batch import assert_same, snapshot
from test_gpu_dynamic import atrium_frames, check_layer
from test_gpu_parity import check_cascades, check_clipmap

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def run(scene, vox, rc, frame):
    w = rtgi.World(vox, dataclasses.replace(rc, recordHits=True))
    w.upload_scene(scene)
    w.set_frame(frame)
    w.encode()
    w.sync()
    return w


@pytest.mark.parametrize("size", [64, 1])
def test_textured_atrium_matches_oracle(require_gpu, size):
    """The atrium with three texture layers, alpha-masked bars on its x < 0 half and toon materials: clipmap,
    cascades and SH equal the oracle's; the masks remove voxels there."""
    scene, vox, rc, frame = rtgi.config("atrium")
    ts = textured(scene, size=size)
    say(f"case: textured atrium, texture size {size}")
    w = run(ts, vox, rc, frame)
    try:
        o = Oracle(vox, rc, frame)
        o.build(ts)
        check_clipmap(w, o, vox)
        check_cascades(w, o, rc)
        plain = Oracle(vox, rc, frame)
        plain.build(scene)
        if size > 1:
            assert o.counters()["voxels"] < plain.counters()["voxels"]
    finally:
        w.destroy()


def test_surface_off_and_on_again(require_gpu):
    """set_surface(None) gives the untextured build bit for bit; setting the surface again gives the textured one;
    a new scene upload drops the surface. (No SH history: each update is compared with another world's first.)"""
    scene, vox, rc, frame = rtgi.config("atrium", history=False)
    ts = textured(scene)
    say("case: surface on, off, on, dropped by a new upload")
    w = run(ts, vox, rc, frame)
    plain = run(scene, vox, rc, frame)
    try:
        tex = snapshot(w)
        w.set_surface(None)
        w.encode()
        w.sync()
        assert_same(snapshot(w), snapshot(plain), "surface off")
        w.set_surface(ts.surface)
        w.encode()
        w.sync()
        assert_same(snapshot(w), tex, "surface on again")
        w.upload_scene(scene)
        w.encode()
        w.sync()
        assert_same(snapshot(w), snapshot(plain), "new upload")
    finally:
        w.destroy()
        plain.destroy()


def test_textured_and_plain_worlds_in_one_batch(require_gpu):
    """A batch holding textured and untextured worlds (the textured instantiation of the voxel passes for all of
    them) equals their solo updates; turning a world's surface on after the batch was made re-plans it. (No SH
    history: the world whose surface changes is compared with a new world's first update.)"""
    scene, vox, rc, frame = rtgi.config("atrium", history=False)
    rng = np.random.default_rng(2)
    specs = [(textured(scene, seed=k), frame) if k % 2 == 0 else (scene, dataclasses.replace(
        frame, eye=tuple(np.asarray(frame.eye) + rng.uniform(-2, 2, 3)))) for k in range(4)]
    say("case: 2 textured + 2 plain atrium worlds in one graph batch")
    worlds = [run(sc, vox, rc, f) for sc, f in specs]
    try:
        solo = [snapshot(w) for w in worlds]
        with rtgi.Batch(worlds, graph=True) as b:
            for rep in range(2):
                b.encode()
                b.sync()
                for i, w in enumerate(worlds):
                    assert_same(snapshot(w), solo[i], f"replay {rep} world {i}")
            worlds[1].set_surface(textured(scene, seed=9).surface)
            b.encode()
            b.sync()
            got = snapshot(worlds[1])
        ref = run(dataclasses.replace(scene, surface=textured(scene, seed=9).surface), vox, rc, specs[1][1])
        assert_same(got, snapshot(ref), "surface set after the batch was made")
        ref.destroy()
    finally:
        for w in worlds:
            w.destroy()


def test_textured_dynamic_layer_matches_oracle(require_gpu):
    """The atrium's moving boxes in the dynamic layer with textured, partly masked materials: every layer word equals
    the oracle's over two updates, and the cascades that read the layer too."""
    vox, rc, frames = atrium_frames(n=2)
    first = frames[0][0].firstDynamicTri
    rcf = dataclasses.replace(rc, recordHits=True)
    surf = textured(frames[0][0]).surface
    # mask half of the boxes' materials whatever their position, so that the layer sees the alpha test
    nm = frames[0][0].materials.shape[0]
    boxes = np.unique(frames[0][0].tri[first:, 3])
    flags = surf.flags.copy()
    flags[boxes[::2]] |= rtgi.SURF_ALPHA_MASK
    layer = surf.layer.copy()
    layer[boxes[::2]] = 2
    surf = dataclasses.replace(surf, flags=flags, layer=layer)
    assert nm == len(flags)
    w = rtgi.World(vox, rcf)
    try:
        w.upload_scene(dataclasses.replace(frames[0][0], surface=surf))
        w.set_dynamic(first, 512)
        d = DynState(vox, 512, first)
        hist = History(rc.probes0)
        for k, (scene, f) in enumerate(frames):
            say(f"case: textured dynamic layer, atrium, update {k + 1}")
            if k:
                w.update_vertices(0, scene.pos, scene.nrm)
            w.set_frame(f)
            w.encode()
            w.sync()
            o = Oracle(vox, rcf, f)
            o.build(dataclasses.replace(scene, surface=surf), dyn=d)
            assert check_layer(w, d) > 0
            check_cascades(w, o, rc, hist=hist)
    finally:
        w.destroy()


def test_set_surface_refuses_bad_input(require_gpu):
    scene, vox, rc, frame = rtgi.config("atrium")
    ts = textured(scene)
    w = rtgi.World(vox, rc)
    try:
        with pytest.raises(rtgi.RtgiError):  # no scene yet
            w.set_surface(ts.surface)
        w.upload_scene(scene)
        s = ts.surface
        bad = [dataclasses.replace(s, images=s.images[:, :48, :48]),                  # not a power of two
               dataclasses.replace(s, uvs=s.uvs[:-1]),                                  # vertex count
               dataclasses.replace(s, layer=np.full_like(s.layer, 4)),                  # layer past the array
               dataclasses.replace(s, layer=s.layer[:-1], flags=s.flags[:-1], cutoff=s.cutoff[:-1])]  # materials
        for b in bad:
            with pytest.raises(rtgi.RtgiError):
                w.set_surface(b)
        w.set_surface(s)
    finally:
        w.destroy()
