"""The Node-API addon (the drop-in binding for the reference's TypeScript host) over the same C ABI."""
import dataclasses
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import rtgi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ADDON = os.path.join(ROOT, "real-time-gi-lab_amd", "napi", "rtgi_napi.node")
NODE = shutil.which("node")
pytestmark = pytest.mark.skipif(NODE is None, reason="node is not installed")


def node(script, *args, timeout=120):
    return subprocess.run([NODE, "-e", script, *args] if not script.endswith(".js") else [NODE, script, *args],
                          capture_output=True, text=True, timeout=timeout, cwd=ROOT)


def test_addon_loads_and_throws_like_the_reference():
    assert os.path.exists(ADDON), "build the addon (make -C real-time-gi-lab_amd/napi)"
    r = node("const r=require('./real-time-gi-lab_amd/napi/rtgi.js');"
             "console.log(JSON.stringify({v:r.version(),n:r.deviceCount()}));"
             "try{r.voxParams({bogus:1})}catch(e){console.log(e.message)}"
             "try{r.addon.createWorld({levels:9},{})}catch(e){console.log(e.message)}")
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()
    assert "gfx950" in json.loads(lines[0])["v"]
    assert lines[1] == "unknown voxel parameter bogus"
    assert "at most 8 levels" in lines[2]


def test_addon_scene_matches_library():
    r = node("const r=require('./real-time-gi-lab_amd/napi/rtgi.js');const s=r.generateScene('atrium',3,2);"
             "let t=0;for(const x of s.tri)t=(t*31+x)>>>0;console.log(s.tri.length/4, t, s.firstDynamicTri)")
    assert r.returncode == 0, r.stderr
    n, h, first = map(int, r.stdout.split())
    sc = rtgi.generate_scene("atrium", seed=3, frame=2)
    assert first == sc.firstDynamicTri < n
    t = 0
    for x in sc.tri.ravel().tolist():
        t = (t * 31 + x) & 0xFFFFFFFF
    assert n == sc.ntris and h == t


@pytest.mark.gpu
def test_addon_update_matches_python_path(require_gpu):
    print("case: node addon, atrium seed 1", flush=True)
    r = node(os.path.join(ROOT, "tests", "napi_world.js"), "atrium", "1", "0.125", "4", "8", "4", "32", "0.625",
             timeout=240)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout)
    scene = rtgi.generate_scene("atrium", seed=1)
    vox = rtgi.VoxParams(v0=0.125, levels=4, G=8, maxBricks=4 * 8 ** 3, maxVoxels=1 << 20)
    rc = rtgi.RcParams(cascades=4, probes0=32, s0=0.625)
    with rtgi.World(vox, rc) as w:
        w.upload_scene(scene)
        w.set_frame(rtgi.Frame(eye=scene.eye, sunDir=scene.sunDir))
        w.encode()
        c = w.read_counters()
        sh = w.read_sh()
    for k in ("bricks", "voxels", "brickOverflow", "voxelOverflow", "bricksPerLevel"):
        assert out["counters"][k] == c[k]
    np.testing.assert_array_equal(np.float32(out["sh0"]), sh[0])
    assert abs(out["shSum"] - float(sh.astype(np.float64).sum())) <= 1e-6 * abs(float(sh.sum()))


@pytest.mark.gpu
def test_addon_textured_scene_matches_python_path(require_gpu):
    """setSurface through the addon (the textured scene built in JS) equals the Python path's rtgi_world_set_surface
    with the same UVs, texels and material words."""
    print("case: node addon, textured atrium seed 1", flush=True)
    r = node(os.path.join(ROOT, "tests", "napi_world.js"), "atrium", "1", "0.125", "4", "8", "4", "32", "0.625",
             "textured", timeout=240)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout)
    scene = rtgi.generate_scene("atrium", seed=1)
    ax = np.argmax(np.abs(scene.nrm), axis=1)
    uv = np.zeros((scene.pos.shape[0], 2), np.float32)
    for a in range(3):
        m = ax == a
        uv[m] = scene.pos[m][:, [k for k in range(3) if k != a]] / np.float32(4)
    y, x = np.mgrid[0:64, 0:64]
    img = np.stack([x * 4, y * 4, np.full_like(x, 128), np.where(x < 32, 0, 255)], -1).astype(np.uint8)[None]
    nm = scene.materials.shape[0]
    flags = rtgi.SURF_ALPHA_MASK | np.where(np.arange(nm) % 2 == 1, rtgi.SURF_TOON, 0)
    surf = rtgi.Surface(uvs=uv, images=img, layer=np.ones(nm, np.uint32), flags=flags.astype(np.uint32),
                        cutoff=np.full(nm, 0.5, np.float32))
    vox = rtgi.VoxParams(v0=0.125, levels=4, G=8, maxBricks=4 * 8 ** 3, maxVoxels=1 << 20)
    rc = rtgi.RcParams(cascades=4, probes0=32, s0=0.625)
    with rtgi.World(vox, rc) as w:
        w.upload_scene(dataclasses.replace(scene, surface=surf))
        w.set_frame(rtgi.Frame(eye=scene.eye, sunDir=scene.sunDir))
        w.encode()
        c = w.read_counters()
        sh = w.read_sh()
        w.set_surface(None)
        w.encode()
        plain = w.read_counters()
    assert c["voxels"] < plain["voxels"]  # the transparent half of the layer empties voxels
    for k in ("bricks", "voxels", "brickOverflow", "voxelOverflow", "bricksPerLevel"):
        assert out["counters"][k] == c[k]
    np.testing.assert_array_equal(np.float32(out["sh0"]), sh[0])
    assert abs(out["shSum"] - float(sh.astype(np.float64).sum())) <= 1e-6 * abs(float(sh.sum()))


@pytest.mark.gpu
def test_addon_dynamic_layer_matches_python_path(require_gpu):
    """createVoxWorld with firstDynamicTri (the atrium's 16 boxes): counters, layer counters and SH equal the Python
    path's with World.set_dynamic."""
    print("case: node addon, atrium seed 4, dynamic layer", flush=True)
    r = node(os.path.join(ROOT, "tests", "napi_world.js"), "atrium", "4", "0.125", "4", "8", "4", "32", "0.625",
             "dynamic", timeout=240)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout)
    scene = rtgi.generate_scene("atrium", seed=4)
    vox = rtgi.VoxParams(v0=0.125, levels=4, G=8, maxBricks=4 * 8 ** 3, maxVoxels=1 << 20)
    rc = rtgi.RcParams(cascades=4, probes0=32, s0=0.625)
    with rtgi.World(vox, rc) as w:
        w.upload_scene(scene)
        w.set_dynamic(scene.firstDynamicTri, 512)
        w.set_frame(rtgi.Frame(eye=scene.eye, sunDir=scene.sunDir))
        w.encode()
        c = w.read_counters()
        dc = w.read_dyn_counters()
        sh = w.read_sh()
    assert out["counters"]["bricks"] == c["bricks"] and out["dyn"] == dc and dc["bricks"] > 0
    np.testing.assert_array_equal(np.float32(out["sh0"]), sh[0])
    assert abs(out["shSum"] - float(sh.astype(np.float64).sum())) <= 1e-6 * abs(float(sh.sum()))


@pytest.mark.gpu
def test_addon_motion_matches_python_path(require_gpu):
    """The runner's per-frame pose through the addon (dynamicSpec(scene).at(f), Math.cos / Math.sin of rotY, the
    objects posed on the device) over two frames: counters and SH equal the Python path's sent the same poses."""
    print("case: node addon, atrium seed 5, 16 objects posed at frames 3 and 11", flush=True)
    r = node(os.path.join(ROOT, "tests", "napi_world.js"), "atrium", "5", "0.125", "4", "8", "4", "32", "0.625",
             "motion", timeout=240)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout)
    scene = rtgi.generate_scene("atrium", seed=5)
    vox = rtgi.VoxParams(v0=0.125, levels=4, G=8, maxBricks=4 * 8 ** 3, maxVoxels=1 << 20)
    rc = rtgi.RcParams(cascades=4, probes0=32, s0=0.625)
    poses = [np.array(p, np.float64) for p in out["poses"]]
    assert len(poses) == 2 and poses[0].shape == (16, 5)
    # Math.cos / Math.sin and the C library agree to an ulp or two.
    np.testing.assert_allclose(poses[1], scene.poses_at(11), rtol=0, atol=1e-15)
    with rtgi.World(vox, rc) as w:
        w.upload_scene(scene)
        w.set_scene_objects(scene)
        w.set_frame(rtgi.Frame(eye=scene.eye, sunDir=scene.sunDir))
        for p in poses:
            w.set_pose(p)
            w.encode()
        c = w.read_counters()
        sh = w.read_sh()
    for k in ("bricks", "voxels", "brickOverflow", "voxelOverflow", "bricksPerLevel"):
        assert out["counters"][k] == c[k]
    np.testing.assert_array_equal(np.float32(out["sh0"]), sh[0])
    assert abs(out["shSum"] - float(sh.astype(np.float64).sum())) <= 1e-6 * abs(float(sh.sum()))


@pytest.mark.gpu
def test_addon_full_default_matches_python_path(require_gpu):
    """The addon with voxel relocation, traced sun visibility and multi-bounce over two updates: SH, gather,
    relocated probes and the bounce volume equal the Python path's."""
    print("case: node addon, atrium seed 2, relocation + sun + bounce, 2 updates", flush=True)
    r = node(os.path.join(ROOT, "tests", "napi_world.js"), "atrium", "2", "0.125", "4", "8", "4", "32", "0.625", "full",
             timeout=240)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout)
    scene = rtgi.generate_scene("atrium", seed=2)
    vox = rtgi.VoxParams(v0=0.125, levels=4, G=8, maxBricks=4 * 8 ** 3, maxVoxels=1 << 20)
    rc = rtgi.RcParams(cascades=4, probes0=32, s0=0.625, relocate=True, sunTrace=True, bounce=True)
    eye = np.asarray(scene.eye, np.float32)
    g = np.arange(64)
    pos = np.stack([eye[0] + (g % 8) - 3.5, np.full(64, eye[1] - 1), eye[2] + g // 8 - 3.5], 1).astype(np.float32)
    nrm = np.tile(np.array([[0, 1, 0]], np.float32), (64, 1))
    with rtgi.World(vox, rc) as w:
        w.upload_scene(scene)
        w.set_frame(rtgi.Frame(eye=scene.eye, sunDir=scene.sunDir))
        w.encode()
        w.encode()
        sh = w.read_sh()
        irr = w.gather(pos, nrm, occlusion=True)
        pp, pn, par = w.read_probes(0)
        vol, org = w.read_bounce_volume()
    np.testing.assert_array_equal(np.float32(out["sh0"]), sh[0])
    np.testing.assert_array_equal(np.float32(out["irr"]), irr.ravel())
    assert abs(out["probeSum"] - (float(pp.astype(np.float64).sum()) + float(pn.astype(np.float64).sum()))) <= 1e-3
    assert out["parentsSum"] == int(par.astype(np.int64).sum())
    assert out["texelSum"] == int(vol.view(np.uint16).astype(np.int64).sum())
    assert out["origin"] == org.tolist()
