"""CPU-side checks of the boundary: the library loads, exports exactly the declared C ABI, validates arguments and
reports errors the reference's way (synchronous failure with a message), and the scene generator is deterministic."""
import ctypes as C
import dataclasses
import os
import re
import subprocess

import numpy as np
import pytest

import rtgi
from rtgi import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "rtgi.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(rtgi_[a-z0-9_]+)\s*\(", src)))


def test_header_matches_binding():
    assert declared_functions() == sorted(_lib.SIGNATURES)


def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True, check=True)
    exported = set(re.findall(r" T (rtgi_[a-z0-9_]+)$", out.stdout, flags=re.M))
    missing = set(declared_functions()) - exported
    assert not missing, missing
    for name in declared_functions():
        assert hasattr(rtgi.lib(), name)


def test_version_and_errors_without_gpu():
    assert b"gfx950" in rtgi.lib().rtgi_version()
    n = C.c_int32(-1)
    assert rtgi.lib().rtgi_device_count(C.byref(n)) == 0 and n.value >= 0
    # Invalid configurations fail synchronously with a message (voxel.ts:749 'at most 8 levels', rc.ts:1022).
    with pytest.raises(rtgi.RtgiError, match="at most 8 levels"):
        rtgi.World(rtgi.VoxParams(v0=0.25, levels=9, G=4, maxBricks=10, maxVoxels=10), rtgi.RcParams())
    with pytest.raises(rtgi.RtgiError, match="at most 8 cascades"):
        rtgi.World(rtgi.VoxParams(v0=0.25, levels=4, G=4, maxBricks=10, maxVoxels=10), rtgi.RcParams(cascades=9))
    with pytest.raises(rtgi.RtgiError, match="multiple of 64"):
        rtgi.World(rtgi.VoxParams(v0=0.25, levels=4, G=5, maxBricks=10, maxVoxels=10), rtgi.RcParams())
    with pytest.raises(rtgi.RtgiError, match="multiple of 2"):
        rtgi.World(rtgi.VoxParams(v0=0.25, levels=4, G=4, maxBricks=10, maxVoxels=10),
                   rtgi.RcParams(cascades=4, probes0=12))
    with pytest.raises(rtgi.RtgiError, match="null"):
        rtgi._lib.check(rtgi.lib().rtgi_world_sync(None))
    assert rtgi.lib().rtgi_last_error() != b""


def test_estimate_bytes_and_server_without_gpu():
    """The byte estimate is pure host arithmetic (the admission test of the server's budget); creating a server needs
    a device and says so."""
    vox, rc = rtgi.CONFIGS["city"]["vox"], rtgi.CONFIGS["city"]["rc"]
    b0 = rtgi.estimate_bytes(vox, rc, 0, 0, 0)
    b1 = rtgi.estimate_bytes(vox, rc, 1000, 2000, 2000)
    # Scene buffers: 2 x 12 B per vertex, 16 B + 2 x 4 levels x 4 B per triangle, 16 B per material.
    assert b1 - b0 == 1000 * 24 + 2000 * (16 + 2 * 4 * 4) + 2000 * 16
    # Cascades: 8,126,464 rays of f32 rgba; f16 halves them, and the SH planes (two with the c0 history: this
    # update's and the previous one's).
    f16 = rtgi.estimate_bytes(vox, dataclasses.replace(rc, fp16=True), 0, 0, 0)
    assert b0 - f16 == 8126464 * 8 + 2 * 64 ** 3 * 27 * 2
    nohist = rtgi.estimate_bytes(vox, dataclasses.replace(rc, history=False), 0, 0, 0)
    assert b0 - nohist == 64 ** 3 * 27 * 4
    with pytest.raises(rtgi.RtgiError, match="at most 8 levels"):
        rtgi.estimate_bytes(dataclasses.replace(vox, levels=9), rc, 0, 0, 0)
    # A cascade's rays are indexed in 32 bits: 1024^3 c0 probes x 16 directions is refused before any allocation.
    with pytest.raises(rtgi.RtgiError, match="below 2\\^32"):
        rtgi.estimate_bytes(vox, dataclasses.replace(rc, probes0=1024, fp16=True), 0, 0, 0)
    # The merge addresses a cascade's 2x2 averages in 32 bits: 512^3 c0 probes (2^31 rays) give c1 256^3 x 16
    # averages of 16 B = 4 GiB, refused; 256^3 c0 probes are accepted.
    with pytest.raises(rtgi.RtgiError, match="averages exceed 4 GiB"):
        rtgi.estimate_bytes(vox, dataclasses.replace(rc, probes0=512, fp16=True), 0, 0, 0)
    assert rtgi.estimate_bytes(vox, dataclasses.replace(rc, probes0=256, fp16=True), 0, 0, 0) > 0
    with pytest.raises(rtgi.RtgiError, match="no HIP device"):
        rtgi.Server()


def test_vox_params_rejects_unknown_keys():
    assert rtgi.vox_params(G=16).G == 16
    with pytest.raises(ValueError, match="unknown voxel parameter"):
        rtgi.vox_params(bogus=1)


@pytest.mark.parametrize("kind,lo,hi", [("atrium", 45_000, 65_000), ("city", 450_000, 560_000)])
def test_scene_generator(kind, lo, hi):
    a = rtgi.generate_scene(kind, seed=3, frame=2)
    b = rtgi.generate_scene(kind, seed=3, frame=2)
    assert lo <= a.ntris <= hi
    for k in ("pos", "nrm", "tri", "materials"):
        assert np.array_equal(getattr(a, k), getattr(b, k))
    assert a.tri[:, :3].max() < a.pos.shape[0]
    assert np.all(a.tri[:, 3] == np.arange(a.ntris))
    alb = a.materials[:, :3]
    assert alb.min() >= 0.1 and alb.max() <= 0.9
    c = rtgi.generate_scene(kind, seed=3, frame=5)
    assert not np.array_equal(a.pos, c.pos)  # moving parts are posed per frame
    assert np.array_equal(a.tri, c.tri)


def test_irradiance_dump_layout(tmp_path):
    """The .irr.f32 dump (packIrradiance, src/runner/report.ts:54-59): 16 bytes per pixel, r g b tag as little-endian
    float32, in pixel order; read back as the reference's comparator unpacks it (tag > 0 valid)."""
    rgb = np.array([[0.5, 1.25, -0.0], [3.0, 0.0, 7.5], [1e-3, 2e-3, 3e-3]], np.float32)
    tag = np.array([1, 0, 2], np.uint8)
    f = tmp_path / "city.static.irr.f32"
    rtgi.write_irradiance(f, rgb, tag)
    raw = f.read_bytes()
    assert len(raw) == 3 * 16
    import struct
    assert struct.unpack("<4f", raw[:16]) == (0.5, 1.25, -0.0, 1.0)
    assert struct.unpack("<4f", raw[16:32]) == (3.0, 0.0, 7.5, 0.0)
    got_rgb, got_tag, valid = rtgi.read_irradiance(f)
    assert np.array_equal(got_rgb, rgb) and np.array_equal(got_tag, tag)
    assert valid.tolist() == [True, False, True]
    with pytest.raises(ValueError):
        rtgi.pack_irradiance(rgb, tag[:2])
