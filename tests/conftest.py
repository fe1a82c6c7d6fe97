import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "real-time-gi-lab_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def gpu_available():
    import rtgi
    try:
        return rtgi.device_count() > 0
    except OSError:
        return False


@pytest.fixture(scope="session")
def require_gpu():
    import rtgi
    n = rtgi.device_count()
    assert n > 0, "no HIP device visible: the GPU tests need one (the library has no CPU fallback)"
    return n


def random_scene(rng, ntris, lo, hi, size, nmat=None):
    """Random triangles (each with its own vertices) inside [lo, hi], edge scale `size`."""
    import rtgi
    c = rng.uniform(lo, hi, (ntris, 1, 3))
    v = (c + rng.normal(0.0, size, (ntris, 3, 3))).astype(np.float32)
    pos = v.reshape(-1, 3)
    nrm = rng.normal(0, 1, (ntris * 3, 3)).astype(np.float32)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    tri = np.zeros((ntris, 4), np.uint32)
    tri[:, 0] = np.arange(ntris) * 3
    tri[:, 1] = tri[:, 0] + 1
    tri[:, 2] = tri[:, 0] + 2
    nm = nmat or ntris
    tri[:, 3] = np.arange(ntris) % nm
    mat = np.ones((nm, 4), np.float32)
    mat[:, :3] = rng.uniform(0.1, 0.9, (nm, 3))
    return rtgi.Scene(pos=pos, nrm=nrm.astype(np.float32), tri=tri, materials=mat)


def surface_points(scene, n, seed):
    """n points on random triangles of the scene (uniform barycentrics) with unit face normals: a stand-in for the
    G-buffer pixels the reference's gather reads."""
    rng = np.random.default_rng(seed)
    t = rng.integers(0, scene.tri.shape[0], n)
    tri = scene.tri[t, :3].astype(np.int64)
    a, b, c = scene.pos[tri[:, 0]], scene.pos[tri[:, 1]], scene.pos[tri[:, 2]]
    u = rng.uniform(0, 1, (n, 2))
    flip = u.sum(1) > 1
    u[flip] = 1 - u[flip]
    p = a + (b - a) * u[:, :1] + (c - a) * u[:, 1:]
    nrm = np.cross(b - a, c - a).astype(np.float64)
    ln = np.linalg.norm(nrm, axis=1, keepdims=True)
    ok = ln[:, 0] > 1e-12
    nrm = np.where(ok[:, None], nrm / np.maximum(ln, 1e-12), np.array([0.0, 1.0, 0.0]))
    return p.astype(np.float32), nrm.astype(np.float32)


def textured(scene, size=64, seed=0):
    """The scene with a textured surface: UVs by planar projection along each vertex normal's dominant axis (a 4 m
    tile), three seeded base-colour layers of size x size texels (a colour checker; two bars, alpha 0 and 255; a
    smooth ramp with a gradient alpha) and per-material words: the materials of triangles
    whose first vertex lies at x < 0 are alpha-masked (cutoff 0.5) bars (layer 2), the others sample layer m % 4
    (0 the white default); toon when m % 5 == 0."""
    import dataclasses

    import rtgi
    rng = np.random.default_rng(seed)
    pos = np.asarray(scene.pos, np.float64)
    ax = np.argmax(np.abs(np.asarray(scene.nrm, np.float64)), axis=1)
    uv = np.zeros((pos.shape[0], 2), np.float32)
    for a in range(3):
        m = ax == a
        o = [k for k in range(3) if k != a]
        uv[m] = (pos[m][:, o] / 4.0).astype(np.float32)
    y, x = np.mgrid[0:size, 0:size]
    half, span = max(size // 2, 1), max(size - 1, 1)
    img = np.zeros((3, size, size, 4), np.uint8)
    pal = rng.integers(0, 256, (4, 3))
    cell = ((x // 8) + (y // 8)) % 4
    img[0, ..., :3] = pal[cell]
    img[0, ..., 3] = 255
    img[1, ..., :3] = rng.integers(0, 256, 3)
    img[1, ..., 3] = np.where((x // half) % 2 == 0, 0, 255)
    img[2, ..., 0] = (x * 255 // span).astype(np.uint8)
    img[2, ..., 1] = (y * 255 // span).astype(np.uint8)
    img[2, ..., 2] = 128
    img[2, ..., 3] = ((x + y) * 255 // (2 * span)).astype(np.uint8)
    nm = scene.materials.shape[0]
    mi = np.arange(nm)
    masked = np.zeros(nm, bool)
    tri = np.asarray(scene.tri)
    masked[tri[pos[tri[:, 0], 0] < 0, 3]] = True
    surf = rtgi.Surface(uvs=uv, images=img, layer=np.where(masked, 2, mi % 4).astype(np.uint32),
                        flags=np.where(masked, rtgi.SURF_ALPHA_MASK, 0).astype(np.uint32)
                        | np.where(mi % 5 == 0, rtgi.SURF_TOON, 0).astype(np.uint32),
                        cutoff=np.full(nm, 0.5, np.float32))
    return dataclasses.replace(scene, surface=surf)
