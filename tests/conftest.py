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

