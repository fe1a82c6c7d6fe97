"""The allocation-free gather (rtgi_world_gather_noalloc, staging in a fixed device array) against rtgi_world_gather
through the C ABI: the same irradiance bit for bit on two calls in a row, the second with fewer points (the array
keeps the first call's contents, which must not leak into the second), with and without occlusion; a call with more
points than the array holds is refused before any device work."""
import dataclasses

import numpy as np
import pytest

import rtgi
from conftest import surface_points
from rtgi import _lib

pytestmark = pytest.mark.gpu


def gather(fn, w, pos, nrm, flags):
    pos = np.ascontiguousarray(pos, np.float32)
    nrm = np.ascontiguousarray(nrm, np.float32)
    irr = np.full((pos.shape[0], 3), np.nan, np.float32)
    _lib.check(fn(w._h, _lib.ptr(pos), _lib.ptr(nrm), pos.shape[0], flags, _lib.ptr(irr)))
    return irr


def test_gather_noalloc_matches_gather(require_gpu):
    scene, vox, rc, frame = rtgi.config("atrium")
    print("case: gather_noalloc vs gather, atrium with relocation, 200,000 then 5,000 points", flush=True)
    L = rtgi.lib()
    with rtgi.World(vox, dataclasses.replace(rc, relocate=True)) as w:
        w.upload_scene(scene)
        w.set_frame(frame)
        w.encode()
        w.sync()
        for n, seed in ((200000, 1), (5000, 2)):
            pos, nrm = surface_points(scene, n, seed)
            for flags in (0, _lib.GATHER_OCCLUSION):
                ref = gather(L.rtgi_world_gather, w, pos, nrm, flags)
                got = gather(L.rtgi_world_gather_noalloc, w, pos, nrm, flags)
                assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), (n, flags)
                assert np.count_nonzero(np.any(ref > 0, axis=1)) > n // 2
        big = np.zeros((262145, 3), np.float32)
        with pytest.raises(rtgi.RtgiError, match="at most 262144 points"):
            gather(L.rtgi_world_gather_noalloc, w, big, big, 0)
        # the world is still usable after the refusal
        pos, nrm = surface_points(scene, 64, 3)
        assert np.array_equal(gather(L.rtgi_world_gather_noalloc, w, pos, nrm, 0),
                              gather(L.rtgi_world_gather, w, pos, nrm, 0))
