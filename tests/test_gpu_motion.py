"""Moving objects posed on the device (rtgi_world_set_objects / rtgi_world_set_pose; DynamicSpec and posedVertices,
src/core/scenes.ts:41-53, 80-102, as the runner poses every frame, src/core/rcrun.ts:194-217).

A world uploads its scene once, registers the objects' rest mesh and then only sends poses: each update poses the
objects on the device before the voxel build. Against the oracle built from the scene the generator poses on the host
for the same frame (which tests/test_oracle.py pins to the oracle's own posedVertices), the clipmap is bit-exact and
the cascades and SH meet the parity bars, over several frames; batched hipGraph replays with poses changing between
them equal solo updates bit for bit."""
import dataclasses

import numpy as np
import pytest

import rtgi
from oracle.oracle import History, Oracle
from test_gpu_batch import assert_same, snapshot
from test_gpu_parity import check_cascades, check_clipmap

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("case", ["atrium", "city"])
def test_device_pose_matches_oracle(require_gpu, case):
    scene0, vox, rc, frame = rtgi.config(case)
    w = rtgi.World(vox, dataclasses.replace(rc, recordHits=True))
    try:
        w.upload_scene(scene0)
        w.set_scene_objects(scene0)
        w.set_frame(frame)
        hist = History(rc.probes0)
        for k, f in enumerate((5, 23, 24)):
            say(f"case: {case}, {len(scene0.objects)} objects posed on the device at frame {f}")
            w.set_pose(scene0.poses_at(f))
            w.encode()
            w.sync()
            sf = rtgi.generate_scene(**dict(rtgi.CONFIGS[case]["scene"], frame=f))
            o = Oracle(vox, rc, frame)
            o.build(sf)
            check_clipmap(w, o, vox)
            check_cascades(w, o, rc, hist=hist)
    finally:
        w.destroy()


def test_batched_motion_matches_solo(require_gpu):
    """Two cities and an atrium with objects plus an atrium without, in one hipGraph batch: three replays with new
    poses each, every world equal to a solo world posed the same way."""
    specs = [("city", 1), ("city", 2), ("atrium", 3), ("atrium", 4)]
    batched, solo = [], []
    for kind, seed in specs:
        sc = rtgi.generate_scene(**dict(rtgi.CONFIGS[kind]["scene"], seed=seed))
        _, vox, rc, frame = rtgi.config(kind)
        for group in (batched, solo):
            w = rtgi.World(vox, dataclasses.replace(rc, recordHits=True))
            w.upload_scene(sc)
            w.set_frame(dataclasses.replace(frame, eye=sc.eye))
            if seed != 4:
                w.set_scene_objects(sc)
            group.append((w, sc))
    b = rtgi.Batch([w for w, _ in batched], graph=True)
    try:
        for f in (2, 9, 10):
            say(f"case: batch of 4 (3 with objects), poses of frame {f}")
            for group in (batched, solo):
                for (w, sc), (_, seed) in zip(group, specs):
                    if seed != 4:
                        w.set_pose(sc.poses_at(f))
            for w, _ in solo:
                w.encode()
                w.sync()
            b.encode()
            b.sync()
            for i, ((wb, _), (ws, _)) in enumerate(zip(batched, solo)):
                assert_same(snapshot(wb), snapshot(ws), f"world {i}, frame {f}")
    finally:
        b.destroy()
        for w, _ in batched + solo:
            w.destroy()


def test_object_registration_errors(require_gpu):
    """Bad object tables are refused; a new scene upload removes the objects (a pose then has nothing to move);
    pose counts must match."""
    sc = rtgi.generate_scene("atrium", seed=1)
    _, vox, rc, frame = rtgi.config("atrium")
    w = rtgi.World(vox, rc)
    try:
        with pytest.raises(rtgi.RtgiError):  # no scene yet
            w.set_scene_objects(sc)
        w.upload_scene(sc)
        o = sc.objects.copy()
        o[1]["first_vertex"] = o[0]["first_vertex"] + 1  # overlaps object 0
        with pytest.raises(rtgi.RtgiError, match="overlaps"):
            w.set_objects(o, sc.rest_pos, sc.rest_nrm)
        o = sc.objects.copy()
        o[-1]["vertex_count"] += 1  # past the last vertex
        with pytest.raises(rtgi.RtgiError, match="out of range"):
            w.set_objects(o, sc.rest_pos, sc.rest_nrm)
        with pytest.raises(rtgi.RtgiError, match="no objects"):
            w.set_pose(np.zeros((0, 5)))
        w.set_scene_objects(sc)
        with pytest.raises(ValueError):
            w.set_pose(sc.poses_at(0)[:3])
        bad = sc.poses_at(0)
        bad[2, 0] = np.nan
        with pytest.raises(rtgi.RtgiError, match="not finite"):
            w.set_pose(bad)
        w.set_pose(sc.poses_at(1))
        w.set_frame(frame)
        w.encode()
        w.sync()
        w.upload_scene(sc)
        with pytest.raises(rtgi.RtgiError, match="no objects"):
            w.set_pose(sc.poses_at(2))
    finally:
        w.destroy()
