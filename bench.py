"""GI-update benchmark on one MI355X.

Workload: W resident city-block worlds (about 500k triangles each, 4 x 128^3 clipmap at 0.25 m, 64^3 c0 probes,
5 cascades, FP32), seeds 1..W, whose 64 buildings each turn and move every update (--static: they stay put). One
update = the buildings posed at the next frame (poses sent per world, rest meshes posed on the device), voxel rebuild
from every triangle, probe-ray trace of all cascades, cascade merge and SH projection of one world.

Two paths are timed on the same worlds:
  plain   one world at a time: its update is enqueued, then waited for (the single-instance path);
  batched all W worlds with one launch per stage per batch, replayed as hipGraphs; the batch width (worlds per
          launch; batches run on their own streams, concurrently), probe tile and trace/merge fusion come from the
          autotuner (rtgi.autotune) unless --batch fixes the width.
A step updates every world once. The value is the batched path's aggregate GI updates per second; the latency of an
update is the time from the step's submission to the completion of the batch that holds its world (p50/p99 over all
updates of the timed steps). Inputs are resident in HBM before timing starts.

Batched mix (the "batched_mix" field; --mix-worlds 0 skips it): --mix-worlds seeded worlds of different sizes and
cascade counts (rtgi.workloads, FP16 storage) updated in batched hipGraph replays with the tuning and batch width the
autotuner (rtgi.autotune) picks, beside the same worlds' plain FP32 updates one at a time.

Serving (the "serving" field; --serve-seconds 0 skips it): a Poisson stream of update requests (--serve-rate per
second) over a resident mix of --serve-worlds seeded worlds of different sizes (rtgi.workloads), FP16 storage, through
the request server (continuous batching under a budget of 128 MiB per world); p50/p99 latency from submission to
completion, beside the same request trace run serially on the plain FP32 path.

Gather at frame rate (the "gather_1080p" field; --gather-frames 0 skips it): 1920 x 1080 surface points in device
memory gathered from world 1's field each frame (rtgi_world_gather_async), with and without the occlusion tests.

Prints one JSON line. Only --gpus 1 is supported: this path runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "real-time-gi-lab_amd"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(scene, vox, rc, frame, repeats):
    """The oracle (CPU restatement, f32) timed on the host: build + trace + merge + SH of one world."""
    from oracle.oracle import Oracle

    o = Oracle(vox, rc, frame)
    t0 = time.perf_counter()
    for _ in range(repeats):
        o.build(scene)
        rgba, _ = o.trace(threads=0)
        o.merge_sh(rgba)
    dt = time.perf_counter() - t0
    return repeats / dt, dt


def pct(x, q):
    return round(float(np.percentile(np.asarray(x) * 1e3, q)), 3)


def mix_bench(a):
    """Batched instances: a seeded mix of worlds of different sizes (FP16 storage), one launch per stage per batch,
    hipGraph replays, tile / fusion / batch width picked by the autotuner; beside it the same worlds' plain FP32
    updates one at a time."""
    import dataclasses

    import rtgi
    from rtgi.autotune import autotune
    from rtgi.workloads import world_mix

    mix = world_mix(a.mix_worlds, seed=5, tail_every=0)
    # Batches are cut from the worlds in order of shape (c0 probes, cascades, clipmap edge), so that a batch's worlds
    # need about equal block counts per launch and its launches take the XCD-contiguous order (32 mixed worlds in
    # batches of 8: 5,836 -> 5,998 updates/s against seed order).
    mix = sorted(mix, key=lambda m: (m.rc.probes0, m.rc.cascades, m.vox.G))
    worlds = []
    for m in mix:
        w = rtgi.World(m.vox, m.rc)
        w.upload_scene(m.scene)
        w.set_frame(m.frame())
        worlds.append(w)
    tune = autotune(worlds, repeats=2, log=log)
    best = tune["best"]
    width = best["width"]
    batches = [rtgi.Batch(worlds[i:i + width], graph=True) for i in range(0, len(worlds), width)]
    for _ in range(2):
        for b in batches:
            b.encode()
        for b in batches:
            b.sync()
    t = time.perf_counter()
    for _ in range(a.steps):
        for b in batches:
            b.encode()
        for b in batches:
            b.sync()
    value = a.steps * len(worlds) / (time.perf_counter() - t)
    for b in batches:
        b.destroy()
    for w in worlds:
        w.destroy()
    plain, plain_worlds = [], []
    for m in mix:
        w = rtgi.World(m.vox, dataclasses.replace(m.rc, fp16=False))
        w.upload_scene(m.scene)
        w.set_frame(m.frame())
        w.encode()
        w.sync()
        plain_worlds.append(w)
    t = time.perf_counter()
    for _ in range(2):
        for w in plain_worlds:
            w.encode()
            w.sync()
    plain_value = 2 * len(plain_worlds) / (time.perf_counter() - t)
    for w in plain_worlds:
        w.destroy()
    log(f"mix: {value:.1f} updates/s batched ({best}), plain FP32 {plain_value:.1f} updates/s")
    return {
        "workload": f"{len(mix)} seeded mixed worlds (atrium/city, 10k-500k triangles log-uniform, clipmap edge "
                    f"32-128 x 4 levels, 3-5 cascades), FP16 storage, batched hipGraph replays, batches cut in "
                    f"order of world shape",
        "updates_per_s": round(value, 2), "tuning": best, "autotune_table": tune["table"],
        "triangles_total_m": round(sum(m.scene.ntris for m in mix) / 1e6, 2),
        "plain_fp32_serial_updates_per_s": round(plain_value, 2), "speedup_vs_plain": round(value / plain_value, 3),
    }


def serve_bench(a):
    """Serving: a Poisson stream of update requests over a resident mix of worlds (10% tail of 2M-triangle blocks),
    FP16 storage, admitted under a budget of 128 MiB per world; then a prefix of the same request trace through the
    plain FP32 path, one update at a time."""
    import dataclasses

    import rtgi
    from rtgi.workloads import world_mix

    t0 = time.perf_counter()
    mix = world_mix(a.serve_worlds, seed=7)
    budget = 128 * 2 ** 20 * a.serve_worlds
    srv = rtgi.Server(budget=budget, max_batch=a.serve_batch, slots=2)
    ids = [srv.add_world(w.scene, w.vox, w.rc) for w in mix]
    rng = np.random.default_rng(11)
    for i in ids:  # warm-up: every world once (admission, first build)
        srv.submit(i, mix[i].frame(), counters=True)
    srv.drain()
    warm = srv.poll(4096)
    overflowing = sum(1 for c in warm if c["counters"]["brickOverflow"] or c["counters"]["voxelOverflow"]
                      or c["counters"]["bigOverflow"])
    st0 = srv.stats()
    log(f"serving setup: {len(mix)} worlds ({sum(w.tail for w in mix)} tail), "
        f"{sum(w.scene.ntris for w in mix) / 1e6:.1f} M triangles, budget {budget / 2 ** 30:.1f} GiB, "
        f"{time.perf_counter() - t0:.1f} s")
    n_req = int(a.serve_rate * a.serve_seconds)
    gaps = rng.exponential(1.0 / a.serve_rate, n_req)
    arrive = np.cumsum(gaps)
    worlds = rng.integers(0, len(mix), n_req)
    trace = [(int(worlds[k]), mix[int(worlds[k])].frame(rng)) for k in range(n_req)]
    done = []
    ts = time.perf_counter()
    for k, (wi, fr) in enumerate(trace):
        while True:
            now = time.perf_counter() - ts
            if now >= arrive[k]:
                break
            if arrive[k] - now > 5e-4:
                time.sleep(min(arrive[k] - now - 4e-4, 2e-3))
        srv.submit(ids[wi], fr)
        if k % 64 == 63:
            done.extend(srv.poll(4096))
            log(f"serving: {k + 1}/{n_req} submitted, {len(done)} done") if k % 2048 == 2047 else None
    srv.drain()
    done.extend(srv.poll(n_req + 16))
    st = srv.stats()
    # Capacity: the same trace submitted at once (closed loop), completions per second until the queue drains.
    tc = time.perf_counter()
    for wi, fr in trace[:a.serve_capacity]:
        srv.submit(ids[wi], fr)
    srv.drain()
    cap_value = a.serve_capacity / (time.perf_counter() - tc)
    cap_done = srv.poll(a.serve_capacity + 16)
    cap_batch = a.serve_capacity / max(1, srv.stats()["batches"] - st["batches"])
    srv.destroy()
    lat = [c["t_done"] - c["t_submit"] for c in done]
    span = max(c["t_done"] for c in done) - min(c["t_submit"] for c in done)
    value = len(done) / span
    # Plain FP32 path, serially: the first requests of the same trace, one world update at a time.
    k_plain = min(n_req, a.serve_plain)
    plain_worlds, plain_lat = {}, []
    for wi, fr in trace[:k_plain]:
        if wi not in plain_worlds:
            m = mix[wi]
            w = rtgi.World(m.vox, dataclasses.replace(m.rc, fp16=False))
            w.upload_scene(m.scene)
            w.set_frame(fr)
            w.encode()
            w.sync()
            plain_worlds[wi] = w
    tp = time.perf_counter()
    for wi, fr in trace[:k_plain]:
        t = time.perf_counter()
        w = plain_worlds[wi]
        w.set_frame(fr)
        w.encode()
        w.sync()
        plain_lat.append(time.perf_counter() - t)
    plain_value = k_plain / (time.perf_counter() - tp)
    for w in plain_worlds.values():
        w.destroy()
    log(f"serving: {value:.1f} updates/s (offered {a.serve_rate}/s), p50 {pct(lat, 50)} ms, p99 {pct(lat, 99)} ms; "
        f"capacity {cap_value:.1f} updates/s; plain FP32 serial {plain_value:.1f} updates/s")
    return {
        "workload": f"{len(mix)} resident mixed worlds (atrium/city, 10k-500k triangles log-uniform, clipmap edge "
                    f"32-128 x 4 levels, 3-5 cascades, FP16 storage; every 10th a 2M-triangle 400 m block), Poisson "
                    f"arrivals at {a.serve_rate:g}/s for {a.serve_seconds:g} s, budget 128 MiB per world",
        "requests": len(done), "updates_per_s": round(value, 2), "offered_per_s": a.serve_rate,
        "latency_ms": {"p50": pct(lat, 50), "p99": pct(lat, 99), "max": pct(lat, 100)},
        "mean_batch": round(len(done) / max(1, st["batches"] - st0["batches"]), 2),
        "admissions": st["admissions"] - st0["admissions"], "evictions": st["evictions"] - st0["evictions"],
        "peak_resident_gib": round(st["peak_resident_bytes"] / 2 ** 30, 3), "budget_gib": round(budget / 2 ** 30, 3),
        "triangles_total_m": round(sum(w.scene.ntris for w in mix) / 1e6, 2), "worlds_overflowing_pools": overflowing,
        "plain_fp32_serial": {"updates_per_s": round(plain_value, 2), "requests": k_plain,
                              "latency_ms": {"p50": pct(plain_lat, 50), "p99": pct(plain_lat, 99)}},
        "capacity": {"updates_per_s": round(cap_value, 2), "requests": len(cap_done), "mean_batch": round(cap_batch, 2)},
        "capacity_vs_plain": round(cap_value / plain_value, 3),
    }


def serve_binding_bench(a):
    """Serving under a binding budget: the serving leg's mix of worlds with a device budget of half their bytes, so
    requests for worlds that are not resident evict idle ones (their SH history and bounce state are kept in host
    memory) and rebuild them on admission. Poisson arrivals at the same rate, half as long."""
    import rtgi
    from rtgi.workloads import world_mix

    mix = world_mix(a.serve_worlds, seed=7)
    sizes = [rtgi.estimate_bytes(w.vox, w.rc, w.scene.pos.shape[0], w.scene.tri.shape[0], w.scene.materials.shape[0])
             for w in mix]
    budget = max(int(0.5 * sum(sizes)), max(sizes))
    srv = rtgi.Server(budget=budget, max_batch=a.serve_batch, slots=2)
    ids = [srv.add_world(w.scene, w.vox, w.rc) for w in mix]
    rng = np.random.default_rng(13)
    n_req = int(a.serve_rate * a.serve_seconds / 2)
    arrive = np.cumsum(rng.exponential(1.0 / a.serve_rate, n_req))
    worlds = rng.integers(0, len(mix), n_req)
    st0 = srv.stats()
    done = []
    ts = time.perf_counter()
    for k in range(n_req):
        while time.perf_counter() - ts < arrive[k]:
            gap = arrive[k] - (time.perf_counter() - ts)
            if gap > 5e-4:
                time.sleep(min(gap - 4e-4, 2e-3))
        srv.submit(ids[int(worlds[k])], mix[int(worlds[k])].frame(rng))
        if k % 64 == 63:
            done.extend(srv.poll(4096))
    srv.drain()
    done.extend(srv.poll(n_req + 16))
    st = srv.stats()
    srv.destroy()
    lat = [c["t_done"] - c["t_submit"] for c in done]
    span = max(c["t_done"] for c in done) - min(c["t_submit"] for c in done)
    fails = sum(1 for c in done if c["status"] != 0)
    log(f"serving, binding budget: {len(done) / span:.1f} updates/s, p50 {pct(lat, 50)} ms, p99 {pct(lat, 99)} ms, "
        f"{st['admissions'] - st0['admissions']} admissions, {st['evictions'] - st0['evictions']} evictions")
    return {"budget_gib": round(budget / 2 ** 30, 3), "resident_bytes_all_gib": round(sum(sizes) / 2 ** 30, 3),
            "requests": len(done), "failed": fails, "updates_per_s": round(len(done) / span, 2),
            "latency_ms": {"p50": pct(lat, 50), "p99": pct(lat, 99), "max": pct(lat, 100)},
            "admissions": st["admissions"] - st0["admissions"], "evictions": st["evictions"] - st0["evictions"],
            "peak_resident_gib": round(st["peak_resident_bytes"] / 2 ** 30, 3)}


class Motion:
    """The worlds' moving objects (every city building) re-posed per update, as the reference's runner poses each
    frame (src/core/rcrun.ts:194-217): the poses of every frame are evaluated before timing (the runner also poses
    ahead of its timed frames, rcrun.ts:200-203), and each update sends its world's poses (rtgi_world_set_pose); the
    device then poses the rest mesh before the voxel build."""

    def __init__(self, worlds, scenes, frames):
        for w, sc in zip(worlds, scenes):
            w.set_scene_objects(sc)
        self.poses = {id(w): [sc.poses_at(f) for f in range(frames)] for w, sc in zip(worlds, scenes)}
        self.worlds = list(worlds)
        self.frames = frames
        self.frame = 0
        self.objects = int(sum(len(sc.objects) for sc in scenes))

    def advance(self, worlds=None):
        """Sends the next frame's poses to the given worlds (all by default)."""
        f = self.frame % self.frames
        self.frame += 1
        for w in self.worlds if worlds is None else worlds:
            w.set_pose(self.poses[id(w)][f])


def gather_bench(a, w, scene):
    """The gather at frame rate (gatherWorld per frame, rc.ts:1218-1220): a 1920 x 1080 G-buffer's worth of surface
    points (uniform on the scene's triangles inside the c0 probe grid, unit face normals) in device memory, in Morton
    order of their c0 cell (neighbouring pixels of a G-buffer are neighbouring surface points), gathered from the
    world's last field with rtgi_world_gather_async, without and with the occlusion tests. HIP events on the gather's
    stream time --gather-frames frames of each mode, the modes alternating twice (the second round is reported)."""
    from rtgi.hip import DeviceArray, Event, Stream

    n = 1920 * 1080
    rng = np.random.default_rng(11)
    eye = np.asarray(w.frame.eye)
    tri = scene.tri[:, :3].astype(np.int64)
    cen = (scene.pos[tri[:, 0]] + scene.pos[tri[:, 1]] + scene.pos[tri[:, 2]]) / 3
    inside = np.all(np.abs(cen - eye) < 30.0, axis=1)
    t = rng.choice(np.nonzero(inside)[0], n)
    A, B, Cc = scene.pos[tri[t, 0]], scene.pos[tri[t, 1]], scene.pos[tri[t, 2]]
    u = rng.uniform(0, 1, (n, 2))
    u[u.sum(1) > 1] = 1 - u[u.sum(1) > 1]
    pos = (A + (B - A) * u[:, :1] + (Cc - A) * u[:, 1:]).astype(np.float32)
    nr = np.cross(B - A, Cc - A).astype(np.float64)
    nr = (nr / np.maximum(np.linalg.norm(nr, axis=1, keepdims=True), 1e-12)).astype(np.float32)
    cell = np.clip(np.floor(pos - pos.min(0)).astype(np.int64), 0, 1023)  # 1 m c0 cells
    key = np.zeros(n, np.int64)
    for bit in range(10):
        for ax in range(3):
            key |= ((cell[:, ax] >> bit) & 1) << (3 * bit + ax)
    order = np.argsort(key, kind="stable")
    pos, nr = np.ascontiguousarray(pos[order]), np.ascontiguousarray(nr[order])
    dp, dn = DeviceArray.from_numpy(pos), DeviceArray.from_numpy(nr)
    di = DeviceArray((n, 3))
    s = Stream()
    out = {"workload": f"{n} surface points (a 1920 x 1080 G-buffer) inside the c0 probe grid of world 1, device "
                       "arrays, rtgi_world_gather_async", "points": n,
           "algorithmic_bytes_per_point": 36, "bytes_note": "positions and normals read, irradiance written (SH reads "
                                                            "of the 8 surrounding probes are shared and mostly L2)"}
    for occl in (False, True, False, True):
        for _ in range(2):
            w.gather_async(dp.ptr, dn.ptr, n, di.ptr, occlusion=occl, stream=s.handle)
        s.synchronize()
        e0, e1 = Event(), Event()
        e0.record(s)
        for _ in range(a.gather_frames):
            w.gather_async(dp.ptr, dn.ptr, n, di.ptr, occlusion=occl, stream=s.handle)
        e1.record(s)
        s.synchronize()
        ms = e0.elapsed_ms(e1) / a.gather_frames
        key = "occlusion" if occl else "plain"
        irr = di.numpy()
        out[key] = {"ms_per_frame": round(ms, 4), "mpoints_per_s": round(n / ms / 1e3, 1),
                    "gbps_algorithmic": round(36 * n / ms / 1e6, 1),
                    "nonzero_fraction": round(float((np.abs(irr).sum(1) > 0).mean()), 4)}
        log(f"gather {key}: {ms:.3f} ms per 1080p frame")
        e0.destroy()
        e1.destroy()
    s.destroy()
    for d in (dp, dn, di):
        d.free()
    return out


def width_bench(a, worlds, widths=(8, 16, 32, 64), motion=None):
    """Latency against throughput at the headline configuration: the resident worlds updated in batches of width W,
    one batch at a time (a batch is launched when the previous one has completed, as a server with one slot launches
    them). An update's latency is its batch's time from launch to completion on the otherwise idle GPU (host clock:
    from sending the batch's poses (motion) and enqueueing the graph replay to the end of the wait); throughput counts
    every update over the whole run."""
    import rtgi

    out = []
    for W in widths:
        if W > len(worlds):
            continue
        batches = [rtgi.Batch(worlds[i:i + W], graph=True) for i in range(0, len(worlds), W)]
        for b in batches:  # capture and warm up
            b.encode()
            b.sync()
        lat, n = [], 0
        t = time.perf_counter()
        for _ in range(a.steps):
            for b in batches:
                t0 = time.perf_counter()
                if motion:
                    motion.advance(b.worlds)
                b.encode()
                b.sync()
                lat.extend([time.perf_counter() - t0] * len(b.worlds))
                n += len(b.worlds)
        rate = n / (time.perf_counter() - t)
        for b in batches:
            b.destroy()
        out.append({"width": W, "updates_per_s": round(rate, 2), "latency_ms": {"p50": pct(lat, 50), "p99": pct(lat, 99)}})
        log(f"width {W}: {rate:.1f} updates/s, p50 {pct(lat, 50)} ms, p99 {pct(lat, 99)} ms")
    ok = [r for r in out if r["latency_ms"]["p99"] <= 4.0]
    best = max(ok, key=lambda r: r["width"]) if ok else None
    return {"workload": f"{len(worlds)} resident worlds of the headline, batches of W run one at a time (hipGraph)",
            "curve": out, "widest_p99_le_4ms": best}


def full_bench(a, scenes, vox, rc, frame0, tuning, bsz):
    """The same worlds with the reference's default extras on: voxel relocation with occlusion-aware parents, traced
    per-voxel sun visibility and the multi-bounce volume. Batched hipGraph updates with the headline's tuning and
    width, and the plain path one update at a time."""
    import dataclasses

    import rtgi

    rcf = dataclasses.replace(rc, relocate=True, sunTrace=True, bounce=True)
    worlds = []
    for sc in scenes[:a.full_worlds]:
        w = rtgi.World(vox, rcf)
        w.upload_scene(sc)
        w.set_frame(dataclasses.replace(frame0, eye=sc.eye, sunDir=sc.sunDir))
        w.set_tuning(**tuning)
        worlds.append(w)
    for w in worlds:
        w.encode()
        w.sync()
    t = time.perf_counter()
    for w in worlds:
        w.encode()
        w.sync()
    plain = len(worlds) / (time.perf_counter() - t)
    groups = [worlds[i:i + bsz] for i in range(0, len(worlds), bsz)]
    batches = [rtgi.Batch(g, graph=True) for g in groups]
    for _ in range(2):
        for b in batches:
            b.encode()
        for b in batches:
            b.sync()
    steps = max(1, a.steps // 2)
    t = time.perf_counter()
    for _ in range(steps):
        for b in batches:
            b.encode()
        for b in batches:
            b.sync()
    value = steps * len(worlds) / (time.perf_counter() - t)
    tb = rtgi.Batch(groups[0])
    tb.set_timing(True)
    tb.encode()
    tb.sync()
    stage = tb.stage_ms()
    tb.destroy()
    active = int(np.count_nonzero(worlds[0].read_bounce_volume()[0][..., 0, 3] > 0))
    for b in batches:
        b.destroy()
    for w in worlds:
        w.destroy()
    log(f"full default: {value:.1f} updates/s batched, {plain:.1f} plain")
    return {"value": round(value, 3), "unit": "GI updates/s", "worlds": len(worlds), "batch": bsz,
            "plain_fp32": round(plain, 3), "speedup_vs_plain": round(value / plain, 3),
            "stage_ms_per_batch": {k: round(v, 4) for k, v in stage.items()},
            "bounce_probes_active_world1": active,
            "extras": "relocation + occlusion-aware parents, traced sun visibility, bounce volume"}


def dyn_bench(a):
    """The atrium, whose 16 moving boxes are the tail of its triangle table, with the dynamic coverage layer: each
    update splats the boxes into the layer (mark, parents, allocate, clear, splat, resolve) and the cascade rays read
    it in front of their static hits. Against the same worlds with the boxes voxelised with the rest; and, with the
    bounce volume on, with and without token obj (the boxes also receive the volume's light and block and reflect its
    rays). Batched hipGraph updates of a.dyn_worlds atrium worlds (seeds 1..n), fused per-cascade trace+merge."""
    import dataclasses

    import rtgi

    _, vox, rc, f0 = rtgi.config("atrium")
    scenes = [rtgi.generate_scene("atrium", seed=s) for s in range(1, a.dyn_worlds + 1)]
    out = {}
    variants = (("boxes_in_bricks", False, rc), ("dynamic_layer", True, rc),
                ("layer_bounce", True, dataclasses.replace(rc, bounce=True)),
                ("layer_bounce_objects", True, dataclasses.replace(rc, bounce=True, objects=True)))
    for key, layer, rcv in variants:
        worlds = []
        for sc in scenes:
            w = rtgi.World(vox, rcv)
            w.upload_scene(sc)
            if layer:
                w.set_dynamic(sc.firstDynamicTri, 512)
            w.set_frame(dataclasses.replace(f0, eye=sc.eye, sunDir=sc.sunDir))
            w.set_tuning(fused=1)
            worlds.append(w)
        b = rtgi.Batch(worlds, graph=True)
        for _ in range(2):
            b.encode()
            b.sync()
        # An update of 16 atrium worlds takes about a millisecond: 5x the steps keep host time out of the window.
        steps = max(1, 5 * a.steps)
        t = time.perf_counter()
        for _ in range(steps):
            b.encode()
        b.sync()
        value = steps * len(worlds) / (time.perf_counter() - t)
        b.set_timing(True)
        ms = {}
        ts = max(1, a.steps)
        for _ in range(ts):
            b.encode()
            b.sync()
            for k, v in b.stage_ms().items():
                ms[k] = ms.get(k, 0.0) + v / ts
        out[key] = {"updates_per_s": round(value, 3), "stage_ms_per_batch": {k: round(v, 4) for k, v in ms.items()}}
        if layer:
            out[key]["layer_bricks_world1"] = worlds[0].read_dyn_counters()["bricks"]
        b.destroy()
        for w in worlds:
            w.destroy()
    log(f"dynamic layer: {out['dynamic_layer']['updates_per_s']:.1f} updates/s, boxes in bricks "
        f"{out['boxes_in_bricks']['updates_per_s']:.1f}; with bounce {out['layer_bounce']['updates_per_s']:.1f}, "
        f"+ token obj {out['layer_bounce_objects']['updates_per_s']:.1f}")
    return {"value": out["dynamic_layer"]["updates_per_s"], "unit": "GI updates/s", "worlds": len(scenes),
            "config": "atrium (55,724 triangles, 1,292 of them the 16 moving boxes), FP32, 512 layer bricks",
            **out}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--worlds", type=int, default=64)
    ap.add_argument("--batch", type=int, default=0, help="worlds per batched launch (0: all)")
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--config", default="city")
    ap.add_argument("--cpu-repeats", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mix-worlds", type=int, default=32, help="batched mixed-world run (0: skip it)")
    ap.add_argument("--full-worlds", type=int, default=64,
                    help="run with relocation, traced sun visibility and multi-bounce on (0: skip it)")
    ap.add_argument("--dyn-worlds", type=int, default=16, help="atrium worlds of the dynamic-layer run (0: skip)")
    ap.add_argument("--no-widths", dest="widths", action="store_false", help="skip the latency-vs-width curve")
    ap.add_argument("--gather-frames", type=int, default=20, help="1080p gather frames per mode (0 skips the leg)")
    ap.add_argument("--static", dest="motion", action="store_false",
                    help="headline without motion (the buildings stay at the pose of frame 0)")
    ap.add_argument("--serve-seconds", type=float, default=4.0, help="serving run length (0: skip serving)")
    ap.add_argument("--serve-rate", type=float, default=2000.0, help="mean Poisson request rate (1/s)")
    ap.add_argument("--serve-worlds", type=int, default=64)
    ap.add_argument("--serve-batch", type=int, default=64, help="server max worlds per batched launch")
    ap.add_argument("--serve-plain", type=int, default=400, help="requests of the trace run on the plain path")
    ap.add_argument("--serve-capacity", type=int, default=4000, help="requests of the closed-loop capacity run")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MB",
                    help="only run a plain device copy of MB MiB (calibrates the FETCH_SIZE / WRITE_SIZE counters)")
    a = ap.parse_args()
    if a.copy_calibration:
        import torch

        n = a.copy_calibration << 18  # float32 elements
        src = torch.ones(n, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_each_way": n * 4, "launches": 5}), flush=True)
        return
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported (use --gpus 1)", flush=True)
        sys.exit(2)

    import dataclasses

    import rtgi

    if rtgi.device_count() < 1:
        raise SystemExit("bench.py needs a GPU: no HIP device visible")

    base = rtgi.CONFIGS[a.config]
    worlds, first, scenes = [], None, []
    t0 = time.perf_counter()
    for k in range(a.worlds):
        sc = rtgi.generate_scene(**dict(base["scene"], seed=k + 1))
        scenes.append(sc)
        _, vox, rc, frame = rtgi.config(a.config) if k == 0 else (None, base["vox"], base["rc"], first[3])
        frame = dataclasses.replace(frame, eye=sc.eye, sunDir=sc.sunDir)
        w = rtgi.World(vox, rc)
        w.upload_scene(sc)
        w.set_frame(frame)
        worlds.append(w)
        if k == 0:
            first = (sc, vox, rc, frame)
    # Moving buildings: every world's objects re-posed on the device by each update (plain and batched alike).
    motion = Motion(worlds, scenes, 64) if a.motion else None
    lay = worlds[0].layout()
    rays = lay["total_rays"]
    log(f"setup: {a.worlds} worlds of '{a.config}' in {time.perf_counter() - t0:.1f} s, {first[0].ntris} triangles "
        f"in world 1, {rays} rays per update, {lay['device_bytes'] / 2 ** 20:.0f} MiB per world")

    # Plain path: one world at a time.
    if motion:
        motion.advance()
    for w in worlds:
        w.encode()
        w.sync()
    plain_lat = []
    t_start = time.perf_counter()
    for s in range(a.plain_steps):
        f = motion.frame if motion else 0
        for w in worlds:
            t = time.perf_counter()
            if motion:
                w.set_pose(motion.poses[id(w)][f % motion.frames])
            w.encode()
            w.sync()
            plain_lat.append(time.perf_counter() - t)
        if motion:
            motion.frame += 1
    plain_value = a.plain_steps * len(worlds) / (time.perf_counter() - t_start)
    log(f"plain: {plain_value:.1f} updates/s")

    # Batched path, hipGraph replays; batch width, probe tile and fusion from the autotuner unless --batch is given.
    tune = None
    if a.batch > 0:
        bsz = a.batch
    else:
        from rtgi.autotune import autotune

        # One batch of every world: its stage events then time each launch alone, so the roofline below is per launch.
        # (Four concurrent batches of 16 were about 1 % faster in throughput, with launches overlapping in time.)
        tune = autotune(worlds, tiles=(1, 4), widths=(len(worlds),), repeats=2, log=log)
        bsz = tune["best"]["width"]
    worlds_tuning = worlds[0].tuning()
    groups = [worlds[i:i + bsz] for i in range(0, len(worlds), bsz)]
    batches = [rtgi.Batch(g, graph=True) for g in groups]
    # The timed steps replay one graph per batch; the stage durations come from further steps with HIP events
    # recorded on each batch's stream between its per-stage graphs (four graphs per update instead of one).
    for _ in range(a.warmup):
        for b in batches:
            b.encode()
        for b in batches:
            b.sync()
    lat, step_ms = [], []
    st = {"vox": [], "trace": [], "merge": [], "sh": []}
    t_start = time.perf_counter()
    for s in range(a.steps):
        ts = time.perf_counter()
        if motion:
            motion.advance()
        for b in batches:
            b.encode()
        for b, g in zip(batches, groups):
            b.sync()
            lat.extend([time.perf_counter() - ts] * len(g))
        step_ms.append((time.perf_counter() - ts) * 1e3)
        log(f"step {s + 1}/{a.steps}: {step_ms[-1]:.2f} ms")
    elapsed = time.perf_counter() - t_start
    value = a.steps * len(worlds) / elapsed
    for b in batches:
        b.set_timing(True)
    for _ in range(max(3, a.steps // 2)):
        if motion:
            motion.advance()
        for b in batches:
            b.encode()
        for b, g in zip(batches, groups):
            b.sync()
            if len(g) == len(groups[0]):
                for k, v in b.stage_ms().items():
                    st[k].append(v)
    stage = {k: float(np.mean(v)) for k, v in st.items()}
    # Roofline of the dominant kernel, the trace (fused with the merge: one launch per cascade step of the batch, 5
    # per update). Algorithmic bytes per world update are the trace's and merge's of the per-unit figures: 16 B of f32
    # rgba written per ray, and per merge level c < N-1 (2 R_c + R_{c+1}) x 16 B (own radiance read and written, the
    # level above read once), 2 R_{N-1} x 16 B for the last. Achieved = those bytes per launch / the mean launch
    # duration (the trace stage's HIP-event time over its launches). The fused kernel itself needs fewer bytes: the
    # traced value is merged in registers, so it writes 16 B per ray plus each level c >= 1's 2x2 averages (one float4
    # per 4 rays, written and read once): frac_fused_bytes relates those to the same time.
    per_c = lay["rays"]
    fused = worlds_tuning["fused"] != 0
    merge_world = 16 * (sum(2 * per_c[c] + per_c[c + 1] for c in range(len(per_c) - 1)) + 2 * per_c[-1])
    alg_world = 16 * rays + merge_world
    fused_world = 16 * rays + 2 * 16 * sum(r // 4 for r in per_c[1:])
    launches = len(per_c) if fused else 1
    alg_launch = alg_world * len(groups[0]) / launches
    launch_s = (stage["trace"] + (0.0 if fused else stage["merge"])) * 1e-3 / launches
    achieved = alg_launch / launch_s / 1e9
    achieved_fused = fused_world * len(groups[0]) / launches / launch_s / 1e9
    # HBM traffic per launch, L2 hit rate, VALU issue and active lanes from the committed counter profile of the
    # fused city trace (FETCH_SIZE / WRITE_SIZE copy-calibrated), scaled to this batch width; null for other
    # configurations.
    traffic, prof = None, {}
    tf = os.path.join(ROOT, "profiles", "trace_traffic_city.json")
    if a.config == "city" and fused and os.path.exists(tf):
        with open(tf) as fh:
            prof = json.load(fh)
        traffic = int(prof["hbm_bytes_per_world_update"] * len(groups[0]) / launches)
    cnt = worlds[0].read_counters()

    for b in batches:
        b.destroy()
    gather = gather_bench(a, worlds[0], scenes[0]) if a.gather_frames > 0 else None
    widths = width_bench(a, worlds, motion=motion) if a.widths else None
    for w in worlds:
        w.destroy()
    full = (full_bench(a, scenes, first[1], first[2], first[3], worlds_tuning, bsz) if a.full_worlds > 0 else None)
    mixed = mix_bench(a) if a.mix_worlds > 0 else None
    dynamic = dyn_bench(a) if a.dyn_worlds > 0 else None
    serving = serve_bench(a) if a.serve_seconds > 0 else None
    if serving is not None:
        serving["binding_budget"] = serve_binding_bench(a)

    cpu = None
    if not a.no_cpu:
        sc, vox, rc, frame = first
        v, dt = cpu_baseline(sc, vox, rc, frame, a.cpu_repeats)
        cpu = {"value": round(v, 4), "unit": "GI updates/s", "cores": os.cpu_count(), "kind": "port",
               "sample": f"{a.cpu_repeats} full updates of world 1 ({sc.ntris} triangles, {rays} rays) by the "
                         f"oracle in {dt:.1f} s; trace on {os.cpu_count()} OpenMP threads, build/merge 1 thread"}

    out = {
        "metric": "GI updates/sec @64 worlds + p50/p99 ms vs plain FP32; Mrays/s; HBM GB/s",
        "value": round(value, 3),
        "unit": "GI updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(elapsed / a.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural city blocks)",
        "config": {"workload": f"{a.worlds} resident '{a.config}' worlds, batched FP32 update (hipGraph), "
                               f"{len(groups)} batch(es) of {bsz}",
                   "worlds": a.worlds, "batch": bsz, "tuning": worlds_tuning,
                   "triangles_world1": int(first[0].ntris),
                   "clipmap": "4 x 128^3 @ 0.25 m", "probes_c0": "64^3", "cascades": 5,
                   "rays_per_update": int(rays),
                   "motion": (f"{motion.objects} buildings ({motion.objects // len(worlds)} per world) turned and "
                              f"moved every update: poses sent per update, rest meshes posed on the device"
                              if motion else None)},
        "latency_ms": {"p50": pct(lat, 50), "p99": pct(lat, 99)},
        "plain_fp32": {"value": round(plain_value, 3), "unit": "GI updates/s",
                       "latency_ms": {"p50": pct(plain_lat, 50), "p99": pct(plain_lat, 99)}},
        "speedup_vs_plain": round(value / plain_value, 3),
        "mrays_per_s": round(rays * value / 1e6, 1),
        "stage_ms_per_batch": {k: round(v, 4) for k, v in stage.items()},
        "stage_timing": "HIP events between each batch's stage graphs, mean over further updates after the timed "
                        "steps (the timed steps replay one graph per batch)",
        "trace_merge_fused": worlds_tuning["fused"] != 0,
        "autotune_table": tune["table"] if tune else None,
        "counters_world1": {k: cnt[k] for k in ("bricks", "voxels", "brickOverflow", "voxelOverflow")},
        "roofline": {"kernel": ("k_trace_merge_b / k_trace_merge_b1 (one launch per cascade step of the batch: "
                                "queued steps / one-wavefront workgroups)" if fused
                                else "k_trace_b + k_merge_step_b"),
                     "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "algorithmic_bytes_per_launch": int(alg_launch), "launch_ms": round(launch_s * 1e3, 4),
                     "bytes_per_world_update": int(alg_world),
                     "fused_kernel_bytes_per_world_update": int(fused_world),
                     "frac_fused_bytes": round(achieved_fused / HBM_PEAK_GBS, 5),
                     "limiter": "valu" if prof.get("valu_issue_busy", 0) > 0.9 else None,
                     "valu_issue_busy_profiled": prof.get("valu_issue_busy"),
                     "lanes_per_valu_profiled": prof.get("lanes_per_valu"),
                     "l2_hit_profiled": prof.get("l2_hit"),
                     "wavefront_cycles_profiled": {k: prof.get(k) for k in ("wait_dependency_frac", "wait_issue_frac",
                                                                            "issuing_frac")},
                     "limiter_note": "VALU issue is busy for about every cycle of the launch (4-cycle model) with "
                                     "about half the lanes active, while HBM moves well below peak; a wavefront waits "
                                     "on its own dependent reads about half its cycles, so the DDA is bound by "
                                     "divergent vector issue and dependent-read latency; frac is against the HBM "
                                     "roofline because nothing on the path is an MFMA contraction"},
        "cpu_baseline": cpu,
        "latency_vs_width": widths,
        "gather_1080p": gather,
        "full_default": full,
        "dynamic_layer": dynamic,
        "batched_mix": mixed,
        "serving": serving,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
