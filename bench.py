"""GI-update benchmark on one MI355X.

Workload: W resident city-block worlds (about 500k triangles each, 4 x 128^3 clipmap at 0.25 m, 64^3 c0 probes,
5 cascades, FP32), seeds 1..W. One update = voxel rebuild from every triangle, probe-ray trace of all cascades,
cascade merge and SH projection of one world.

Two paths are timed on the same worlds:
  plain   one world at a time: its update is enqueued, then waited for (the single-instance path);
  batched all W worlds with one launch per stage (worlds in batches of --batch), replayed as a hipGraph.
A step updates every world once. The value is the batched path's aggregate GI updates per second; the latency of an
update is the time from the step's submission to the completion of the batch that holds its world (p50/p99 over all
updates of the timed steps). Inputs are resident in HBM before timing starts.

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
    worlds, first = [], None
    t0 = time.perf_counter()
    for k in range(a.worlds):
        sc = rtgi.generate_scene(**dict(base["scene"], seed=k + 1))
        _, vox, rc, frame = rtgi.config(a.config) if k == 0 else (None, base["vox"], base["rc"], first[3])
        frame = dataclasses.replace(frame, eye=sc.eye, sunDir=sc.sunDir)
        w = rtgi.World(vox, rc)
        w.upload_scene(sc)
        w.set_frame(frame)
        worlds.append(w)
        if k == 0:
            first = (sc, vox, rc, frame)
    lay = worlds[0].layout()
    rays = lay["total_rays"]
    log(f"setup: {a.worlds} worlds of '{a.config}' in {time.perf_counter() - t0:.1f} s, {first[0].ntris} triangles "
        f"in world 1, {rays} rays per update, {lay['device_bytes'] / 2 ** 20:.0f} MiB per world")

    # Plain path: one world at a time.
    for w in worlds:
        w.encode()
        w.sync()
    plain_lat = []
    t_start = time.perf_counter()
    for s in range(a.plain_steps):
        for w in worlds:
            t = time.perf_counter()
            w.encode()
            w.sync()
            plain_lat.append(time.perf_counter() - t)
    plain_value = a.plain_steps * len(worlds) / (time.perf_counter() - t_start)
    log(f"plain: {plain_value:.1f} updates/s")

    # Batched path, hipGraph replays.
    bsz = a.batch if a.batch > 0 else len(worlds)
    groups = [worlds[i:i + bsz] for i in range(0, len(worlds), bsz)]
    batches = [rtgi.Batch(g, graph=True) for g in groups]
    for _ in range(a.warmup):
        for b in batches:
            b.encode()
        for b in batches:
            b.sync()
    lat, step_ms = [], []
    t_start = time.perf_counter()
    for s in range(a.steps):
        ts = time.perf_counter()
        for b in batches:
            b.encode()
        for b, g in zip(batches, groups):
            b.sync()
            lat.extend([time.perf_counter() - ts] * len(g))
        step_ms.append((time.perf_counter() - ts) * 1e3)
        log(f"step {s + 1}/{a.steps}: {step_ms[-1]:.2f} ms")
    elapsed = time.perf_counter() - t_start
    value = a.steps * len(worlds) / elapsed

    # Per-stage device time of one batched update (HIP events, direct launches on the batch's stream).
    tb = rtgi.Batch(groups[0])
    tb.set_timing(True)
    st = {"vox": [], "trace": [], "merge": [], "sh": []}
    for _ in range(3):
        tb.encode()
        tb.sync()
        for k, v in tb.stage_ms().items():
            st[k].append(v)
    tb.destroy()
    stage = {k: float(np.median(v)) for k, v in st.items()}
    # Roofline of the trace launch (the dominant kernel): algorithmic bytes = 16 B of f32 rgba written per ray.
    alg_bytes = rays * 16 * len(groups[0])
    achieved = alg_bytes / (stage["trace"] * 1e-3) / 1e9
    cnt = worlds[0].read_counters()

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
                   "worlds": a.worlds, "batch": bsz, "triangles_world1": int(first[0].ntris),
                   "clipmap": "4 x 128^3 @ 0.25 m", "probes_c0": "64^3", "cascades": 5,
                   "rays_per_update": int(rays)},
        "latency_ms": {"p50": pct(lat, 50), "p99": pct(lat, 99)},
        "plain_fp32": {"value": round(plain_value, 3), "unit": "GI updates/s",
                       "latency_ms": {"p50": pct(plain_lat, 50), "p99": pct(plain_lat, 99)}},
        "speedup_vs_plain": round(value / plain_value, 3),
        "mrays_per_s": round(rays * value / 1e6, 1),
        "stage_ms_per_batch": {k: round(v, 4) for k, v in stage.items()},
        "counters_world1": {k: cnt[k] for k in ("bricks", "voxels", "brickOverflow", "voxelOverflow")},
        "roofline": {"kernel": "k_trace (batched launch)", "bound": "hbm", "achieved": round(achieved, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": None},
        "cpu_baseline": cpu,
    }
    for b in batches:
        b.destroy()
    for w in worlds:
        w.destroy()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
