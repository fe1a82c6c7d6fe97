"""GI-update benchmark on one MI355X.

Workload: W resident city-block worlds (500k triangles each, 4 x 128^3 clipmap at 0.25 m, 64^3 c0 probes, 5
cascades, FP32), seeds 1..W. One step = one full GI update (voxel rebuild from every triangle, probe-ray trace of all
cascades, cascade merge, SH projection) of every world. The value is aggregate GI updates per second; per-update
latency percentiles are reported beside it. Inputs are resident in HBM before timing starts.

Prints one JSON line (rank 0). Only --gpus 1 is supported: this path runs on one GPU.
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--worlds", type=int, default=64)
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

    import rtgi

    if rtgi.device_count() < 1:
        raise SystemExit("bench.py needs a GPU: no HIP device visible")
    import dataclasses

    base = rtgi.CONFIGS[a.config]
    worlds, scenes = [], []
    t0 = time.perf_counter()
    for k in range(a.worlds):
        sc = rtgi.generate_scene(**dict(base["scene"], seed=k + 1))
        _, vox, rc, frame = rtgi.config(a.config)
        frame = dataclasses.replace(frame, eye=sc.eye, sunDir=sc.sunDir)
        w = rtgi.World(vox, rc)
        w.upload_scene(sc)
        w.set_frame(frame)
        worlds.append(w)
        if k == 0:
            scenes.append((sc, vox, rc, frame))
    log(f"setup: {a.worlds} worlds of '{a.config}' in {time.perf_counter() - t0:.1f} s, "
        f"{scenes[0][0].ntris} triangles in world 1, {worlds[0].layout()['total_rays']} rays per update")

    for _ in range(a.warmup):
        for w in worlds:
            w.encode()
            w.sync()
    # Timed region: every step updates every world, one update at a time (the plain single-instance path).
    lat = []
    t_start = time.perf_counter()
    for s in range(a.steps):
        for w in worlds:
            t = time.perf_counter()
            w.encode()
            w.sync()
            lat.append(time.perf_counter() - t)
        log(f"step {s + 1}/{a.steps} done")
    elapsed = time.perf_counter() - t_start
    n_upd = a.steps * len(worlds)
    value = n_upd / elapsed
    lat_ms = np.array(lat) * 1e3

    # Per-stage device time (HIP events on the world's stream) over a few extra updates of world 1.
    w0 = worlds[0]
    w0.set_timing(True)
    st = {"vox": [], "trace": [], "merge": [], "sh": []}
    for _ in range(max(3, a.steps)):
        w0.encode()
        w0.sync()
        for k, v in w0.stage_ms().items():
            st[k].append(v)
    w0.set_timing(False)
    stage = {k: float(np.mean(v)) for k, v in st.items()}
    lay = w0.layout()
    rays = lay["total_rays"]
    # Roofline of the trace kernel: algorithmic bytes = 16 B (f32 rgba) written per ray.
    alg_bytes = rays * 16
    achieved = alg_bytes / (stage["trace"] * 1e-3) / 1e9
    cnt = w0.read_counters()

    cpu = None
    if not a.no_cpu:
        sc, vox, rc, frame = scenes[0]
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
        "config": {"workload": f"{a.worlds} resident '{a.config}' worlds, plain FP32 path, one update at a time",
                   "worlds": a.worlds, "triangles_world1": int(scenes[0][0].ntris), "clipmap": "4 x 128^3 @ 0.25 m",
                   "probes_c0": "64^3", "cascades": 5, "rays_per_update": int(rays)},
        "latency_ms": {"p50": round(float(np.percentile(lat_ms, 50)), 3),
                       "p99": round(float(np.percentile(lat_ms, 99)), 3)},
        "mrays_per_s": round(rays * value / 1e6, 1),
        "stage_ms": {k: round(v, 4) for k, v in stage.items()},
        "counters_world1": {k: cnt[k] for k in ("bricks", "voxels", "brickOverflow", "voxelOverflow")},
        "roofline": {"kernel": "k_trace", "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None},
        "cpu_baseline": cpu,
    }
    for w in worlds:
        w.destroy()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
