"""Precision of the GPU results against the FP64 oracle tracer (ray-voxel traversal, shading, merge and SH in double
over the same clipmap), on the atrium configuration (50k triangles, 4 x 64^3 clipmap, 32^3 c0 probes, 4 cascades).

Per cascade, 10k seeded rays (origins at probe positions drawn uniformly from the cascade's grid, directions uniform
on the sphere) go through the GPU ray query, in the plain FP32 form and in the precision forms (FP16 radiance
storage; FP16 radiance with BF16 albedo; FP32 arithmetic throughout); the irradiance (SH) of all c0 probes comes from a
full update. Rays that pass a voxel
edge, corner, window face or skip distance within 1e-6 (relative) are ties: their first voxel is not fixed by the
arithmetic's neighbourhood, and they are counted apart.

Bounds: on non-tie rays the first-hit voxel equals the FP64 walk's, and radiance is within 1e-4 relative (FP32) and
1e-3 relative (FP16: one f16 rounding, <= 2^-11). The ray query itself matches the f32 oracle (hit ids bit-exact,
radiance within 1e-4; the FP16 values exactly). RTGI_PRECISION_REPORT=<path> writes the measured errors as JSON.
"""
import dataclasses
import json
import os

import numpy as np
import pytest

import rtgi
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu
NRAYS = 10_000


def say(msg):
    print(msg, flush=True)


def sphere_rays(rc, G0, c, n, seed):
    rng = np.random.default_rng(seed)
    N = rc.probes0 >> c
    g = rng.integers(0, N, (n, 3)) + np.asarray(G0) // (1 << c)
    org = (g.astype(np.float32) * np.float32(rc.s0 * 2 ** c)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    return org, (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)


def rel_err(x, ref, floor=1e-3):
    return np.abs(x - ref) / np.maximum(np.abs(ref), floor)


FORMS = {"fp32": dict(fp16=False), "fp16": dict(fp16=True), "fp16+bf16": dict(fp16=True, bf16=True)}


def world(form):
    scene, vox, rc, frame = rtgi.config("atrium", **FORMS[form])
    w = rtgi.World(vox, rc)
    w.upload_scene(scene)
    w.set_frame(frame)
    w.encode()
    w.sync()
    o = Oracle(vox, rc, frame)
    o.build(scene)
    return w, o, rc


@pytest.mark.parametrize("form", list(FORMS))
def test_ray_query_matches_oracle(require_gpu, form):
    say(f"case: GPU ray query vs f32 oracle, atrium, {form}")
    fp16 = FORMS[form]["fp16"]
    w, o, rc = world(form)
    G0 = w.layout()["grid_origin"]
    for c in range(rc.cascades):
        org, d = sphere_rays(rc, G0, c, 4000, seed=100 + c)
        g, gh = w.trace_rays(c, org, d)
        r, rh = o.trace_rays(c, org, d)
        assert np.array_equal(gh, rh), f"cascade {c}: {np.count_nonzero(gh != rh)} hit ids differ"
        if fp16:
            assert np.array_equal(g, r)
        else:
            np.testing.assert_allclose(g, r, rtol=1e-4, atol=1e-6)
    w.destroy()


def test_precision_report(require_gpu):
    say(f"case: precision report, atrium, {NRAYS} rays per cascade, FP32, FP16 and FP16 + BF16 albedo forms")
    report = {"config": "atrium scene, 10k seeded sphere rays per cascade; FP64 oracle over the same clipmap "
                        "and the same stored albedo", "forms": {}}
    for name in FORMS:
        fp16 = FORMS[name]["fp16"]
        w, o, rc = world(name)
        G0 = w.layout()["grid_origin"]
        bound = 1e-3 if fp16 else 1e-4
        form = {"cascades": []}
        for c in range(rc.cascades):
            org, d = sphere_rays(rc, G0, c, NRAYS, seed=c)
            g, gh = w.trace_rays(c, org, d)
            r64, h64, tie = o.trace_rays_f64(c, org.astype(np.float64), d.astype(np.float64), eps=1e-6)
            clean = tie == 0
            assert np.array_equal(gh[clean], h64[clean]), f"cascade {c}: first hits differ off ties"
            hit = clean & (gh != 0xFFFFFFFF)
            e = rel_err(g[hit, :3], r64[hit, :3])
            assert hit.sum() > 100 and e.max() < bound, (c, e.max())
            form["cascades"].append({
                "cascade": c, "rays": NRAYS, "non_tie": int(clean.sum()), "non_tie_hits": int(hit.sum()),
                "hit_mismatch_all_rays": int((gh != h64).sum()),
                "radiance_rel_err": {"max": float(e.max()), "p99": float(np.percentile(e, 99)),
                                     "mean": float(e.mean())}})
        # Irradiance: the GPU's SH of all c0 probes against FP64 trace + merge + SH of the same clipmap.
        rgbt, _, tie = o.trace_f64(eps=1e-6)
        _, sh64 = o.merge_sh_f64(rgbt)
        sh = w.read_sh()
        rel = np.abs(sh - sh64).max(axis=1) / np.maximum(np.abs(sh64).max(axis=1), 1e-6)
        form["sh_probe_rel_err"] = {"median": float(np.median(rel)), "p99": float(np.percentile(rel, 99)),
                                    "max": float(rel.max()), "rays_tied": float(tie.mean())}
        report["forms"][name] = form
        w.destroy()
    say(json.dumps(report))
    path = os.environ.get("RTGI_PRECISION_REPORT")
    if path:
        with open(path, "w") as f:
            json.dump(report, f, indent=1)


@pytest.mark.parametrize("case,form", [("atrium", "fp32"), ("atrium", "fp16"), ("atrium", "fp16+bf16"), ("city", "fp32"),
                                       ("city", "fp16"), ("city", "fp16+bf16")])
def test_irradiance_error_vs_fp64(require_gpu, case, form):
    """Irradiance (the SH of every c0 probe) against FP64 trace + merge + SH over the same clipmap, bound
    1e-3 relative per probe (the largest coefficient difference over the probe's largest coefficient). The FP64 rays
    take the f32 walk's first-hit voxel wherever the two walks may disagree (tie rays and rays whose FP64 first hit
    differs: there both voxels are right answers of the reference algorithm, and the f32 one is what the reference's
    own f32 arithmetic picks); every other ray is the FP64 walk's own. The tie-inclusive error (the plain FP64 walk
    everywhere) is reported beside it."""
    scene, vox, rc, frame = rtgi.config(case, **FORMS[form])
    say(f"case: irradiance vs FP64, {case}, {form}")
    w = rtgi.World(vox, rc)
    try:
        w.upload_scene(scene)
        w.set_frame(frame)
        w.encode()
        w.sync()
        sh = w.read_sh()
    finally:
        w.destroy()
    o = Oracle(vox, rc, frame)
    o.build(scene)
    _, hits = o.trace()
    take, taken = o.trace_f64_take(hits, eps=1e-6)
    _, sh64 = o.merge_sh_f64(take)

    def per_probe(ref):
        return np.abs(sh - ref).max(axis=1) / np.maximum(np.abs(ref).max(axis=1), 1e-6)

    rel = per_probe(sh64)
    rgbt, _, _ = o.trace_f64(eps=1e-6)
    _, sh_ties = o.merge_sh_f64(rgbt)
    rel_ties = per_probe(sh_ties)
    say(f"irradiance rel. error: max {rel.max():.3g}, p99 {np.percentile(rel, 99):.3g}, median {np.median(rel):.3g} "
        f"({taken.mean() * 100:.2f} % of rays take the f32 voxel); tie-inclusive max {rel_ties.max():.3g}, "
        f"p99 {np.percentile(rel_ties, 99):.3g}, median {np.median(rel_ties):.3g}")
    assert rel.max() <= 1e-3, rel.max()
