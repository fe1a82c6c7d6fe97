"""The batched and serving workloads at their sizes.

Batched instances (32 seeded mixed worlds: atrium / city, 10k-500k triangles, clipmap edges 32-128, 3-5 cascades,
FP16 storage): one autotuned hipGraph batch gives every world's solo result bit for bit, and three of them (the
largest city, a G = 12 world and an atrium) meet the oracle bit for bit.

Serving (64 seeded mixed worlds with a 10 % tail of 2M-triangle 400 m blocks, a budget of half their bytes, requests
at about 2,000 per second): every completion's SH and counters equal those of a plain world that runs the same
world's requests in submission order, the budget forces admissions and evictions, and a tail world meets the oracle.
"""
import dataclasses
import hashlib
import time

import numpy as np
import pytest

import rtgi
from oracle.oracle import Oracle
from rtgi.autotune import autotune
from rtgi.workloads import world_mix

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def digest(*arrays):
    h = hashlib.sha1()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def snapshot(w):
    """SH, counters, every cascade's radiance and hit ids of the world's last update, as one digest."""
    parts = [w.read_sh()]
    for c in range(w.rc.cascades):
        parts.extend(w.read_cascade(c, hits=True))
    cnt = w.read_counters()
    return digest(*parts), cnt


def check_oracle(w, m, frame):
    """The world's last update (its first, or a repeat of it at the same frame) against the oracle, FP16 storage:
    hit ids, stored radiance and SH bit for bit."""
    o = Oracle(m.vox, m.rc, frame)
    o.build(m.scene)
    rgba, hits = o.trace()
    merged, sh = o.merge_sh(rgba)
    cnt, oc = w.read_counters(), o.counters()
    for k in ("bricks", "voxels", "brickOverflow", "voxelOverflow"):
        assert cnt[k] == oc[k], k
    for c, (mr, hr) in enumerate(zip(o.split(merged), o.split(hits))):
        g, h = w.read_cascade(c, hits=True)
        assert np.array_equal(h, hr), f"cascade {c}: {np.count_nonzero(h != hr)} hit ids differ"
        assert np.array_equal(g, mr), f"cascade {c}: radiance differs"
    assert np.array_equal(w.read_sh(), sh)


def test_batched_mix_of_32_worlds(require_gpu):
    mix = world_mix(32, seed=5, tail_every=0)
    say(f"case: batched mix, 32 worlds, {sum(m.scene.ntris for m in mix) / 1e6:.2f} M triangles")
    worlds = []
    try:
        for m in mix:
            w = rtgi.World(m.vox, dataclasses.replace(m.rc, recordHits=True))
            w.upload_scene(m.scene)
            w.set_frame(m.frame())
            worlds.append(w)
        solo = []
        for w in worlds:
            w.encode()
            w.sync()
            solo.append(snapshot(w))
        res = autotune(worlds, repeats=1)
        width = res["best"]["width"]
        say(f"autotuned: {res['best']}")
        batches = [rtgi.Batch(worlds[i:i + width], graph=True) for i in range(0, len(worlds), width)]
        try:
            for rep in range(2):  # the capture, then a replay
                for b in batches:
                    b.encode()
                for b in batches:
                    b.sync()
                for i, w in enumerate(worlds):
                    assert snapshot(w) == solo[i], f"replay {rep}: world {i} differs from its solo update"
        finally:
            for b in batches:
                b.destroy()
        city = max((i for i, m in enumerate(mix) if m.kind == "city"), key=lambda i: mix[i].scene.ntris)
        g12 = next(i for i, m in enumerate(mix) if m.vox.G == 12 and i != city)
        atrium = next(i for i, m in enumerate(mix) if m.kind == "atrium" and i not in (city, g12))
        for i in (city, g12, atrium):
            m = mix[i]
            say(f"oracle: world {i} ({m.kind}, {m.scene.ntris} triangles, G = {m.vox.G}, {m.rc.cascades} cascades)")
            check_oracle(worlds[i], m, m.frame())
    finally:
        for w in worlds:
            w.destroy()


def test_serving_mix_of_64_worlds_with_tail(require_gpu):
    mix = world_mix(64, seed=7)
    tails = [i for i, m in enumerate(mix) if m.tail]
    sizes = [rtgi.estimate_bytes(m.vox, m.rc, m.scene.pos.shape[0], m.scene.tri.shape[0], m.scene.materials.shape[0])
             for m in mix]
    budget = max(int(0.5 * sum(sizes)), max(sizes))
    rng = np.random.default_rng(3)
    n_req = 160
    plan = [int(x) for x in rng.integers(0, len(mix), n_req)]
    plan[:4] = [tails[0], tails[1], tails[0], tails[1]]  # two tail worlds, twice each, early in the trace
    frames = [mix[wi].frame(rng) for wi in plan]
    gaps = rng.exponential(1.0 / 2000.0, n_req)
    say(f"case: serving, 64 mixed worlds ({len(tails)} tail), {n_req} requests, budget {budget / 2 ** 30:.2f} of "
        f"{sum(sizes) / 2 ** 30:.2f} GiB")
    got = {}
    with rtgi.Server(budget=budget, max_batch=16, slots=2) as srv:
        ids = [srv.add_world(m.scene, m.vox, m.rc) for m in mix]
        tags = {}
        t0 = time.perf_counter()
        for k, (wi, fr) in enumerate(zip(plan, frames)):
            t0 += gaps[k]
            while time.perf_counter() < t0:
                pass
            tags[srv.submit(ids[wi], fr, sh=True, counters=True)] = k
            for c in srv.poll(64):
                assert c["status"] == 0, c
                got[tags[c["tag"]]] = (digest(c["sh"]), c["counters"])
        srv.drain()
        for c in srv.poll(1024):
            assert c["status"] == 0, c
            got[tags[c["tag"]]] = (digest(c["sh"]), c["counters"])
        st = srv.stats()
    say(f"stats: {st}")
    assert len(got) == n_req
    assert st["admissions"] > 0 and st["evictions"] > 0 and st["peak_resident_bytes"] <= budget
    # The reference: each world's requests in submission order on a plain world of its own.
    by_world = {}
    for k, wi in enumerate(plan):
        by_world.setdefault(wi, []).append(k)
    for wi, ks in by_world.items():
        m = mix[wi]
        with rtgi.World(m.vox, dataclasses.replace(m.rc, recordHits=True)) as w:
            w.upload_scene(m.scene)
            for j, k in enumerate(ks):
                w.set_frame(frames[k])
                w.encode()
                w.sync()
                if wi == tails[0] and j == 0:
                    say(f"oracle: tail world {wi} ({m.scene.ntris} triangles), its first request")
                    check_oracle(w, m, frames[k])
                ref = (digest(w.read_sh()), w.read_counters())
                assert got[k] == ref, f"request {k} (world {wi}, its update {j + 1}) differs from the plain world"
