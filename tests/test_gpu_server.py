"""Request server (rtgi_server_*): continuous batching over resident worlds under a memory budget. Each request's
result (SH and clipmap counters) is bit-identical to the same update run on a plain world of its own that runs its
world's requests in submission order (the c0 SH history and the bounce volume carry from one update of a world to the
next), whether it ran in a mix of uneven worlds, in any batch, with its world evicted and rebuilt in between or not."""
import dataclasses
import time

import numpy as np
import pytest

import rtgi
from conftest import random_scene

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def world_specs(seed=0):
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(6):
        G = (4, 8, 12)[i % 3]
        specs.append((random_scene(rng, int(rng.integers(100, 1500)), -5, 5, 0.6),
                      rtgi.VoxParams(0.25, 3, G, 3 * G ** 3, 1 << 16),
                      rtgi.RcParams(cascades=3, probes0=8 * (1 + i % 2), s0=0.5, fp16=(i == 4))))
    at = rtgi.generate_scene("atrium", seed=5)
    specs.append((at, rtgi.VoxParams(0.125, 4, 8, 2048, 1 << 20), rtgi.RcParams(cascades=4, probes0=32, s0=0.625)))
    # multi-bounce: the bounce volume of each update feeds the next one's hits
    specs.append((random_scene(rng, 800, -4, 4, 0.6), rtgi.VoxParams(0.25, 3, 8, 3 * 512, 1 << 16),
                  rtgi.RcParams(cascades=3, probes0=16, s0=0.5, bounce=True)))
    return specs


def frames(rng, k):
    out = []
    for _ in range(k):
        eye = tuple(float(x) for x in rng.uniform(-2, 2, 3))
        out.append(rtgi.Frame(eye=eye, sunDir=tuple(float(x) for x in rng.uniform(-1, 1, 3) + (0, 2, 0))))
    return out


def sequential(specs, plan):
    """Each request's SH and counters from a plain world of its own per server world, updated through that world's
    requests in submission order."""
    worlds, out = {}, []
    try:
        for wi, fr in plan:
            if wi not in worlds:
                sc, vox, rc = specs[wi]
                worlds[wi] = rtgi.World(vox, rc)
                worlds[wi].upload_scene(sc)
            w = worlds[wi]
            w.set_frame(fr)
            w.encode()
            w.sync()
            out.append((w.read_sh(), w.read_counters()))
    finally:
        for w in worlds.values():
            w.destroy()
    return out


def test_estimate_matches_allocation(require_gpu):
    for sc, vox, rc in world_specs():
        with rtgi.World(vox, rc) as w:
            w.upload_scene(sc)
            assert w.layout()["device_bytes"] == rtgi.estimate_bytes(vox, rc, sc.pos.shape[0], sc.tri.shape[0],
                                                                     sc.materials.shape[0])


@pytest.mark.parametrize("budget_worlds", [0, 3])
def test_server_mix_equals_alone(require_gpu, budget_worlds):
    """Requests over 8 uneven worlds (one with multi-bounce), submitted in bursts and spaced arrivals, several per
    world with different frames. budget_worlds = 3: the budget holds about 3 of them, so worlds are evicted and
    rebuilt between updates, and their history and bounce volume must survive that."""
    specs = world_specs()
    rng = np.random.default_rng(1)
    plan = []  # (world, frame)
    for k, fr in enumerate(frames(rng, 32)):
        plan.append((int(rng.integers(0, len(specs))), fr))
    assert sum(1 for wi, _ in plan if wi == len(specs) - 1) >= 2  # the bounce world runs more than once
    say(f"case: server, {len(specs)} worlds, {len(plan)} requests, budget for {budget_worlds or 'all'} worlds")
    ref = sequential(specs, plan)
    sizes = [rtgi.estimate_bytes(vox, rc, sc.pos.shape[0], sc.tri.shape[0], sc.materials.shape[0])
             for sc, vox, rc in specs]
    budget = 0 if budget_worlds == 0 else int(sum(sorted(sizes)[-budget_worlds:]))
    with rtgi.Server(budget=budget, max_batch=4, slots=2) as srv:
        ids = [srv.add_world(sc, vox, rc) for sc, vox, rc in specs]
        tags = {}
        for k, (wi, fr) in enumerate(plan):
            tags[srv.submit(ids[wi], fr, sh=True, counters=True)] = k
            if k % 7 == 6:
                time.sleep(0.002)  # spaced arrivals between bursts
        srv.drain()
        got = srv.poll(1024)
        st = srv.stats()
    assert len(got) == len(plan)
    order = {}
    for c in got:
        assert c["status"] == 0, c
        k = tags[c["tag"]]
        sh, cnt = ref[k]
        assert c["counters"] == cnt, k
        assert np.array_equal(c["sh"].view(np.uint32), sh.view(np.uint32)), f"request {k}: SH differs"
        order.setdefault(c["world"], []).append(k)
    for wi, ks in order.items():
        assert ks == sorted(ks), f"world {wi}: requests completed out of order {ks}"
    assert st["completed"] == len(plan) and st["queued"] == 0
    assert st["max_batch_seen"] > 1
    if budget_worlds:
        assert st["evictions"] > 0 and st["peak_resident_bytes"] <= budget
    say(f"stats: {st}")


def test_server_errors(require_gpu):
    sc, vox, rc = world_specs()[6]
    need = rtgi.estimate_bytes(vox, rc, sc.pos.shape[0], sc.tri.shape[0], sc.materials.shape[0])
    with rtgi.Server(budget=need - 1) as srv:
        with pytest.raises(rtgi.RtgiError, match="over the whole budget"):
            srv.add_world(sc, vox, rc)
        with pytest.raises(KeyError):
            srv.submit(0, rtgi.Frame(eye=(0.0, 0.0, 0.0)))
        with pytest.raises(rtgi.RtgiError, match="unknown world"):
            rtgi._lib.check(rtgi.lib().rtgi_server_submit(srv._h, rtgi._lib.RequestC(5, 1, rtgi.Frame(eye=(0.0, 0.0, 0.0)).to_c())))


def test_server_refuses_a_far_eye_at_submission(require_gpu):
    """A frame the world would refuse (its window origin past 1e6 bricks) is refused by submit, so it cannot fail the
    batch it would have joined; the other requests complete normally."""
    specs = world_specs()[:3]
    with rtgi.Server(max_batch=4, slots=1) as srv:
        ids = [srv.add_world(sc, vox, rc) for sc, vox, rc in specs]
        t0 = srv.submit(ids[0], rtgi.Frame(eye=(0.1, 0.0, 0.0)), sh=True)
        with pytest.raises(rtgi.RtgiError, match="too far"):
            srv.submit(ids[1], rtgi.Frame(eye=(1e9, 0.0, 0.0)))
        t2 = srv.submit(ids[2], rtgi.Frame(eye=(0.0, 0.2, 0.0)), sh=True)
        srv.drain()
        got = {c["tag"]: c for c in srv.poll(16)}
    assert set(got) == {t0, t2} and all(c["status"] == 0 for c in got.values())
