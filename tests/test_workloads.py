"""The seeded world mix of the batched and serving workloads: deterministic, 10% tail worlds, sizes in the stated
ranges, the nominal memory of 128 MiB per world on average, and pools that the scenes do not overflow (clipmap
counters from the oracle's build)."""
import numpy as np

import rtgi
from rtgi.workloads import world_mix
from oracle.oracle import Oracle


def test_mix_is_seeded_and_in_range():
    a, b = world_mix(20, seed=3), world_mix(20, seed=3)
    assert [w.scene.ntris for w in a] == [w.scene.ntris for w in b]
    assert [w.tail for w in a] == [i % 10 == 9 for i in range(20)]
    for w in a:
        assert w.rc.fp16 and 3 <= w.rc.cascades <= 5 and w.vox.levels == 4
        if w.tail:
            assert w.scene.ntris > 1_500_000 and w.vox.G == 16
        else:
            assert 5_000 < w.scene.ntris < 700_000 and w.vox.G in (4, 8, 12, 16)
    sizes = [rtgi.estimate_bytes(w.vox, w.rc, w.scene.pos.shape[0], w.scene.tri.shape[0], w.scene.materials.shape[0])
             for w in world_mix(64)]
    assert np.mean(sizes) < 128 * 2 ** 20


def test_mix_pools_do_not_overflow():
    mix = world_mix(64)
    pick = [i for i, w in enumerate(mix) if not w.tail and w.scene.ntris < 150_000][:6]
    pick += [next(i for i, w in enumerate(mix) if w.vox.G == 4)]
    for i in pick:
        w = mix[i]
        o = Oracle(w.vox, w.rc, w.frame())
        o.build(w.scene)
        c = o.counters()
        assert c["brickOverflow"] == 0 and c["voxelOverflow"] == 0 and c["bigOverflow"] == 0, (i, c)
