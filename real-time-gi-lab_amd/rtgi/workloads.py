"""Seeded world mixes for the batched and serving workloads.

A mixed world is a seeded atrium or city block whose triangle count is drawn log-uniform in 10k..500k (through the
generator's tessellation detail), with a clipmap of edge E in {32, 64, 96, 128} voxels (G = E / 8) x 4 levels, 3..5
cascades and FP16 radiance storage (FP32 accumulation). The voxel size follows the edge so that the coarsest level
covers the scene: atrium v0 = 8 / E m (a 64 m window at level 3), city v0 = 32 / E m (256 m). A tail world is a
400 x 400 m city block of 256 buildings (about 2M triangles) at E = 128, v0 = 0.5 m. Pools are sized for the scenes
(bricks for every pointer of a clipmap up to G = 8, half of them above; 2^19 / 2^20 / 2^21 voxel attributes for atrium /
city / tail), which none of the mix overflows.
"""
import dataclasses
import math

import numpy as np

from . import Frame, RcParams, VoxParams, generate_scene

ATRIUM_TRIS_AT_DETAIL_1 = 55_700
CITY_TRIS_AT_DETAIL_1 = 495_000


@dataclasses.dataclass
class MixedWorld:
    scene: object
    vox: VoxParams
    rc: RcParams
    kind: str
    tail: bool

    def frame(self, rng=None, jitter=4.0):
        """A frame at the scene's eye, moved by up to `jitter` m horizontally when rng is given."""
        eye = np.asarray(self.scene.eye, np.float64)
        if rng is not None:
            eye = eye + np.array([rng.uniform(-jitter, jitter), 0.0, rng.uniform(-jitter, jitter)])
        return Frame(eye=tuple(float(x) for x in eye), sunDir=self.scene.sunDir)


def sample_world(rng, tail=False, fp16=True, seed=None):
    """One mixed world drawn from rng (a numpy Generator)."""
    seed = int(rng.integers(1, 2 ** 31)) if seed is None else seed
    if tail:
        sc = generate_scene("city", seed=seed, extent=400.0, buildings=256)
        E, kind = 128, "city"
        vox = VoxParams(v0=64.0 / E, levels=4, G=E // 8, maxBricks=2 * (E // 8) ** 3, maxVoxels=1 << 21)
        rc = RcParams(cascades=5, probes0=64, s0=2.0, fp16=fp16)
        return MixedWorld(sc, vox, rc, kind, True)
    kind = "atrium" if rng.uniform() < 0.5 else "city"
    target = math.exp(rng.uniform(math.log(10_000), math.log(500_000)))
    base = ATRIUM_TRIS_AT_DETAIL_1 if kind == "atrium" else CITY_TRIS_AT_DETAIL_1
    detail = math.sqrt(target / base)
    sc = generate_scene(kind, seed=seed, detail=detail)
    E = int(rng.choice([32, 64, 96, 128]))
    G = E // 8
    v0 = (8.0 if kind == "atrium" else 32.0) / E
    vox = VoxParams(v0=v0, levels=4, G=G, maxBricks=(4 if G <= 8 else 2) * G ** 3,
                    maxVoxels=(1 << 19) if kind == "atrium" else 1 << 20)
    cascades = int(rng.integers(3, 6))
    rc = RcParams(cascades=cascades, probes0=32 if kind == "atrium" else 64, s0=0.625 if kind == "atrium" else 1.0,
                  fp16=fp16)
    return MixedWorld(sc, vox, rc, kind, False)


def world_mix(n, seed=7, tail_every=10, fp16=True):
    """n mixed worlds; every tail_every-th one (10%) is a tail world (tail_every <= 0: none)."""
    rng = np.random.default_rng(seed)
    return [sample_world(rng, tail=tail_every > 0 and i % tail_every == tail_every - 1, fp16=fp16) for i in range(n)]
