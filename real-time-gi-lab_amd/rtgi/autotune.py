"""Autotuner for batched updates: picks the launch tuning and the batch width of a set of resident worlds by timing.

Candidates are every combination of
  probe_tile   1, 2, 4   probes per axis of a trace block (the ray tile: which rays share a wavefront);
  fused        0, 1      one trace launch then merge steps, or per-cascade trace+merge launches (the ray-sort
                         order: all cascades at once, or one cascade at a time from the coarsest);
  width        worlds per batched launch (divisors of the world count in 8, 16, 32, all).
Each candidate runs `warmup` untimed and `repeats` timed batched updates of every world (hipGraph replays, as the
serving and benchmark paths run them); the fastest aggregate updates per second wins and is applied to the worlds.
No candidate changes a result: the tuning only changes speed and device storage order.
"""
import itertools
import time

from . import Batch


def _step(batches):
    for b in batches:
        b.encode()
    for b in batches:
        b.sync()


def autotune(worlds, tiles=(1, 2, 4), fused=(0, 1), widths=None, warmup=1, repeats=3, log=None):
    """Times every candidate on `worlds` and applies the best. Returns dict(best=..., table=[...])."""
    n = len(worlds)
    if widths is None:
        widths = sorted({w for w in (8, 16, 32) if w < n} | {n})
    table = []
    for tile, fu, width in itertools.product(tiles, fused, widths):
        for w in worlds:
            w.set_tuning(probe_tile=tile, fused=fu)
        batches = [Batch(worlds[i:i + width], graph=True) for i in range(0, n, width)]
        try:
            for _ in range(warmup):
                _step(batches)
            t = time.perf_counter()
            for _ in range(repeats):
                _step(batches)
            ups = repeats * n / (time.perf_counter() - t)
        finally:
            for b in batches:
                b.destroy()
        row = dict(probe_tile=tile, fused=fu, width=width, updates_per_s=round(ups, 2))
        table.append(row)
        if log:
            log(f"autotune: {row}")
    best = max(table, key=lambda r: r["updates_per_s"])
    for w in worlds:
        w.set_tuning(probe_tile=best["probe_tile"], fused=best["fused"])
    return dict(best=best, table=table)
