"""Tile-size autotuner of the parallel ICD sweep.

The sweep's tile edge T sets both the work of one launch (T^2 columns walked
in sequence by each workgroup) and how many workgroups contend for the
shared error sinogram at once, and the best value depends on the stage's
grid, window width and views per frame (measured on the 4D benchmark
workload: 64x64x2 prefers T = 4 at 119 us per launch, 128x128x4 T = 4 at
221 us, 256x256x8 T = 8 at 337 us; T = 4 costs 524 us there).  The
autotuner times a few equits of the stage itself at each candidate tile
with the library's per-launch HIP events and keeps the fastest; the object
and the error sinogram it changes are restored afterwards.
"""
from __future__ import annotations

import numpy as np

from .lib import K_SWEEP


def tune_tile(stage, candidates=(4, 8, 16), equits=2, seed=1):
    """Sets the stage's tile to the candidate with the shortest mean sweep
    launch over `equits` equits from the stage's current object; returns
    (best tile, {tile: ms per launch}).  A must be built."""
    x0 = stage.get_object()
    timings = {}
    for t in candidates:
        try:
            stage.option("tile", t)
            stage.set_object(x0)
            stage.init_error()
            stage.icd(1, 1, seed)            # tile lists built, caches warm
            stage.synchronize()
            stage.reset_kernel_time()
            for it in range(2, equits + 2):
                stage.icd(1, it, seed)
            stage.synchronize()
            ms, n = stage.kernel_time(K_SWEEP)
            timings[t] = ms / max(n, 1)
        except RuntimeError:
            continue                         # tile slab too large for this N_z
    if not timings:
        raise RuntimeError("no candidate tile fits this stage")
    best = min(timings, key=timings.get)
    stage.option("tile", best)
    stage.set_object(x0)
    stage.init_error()
    stage.reset_kernel_time()
    return best, timings
