"""Tile-shape autotuner for the particle kernels (sort and deposit).

The tile shape is fixed when a context is created (rh_grid_desc.tile).  The
autotuner creates one context per candidate shape -- power-of-two sides that
divide the grid, at most 64 cells per tile (the deposit runs one thread per
subcell; the context also rejects shapes whose LDS exceeds the device's),
enough tiles to fill the chip -- runs a few full steps of the
caller's problem on each and keeps the fastest.  Timing is the wall clock of
the library's own steps after one warm-up step, HIP-synchronised.
"""
from __future__ import annotations

import itertools
import time
from typing import Callable, Dict, List, Optional, Sequence, Tuple

from . import problem as P
from .gpu import RH_FP32, Context

Tile = Tuple[int, int, int]


def candidates(grid: P.Grid, precision: int = RH_FP32, min_tiles: int = 2048,
               max_cells: Optional[int] = None) -> List[Tile]:
    """Tile shapes worth timing on `grid`."""
    if max_cells is None:
        max_cells = 64
    sides = (1, 2, 4, 8, 16)
    out = []
    for tx, ty, tz in itertools.product(sides, sides, sides):
        tc = tx * ty * tz
        if tc > max_cells or tc < 8:
            continue
        if grid.nx % tx or grid.ny % ty or grid.nz % tz:
            continue
        if (grid.nx // tx) * (grid.ny // ty) * (grid.nz // tz) < min_tiles:
            continue
        out.append((tx, ty, tz))
    # larger tiles first (fewer halo slots per particle), then the most
    # cubic (smallest extended region for the cells)
    out.sort(key=lambda t: (-t[0] * t[1] * t[2], (t[0] + 2) * (t[1] + 2) * (t[2] + 2), t))
    return out


def tune(problem: P.Problem, precision: int = RH_FP32, steps: int = 3,
         shapes: Optional[Sequence[Tile]] = None, max_shapes: int = 8,
         prepare: Optional[Callable[[Context], None]] = None) -> Dict[str, object]:
    """Time `steps` full steps per candidate shape (after `prepare(ctx)`, e.g.
    a device fill); returns the best shape and every shape's milliseconds per
    step."""
    if shapes is None:
        shapes = candidates(problem.grid, precision)[:max_shapes]
    if not shapes:
        raise ValueError("no tile shape divides the grid")
    times = {}
    for t in shapes:
        ctx = Context(problem, precision=precision, tile=t)
        if prepare is not None:
            prepare(ctx)
        ctx.step(1)
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.step(steps)
        ctx.synchronize()
        times[t] = (time.perf_counter() - t0) / steps * 1e3
        ctx.close()
    best = min(times, key=times.get)
    return {"best": best, "ms_per_step": {"%dx%dx%d" % k: round(v, 4) for k, v in times.items()}}
