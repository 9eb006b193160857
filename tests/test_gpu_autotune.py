"""Tile-shape autotuner (rhybrid_amd/autotune.py)."""
import numpy as np
import pytest

from rhybrid_amd import problem as P

def test_candidates_are_valid_shapes():
    g = P.Grid(32, 24, 16, (0, 0, 0), 1e5, 0, 0, 0)
    from rhybrid_amd import autotune
    cs = autotune.candidates(g, min_tiles=16)
    assert cs and all(32 % t[0] == 0 and 24 % t[1] == 0 and 16 % t[2] == 0 for t in cs)
    assert all(8 <= t[0] * t[1] * t[2] <= 64 for t in cs)


@pytest.mark.gpu
def test_tune_picks_the_fastest_measured_shape(gpu_mod):
    from rhybrid_amd import autotune
    pr = P.config_obstacle(nx=32, ny=24, nz=16, ppc=16, R_cells=3, seed=3)
    shapes = autotune.candidates(pr.grid, min_tiles=16)[:4]
    r = autotune.tune(pr, steps=2, shapes=shapes)
    ms = r["ms_per_step"]
    assert len(ms) == len(shapes) and all(np.isfinite(v) and v > 0 for v in ms.values())
    best = "%dx%dx%d" % r["best"]
    assert ms[best] == min(ms.values())
