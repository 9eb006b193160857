"""Per-kernel timing through the C ABI (rh_set_timing, rh_set_timing_only,
rh_kernel_times; the profile:: regions of src/hybrid_propagator.cpp:33-38):
with a filter only the named kernel's launches are timed."""
import pytest

from rhybrid_amd import problem as P

pytestmark = pytest.mark.gpu


def test_timing_filter(gpu_mod):
    pr = P.config_obstacle(nx=16, ny=12, nz=12, ppc=4, R_cells=3, seed=30)
    ctx = gpu_mod.Context(pr, precision=gpu_mod.RH_FP32)
    ns = len(ctx.species)
    ctx.step(1)
    ctx.reset_timing()
    ctx.set_timing_only("move")
    ctx.set_timing(True)
    ctx.step(2)
    ctx.set_timing(False)
    ms, n = ctx.kernel_time("move")
    assert n == 2 * ns and ms > 0
    assert ctx.kernel_time("push") == (0.0, 0)
    assert ctx.kernel_time("deposit") == (0.0, 0)
    # no filter: every launch again
    ctx.reset_timing()
    ctx.set_timing_only(None)
    ctx.set_timing(True)
    ctx.step(1)
    ctx.set_timing(False)
    assert ctx.kernel_time("push")[1] == ns
    assert ctx.kernel_time("move")[1] == ns
    assert ctx.kernel_time("deposit")[1] == ns
    assert ctx.kernel_time("move")[0] > 0


def test_per_species_times(gpu_mod):
    # a launch made for one species is also counted under "<name>:<species>"
    pr = P.config_obstacle(nx=16, ny=12, nz=12, ppc=4, R_cells=3, seed=31)
    ctx = gpu_mod.Context(pr, precision=gpu_mod.RH_FP32)
    ns = len(ctx.species)
    assert ns == 2
    ctx.step(1)
    ctx.reset_timing()
    ctx.set_timing(True)
    ctx.step(3)
    ctx.set_timing(False)
    for nm in ("push", "move", "deposit", "sort_slots", "boundary"):
        ms, n = ctx.kernel_time(nm)
        parts = [ctx.kernel_time("%s:%d" % (nm, s)) for s in range(ns)]
        assert n == 3 * ns and all(p[1] == 3 for p in parts), nm
        assert abs(sum(p[0] for p in parts) - ms) <= 1e-6 * max(ms, 1.0), nm
    # grid kernels belong to no species
    assert ctx.kernel_time("fold")[1] == 3
    assert ctx.kernel_time("fold:0") == (0.0, 0)
    ctx.close()
