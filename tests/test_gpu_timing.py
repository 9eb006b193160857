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
