"""The reference's log lines (field.log, pop*.log; src/diagnostics.h:187-816)
from the HIP path against the FP64 oracle, through the C ABI, and the files
the library writes (headers of src/user.cpp:2147-2231)."""
import os

import numpy as np
import pytest

from rhybrid_amd import problem as P

pytestmark = pytest.mark.gpu


def inflow_problem(seed=16):
    pr = P.config_obstacle(nx=16, ny=12, nz=12, ppc=8, R_cells=3, seed=seed, inflow=True)
    for sp in pr.species:
        sp.capacity = 3 * pr.particles[0].n
    return pr


def close(a, b, rtol):
    a, b = np.asarray(a), np.asarray(b)
    return np.all(np.abs(a - b) <= rtol * np.maximum(np.abs(b), 1e-300))


def test_log_lines_match_oracle(gpu_mod, oracle_mod):
    pr = inflow_problem()
    print("case log parity np=%d" % pr.n_particles, flush=True)
    ctx = gpu_mod.Context(pr, precision=gpu_mod.RH_FP64)
    ora = oracle_mod.Oracle(pr)
    ctx.log_begin(0.0)
    ora.log_begin(0.0)
    dt = pr.params.dt
    t = 0.0
    for steps in (6, 4):
        ctx.step(steps)
        ora.step(steps)
        t += steps * dt
        fg, pg = ctx.log_values(t)
        fo, po = ora.log_values(t)
        # counters and particle counts are integers: equal
        assert np.array_equal(pg[:, 2], po[:, 2])
        for s in range(2):
            assert po[s, 9] > 0 and po[s, 11] > 0         # escapes and injections happened
        assert close(pg, po, 1e-9), np.max(np.abs(pg - po) / np.maximum(np.abs(po), 1e-300), axis=0)
        # columns 7-9 (div B) are rounding noise on both sides: both at the
        # rounding level of B/dx, not equal
        dx, B = pr.grid.dx, fo[4]
        for a in (fg, fo):
            assert a[6] * dx / B < 1e-13 and a[7] * dx / B < 1e-13 and a[8] < 1e-13
        rest = [q for q in range(41) if q not in (6, 7, 8)]
        rel = np.abs(fg[rest] - fo[rest]) / np.maximum(np.abs(fo[rest]), 1e-300)
        print("field log max rel diff %.2e at column %d" % (rel.max(), rest[rel.argmax()] + 1), flush=True)
        assert rel.max() < 1e-9
    # inner cells included
    fg, _ = ctx.log_values(t, include_inner=True)
    fo, _ = ora.log_values(t, include_inner=True)
    cols = [q for q in range(1, 28) if q not in (6, 7, 8)]
    assert close(fg[cols], fo[cols], 1e-9)


def test_log_files(gpu_mod, oracle_mod, tmp_path):
    pr = inflow_problem(seed=17)
    ctx = gpu_mod.Context(pr, precision=gpu_mod.RH_FP64)
    ora = oracle_mod.Oracle(pr)
    ctx.log_open(str(tmp_path), precision=10)
    ctx.step(3)
    ora.step(3)
    t = 3 * pr.params.dt
    ctx.log_write(t)
    fo, po = ora.log_values(t)
    names = sorted(os.listdir(tmp_path))
    assert names == ["field.log", "pop001_H+.log", "pop002_He++.log"]
    pop = open(tmp_path / "pop001_H+.log").read().splitlines()
    assert pop[:5] == ["% H+", "% m [kg] = +1.6726217160e-27", "% q [C] = +1.6021765300e-19", "% columns = 17",
                       "% 01. Time [s]"]
    assert pop[20] == "% 17. Constraint maxVi rate [#/dt]"
    vals = [float(x) for x in pop[21].split()]
    assert len(vals) == 17 and pop[21].startswith("+")
    assert close(vals, po[0], 1e-9)
    fl = open(tmp_path / "field.log").read().splitlines()
    assert fl[:3] == ["% field", "% columns = 41", "% 01. Time [s]"]
    assert fl[42] == "% 41. Maximum electron fluid speed [m/s]"
    line = fl[43]
    vals = [float(x) for x in line.split()]
    assert len(vals) == 41
    rest = [q for q in range(41) if q not in (6, 7, 8)]
    assert close(np.array(vals)[rest], fo[rest], 1e-9)
    # the extra space before max(|nodeE|), as the reference writes it
    assert "  " in line and line.split("  ")[0].count(" ") == 24


def test_max_B_stop(gpu_mod, tmp_path):
    pr = inflow_problem(seed=18)
    ctx = gpu_mod.Context(pr, precision=gpu_mod.RH_FP32)
    ctx.log_open(str(tmp_path), max_B_stop=1e-12)
    with pytest.raises(gpu_mod.RhError) as e:
        ctx.log_write(0.0)
    assert e.value.code == gpu_mod.RH_ERR_STATE
    assert len(open(tmp_path / "field.log").read().splitlines()) == 44


def test_fp32_log_follows_fp64(gpu_mod, oracle_mod):
    pr = inflow_problem(seed=19)
    ctx = gpu_mod.Context(pr, precision=gpu_mod.RH_FP32)
    ora = oracle_mod.Oracle(pr)
    ctx.step(5)
    ora.step(5)
    t = 5 * pr.params.dt
    fg, pg = ctx.log_values(t)
    fo, po = ora.log_values(t)
    # particle sums: float state, FP64 sums
    # (columns whose values are far from zero: N, avg Vx, avg |V|, energy;
    # avg B along the IMF, |B|, max |B|, B energy, avg and max |Ji|)
    assert close(pg[:, [1, 3, 6, 7]], po[:, [1, 3, 6, 7]], 1e-5)
    assert close(fg[[1, 2, 4, 5, 9, 13, 14]], fo[[1, 2, 4, 5, 9, 13, 14]], 1e-4)
