"""VLSV output (UserDataOP::writeData, src/operator_userdata.cpp:55-311)
through the C ABI, read back the way the reference's reader reads it
(tools/rhybridpy.py:113-126, 263-290: the footer offset at byte 8, an XML
footer whose elements give each array's offset, type and shape).  Values are
checked against the context's own arrays and against numpy on the
downloaded particles."""
import numpy as np
import pytest

from rhybrid_amd import problem as P
from vlsv_read import read_vlsv

pytestmark = pytest.mark.gpu


def test_vlsv_file(gpu_mod, tmp_path):
    pr = P.config_obstacle(nx=16, ny=12, nz=12, ppc=8, R_cells=3, seed=20, inflow=True)
    for sp in pr.species:
        sp.capacity = 3 * pr.particles[0].n
    g = pr.grid
    ctx = gpu_mod.Context(pr, precision=gpu_mod.RH_FP64)
    ctx.step(4)
    fields = ["faceB", "cellRhoQi", "cellB", "cellJ", "cellUe", "cellJi", "cellEp", "nodeRhoQi", "nodeE",
              "nodeB", "nodeJ", "nodeUe", "nodeJi", "nodeEta"]
    flags = ["innerFlagField", "innerFlagNode", "innerFlagParticle", "innerFlagCellEp"]
    path = str(tmp_path / "state00004.vlsv")
    t = 4 * pr.params.dt
    ctx.write_vlsv(path, t, 4, fields + flags + ["ne", "cellDivB", "cellNPles", "n", "v", "T"])
    f = read_vlsv(path)
    assert f[("PARAMETER", "time")][0][0] == t
    assert f[("PARAMETER", "dt")][0][0] == pr.params.dt
    assert list(f[("MESH_BBOX", "")][0]) == [g.nx, g.ny, g.nz, 1, 1, 1]
    assert f[("MESH_BBOX", "")][1]["mesh"] == "SpatialGrid"
    x = f[("MESH_NODE_CRDS_X", "")][0]
    assert x.size == g.nx + 1 and x[0] == g.x0 and np.isclose(x[-1], g.x0 + g.nx * g.dx)
    cid = f[("VARIABLE", "CellID")][0]
    assert np.array_equal(cid, np.arange(g.nc))
    for name in fields:
        data, a = f[("VARIABLE", name)]
        assert a["type"] == "celldata" and a["mesh"] == "SpatialGrid"
        assert np.array_equal(data.ravel(), ctx.field(name)), name
    for name in flags:
        assert np.array_equal(f[("VARIABLE", name)][0], pr.flags[name]), name
    assert np.allclose(f[("VARIABLE", "ne")][0], ctx.field("cellRhoQi") / P.QE, rtol=1e-15)
    # cellDivB: interior cells only
    fb = ctx.field("faceB").reshape(g.nz, g.ny, g.nx, 3)
    div = ((fb[..., 0] - np.roll(fb[..., 0], 1, 2)) + (fb[..., 1] - np.roll(fb[..., 1], 1, 1))
           + (fb[..., 2] - np.roll(fb[..., 2], 1, 0))) / g.dx
    inner = g.all_neighbours()
    got = f[("VARIABLE", "cellDivB")][0]
    assert np.allclose(got[inner], div.ravel()[inner], rtol=0, atol=1e-30)
    assert np.all(got[~inner] == 0)
    # particle moments from the downloaded particles
    cells, states = [], []
    for s in range(2):
        c, st = ctx.particles(s)
        cells.append(c)
        states.append(st)
    npl = np.bincount(np.concatenate(cells), minlength=g.nc)
    assert np.array_equal(f[("VARIABLE", "cellNPles")][0], npl)
    for s, sp in enumerate(pr.species):
        c, st = cells[s], states[s]
        w = st[:, 6]
        W = np.bincount(c, w, minlength=g.nc)
        U = np.stack([np.bincount(c, w * st[:, 3 + l], minlength=g.nc) for l in range(3)], 1)
        U = np.where(W[:, None] > 0, U / np.maximum(W, 1e-300)[:, None], 0.0)
        dv = st[:, 3:6] - U[c]
        T = np.bincount(c, w * (dv ** 2).sum(1), minlength=g.nc)
        T = np.where(W > 0, T * sp.m / (3 * np.maximum(W, 1e-300) * P.KB), 0.0)
        assert np.allclose(f[("VARIABLE", "n_" + sp.name)][0], W / g.dV, rtol=1e-12, atol=0)
        assert np.allclose(f[("VARIABLE", "v_" + sp.name)][0], U, rtol=1e-10, atol=1e-6)
        assert np.allclose(f[("VARIABLE", "T_" + sp.name)][0], T, rtol=1e-8, atol=1e-6)


def test_vlsv_population_groups_and_errors(gpu_mod, tmp_path):
    pr = P.config_obstacle(nx=16, ny=12, nz=12, ppc=4, R_cells=3, seed=21)
    g = pr.grid
    ctx = gpu_mod.Context(pr, precision=gpu_mod.RH_FP32)
    path = str(tmp_path / "ions.vlsv")
    ctx.write_vlsv(path, 0.0, 0, ["n"], populations=["ions", "ions"])
    f = read_vlsv(path)
    W = sum(np.bincount(p.cell, p.state[:, 6], minlength=g.nc) for p in pr.particles)
    assert np.allclose(f[("VARIABLE", "n_ions")][0], W / g.dV, rtol=1e-12)
    assert ("VARIABLE", "n_H+") not in f
    with pytest.raises(gpu_mod.RhError) as e:
        ctx.write_vlsv(path, 0.0, 0, ["faceJ"])
    assert e.value.code == gpu_mod.RH_ERR_ARG


def test_vlsv_temporal_averages(gpu_mod, tmp_path):
    """WRITE_GRID_TEMPORAL_AVERAGES: cellBAverage is the mean of cellB over
    the field updates since the previous file, n_<pop>_ave the mean charge
    density of the population's species over q dV and v_<pop>_ave their
    summed current over their summed charge (src/operator_userdata.cpp:
    156-232); writing restarts the sums.  Checked against the per-step
    arrays in FP64, two species combined into one population, on the cells
    the rho floor and the Neumann copies leave alone."""
    pr = P.config_obstacle(nx=16, ny=12, nz=12, ppc=8, R_cells=3, seed=24, inflow=True)
    for sp in pr.species:
        sp.capacity = 3 * pr.particles[0].n
    print("case temporal averages", flush=True)
    ctx = gpu_mod.Context(pr, precision=gpu_mod.RH_FP64)
    with pytest.raises(gpu_mod.RhError):
        ctx.write_vlsv(str(tmp_path / "off.vlsv"), 0.0, 0, ["cellBAverage"], populations=["ions", "ions"])
    ctx.set_temporal_averages(True)
    inside = pr.grid.all_neighbours()
    q0, dV = pr.species[0].q, pr.grid.dV
    for window, steps in enumerate((4, 3)):
        B, rho, J = [], [], []
        for _ in range(steps):
            ctx.step(1)
            B.append(ctx.field("cellB").reshape(-1, 3))
            rho.append(ctx.field("cellRhoQi"))
            J.append(ctx.field("cellJi").reshape(-1, 3))
        path = str(tmp_path / ("ave%d.vlsv" % window))
        ctx.write_vlsv(path, 0.0, steps, ["cellBAverage", "n_ave", "v_ave"], populations=["ions", "ions"])
        f = read_vlsv(path)
        np.testing.assert_allclose(f[("VARIABLE", "cellBAverage")][0], np.mean(B, axis=0), rtol=1e-12, atol=1e-25)
        ok = inside & (np.min(rho, axis=0) > 1.0001 * pr.params.min_rho)
        assert ok.sum() > 0.5 * inside.sum()
        n_ave = f[("VARIABLE", "n_ions_ave")][0]
        np.testing.assert_allclose(n_ave[ok], (np.mean(rho, axis=0) / q0)[ok], rtol=1e-11)
        v_ave = f[("VARIABLE", "v_ions_ave")][0]
        want = np.sum(J, axis=0) / np.sum(rho, axis=0)[:, None]
        np.testing.assert_allclose(v_ave[ok], want[ok], rtol=1e-9, atol=1e-6)
    # an immediate second file: the sums restarted at the previous one
    path = str(tmp_path / "empty.vlsv")
    ctx.write_vlsv(path, 0.0, 0, ["cellBAverage", "n_ave"], populations=["ions", "ions"])
    f = read_vlsv(path)
    assert not f[("VARIABLE", "cellBAverage")][0].any()
    assert not f[("VARIABLE", "n_ions_ave")][0].any()
    ctx.close()
