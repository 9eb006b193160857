"""Pins the oracle's cloud-in-cell accumulation (src/particle_accumulator.cpp:
43-156, 260-362) against an independent brute-force tent-function deposit,
exact charge and current conservation, and the reference's own
accumulation check (examples/test_accumulation.cfg with
examples/test_accumulation_sum_grid.py: sum n dV = 1e27 exactly)."""
import numpy as np

from rhybrid_amd import problem as P


def brute_cic(g, cell, st, q, periodic=True):
    """Independent CIC: weight of a particle on cell c is prod over axes of
    max(0, 1 - |x_p - x_c|/dx) with cell centres x_c (tent function)."""
    rho = np.zeros(g.nc)
    J = np.zeros((g.nc, 3))
    dims = np.array([g.nx, g.ny, g.nz])
    i = np.stack([cell % g.nx, (cell // g.nx) % g.ny, cell // (g.nx * g.ny)], 1)
    xg = (i + st[:, 0:3] / g.dx)  # global position in cell units, centres at c+0.5
    for p in range(len(cell)):
        base = np.floor(xg[p] - 0.5).astype(int)
        for dz in (0, 1):
            for dy in (0, 1):
                for dxx in (0, 1):
                    c = base + np.array([dxx, dy, dz])
                    w = np.prod(np.maximum(0.0, 1.0 - np.abs(xg[p] - (c + 0.5))))
                    if periodic:
                        c = np.mod(c, dims)
                    elif np.any(c < 0) or np.any(c >= dims):
                        continue
                    n = c[0] + g.nx * (c[1] + g.ny * c[2])
                    rho[n] += w * st[p, 6] * q
                    J[n] += w * st[p, 6] * q * st[p, 3:6]
    return rho / g.dV, J / g.dV


def test_brute_force_cic_periodic(oracle_mod):
    pr = P.config_periodic(n=5, ppc=3, seed=4)
    pr.params.min_rho = 0.0
    o = oracle_mod.Oracle(pr, threads=2)
    o.deposit()
    cell, st = o.particles(0)
    rho, J = brute_cic(pr.grid, cell, st, pr.species[0].q)
    assert np.allclose(o.field("cellRhoQi"), rho, rtol=1e-12, atol=0)
    assert np.allclose(o.field("cellJi").reshape(-1, 3), J, rtol=1e-10, atol=1e-12 * np.abs(J).max())


def test_brute_force_cic_open_boundaries(oracle_mod):
    # deposits falling outside a non-periodic domain are dropped
    pr = P.config_obstacle(nx=6, ny=5, nz=5, ppc=2, R_cells=0.5, seed=2)
    pr.params.min_rho = 0.0
    # put particles into exterior cells as well
    g = pr.grid
    rng = np.random.default_rng(1)
    n = 200
    cell = rng.integers(0, g.nc, n)
    st = np.zeros((n, 7))
    st[:, 0:3] = rng.random((n, 3)) * g.dx
    st[:, 3:6] = rng.standard_normal((n, 3)) * 1e5
    st[:, 6] = pr.species[0].w
    pr.particles = [P.Particles(cell, st), P.Particles(np.zeros(0, np.int64), np.zeros((0, 7)))]
    o = oracle_mod.Oracle(pr, threads=2)
    o.deposit()
    c2, s2 = o.particles(0)
    rho, J = brute_cic(g, c2, s2, pr.species[0].q, periodic=False)
    assert np.allclose(o.field("cellRhoQi"), rho, rtol=1e-12, atol=0)
    assert np.allclose(o.field("cellJi").reshape(-1, 3), J, rtol=1e-10, atol=1e-12 * np.abs(J).max())


def test_single_particle_weights(oracle_mod):
    g = P.Grid(4, 4, 4, (1, 1, 1), 1.0, 0, 0, 0)
    p = P.Params(dt=1.0)
    sp = P.Species("H+", 1.0, 1.0, 1.0)
    for r, expect in [((0.5, 0.5, 0.5), {21: 1.0}),          # centre: all in own cell
                      ((0.0, 0.0, 0.0), None)]:                # corner: 1/8 to 8 cells
        parts = P.Particles(np.array([21]), np.array([[*r, 0, 0, 0, 1.0]]))
        pr = P.Problem("t", g, p, [sp], P.make_fields(g), P.make_flags(g), [parts])
        o = oracle_mod.Oracle(pr, threads=1)
        o.deposit()
        rho = o.field("cellRhoQi")
        if expect:
            assert rho[21] == 1.0 and rho.sum() == 1.0
        else:
            nz = np.nonzero(rho)[0]
            assert len(nz) == 8 and np.allclose(rho[nz], 0.125, rtol=0, atol=0)
            # the 8 cells are (0..1)^3 around the lower corner of cell (1,1,1)
            assert set(nz.tolist()) == {i + 4 * (j + 4 * k) for i in (0, 1) for j in (0, 1) for k in (0, 1)}


def test_conservation_and_translation(oracle_mod):
    pr = P.config_periodic(n=6, ppc=10, seed=9)
    pr.params.min_rho = 0.0
    o = oracle_mod.Oracle(pr, threads=2)
    o.deposit()
    rho = o.field("cellRhoQi").copy()
    J = o.field("cellJi").reshape(-1, 3).copy()
    c, st = o.particles(0)
    q, w, g = pr.species[0].q, pr.species[0].w, pr.grid
    assert abs(rho.sum() * g.dV / (q * w * len(c)) - 1) < 1e-13
    assert np.allclose(J.sum(0) * g.dV, q * w * st[:, 3:6].sum(0), rtol=1e-12)
    # shifting every particle by one cell in x shifts the moments by one cell
    i = c % g.nx
    c2 = (c - i) + (i + 1) % g.nx
    o.set_particles(0, c2, st)
    o.deposit()
    rho2 = o.field("cellRhoQi").reshape(g.nz, g.ny, g.nx)
    assert np.allclose(rho2, np.roll(rho.reshape(g.nz, g.ny, g.nx), 1, axis=2), rtol=1e-12, atol=0)


def test_rho_floor(oracle_mod):
    pr = P.config_periodic(n=4, ppc=1, seed=3)
    pr.params.min_rho = 1e30
    o = oracle_mod.Oracle(pr, threads=1)
    o.deposit()
    assert np.all(o.field("cellRhoQi") == 1e30)
    assert o.counters()["min_cell_rho"] == pr.grid.nc


def test_reference_accumulation_check(oracle_mod):
    """examples/test_accumulation.cfg: 40^3 periodic cells of 250 km, H+ at
    1e6 m^-3, 30 ppc, T = 1e5 K, no acceleration; the reference's checker
    (examples/test_accumulation_sum_grid.py) requires sum(n dV) = 1e27 ions.
    Run here on a 20^3 box of the same cells (sum = 1.25e26) for 5 steps."""
    n, dx = 20, 250e3
    g = P.Grid(n, n, n, (1, 1, 1), dx, -0.5 * n * dx, -0.5 * n * dx, -0.5 * n * dx)
    sp = P.Species("H+", P.QE, P.MP, 1e6 * g.dV / 30, accelerate=0)
    p = P.Params(dt=0.01, min_rho=0.0)
    parts = P.uniform_particles(g, sp, 30, (0, 0, 0), 1e5, np.random.default_rng(6461674))
    pr = P.Problem("acc", g, p, [sp], P.make_fields(g), P.make_flags(g), [parts])
    o = oracle_mod.Oracle(pr, threads=4)
    ref = 1e6 * (n * dx) ** 3
    for _ in range(5):
        o.step(1)
        total = (o.field("cellRhoQi") / P.QE).sum() * g.dV
        assert abs(total / ref - 1) < 1e-12


def test_reference_flow_check(oracle_mod):
    """tests/run_tests.py:56-58,360-403: domain-mean density and bulk velocity
    of a free flow (accelerate = 0) stay within 1/sqrt(ppc) of the inputs."""
    pr = P.config_periodic(n=10, ppc=20, seed=5)
    pr.species[0].accelerate = 0
    o = oracle_mod.Oracle(pr, threads=4)
    o.step(10)
    ppc = 20
    var = 1.0 / np.sqrt(ppc)
    ndens = o.field("cellRhoQi") / P.QE
    U = o.field("cellJi").reshape(-1, 3) / o.field("cellRhoQi")[:, None]
    assert abs(ndens.mean() / 3e6 - 1) < var
    assert abs(U[:, 0].mean() / -400e3 - 1) < var
    assert np.all(np.abs(U[:, 1:].mean(0)) < var * 400e3)
