"""Pins the oracle's log lines (logWriteParticleField, src/diagnostics.h:
187-816; column lists src/user.cpp:2147-2231) against closed forms: a face
field with a known divergence, uniform currents and electric fields,
particles with known velocities, and counters driven by known escapes,
impacts, injections and speed clamps."""
import math

import numpy as np
import pytest

from rhybrid_amd import problem as P

MU0 = 4e-7 * math.pi
ME = 9.10938215e-31
EPS0 = 8.854187817e-12


def box(nx=8, ny=7, nz=6, dx=1e5, dt=0.05):
    g = P.Grid(nx, ny, nz, (0, 0, 0), dx, -0.5 * nx * dx, -0.4 * ny * dx, -0.5 * nz * dx)
    p = P.Params(dt=dt, min_rho=2e-13, max_ion_speed2=(8e5) ** 2)
    sp = P.Species("H+", P.QE, P.MP, 3e18, R2_obstacle=(1.2 * dx) ** 2)
    return g, p, sp


def interior(g):
    return g.all_neighbours()


def test_field_log_closed_forms(oracle_mod):
    g, p, sp = box()
    f = P.make_fields(g)
    i, j, k = g.ijk()
    a, B0 = 3e-15, np.array([2e-9, -1e-9, 4e-9])
    xf = g.x0 + (i + 1) * g.dx            # x of each cell's +x face
    f["faceB"][0::3] = B0[0] + a * xf
    f["faceB"][1::3] = B0[1]
    f["faceB"][2::3] = B0[2]
    Ji, Ep, E, U, rho = [1e-9, 2e-9, -3e-9], [1e-4, 0, -2e-4], [3e-3, -1e-3, 2e-3], [-4e5, 5e4, 1e4], 5e-13
    for l in range(3):
        f["cellJi"][l::3] = Ji[l]
        f["cellEp"][l::3] = Ep[l]
        f["nodeE"][l::3] = E[l]
        f["cellUe"][l::3] = U[l]
    f["cellRhoQi"][:] = rho
    flags = P.make_flags(g)
    flags["innerFlagField"][g.nc // 2] = 1          # an interior cell left out of the field log
    empty = P.Particles(np.zeros(0, dtype=np.int64), np.zeros((0, 7)))
    pr = P.Problem("log", g, p, [sp], f, flags, [empty])
    o = oracle_mod.Oracle(pr, threads=1)
    o.log_begin(1.0)
    field, pop = o.log_values(1.5)
    sel = interior(g) & (flags["innerFlagField"] == 0)
    N = sel.sum()
    xc = g.x0 + (i + 0.5) * g.dx
    Bc = np.stack([B0[0] + a * xc, np.full(g.nc, B0[1]), np.full(g.nc, B0[2])], 1)[sel]
    Bt = np.linalg.norm(Bc, axis=1)
    assert field[0] == 1.5
    assert np.allclose(field[1:4], Bc.mean(0), rtol=1e-12)
    assert np.isclose(field[4], Bt.mean(), rtol=1e-12)
    assert np.isclose(field[5], Bt.max(), rtol=1e-12)
    assert np.isclose(field[6], abs(a), rtol=1e-9) and np.isclose(field[7], abs(a), rtol=1e-9)
    assert np.isclose(field[8], g.dx * abs(a) / Bt.min(), rtol=1e-9)
    assert np.isclose(field[9], (Bt ** 2).sum() * g.dV / (2 * MU0), rtol=1e-12)
    assert np.allclose(field[10:13], Ji, rtol=1e-12) and np.isclose(field[13], np.linalg.norm(Ji))
    assert np.isclose(field[14], np.linalg.norm(Ji))
    assert np.allclose(field[15:18], Ep, rtol=1e-12, atol=1e-30) and np.isclose(field[19], np.linalg.norm(Ep))
    assert np.isclose(field[20], N * np.dot(Ep, Ep) * g.dV * EPS0 / 2, rtol=1e-12)
    assert np.allclose(field[21:24], E, rtol=1e-12) and np.isclose(field[25], np.linalg.norm(E))
    assert np.isclose(field[26], N * np.dot(E, E) * g.dV * EPS0 / 2, rtol=1e-12)
    cE = -np.cross(np.array(U)[None, :], Bc)
    assert np.isclose(field[27], (cE ** 2).sum() * g.dV * EPS0 / 2, rtol=1e-12)
    assert np.all(field[28:34] == 0)
    de = math.sqrt(ME / (MU0 * P.QE * rho))
    assert np.isclose(field[34], de) and np.isclose(field[35], de)
    rhoM = p.min_rho * P.MP / P.QE               # no particles: the floor
    di = math.sqrt(P.MP ** 2 / (MU0 * P.QE ** 2 * rhoM))
    assert np.isclose(field[36], di) and np.isclose(field[37], di)
    assert np.isclose(field[38], 2 * math.pi * P.MP / (P.QE * Bt.max()))
    assert np.isclose(field[39], Bt.max() / math.sqrt(MU0 * rhoM))
    assert np.isclose(field[40], np.linalg.norm(U))
    # with the inner cells included, one more cell enters the averages
    f2, _ = o.log_values(2.0, include_inner=True)
    Bc2 = np.stack([B0[0] + a * xc, np.full(g.nc, B0[1]), np.full(g.nc, B0[2])], 1)[interior(g)]
    assert np.allclose(f2[1:4], Bc2.mean(0), rtol=1e-12)
    assert pop[0, 1] == 0 and pop[0, 2] == 0


def test_particle_log_and_counters(oracle_mod):
    g, p, sp = box()
    f = P.make_fields(g)
    flags = P.make_flags(g, R_particle=math.sqrt(sp.R2_obstacle))
    rng = np.random.default_rng(3)
    i, j, k = g.ijk()
    inner = np.nonzero(interior(g) & (flags["innerFlagParticle"] == 0))[0]
    ext = np.nonzero(~interior(g))[0]
    cells_in = rng.choice(inner, 40)
    cells_ext = rng.choice(ext, 7)
    # particles at the origin's cells are inside the obstacle radius
    ci = np.nonzero(flags["innerFlagParticle"])[0]
    cells_imp = ci[:3]
    cells = np.concatenate([cells_in, cells_ext, cells_imp])
    n = cells.size
    st = np.zeros((n, 7))
    st[:, 0:3] = rng.random((n, 3)) * g.dx
    # impacts: put those particles at the point of the cell nearest the origin
    for q, c in enumerate(cells_imp):
        qi = 47 + q
        cx, cy, cz = g.x0 + i[c] * g.dx, g.y0 + j[c] * g.dx, g.z0 + k[c] * g.dx
        st[qi, 0:3] = [np.clip(-cx, 0, g.dx * 0.999), np.clip(-cy, 0, g.dx * 0.999), np.clip(-cz, 0, g.dx * 0.999)]
    st[:, 3:6] = rng.standard_normal((n, 3)) * 3e5
    st[:, 6] = sp.w
    pr = P.Problem("log", g, p, [sp], f, flags, [P.Particles(cells, st)])
    o = oracle_mod.Oracle(pr, threads=1)
    o.log_begin(0.0)
    _, pop = o.log_values(0.0)                  # Dt = 0: rates are zero
    v = st[:, 3:6]
    assert np.isclose(pop[0, 1], sp.w * n) and pop[0, 2] == n
    assert np.allclose(pop[0, 3:6], v.mean(0), rtol=1e-12)
    assert np.isclose(pop[0, 6], np.linalg.norm(v, axis=1).mean(), rtol=1e-12)
    assert np.isclose(pop[0, 7], 0.5 * sp.m * sp.w * (v ** 2).sum(), rtol=1e-12)
    assert np.isclose(pop[0, 8], np.linalg.norm(v, axis=1).max(), rtol=1e-15)
    assert np.all(pop[0, 9:] == 0)
    # escapes (exterior cells) and impacts (inside the obstacle radius)
    cnt0 = o.num_particles(0)
    o.apply_boundary_conditions(0)
    r_imp = []
    for q, c in enumerate(cells_imp):
        qi = 47 + q
        gx = g.x0 + i[c] * g.dx + st[qi, 0]
        gy = g.y0 + j[c] * g.dx + st[qi, 1]
        gz = g.z0 + k[c] * g.dx + st[qi, 2]
        r_imp.append(gx * gx + gy * gy + gz * gz < sp.R2_obstacle)
    r_imp = np.array(r_imp)
    assert r_imp.any()
    assert o.num_particles(0) == cnt0 - 7 - r_imp.sum()
    # maxVi: a fast particle in an accelerating interior cell is clamped
    Dt = 2.5
    _, pop = o.log_values(Dt)
    ve = v[40:47]
    vi = v[47:][r_imp]
    assert np.isclose(pop[0, 9], sp.w * 7 / Dt)
    assert np.isclose(pop[0, 10], sp.w * r_imp.sum() / Dt)
    assert np.isclose(pop[0, 13], 0.5 * sp.m * sp.w * (ve ** 2).sum() / Dt, rtol=1e-12)
    assert np.isclose(pop[0, 14], 0.5 * sp.m * sp.w * (vi ** 2).sum() / Dt, rtol=1e-12)
    # counters restart after a line
    _, pop2 = o.log_values(Dt + 1.0)
    assert np.all(pop2[0, 9:] == 0)


def test_inject_and_clamp_counters(oracle_mod):
    g = P.Grid(10, 6, 6, (0, 1, 1), 1e5, 0.0, 0.0, 0.0)
    p = P.Params(dt=0.1, use_hall=0, max_ion_speed2=(2e5) ** 2)
    n, ppc, U, T = 1e6, 20.0, 1e5, 1e5
    sp = P.Species("H+", P.QE, P.MP, n * g.dV / ppc)
    # two particles above the speed limit in interior cells
    st = np.array([[1e4, 1e4, 1e4, 5e5, 0, 0, sp.w], [2e4, 2e4, 2e4, 0, -3e5, 1e5, sp.w]])
    parts = P.Particles(np.array([g.nx * 2 + 3, g.nx * 3 + 5]), st)
    inj = P.solar_wind_injector((-U, 0, 0), n, T, ppc, seed=4)
    pr = P.Problem("inj", g, p, [sp], P.make_fields(g), P.make_flags(g), [parts], injectors=[inj])
    o = oracle_mod.Oracle(pr, threads=1)
    o.log_begin(0.0)
    o.push(0)
    o.migrate(0)
    before = o.particles(0)[1]
    o.set_timestep(1)
    o.inject(0)
    after = o.particles(0)[1]
    new = after[~np.isin(after[:, 0] + 1e6 * after[:, 3], before[:, 0] + 1e6 * before[:, 3])]
    ninj = o.injected(0)
    assert new.shape[0] == ninj > 0
    Dt = 0.2
    field, pop = o.log_values(Dt)
    assert np.isclose(pop[0, 11], sp.w * ninj / Dt)
    assert np.isclose(pop[0, 12], ninj / Dt * p.dt)
    assert np.isclose(pop[0, 15], 0.5 * sp.m * sp.w * (new[:, 3:6] ** 2).sum() / Dt, rtol=1e-12)
    assert np.isclose(pop[0, 16], 2 / Dt * p.dt)
    # column 31 (maxVw rate) repeats the node-Ue counter (src/diagnostics.h:589)
    assert field[30] == field[29]
