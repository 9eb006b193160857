"""Pins the oracle's Boris-Buneman push (src/particle_propagator_boris_buneman.h:
64-194) and field gather (src/hybrid_propagator.cpp:1740-1891) against closed
forms: exact gyration angle in a uniform B, the E x B drift, acceleration by E
alone, the speed clamp, free streaming, and the face interpolation."""
import numpy as np
import pytest

from rhybrid_amd import problem as P


def box(n=4, dx=1e5, dt=0.01, **pkw):
    g = P.Grid(n, n, n, (1, 1, 1), dx, 0.0, 0.0, 0.0)
    p = P.Params(dt=dt, **pkw)
    sp = P.Species("H+", P.QE, P.MP, 1.0)
    return g, p, sp


def one_particle_problem(g, p, sp, cell, r, v, B=(0, 0, 0), Ue=(0, 0, 0), Ep=(0, 0, 0)):
    fields = P.make_fields(g, B)
    for l in range(3):
        fields["cellUe"][l::3] = Ue[l]
        fields["cellEp"][l::3] = Ep[l]
    st = np.array([[r[0], r[1], r[2], v[0], v[1], v[2], sp.w]], dtype=np.float64)
    parts = P.Particles(np.array([cell], dtype=np.int64), st)
    return P.Problem("one", g, p, [sp], fields, P.make_flags(g), [parts])


def pushed(O, pr):
    o = O.Oracle(pr, threads=1)
    o.push(0)
    return o.particles(0)[1][0]


def test_gyration_angle(oracle_mod):
    g, p, sp = box()
    B0 = 5e-9
    v0 = np.array([1e5, 0.0, 2e4])
    pr = one_particle_problem(g, p, sp, 21, (3e4, 4e4, 5e4), v0, B=(0, 0, B0))
    st = pushed(oracle_mod, pr)
    t = 0.5 * sp.q * p.dt * B0 / sp.m
    theta = 2 * np.arctan(t)
    # dv/dt = q/m v x B turns a proton clockwise about +z
    exp = np.array([v0[0] * np.cos(theta), -v0[0] * np.sin(theta), v0[2]])
    assert np.allclose(st[3:6], exp, rtol=0, atol=1e-9 * np.linalg.norm(v0))
    assert abs(np.linalg.norm(st[3:6]) - np.linalg.norm(v0)) < 1e-9 * np.linalg.norm(v0)


def test_exb_drift_is_stationary(oracle_mod):
    # E = -Ue x B; a particle moving at the E x B velocity (= Ue perpendicular
    # to B) feels no net force and keeps its velocity
    g, p, sp = box()
    B = np.array([0.0, 0.0, 4e-9])
    Ue = np.array([-3e5, 1e5, 0.0])
    pr = one_particle_problem(g, p, sp, 5, (1e4, 2e4, 3e4), Ue, B=tuple(B), Ue=tuple(Ue))
    st = pushed(oracle_mod, pr)
    assert np.allclose(st[3:6], Ue, rtol=1e-12, atol=1e-6)


def test_exb_drift_average(oracle_mod):
    # a particle starting at rest drifts, on average over gyrations, at E x B / B^2
    g, p, sp = box(dt=0.05)
    B = np.array([0.0, 0.0, 1e-8])
    Ue = np.array([2e5, 0.0, 0.0])
    pr = one_particle_problem(g, p, sp, 0, (5e4, 5e4, 5e4), (0.0, 0.0, 0.0), B=tuple(B), Ue=tuple(Ue))
    o = oracle_mod.Oracle(pr, threads=1)
    E = -np.cross(Ue, B)
    vd = np.cross(E, B) / B.dot(B)
    t = 0.5 * sp.q * p.dt * B[2] / sp.m
    per = int(round(2 * np.pi / (2 * np.arctan(t))))
    vs = []
    for _ in range(per * 10):
        o.push(0)
        o.migrate(0)
        vs.append(o.particles(0)[1][0, 3:6].copy())
    vm = np.mean(vs, axis=0)
    assert np.allclose(vm, vd, atol=0.02 * np.linalg.norm(vd))


def test_electric_acceleration(oracle_mod):
    # with B = 0 the two half kicks add to q E dt / m (E from the e-pressure term)
    g, p, sp = box(use_e_pressure=1)
    Ep = np.array([1e-3, -2e-3, 5e-4])
    v0 = np.array([1e4, 0.0, -3e4])
    pr = one_particle_problem(g, p, sp, 9, (1e4, 9e4, 5e4), v0, Ep=tuple(Ep))
    st = pushed(oracle_mod, pr)
    assert np.allclose(st[3:6], v0 + sp.q * Ep * p.dt / sp.m, rtol=1e-13, atol=1e-9)


def test_speed_clamp(oracle_mod):
    g, p, sp = box(use_e_pressure=1, max_ion_speed2=(2e5) ** 2)
    pr = one_particle_problem(g, p, sp, 9, (1e4, 9e4, 5e4), (3e5, 0, 0), Ep=(1e-4, 0, 0))
    o = oracle_mod.Oracle(pr, threads=1)
    o.push(0)
    st = o.particles(0)[1][0]
    assert abs(np.linalg.norm(st[3:6]) - 2e5) < 1e-6
    assert o.counters()["max_vi"] == 1


def test_position_update_and_no_accel_species(oracle_mod):
    g, p, sp = box()
    sp.accelerate = 0
    v = np.array([1e6, -2e6, 3e5])
    r = np.array([2e4, 7e4, 1e4])
    pr = one_particle_problem(g, p, sp, 30, r, v, B=(0, 0, 1e-6), Ue=(1e5, 0, 0))
    st = pushed(oracle_mod, pr)
    assert np.array_equal(st[3:6], v)
    assert np.allclose(st[0:3], r + p.dt * v, rtol=0, atol=1e-9)


def test_face2r_linear_along_own_axis(oracle_mod):
    g, p, sp = box()
    pr = one_particle_problem(g, p, sp, 0, (0, 0, 0), (0, 0, 0))
    o = oracle_mod.Oracle(pr, threads=1)
    rng = np.random.default_rng(3)
    fb = rng.standard_normal(g.nc * 3)
    o.field("faceB")[:] = fb
    n = 1 + 4 * (2 + 4 * 1)  # cell (1,2,1)
    r = np.array([0.3, 0.6, 0.9]) * g.dx
    B, _, _ = o.get_fields_at(n, r)
    mx, my, mz = n - 1, n - 4, n - 16
    exp = [0.3 * fb[3 * n] + 0.7 * fb[3 * mx], 0.6 * fb[3 * n + 1] + 0.4 * fb[3 * my + 1],
           0.9 * fb[3 * n + 2] + 0.1 * fb[3 * mz + 2]]
    assert np.allclose(B, exp, rtol=1e-14, atol=1e-15)


def test_migration_preserves_global_position(oracle_mod):
    pr = P.config_periodic(n=6, ppc=8, seed=11)
    pr.species[0].accelerate = 0
    g = pr.grid

    def glob(cell, st):
        i, j, k = cell % g.nx, (cell // g.nx) % g.ny, cell // (g.nx * g.ny)
        return np.stack([i * g.dx + st[:, 0], j * g.dx + st[:, 1], k * g.dx + st[:, 2]], 1)

    # free streaming: accelerate off, large velocities so that many cross
    o2 = oracle_mod.Oracle(pr, threads=2)
    c0, s0 = o2.particles(0)
    s0[:, 3:6] = np.random.default_rng(0).uniform(-0.9, 0.9, (len(c0), 3)) * g.dx / pr.params.dt
    o2.set_particles(0, c0, s0)
    c0, s0 = o2.particles(0)
    L = np.array([g.nx, g.ny, g.nz]) * g.dx
    # the oracle drops no particle in a periodic box
    o2.push(0)
    o2.migrate(0)
    c1, s1 = o2.particles(0)
    assert len(c1) == len(c0)
    assert np.all((s1[:, 0:3] >= 0) & (s1[:, 0:3] < g.dx))
    # match particles by their (unchanged) velocities
    exp = np.mod(glob(c0, s0) + pr.params.dt * s0[:, 3:6], L)
    o0 = np.lexsort(s0[:, 3:6].T)
    o1 = np.lexsort(s1[:, 3:6].T)
    got = glob(c1[o1], s1[o1])
    d = np.abs(got - exp[o0])
    d = np.minimum(d, L - d)
    assert d.max() < 1e-6


def test_gravity_kick(oracle_mod):
    """With no fields the Boris rotation leaves v unchanged and the
    gravitational kick is exactly -GMdt r / |r|^3 at the particle's global
    position before the move (src/particle_propagator_boris_buneman.h:
    137-145); the box origin is offset so that the global and cell-local
    positions differ."""
    g = P.Grid(4, 4, 4, (1, 1, 1), 1e5, -1.7e5, 0.6e5, -2.2e5)
    GMdt = 3e13
    p = P.Params(dt=0.01, use_gravity=1, GMdt=GMdt)
    sp = P.Species("H+", P.QE, P.MP, 1.0)
    v0 = np.array([2e4, -1e4, 5e3])
    cell, r = 37, np.array([2.5e4, 7.5e4, 4.0e4])
    pr = one_particle_problem(g, p, sp, cell, r, v0)
    st = pushed(oracle_mod, pr)
    i, j, k = cell % 4, (cell // 4) % 4, cell // 16
    rg = np.array([g.x0 + i * g.dx, g.y0 + j * g.dx, g.z0 + k * g.dx]) + r
    want = v0 - GMdt * rg / np.linalg.norm(rg) ** 3
    assert np.allclose(st[3:6], want, rtol=1e-13, atol=0)
    assert np.linalg.norm(want - v0) > 100.0  # the kick is there
    # an exterior cell (not all neighbours) neither accelerates nor feels gravity
    g2 = P.Grid(4, 4, 4, (0, 1, 1), 1e5, -1.7e5, 0.6e5, -2.2e5)
    pr2 = one_particle_problem(g2, p, sp, 36, r, v0)   # i = 0: -x wall
    assert np.array_equal(pushed(oracle_mod, pr2)[3:6], v0)


def test_initial_flow_through_switches_acceleration_off(oracle_mod):
    """Hybrid::useInitialFlowThrough: no particle is accelerated, whatever
    the fields (src/particle_propagator_boris_buneman.h:66-68)."""
    g, p, sp = box(use_e_pressure=1, initial_flow_through=1)
    v = np.array([1e5, 2e4, -3e4])
    r = np.array([2e4, 7e4, 1e4])
    pr = one_particle_problem(g, p, sp, 30, r, v, B=(0, 0, 1e-6), Ue=(1e5, 0, 0), Ep=(1e-3, 0, 0))
    st = pushed(oracle_mod, pr)
    assert np.array_equal(st[3:6], v)
    assert np.allclose(st[0:3], r + p.dt * v, rtol=0, atol=1e-9)
