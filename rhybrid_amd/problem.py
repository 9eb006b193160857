"""Problem set-up for the hot path: grid, parameters, species, initial fields
and seeded initial particles, in the reference's layouts.

Everything here is host-side numpy.  The same ``Problem`` object is uploaded
to the GPU context (``rhybrid_amd.gpu``) and, in the tests, to the FP64
oracle, so both start from identical bytes.

Derived quantities follow the reference's initialisation code:
  macroparticle weight  w = n*dV/ppc          src/particle_injector.cpp:305,682
  thermal speed         vth = sqrt(kB*T/m)    src/particle_injector.cpp:310,687
  min ion charge dens.  rel * upstream rho_q  src/user.cpp:1779-1789
  e-pressure coeff      kB*Te/e (gamma=1), 2*kB*Te/(ne*e^2) (gamma=2)
                                              src/user.cpp:1579-1589
  ion speed cap         min(max_ion_velocity, 0.9*dx/dt)  src/user.cpp:1774-1777
  inner-boundary flags  src/user.cpp:1882-1912
  initial B (constantB) faces and cells      src/user.cpp:2107-2141,
                                              src/magnetic_field.h:32-36
  uniform fill          src/particle_injector.cpp:187-240 (epsCell = 1e-4)
"""
from __future__ import annotations

import dataclasses
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np

# CODATA values (Corsair's constants table is not part of the reference tree).
KB = 1.3806505e-23
QE = 1.60217653e-19
MP = 1.672621716e-27
MU0 = 4.0e-7 * np.pi
EPS_CELL = 1.0e-4

FIELD_NAMES = ["faceB", "cellB", "cellJ", "cellUe", "cellJi", "cellEp", "cellRhoQi",
               "nodeRhoQi", "nodeE", "nodeB", "nodeJ", "nodeUe", "nodeJi", "nodeEta"]
FIELD_DIM = [3, 3, 3, 3, 3, 3, 1, 1, 3, 3, 3, 3, 3, 1]
FLAG_NAMES = ["innerFlagField", "innerFlagNode", "innerFlagParticle", "innerFlagCellEp"]


@dataclass
class Grid:
    nx: int
    ny: int
    nz: int
    periodic: Tuple[int, int, int]
    dx: float
    x0: float
    y0: float
    z0: float

    @property
    def nc(self) -> int:
        return self.nx * self.ny * self.nz

    @property
    def dV(self) -> float:
        return self.dx ** 3

    def ijk(self):
        n = np.arange(self.nc, dtype=np.int64)
        return n % self.nx, (n // self.nx) % self.ny, n // (self.nx * self.ny)

    def all_neighbours(self) -> np.ndarray:
        """True where all 26 neighbours exist (not an exterior cell)."""
        i, j, k = self.ijk()
        ok = np.ones(self.nc, dtype=bool)
        for d, (c, n) in enumerate(zip((i, j, k), (self.nx, self.ny, self.nz))):
            if not self.periodic[d]:
                ok &= (c >= 1) & (c <= n - 2)
        return ok


@dataclass
class Params:
    dt: float
    use_hall: int = 1
    use_e_pressure: int = 0
    isothermal: int = 1
    e_pressure_coeff: float = 0.0
    max_ion_speed2: float = 1e20
    max_electron_speed2: float = 1e20
    min_rho: float = 0.0
    max_E2: float = 0.0
    max_whistler: float = 0.0
    n_linear_smoothings: int = 0
    use_gaussian: int = 0
    gauss_coeffs: Tuple[float, float, float, float] = (0.0, 0.0, 0.0, 0.0)
    imf: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    imf_cell: Tuple[int, int, int, int, int, int] = (0, 0, 0, 0, 0, 0)
    imf_face: Tuple[int, int, int, int, int, int] = (0, 0, 0, 0, 0, 0)
    obstacle_ue: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    initial_flow_through: int = 0
    use_gravity: int = 0
    GMdt: float = 0.0


@dataclass
class Species:
    name: str
    q: float
    m: float
    w: float
    R2_obstacle: float = -1.0
    accumulate: int = 1
    accelerate: int = 1
    capacity: int = 0       # particle store size to reserve on the GPU (0: 1.5x the first upload)


@dataclass
class Particles:
    """Reference layout: global cell index and AoS state[7] =
    {x,y,z (cell-local, m), vx,vy,vz (m/s), w} (src/particle_definition.h:29)."""
    cell: np.ndarray   # int64 [n]
    state: np.ndarray  # float64 [n, 7]

    @property
    def n(self) -> int:
        return int(self.cell.shape[0])


INJ_NONE, INJ_SOLAR_WIND, INJ_EXOSPHERE, INJ_UNIFORM = 0, 1, 2, 3
ALL_SPACE = (-1e50, 1e50, -1e50, 1e50, -1e50, 1e50)


@dataclass
class Injector:
    """A species' injector (src/particle_injector.cpp), as the C ABI's
    rh_injector_desc: the solar-wind inflow through one outer wall
    (InjectorSolarWind, :380-700) or photoions from a per-cell rate table
    (InjectorExosphere, :1489-1840)."""
    kind: int
    seed: int
    temperature: float
    velocity: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    density: float = 0.0
    ppc: float = 0.0
    cell_rate: Optional[np.ndarray] = None   # exosphere: macroparticles per cell per dt
    exobase_radius: float = 0.0
    box: Tuple[float, float, float, float, float, float] = ALL_SPACE  # uniform: filled cell centres


@dataclass
class Problem:
    name: str
    grid: Grid
    params: Params
    species: List[Species]
    fields: Dict[str, np.ndarray]
    flags: Dict[str, np.ndarray]
    particles: List[Particles]
    steps: int = 1
    meta: Dict[str, object] = field(default_factory=dict)
    injectors: List[Optional[Injector]] = field(default_factory=list)

    def copy(self) -> "Problem":
        return dataclasses.replace(
            self,
            fields={k: v.copy() for k, v in self.fields.items()},
            flags={k: v.copy() for k, v in self.flags.items()},
            particles=[Particles(p.cell.copy(), p.state.copy()) for p in self.particles],
            meta=dict(self.meta), injectors=list(self.injectors))

    @property
    def n_particles(self) -> int:
        return sum(p.n for p in self.particles)


def gaussian_coeffs(sigma: float) -> Tuple[float, float, float, float]:
    """Nodal E smoothing weights normalised over 27 nodes (src/user.cpp:1794-1807)."""
    g = lambda d: np.exp(-0.5 * d * d / (sigma * sigma)) / (sigma * np.sqrt(2 * np.pi))
    c = [g(0.0), g(1.0), g(np.sqrt(2.0)), g(np.sqrt(3.0))]
    s = c[0] + 6 * c[1] + 12 * c[2] + 8 * c[3]
    return tuple(float(x / s) for x in c)


def make_fields(grid: Grid, B0=(0.0, 0.0, 0.0), eta: float = 0.0, intrinsic_B=None, resistivity=None,
                dt: float = 0.0, u_bulk: float = 0.0) -> Dict[str, np.ndarray]:
    """The initial grid arrays: faceB and cellB of the constantB profile B0
    (USE_B_INITIAL, src/user.cpp:2107-2141) or of `intrinsic_B`
    (profiles.IntrinsicB: the reference's magnetic_field.h profiles), and
    nodeEta = eta or the `resistivity` profile (profiles.Resistivity, with its
    unit conversion by dt and u_bulk and the outer-boundary multipliers)."""
    from . import profiles
    f = {n: np.zeros(grid.nc * d, dtype=np.float64) for n, d in zip(FIELD_NAMES, FIELD_DIM)}
    if intrinsic_B is None:
        for l in range(3):
            f["faceB"][l::3] = B0[l]
            f["cellB"][l::3] = B0[l]
    else:
        f["faceB"], f["cellB"] = profiles.initial_B(grid, intrinsic_B)
    if resistivity is None:
        f["nodeEta"][:] = eta
    else:
        if not dt > 0:
            raise ValueError("a resistivity profile needs dt (its grid and td units)")
        f["nodeEta"] = profiles.node_eta(grid, resistivity, dt, u_bulk)
    return f


def make_flags(grid: Grid, R_field: float = -1.0, R_particle: float = -1.0,
               R_zero_pe: float = -1.0) -> Dict[str, np.ndarray]:
    """Inner-boundary flags (src/user.cpp:1882-1912)."""
    i, j, k = grid.ijk()
    dx = grid.dx
    xc = grid.x0 + (i + 0.5) * dx
    yc = grid.y0 + (j + 0.5) * dx
    zc = grid.z0 + (k + 0.5) * dx
    r2 = xc * xc + yc * yc + zc * zc
    rp2 = (np.sqrt(r2) - 0.5 * np.sqrt(3.0) * dx) ** 2
    xn, yn, zn = xc + 0.5 * dx, yc + 0.5 * dx, zc + 0.5 * dx
    rn2 = xn * xn + yn * yn + zn * zn
    R2f = R_field * R_field if R_field > 0 else -1.0
    R2p = R_particle * R_particle if R_particle > 0 else -1.0
    R2e = R_zero_pe * R_zero_pe if R_zero_pe > 0 else 0.0
    return {
        "innerFlagField": (r2 < R2f).astype(np.uint8),
        "innerFlagNode": (rn2 < R2f).astype(np.uint8),
        "innerFlagParticle": (rp2 < R2p).astype(np.uint8),
        "innerFlagCellEp": (rp2 < R2e).astype(np.uint8),
    }


def uniform_particles(grid: Grid, sp: Species, ppc: int, velocity, T: float,
                      rng: np.random.Generator, cells: Optional[np.ndarray] = None,
                      R_exclude: float = -1.0) -> Particles:
    """Uniform Maxwellian fill, InjectorUniform::injectParticles
    (src/particle_injector.cpp:187-240): positions uniform in the cell shrunk by
    epsCell, velocities U + vth*N(0,1) per component."""
    if cells is None:
        cells = np.nonzero(grid.all_neighbours())[0]
    cells = np.asarray(cells, dtype=np.int64)
    n = cells.size * ppc
    cell = np.repeat(cells, ppc)
    dx = grid.dx
    st = np.empty((n, 7), dtype=np.float64)
    st[:, 0:3] = 0.5 * dx + (1.0 - EPS_CELL) * dx * (rng.random((n, 3)) - 0.5)
    vth = np.sqrt(KB * T / sp.m) if T > 0 else 0.0
    st[:, 3:6] = np.asarray(velocity, dtype=np.float64)[None, :] + vth * rng.standard_normal((n, 3))
    st[:, 6] = sp.w
    if R_exclude > 0:
        i = cell % grid.nx
        j = (cell // grid.nx) % grid.ny
        k = cell // (grid.nx * grid.ny)
        gx = grid.x0 + i * dx + st[:, 0]
        gy = grid.y0 + j * dx + st[:, 1]
        gz = grid.z0 + k * dx + st[:, 2]
        keep = gx * gx + gy * gy + gz * gz >= R_exclude * R_exclude
        cell, st = cell[keep], st[keep]
    return Particles(cell, st)


def _constraints(p: Params, dx: float, dt: float, max_vi: float, max_ve: float,
                 rho_up: float, min_rho_rel: float, max_vw: float) -> None:
    vi = max_vi if max_vi >= 0 else 1e10
    if vi > dx / dt:
        vi = 0.9 * dx / dt
    p.max_ion_speed2 = vi * vi
    ve = max_ve if max_ve >= 0 else 1e10
    p.max_electron_speed2 = ve * ve
    p.min_rho = max(min_rho_rel, 0.0) * rho_up
    p.max_whistler = max_vw


def config_periodic(n: int = 32, ppc: int = 64, steps: int = 100, seed: int = 6461674,
                    dx: float = 100e3, dt: float = 0.01) -> Problem:
    """Plain method, small: periodic box of n^3 cells, dx=100 km, dt=0.01 s, H+
    at n=3 cm^-3 with 64 ppc, drifting Maxwellian v=(-400,0,0) km/s, T=1e5 K,
    B=(0,3,0) nT, isothermal electrons Te=1e5 K, eta=0."""
    grid = Grid(n, n, n, (1, 1, 1), dx, -0.5 * n * dx, -0.5 * n * dx, -0.5 * n * dx)
    dens, T, Te = 3e6, 1e5, 1e5
    U = (-400e3, 0.0, 0.0)
    B0 = (0.0, 3e-9, 0.0)
    sp = Species("H+", QE, MP, dens * grid.dV / ppc)
    p = Params(dt=dt, use_hall=1, use_e_pressure=1, isothermal=1,
               e_pressure_coeff=KB * Te / QE, imf=B0)
    # constraints as in examples/empty_box_periodic.cfg [Constraints]
    _constraints(p, dx, dt, 1500e3, 1500e3, dens * QE, 0.05, 1500e3)
    rng = np.random.default_rng(seed)
    parts = uniform_particles(grid, sp, ppc, U, T, rng)
    return Problem("periodic%d_ppc%d" % (n, ppc), grid, p, [sp], make_fields(grid, B0),
                   make_flags(grid), [parts], steps=steps,
                   meta={"ppc": ppc, "density": dens, "U": U, "T": T})


def config_obstacle(nx: int = 128, ny: int = 96, nz: int = 96, ppc: int = 32,
                    steps: int = 500, seed: int = 6461674, dx: float = 300e3,
                    dt: float = 0.04, R_cells: float = 10.0, he_fraction: float = 0.04,
                    n_sw: float = 3e6, inflow: bool = False) -> Problem:
    """Inflow/outflow box with an absorbing spherical obstacle of radius
    R_cells cells at the centre; H+ and He++ (he_fraction by number) solar
    wind, ppc each, drifting Maxwellians at 450 km/s, T=1e5 K, IMF
    B=(-1.5,2.5,0) nT, IMF on the +x wall (as examples/mars.cfg [IMF]).
    Choices the workload description leaves open: n_sw=3 cm^-3 total,
    dt=0.04 s, isothermal electrons at Te=1e5 K, eta=0, uniform preload of
    the interior (the reference's InjectorUniform); with inflow, both
    species also enter through the +x wall every step (InjectorSolarWind,
    as examples/anomaly.cfg pairs a uniform and a solar-wind population)."""
    grid = Grid(nx, ny, nz, (0, 0, 0), dx, -0.5 * nx * dx, -0.5 * ny * dx, -0.5 * nz * dx)
    T, Te = 1e5, 1e5
    U = (-450e3, 0.0, 0.0)
    B0 = (-1.5e-9, 2.5e-9, 0.0)
    R = R_cells * dx
    nH, nHe = n_sw * (1 - he_fraction), n_sw * he_fraction
    spH = Species("H+", QE, MP, nH * grid.dV / ppc, R2_obstacle=R * R)
    spHe = Species("He++", 2 * QE, 3.97 * MP, nHe * grid.dV / ppc, R2_obstacle=R * R)
    rho_up = (nH + 2 * nHe) * QE
    p = Params(dt=dt, use_hall=1, use_e_pressure=1, isothermal=1,
               e_pressure_coeff=KB * Te / QE, imf=B0,
               imf_cell=(1, 0, 0, 0, 0, 0), imf_face=(1, 0, 0, 0, 0, 0))
    _constraints(p, dx, dt, 1500e3, 1500e3, rho_up, 0.05, 1500e3)
    rng = np.random.default_rng(seed)
    flags = make_flags(grid, R_field=R, R_particle=R)
    pH = uniform_particles(grid, spH, ppc, U, T, rng, R_exclude=R)
    pHe = uniform_particles(grid, spHe, ppc, U, T, rng, R_exclude=R)
    injectors = []
    if inflow:
        injectors = [solar_wind_injector(U, nH, T, ppc, seed=seed + 1),
                     solar_wind_injector(U, nHe, T, ppc, seed=seed + 2)]
    return Problem("obstacle%dx%dx%d_ppc%d%s" % (nx, ny, nz, ppc, "_inflow" if inflow else ""), grid, p,
                   [spH, spHe], make_fields(grid, B0), flags, [pH, pHe], steps=steps,
                   meta={"ppc": ppc, "density": n_sw, "U": U, "T": T, "R_cells": R_cells, "inflow": inflow},
                   injectors=injectors)


def interior_box(grid: Grid) -> Tuple[float, float, float, float, float, float]:
    """The extent holding the centres of the cells with all 26 neighbours (a
    periodic axis: all of it)."""
    out = []
    for d, (n, o) in enumerate(((grid.nx, grid.x0), (grid.ny, grid.y0), (grid.nz, grid.z0))):
        if grid.periodic[d]:
            out += [o - grid.dx, o + (n + 1) * grid.dx]
        else:
            out += [o + grid.dx, o + (n - 1) * grid.dx]
    return tuple(out)


def config_obstacle_filled(nx: int = 128, ny: int = 96, nz: int = 96, ppc: int = 32, seed: int = 6461674,
                           R_cells: float = 10.0, **kw) -> Problem:
    """config_obstacle with inflow whose uniform preload is done on the device
    (InjectorUniform at timestep 1, meta["fill"], one injector per species,
    seeded by `seed`) instead of uploaded: for runs at hundreds of particles
    per cell.  The particles inside the obstacle are removed by the first
    step's boundary conditions."""
    base = config_obstacle(nx=nx, ny=ny, nz=nz, ppc=1, seed=seed, R_cells=R_cells, inflow=True, **kw)
    g, U, T = base.grid, base.meta["U"], base.meta["T"]
    n_sw, he = base.meta["density"], kw.get("he_fraction", 0.04)
    dens = (n_sw * (1 - he), n_sw * he)
    species = [dataclasses.replace(sp, w=d * g.dV / ppc, capacity=int(1.25 * ppc * g.nc) + 100000)
               for sp, d in zip(base.species, dens)]
    fill = [uniform_injector(U, d, T, ppc, seed=seed + 11 + q, box=interior_box(g)) for q, d in enumerate(dens)]
    injectors = [solar_wind_injector(U, d, T, ppc, seed=seed + 1 + q) for q, d in enumerate(dens)]
    empty = [Particles(np.zeros(0, dtype=np.int64), np.zeros((0, 7))) for _ in species]
    meta = dict(base.meta, ppc=ppc, fill=fill)
    return Problem("obstacle%dx%dx%d_ppc%d_filled" % (nx, ny, nz, ppc), g, base.params, species, base.fields,
                   base.flags, empty, steps=base.steps, meta=meta, injectors=injectors)


def uniform_injector(velocity, density: float, T: float, ppc: float, seed: int, box=ALL_SPACE) -> Injector:
    """InjectorUniform (src/particle_injector.cpp:159-345): at timestep 1,
    probRound(ppc) particles in every cell whose centre is in the box; the
    species weight must be density*dV/ppc."""
    return Injector(INJ_UNIFORM, seed, T, tuple(float(v) for v in velocity), density, float(ppc), box=tuple(box))


def solar_wind_injector(velocity, density: float, T: float, ppc: float, seed: int) -> Injector:
    """InjectorSolarWind: inflow through the wall the bulk velocity points
    away from; the species weight must be density*dV/ppc."""
    return Injector(INJ_SOLAR_WIND, seed, T, tuple(float(v) for v in velocity), density, float(ppc))


def exosphere_neutrals(grid: Grid, n0: float, r0: float, H: float, R_exobase: float,
                       sub: int = 4) -> np.ndarray:
    """Neutrals per cell (cell volume times the mean density) of the
    Exponential profile n0*exp(-(r-r0)/H) (src/neutral_profiles.h:57-66), zero
    below the exobase.  The reference averages 1000 random points per cell
    (src/particle_injector.cpp:1690-1712); this uses the midpoint rule on
    sub^3 points of the same shrunk cell, a deterministic estimate of the
    same integral."""
    i, j, k = grid.ijk()
    dx = grid.dx
    xc, yc, zc = grid.x0 + (i + 0.5) * dx, grid.y0 + (j + 0.5) * dx, grid.z0 + (k + 0.5) * dx
    offs = ((np.arange(sub) + 0.5) / sub - 0.5) * (1.0 - EPS_CELL) * dx
    acc = np.zeros(grid.nc)
    for ox in offs:
        for oy in offs:
            for oz in offs:
                r = np.sqrt((xc + ox) ** 2 + (yc + oy) ** 2 + (zc + oz) ** 2)
                acc += np.where(r * r < R_exobase * R_exobase, 0.0, n0 * np.exp(-(r - r0) / H))
    return acc / sub ** 3 * grid.dV


def exosphere_injector(grid: Grid, neutrals: np.ndarray, T: float, macro_per_dt: float,
                       R_exobase: float, seed: int, R_shadow: float = 0.0,
                       shadow_factor: float = 0.0) -> Injector:
    """InjectorExosphere's rate table (src/particle_injector.cpp:1726-1792):
    macro_per_dt macroparticles per step spread over the cells in proportion
    to their neutrals times the ionisation factor (shadow_factor in the
    shadow cylinder behind the object, x < 0).  The species weight is then
    total_rate*dt/macro_per_dt with total_rate = ionisation rate * sum(neutrals)."""
    i, j, k = grid.ijk()
    dx = grid.dx
    xc, yc, zc = grid.x0 + (i + 0.5) * dx, grid.y0 + (j + 0.5) * dx, grid.z0 + (k + 0.5) * dx
    factor = np.where((xc < 0) & (yc * yc + zc * zc < R_shadow * R_shadow), shadow_factor, 1.0)
    rate = macro_per_dt * neutrals * factor / neutrals.sum()
    return Injector(INJ_EXOSPHERE, seed, T, cell_rate=rate, exobase_radius=R_exobase)


def config_exosphere(nx: int = 256, ny: int = 192, nz: int = 192, ppc: int = 24, dx: float = 200e3,
                     dt: float = 0.04, R_cells: float = 12.0, H_cells: float = 5.0, n0: float = 1e12,
                     ionization_rate: float = 1e-6, o_relative: float = 4.0, n_sw: float = 3e6,
                     seed: int = 6461674, steps: int = 500, sub: int = 2, o_per_ppc: bool = False) -> Problem:
    """Uneven-load box: solar-wind H+ (ppc per cell, 450 km/s, T=1e5 K, IMF
    on the +x wall) around an absorbing sphere of R_cells cells with an
    exponential oxygen corona (n0 at r0 = R, scale height H_cells cells,
    exobase at R) photoionised at ionization_rate into O+ (16 m_p, +e, 1e4 K).
    The O+ macroparticles per step are o_relative times the solar wind's
    (InjectorExosphere's macroparticles_per_cell_relative x the upstream
    ratio, src/particle_injector.cpp:1680-1688).  Particles start empty: H+
    is filled on the device (meta["fill"], InjectorUniform at timestep 1),
    then enters through the +x wall; O+ enters by photoionisation (with
    o_per_ppc, in proportion to the H+ ppc, so that runs at different ppc
    carry the same relative noise in both species)."""
    grid = Grid(nx, ny, nz, (0, 0, 0), dx, -0.5 * nx * dx, -0.5 * ny * dx, -0.5 * nz * dx)
    T, Te = 1e5, 1e5
    U = (-450e3, 0.0, 0.0)
    B0 = (-1.5e-9, 2.5e-9, 0.0)
    R, H = R_cells * dx, H_cells * dx
    interior = (nx - 2) * (ny - 2) * (nz - 2)
    spH = Species("H+", QE, MP, n_sw * grid.dV / ppc, R2_obstacle=R * R, capacity=int(1.3 * ppc * grid.nc))
    # upstream ratio: solar-wind macroparticles per step over ppc (:698-701)
    macro_per_dt = o_relative * (abs(U[0]) * dt / dx) * (ny - 2) * (nz - 2)
    if o_per_ppc:  # O+ macroparticles in proportion to the H+ ppc (24 ppc: unchanged)
        macro_per_dt *= ppc / 24.0
    neut = exosphere_neutrals(grid, n0, R, H, R_exobase=R, sub=sub)
    total_rate = ionization_rate * neut.sum()
    # O+ accumulates until pick-up carries it out of the box (~1400 steps
    # at the default sizes): room for `steps` steps of production
    spO = Species("O+", QE, 16 * MP, total_rate * dt / macro_per_dt, R2_obstacle=R * R,
                  capacity=int(macro_per_dt * max(1600, steps + 200) + 1e6))
    p = Params(dt=dt, use_hall=1, use_e_pressure=1, isothermal=1, e_pressure_coeff=KB * Te / QE, imf=B0,
               imf_cell=(1, 0, 0, 0, 0, 0), imf_face=(1, 0, 0, 0, 0, 0))
    _constraints(p, dx, dt, 1500e3, 1500e3, n_sw * QE, 0.05, 1500e3)
    empty = [Particles(np.zeros(0, dtype=np.int64), np.zeros((0, 7))) for _ in range(2)]
    injectors = [solar_wind_injector(U, n_sw, T, ppc, seed=seed + 1),
                 exosphere_injector(grid, neut, 1e4, macro_per_dt, R_exobase=R, seed=seed + 3)]
    return Problem("exosphere%dx%dx%d_ppc%d" % (nx, ny, nz, ppc), grid, p, [spH, spO], make_fields(grid, B0),
                   make_flags(grid, R_field=R, R_particle=R), empty, steps=steps,
                   meta={"ppc": ppc, "density": n_sw, "U": U, "T": T, "R_cells": R_cells, "interior": interior,
                         "O_macro_per_dt": macro_per_dt, "O_total_rate": total_rate,
                         "fill": [uniform_injector(U, n_sw, T, ppc, seed=seed + 5), None]},
                   injectors=injectors)


def perturb(problem: Problem, seed: int = 1, scale: float = 0.2) -> Problem:
    """Add a smooth, divergence-free perturbation to faceB (as a curl of a
    random node-centred vector potential, so div B stays zero) and a random
    resistivity, so that every term of the field update is exercised."""
    g = problem.grid
    rng = np.random.default_rng(seed)
    Bmag = max(np.abs(problem.fields["faceB"]).max(), 1e-9)
    # vector potential on cell edges, approximated by node values of a smooth field
    i, j, k = g.ijk()
    A = np.zeros((3, g.nc))
    for l in range(3):
        kx, ky, kz = rng.integers(1, 3, size=3)
        ph = rng.random(3) * 2 * np.pi
        A[l] = np.sin(2 * np.pi * kx * i / g.nx + ph[0]) * np.cos(2 * np.pi * ky * j / g.ny + ph[1]) \
            * np.sin(2 * np.pi * kz * k / g.nz + ph[2])
    A *= scale * Bmag * g.dx

    def sh(a, d, s):
        shape = (g.nz, g.ny, g.nx)
        a3 = a.reshape(shape)
        return np.roll(a3, -s, axis=2 - d).reshape(-1)
    # B = curl A on faces: Bx(face +x of cell) = dAz/dy - dAy/dz with A on x-edges/nodes
    dB = np.zeros((3, g.nc))
    dB[0] = (A[2] - sh(A[2], 1, -1)) / g.dx - (A[1] - sh(A[1], 2, -1)) / g.dx
    dB[1] = (A[0] - sh(A[0], 2, -1)) / g.dx - (A[2] - sh(A[2], 0, -1)) / g.dx
    dB[2] = (A[1] - sh(A[1], 0, -1)) / g.dx - (A[0] - sh(A[0], 1, -1)) / g.dx
    out = problem.copy()
    for l in range(3):
        out.fields["faceB"][l::3] += dB[l]
    return out
