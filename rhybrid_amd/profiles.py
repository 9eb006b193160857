"""Init-time profiles that produce the hot path's inputs: the resistivity
table nodeEta (src/resistivity.h:29-215 and its use in src/user.cpp:
1636-1640,1913-1940) and the initial magnetic field on faces and cell centres
(src/magnetic_field.h:28-252, setInitialB :257-260, applied by
src/user.cpp:2107-2141 under USE_B_INITIAL).  Host code (numpy); both the
HIP path and the oracle receive the arrays these functions build.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Sequence, Tuple

import numpy as np

MU0 = 4.0e-7 * np.pi  # the permeability of the field update (calcNodeJ)

RESISTIVITY_PROFILES = ("resistivityConstant", "resistivitySuperConductingSphere", "resistivitySphericalShells")
RESISTIVITY_UNITS = ("SI", "grid", "td", "Rm", "URm")


def resistivity_si(unit: str, value: float, dx: float, dt: float, u_bulk: float) -> float:
    """A [Resistivity] value in Ohm m (src/resistivity.h:65-102): SI as is;
    grid: value * mu0 dx^2/dt; td (diffusion time over dt): mu0 dx^2/dt /
    value; Rm (magnetic Reynolds number): mu0 dx U/value; URm (U/Rm, m/s):
    value * mu0 dx."""
    grid_unit = MU0 * dx * dx / dt  # src/user.cpp:1637
    if unit == "SI":
        return float(value)
    if unit == "grid":
        return float(value) * grid_unit
    if unit == "td":
        if value == 0:
            raise ValueError("resistivity in units of td_per_dt cannot be zero")
        return grid_unit / float(value)
    if unit == "Rm":
        if value == 0:
            raise ValueError("resistivity in units of Rm cannot be zero")
        return MU0 * dx * u_bulk / float(value)
    if unit == "URm":
        return float(value) * MU0 * dx
    raise ValueError("unknown resistivity unit %r" % unit)


@dataclass
class Resistivity:
    """The [Resistivity] section (src/user.cpp:1609-1634)."""
    profile: str = "resistivityConstant"
    unit: str = "SI"
    value: float = 0.0
    R: float = 0.0                                   # superconducting sphere radius [m]
    value_spherical: Sequence[float] = ()
    R_spherical: Sequence[float] = ()                # shell outer radii [m], increasing
    outer_boundary_cells: Sequence[int] = (0, 0, 0, 0, 0, 0)      # +x -x +y -y +z -z
    outer_boundary_coeffs: Sequence[float] = (0.0, 0.0, 0.0, 0.0, 0.0, 0.0)


def resistivity_at(res: Resistivity, x, y, z, dx: float, dt: float, u_bulk: float) -> np.ndarray:
    """eta(x, y, z) of the profile (resistivityConstant,
    resistivitySuperConductingSphere: zero for r < R, resistivitySphericalShells:
    the value of the first shell whose outer radius is >= r, zero beyond the
    last; src/resistivity.h:29-61)."""
    x, y, z = (np.asarray(a, dtype=np.float64) for a in (x, y, z))
    r2 = x * x + y * y + z * z
    if res.profile == "resistivityConstant":
        return np.full(r2.shape, resistivity_si(res.unit, res.value, dx, dt, u_bulk))
    if res.profile == "resistivitySuperConductingSphere":
        eta = resistivity_si(res.unit, res.value, dx, dt, u_bulk)
        return np.where(r2 < res.R * res.R, 0.0, eta)
    if res.profile == "resistivitySphericalShells":
        vals, radii = list(res.value_spherical), list(res.R_spherical)
        if len(vals) != len(radii) or not vals:
            raise ValueError("value_spherical and R_spherical need the same non-zero length")
        R2 = []
        for i, R in enumerate(radii):
            if R < 0:
                raise ValueError("R_spherical < 0")
            R2.append(R * R)
            if i > 0 and R2[i - 1] >= R2[i]:
                raise ValueError("R_spherical must increase")
        etas = [resistivity_si(res.unit, v, dx, dt, u_bulk) for v in vals]
        out = np.zeros(r2.shape)
        out[r2 <= R2[0]] = etas[0]
        for i in range(1, len(R2)):
            out[(r2 > R2[i - 1]) & (r2 <= R2[i])] = etas[i]
        return out
    raise ValueError("unknown resistivity profile %r" % res.profile)


def node_eta(grid, res: Resistivity, dt: float, u_bulk: float) -> np.ndarray:
    """nodeEta of every cell (its node is the cell's upper corner), with the
    outer-boundary multipliers: within outer_boundary_cells of a wall the
    profile's value times that wall's coefficient, the walls tried in the
    order +x -x +y -y +z -z and the last one that applies kept
    (src/user.cpp:1913-1940; the 0.99 / 1.01 margins as there)."""
    i, j, k = grid.ijk()
    dx = grid.dx
    xn, yn, zn = grid.x0 + (i + 1.0) * dx, grid.y0 + (j + 1.0) * dx, grid.z0 + (k + 1.0) * dx
    base = resistivity_at(res, xn, yn, zn, dx, dt, u_bulk)
    eta = base.copy()
    lo = (grid.x0, grid.y0, grid.z0)
    hi = (grid.x0 + grid.nx * dx, grid.y0 + grid.ny * dx, grid.z0 + grid.nz * dx)
    pos = (xn, yn, zn)
    cells, coeffs = list(res.outer_boundary_cells), list(res.outer_boundary_coeffs)
    if len(cells) != 6 or len(coeffs) != 6:
        raise ValueError("outer_boundary_cells and outer_boundary_coeffs need six values (+x -x +y -y +z -z)")
    for w in range(6):
        a = w // 2
        if cells[w] <= 0:
            continue
        if w % 2 == 0:
            sel = pos[a] > hi[a] - dx * cells[w] * 0.99
        else:
            sel = pos[a] < lo[a] + dx * cells[w] * 1.01
        eta[sel] = base[sel] * coeffs[w]
    return eta


MAGNETIC_PROFILES = ("constantBx", "constantB", "laminarFlowAroundSphereBx", "hemisphericDipoleB",
                     "translateDipoleB", "generalDipoleB", "lineDipoleB",
                     "translateDipoleBAndLaminarFlowAroundSphereBx",
                     "translateDipoleBWithMirrorAndLaminarFlowAroundSphereBx", "generalDipoleBAndConstantBx",
                     "marsCrustalB")


@dataclass
class IntrinsicB:
    """The [IMF] field and the [IntrinsicB] section (src/user.cpp:499-540):
    dipole moment coefficient 3 B_surface R_surface^3, squared radii.
    randomB draws from Corsair's generator, which is not in the reference
    tree, and is not restated."""
    profile: str = "constantB"
    imf: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    laminar_R: float = 0.0
    coeff_dipole: float = 0.0
    coeff_quadrupole: float = 0.0
    surface_B: float = 0.0
    surface_R: float = 0.0
    minimum_R: float = 0.0
    x: float = 0.0
    y: float = 0.0
    z: float = 0.0
    x_mirror: Sequence[float] = ()
    y_mirror: Sequence[float] = ()
    z_mirror: Sequence[float] = ()
    theta: float = 0.0     # degrees
    phi: float = 0.0       # degrees

    @property
    def mom(self) -> float:
        return 3.0 * self.surface_B * self.surface_R ** 3


def _dipole(ib: IntrinsicB, x1, y1, z1, sign: float, B):
    # translateDipoleB / translateMirrorDipoleB (src/magnetic_field.h:88-117)
    r2 = x1 * x1 + y1 * y1 + z1 * z1
    ok = r2 >= ib.minimum_R ** 2
    rr = np.sqrt(r2)
    with np.errstate(divide="ignore", invalid="ignore"):
        c = np.where(ok, ib.mom / (r2 * r2 * rr), 0.0)
    B[0] += sign * c * x1 * z1
    B[1] += sign * c * y1 * z1
    B[2] += sign * c * (z1 * z1 - r2 / 3.0)


def _rotated(ib: IntrinsicB, x, y, z):
    # the point in the dipole's frame: rotation by theta about y, then phi about x
    th, ph = np.deg2rad(ib.theta), np.deg2rad(ib.phi)
    x1, y1, z1 = x - ib.x, y - ib.y, z - ib.z
    x2, y2, z2 = x1 * np.cos(th) + z1 * np.sin(th), y1, -x1 * np.sin(th) + z1 * np.cos(th)
    return x2, y2 * np.cos(ph) - z2 * np.sin(ph), y2 * np.sin(ph) + z2 * np.cos(ph)


def _unrotate(ib: IntrinsicB, b1, B):
    # the field back to the simulation frame (src/magnetic_field.h:80-85,147-152)
    th, ph = np.deg2rad(ib.theta), np.deg2rad(ib.phi)
    bx2 = b1[0]
    by2 = b1[1] * np.cos(ph) + b1[2] * np.sin(ph)
    bz2 = -b1[1] * np.sin(ph) + b1[2] * np.cos(ph)
    B[0] += bx2 * np.cos(th) - bz2 * np.sin(th)
    B[1] += by2
    B[2] += bx2 * np.sin(th) + bz2 * np.cos(th)


def _general(ib: IntrinsicB, x, y, z, B, quadrupole: bool):
    # generalDipoleB (:119-153) and hemisphericDipoleB (:54-86)
    x3, y3, z3 = _rotated(ib, x, y, z)
    rr = x3 * x3 + y3 * y3 + z3 * z3
    ok = rr >= ib.minimum_R ** 2
    r = np.sqrt(rr)
    with np.errstate(divide="ignore", invalid="ignore"):
        c = np.where(ok, ib.mom / (rr * rr * r), 0.0)
        if quadrupole:
            q = 0.5 * ib.coeff_quadrupole * ib.surface_R
            bx1 = c * x3 * (ib.coeff_dipole * z3 + q * (5.0 * z3 * z3 / rr - 1.0))
            by1 = c * y3 * (ib.coeff_dipole * z3 + q * (5.0 * z3 * z3 / rr - 1.0))
            # the reference's Bz1 takes the unrotated z in its quadrupole term
            bz1 = c * (ib.coeff_dipole * (z3 * z3 - rr / 3.0) + q * z * (5.0 * z3 * z3 - 3.0 * rr) / rr)
        else:
            bx1, by1, bz1 = c * x3 * z3, c * y3 * z3, c * (z3 * z3 - rr / 3.0)
    b1 = [np.where(ok, v, 0.0) for v in (bx1, by1, bz1)]
    _unrotate(ib, b1, B)


def _laminar(ib: IntrinsicB, x, y, z, B):
    # laminarFlowAroundSphereBx (:44-52): potential flow of the IMF's Bx around a sphere
    r2 = x * x + y * y + z * z
    R2, R3 = ib.laminar_R ** 2, ib.laminar_R ** 3
    ok = r2 >= R2
    r = np.sqrt(r2)
    with np.errstate(divide="ignore", invalid="ignore"):
        c = np.where(ok, -1.5 * ib.imf[0] * R3 * x / r ** 5, 0.0)
        bx = np.where(ok, c * x + ib.imf[0] * (1.0 + 0.5 * R3 / r ** 3), 0.0)
    B[0] += bx
    B[1] += c * y
    B[2] += c * z


def magnetic_field_at(ib: IntrinsicB, x, y, z) -> np.ndarray:
    """B (3, ...) of the profile at the points (setInitialB: zero plus the
    profile, src/magnetic_field.h:257-260)."""
    x, y, z = (np.asarray(a, dtype=np.float64) for a in (x, y, z))
    B = np.zeros((3,) + x.shape)
    p = ib.profile
    if p == "constantBx":
        B[0] += ib.imf[0]
    elif p == "constantB":
        for l in range(3):
            B[l] += ib.imf[l]
    elif p == "laminarFlowAroundSphereBx":
        _laminar(ib, x, y, z, B)
    elif p == "hemisphericDipoleB":
        _general(ib, x, y, z, B, quadrupole=True)
    elif p == "translateDipoleB":
        _dipole(ib, x - ib.x, y - ib.y, z - ib.z, 1.0, B)
    elif p == "generalDipoleB":
        _general(ib, x, y, z, B, quadrupole=False)
    elif p == "lineDipoleB":
        # (:155-165): the radius from the dipole's position, the field from the unshifted x, z
        x1, y1, z1 = x - ib.x, y - ib.y, z - ib.z
        r2 = x1 * x1 + y1 * y1 + z1 * z1
        ok = r2 >= ib.minimum_R ** 2
        D = -ib.coeff_dipole
        with np.errstate(divide="ignore", invalid="ignore"):
            B[0] += np.where(ok, D * 2.0 * x * z / (r2 * r2), 0.0)
            B[2] += np.where(ok, D * (z * z - x * x) / (r2 * r2), 0.0)
    elif p == "translateDipoleBAndLaminarFlowAroundSphereBx":
        _dipole(ib, x - ib.x, y - ib.y, z - ib.z, 1.0, B)
        _laminar(ib, x, y, z, B)
    elif p == "translateDipoleBWithMirrorAndLaminarFlowAroundSphereBx":
        _dipole(ib, x - ib.x, y - ib.y, z - ib.z, 1.0, B)
        if not (len(ib.x_mirror) == len(ib.y_mirror) == len(ib.z_mirror)):
            raise ValueError("mirror dipole coordinates need the same lengths")
        for xm, ym, zm in zip(ib.x_mirror, ib.y_mirror, ib.z_mirror):
            _dipole(ib, x - xm, y - ym, z - zm, -1.0, B)
        _laminar(ib, x, y, z, B)
    elif p == "generalDipoleBAndConstantBx":
        _general(ib, x, y, z, B, quadrupole=False)
        B[0] += ib.imf[0]
    elif p == "marsCrustalB":
        pass  # the reference's crustal field adds zero (:200-207)
    else:
        raise ValueError("unknown magnetic field profile %r" % p)
    return B


def initial_B(grid, ib: IntrinsicB) -> Tuple[np.ndarray, np.ndarray]:
    """faceB and cellB (interleaved n*3+l) of USE_B_INITIAL: component l of
    faceB at the centre of the cell's +l face, cellB at the cell centre
    (src/user.cpp:2107-2141)."""
    i, j, k = grid.ijk()
    dx = grid.dx
    xc, yc, zc = grid.x0 + (i + 0.5) * dx, grid.y0 + (j + 0.5) * dx, grid.z0 + (k + 0.5) * dx
    face = np.zeros((grid.nc, 3))
    face[:, 0] = magnetic_field_at(ib, xc + 0.5 * dx, yc, zc)[0]
    face[:, 1] = magnetic_field_at(ib, xc, yc + 0.5 * dx, zc)[1]
    face[:, 2] = magnetic_field_at(ib, xc, yc, zc + 0.5 * dx)[2]
    cell = magnetic_field_at(ib, xc, yc, zc).T.copy()
    return face.ravel(), cell.ravel()
