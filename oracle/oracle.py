"""ctypes wrapper of the FP64 oracle (oracle/rh_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product path never loads
this module or liboracle.so.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# RH_ORACLE_SANITIZED=1: the ASan/UBSan build (the caller preloads the ASan runtime)
_SAN = os.environ.get("RH_ORACLE_SANITIZED") == "1"
_LIB = os.path.join(_HERE, "liboracle_san.so" if _SAN else "liboracle.so")

N_FIELDS = 14
N_FLAGS = 4
N_COUNTERS = 7
FIELD_IDS = {n: i for i, n in enumerate(
    ["faceB", "cellB", "cellJ", "cellUe", "cellJi", "cellEp", "cellRhoQi", "nodeRhoQi",
     "nodeE", "nodeB", "nodeJ", "nodeUe", "nodeJi", "nodeEta"])}
FLAG_IDS = {n: i for i, n in enumerate(
    ["innerFlagField", "innerFlagNode", "innerFlagParticle", "innerFlagCellEp"])}
COUNTER_NAMES = ["max_vi", "min_cell_rho", "max_cell_ue", "min_node_rho", "max_node_ue",
                 "max_E", "max_vw"]


class OrcGrid(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("periodic", C.c_int32 * 3), ("dx", C.c_double),
                ("x0", C.c_double), ("y0", C.c_double), ("z0", C.c_double)]


class OrcParams(C.Structure):
    _fields_ = [("dt", C.c_double), ("use_hall", C.c_int32), ("use_e_pressure", C.c_int32),
                ("isothermal", C.c_int32), ("e_pressure_coeff", C.c_double),
                ("max_ion_speed2", C.c_double), ("max_electron_speed2", C.c_double),
                ("min_rho", C.c_double), ("max_E2", C.c_double), ("max_whistler", C.c_double),
                ("n_linear_smoothings", C.c_int32), ("use_gaussian", C.c_int32),
                ("gauss_coeffs", C.c_double * 4), ("imf", C.c_double * 3),
                ("imf_cell", C.c_int32 * 6), ("imf_face", C.c_int32 * 6),
                ("obstacle_ue", C.c_double * 3), ("initial_flow_through", C.c_int32),
                ("use_gravity", C.c_int32), ("GMdt", C.c_double)]


class OrcInjector(C.Structure):
    _fields_ = [("kind", C.c_int32), ("seed", C.c_uint64), ("temperature", C.c_double),
                ("velocity", C.c_double * 3), ("density", C.c_double), ("ppc", C.c_double),
                ("cell_rate", C.c_void_p), ("exobase_radius", C.c_double), ("box", C.c_double * 6)]


INJ_NONE, INJ_SOLAR_WIND, INJ_EXOSPHERE, INJ_UNIFORM = 0, 1, 2, 3


class OrcSpecies(C.Structure):
    _fields_ = [("q", C.c_double), ("m", C.c_double), ("w", C.c_double),
                ("R2_obstacle", C.c_double), ("accumulate", C.c_int32), ("accelerate", C.c_int32)]


def build(force: bool = False) -> str:
    src = os.path.join(_HERE, "rh_oracle.c")
    if force or not os.path.exists(_LIB) or os.path.getmtime(_LIB) < os.path.getmtime(src):
        subprocess.check_call(["make", "-s", "-B" if force else "-s", "-C", _HERE, os.path.basename(_LIB)])
    return _LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        vp = C.c_void_p
        L.orc_create.restype = vp
        L.orc_create.argtypes = [C.POINTER(OrcGrid), C.POINTER(OrcParams)]
        L.orc_destroy.argtypes = [vp]
        L.orc_add_species.restype = C.c_int
        L.orc_add_species.argtypes = [vp, C.POINTER(OrcSpecies)]
        L.orc_num_cells.argtypes = [vp]
        L.orc_field.restype = C.POINTER(C.c_double)
        L.orc_field.argtypes = [vp, C.c_int]
        L.orc_flag.restype = C.POINTER(C.c_uint8)
        L.orc_flag.argtypes = [vp, C.c_int]
        L.orc_set_threads.argtypes = [vp, C.c_int]
        L.orc_set_particles.argtypes = [vp, C.c_int, C.c_int64, vp, vp]
        L.orc_num_particles.restype = C.c_int64
        L.orc_num_particles.argtypes = [vp, C.c_int]
        L.orc_get_particles.argtypes = [vp, C.c_int, vp, vp]
        for name in ["orc_push", "orc_migrate", "orc_accumulate", "orc_apply_boundary_conditions"]:
            getattr(L, name).argtypes = [vp, C.c_int]
        for name in ["orc_clear_accumulation", "orc_add_remote_updates", "orc_propagate_B",
                     "orc_step", "orc_reset_counters"]:
            getattr(L, name).argtypes = [vp]
        L.orc_diagnostics.argtypes = [vp, C.c_int, vp]
        L.orc_philox4x32_10.argtypes = [vp, vp, vp]
        L.orc_rng_sample.argtypes = [C.c_uint64, C.c_int, C.c_double, C.c_int64, vp]
        L.orc_set_injector.argtypes = [vp, C.c_int, C.POINTER(OrcInjector)]
        L.orc_inject.argtypes = [vp, C.c_int]
        L.orc_set_timestep.argtypes = [vp, C.c_int64]
        L.orc_timestep.restype = C.c_int64
        L.orc_timestep.argtypes = [vp]
        L.orc_injected.restype = C.c_int64
        L.orc_injected.argtypes = [vp, C.c_int]
        L.orc_log_begin.argtypes = [vp, C.c_double]
        L.orc_log_values.argtypes = [vp, C.c_double, C.c_int, vp, vp]
        L.orc_counter.restype = C.c_double
        L.orc_counter.argtypes = [vp, C.c_int]
        L.orc_get_fields.argtypes = [vp, C.c_int64, vp, vp, vp, vp]
        L.orc_set_exterior_order_reversed.argtypes = [vp, C.c_int]
        L.orc_face2cell.argtypes = [vp, vp, vp]
        L.orc_cell2node.argtypes = [vp, vp, vp, C.c_int]
        L.orc_node2cell.argtypes = [vp, vp, vp]
        L.orc_neumann_cell.argtypes = [vp, vp, C.c_int]
        L.orc_neumann_face.argtypes = [vp, vp]
        L.orc_face_curl.argtypes = [vp, vp, vp, C.c_double]
        L.orc_set_imf_cell.argtypes = [vp, vp]
        L.orc_set_imf_face.argtypes = [vp, vp]
        L.orc_node_avg.argtypes = [vp, vp, vp]
        for name in ["orc_calc_node_J", "orc_calc_cell_Ue", "orc_calc_node_Ue", "orc_upwind_node_B",
                     "orc_calc_node_E", "orc_calc_cell_Ep"]:
            getattr(L, name).argtypes = [vp]
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def philox4x32_10(ctr, key) -> np.ndarray:
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(_ptr(c), _ptr(k), _ptr(out))
    return out


RNG_UNIFORM, RNG_GAUSS, RNG_DERIV_GAUSS, RNG_PROB_ROUND = 0, 1, 2, 3


def rng_sample(seed: int, which: int, n: int, x0: float = 0.0) -> np.ndarray:
    """n draws of one generator stream: uniform, sampleGaussian,
    sampleDerivGaussian(x0) or probRound(x0)."""
    out = np.empty(n, dtype=np.float64)
    lib().orc_rng_sample(int(seed), int(which), float(x0), int(n), _ptr(out))
    return out


def injector_struct(inj) -> OrcInjector:
    """An OrcInjector from a problem.Injector (keeps the rate array alive on
    the returned object)."""
    ci = OrcInjector()
    ci.kind = inj.kind
    ci.seed = inj.seed
    ci.temperature = inj.temperature
    for l in range(3):
        ci.velocity[l] = inj.velocity[l]
    ci.density = inj.density
    ci.ppc = inj.ppc
    if inj.cell_rate is not None:
        ci._rate = np.ascontiguousarray(inj.cell_rate, dtype=np.float64)
        ci.cell_rate = ci._rate.ctypes.data
    ci.exobase_radius = inj.exobase_radius
    for l in range(6):
        ci.box[l] = inj.box[l]
    return ci


class Oracle:
    """FP64 oracle simulation holding one Problem's state."""

    def __init__(self, problem, threads: Optional[int] = None):
        L = lib()
        g, p = problem.grid, problem.params
        self.grid = g
        cg = OrcGrid(g.nx, g.ny, g.nz, (C.c_int32 * 3)(*g.periodic), g.dx, g.x0, g.y0, g.z0)
        cp = OrcParams()
        for name, _ in OrcParams._fields_:
            v = getattr(p, name)
            if isinstance(v, (tuple, list)):
                arr = getattr(cp, name)
                for i, x in enumerate(v):
                    arr[i] = x
            else:
                setattr(cp, name, v)
        self._h = L.orc_create(C.byref(cg), C.byref(cp))
        if not self._h:
            raise ValueError("oracle rejected the grid")
        if threads is not None:
            L.orc_set_threads(self._h, int(threads))
        self.nspecies = 0
        for sp in problem.species:
            cs = OrcSpecies(sp.q, sp.m, sp.w, sp.R2_obstacle, sp.accumulate, sp.accelerate)
            L.orc_add_species(self._h, C.byref(cs))
            self.nspecies += 1
        for name, arr in problem.fields.items():
            self.field(name)[:] = arr
        for name, arr in problem.flags.items():
            self.flag(name)[:] = arr
        for s, parts in enumerate(problem.particles):
            self.set_particles(s, parts.cell, parts.state)
        for s, inj in enumerate(getattr(problem, "injectors", []) or []):
            if inj is not None:
                self.set_injector(s, inj)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_destroy(h)
            self._h = None

    def field(self, name: str) -> np.ndarray:
        fid = FIELD_IDS[name]
        dim = 1 if name in ("cellRhoQi", "nodeRhoQi", "nodeEta") else 3
        ptr = lib().orc_field(self._h, fid)
        return np.ctypeslib.as_array(ptr, shape=(self.grid.nc * dim,))

    def flag(self, name: str) -> np.ndarray:
        ptr = lib().orc_flag(self._h, FLAG_IDS[name])
        return np.ctypeslib.as_array(ptr, shape=(self.grid.nc,))

    def set_particles(self, s: int, cell: np.ndarray, state: np.ndarray) -> None:
        cell = np.ascontiguousarray(cell, dtype=np.int64)
        state = np.ascontiguousarray(state, dtype=np.float64)
        rc = lib().orc_set_particles(self._h, s, cell.shape[0], _ptr(cell), _ptr(state))
        if rc != 0:
            raise ValueError("orc_set_particles failed: %d" % rc)

    def num_particles(self, s: int) -> int:
        return int(lib().orc_num_particles(self._h, s))

    def particles(self, s: int):
        n = self.num_particles(s)
        cell = np.empty(n, dtype=np.int64)
        st = np.empty((n, 7), dtype=np.float64)
        lib().orc_get_particles(self._h, s, _ptr(cell), _ptr(st))
        return cell, st

    def push(self, s: int) -> None: lib().orc_push(self._h, s)
    def migrate(self, s: int) -> None: lib().orc_migrate(self._h, s)
    def clear_accumulation(self) -> None: lib().orc_clear_accumulation(self._h)
    def accumulate(self, s: int) -> None: lib().orc_accumulate(self._h, s)
    def add_remote_updates(self) -> None: lib().orc_add_remote_updates(self._h)
    def apply_boundary_conditions(self, s: int) -> None: lib().orc_apply_boundary_conditions(self._h, s)
    def propagate_B(self) -> None: lib().orc_propagate_B(self._h)

    def step(self, n: int = 1) -> None:
        for _ in range(n):
            lib().orc_step(self._h)

    def deposit(self) -> None:
        """clear + accumulate all species + finalise (moments of the current particles)."""
        self.clear_accumulation()
        for s in range(self.nspecies):
            self.accumulate(s)
        self.add_remote_updates()

    def diagnostics(self, s: int) -> np.ndarray:
        out = np.zeros(5)
        lib().orc_diagnostics(self._h, s, _ptr(out))
        return out

    def counters(self) -> dict:
        return {n: lib().orc_counter(self._h, i) for i, n in enumerate(COUNTER_NAMES)}

    def get_fields_at(self, cell: int, r) -> tuple:
        r = np.ascontiguousarray(r, dtype=np.float64)
        B, Ue, Ep = np.zeros(3), np.zeros(3), np.zeros(3)
        lib().orc_get_fields(self._h, int(cell), _ptr(r), _ptr(B), _ptr(Ue), _ptr(Ep))
        return B, Ue, Ep

    def set_injector(self, s: int, inj) -> None:
        rc = lib().orc_set_injector(self._h, s, C.byref(injector_struct(inj)))
        if rc != 0:
            raise ValueError("orc_set_injector failed: %d" % rc)

    def inject(self, s: int) -> None: lib().orc_inject(self._h, s)
    def set_timestep(self, t: int) -> None: lib().orc_set_timestep(self._h, int(t))
    def timestep(self) -> int: return int(lib().orc_timestep(self._h))
    def injected(self, s: int) -> int: return int(lib().orc_injected(self._h, s))

    def log_begin(self, t: float) -> None:
        lib().orc_log_begin(self._h, float(t))

    def log_values(self, t: float, include_inner: bool = False):
        """(field[41], pop[nspecies][17]) of one log line at time t; the log
        counters restart."""
        field = np.zeros(41)
        pop = np.zeros((self.nspecies, 17))
        lib().orc_log_values(self._h, float(t), int(include_inner), _ptr(field), _ptr(pop))
        return field, pop

    def set_exterior_order_reversed(self, r: bool) -> None:
        lib().orc_set_exterior_order_reversed(self._h, int(r))

    def raw(self):
        return self._h
