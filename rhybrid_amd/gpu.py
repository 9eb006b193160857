"""Python binding of the C ABI (include/rh_gpu.h) and the reference's
operator surface on top of it.

The reference drives its hot path through Corsair operator objects
(src/register_objects.cpp:46-61): a "BorisBuneman" particle propagator, a
"HybridAccumulator" and the field propagator functions propagateB /
setupGetFields (src/hybrid_propagator.h:27,62), called in the order of
propagate() (src/user.cpp:118-134).  The classes below keep those names and
that order; every call goes to the HIP library, which must be built
(``python -c "import __graft_entry__ as g; g.build()"``).  There is no CPU
fallback: a missing library raises.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict, List, Optional

import numpy as np

from . import problem as P
from .problem import FIELD_DIM, FIELD_NAMES, FLAG_NAMES, Problem

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librhybrid_amd.so")

RH_OK, RH_ERR_ARG, RH_ERR_HIP, RH_ERR_STATE, RH_ERR_CAPACITY, RH_ERR_NO_DEVICE = 0, -1, -2, -3, -4, -5
RH_FP32, RH_FP64 = 0, 1
COUNTER_NAMES = ["max_vi", "min_cell_rho", "max_cell_ue", "min_node_rho", "max_node_ue",
                 "max_E", "max_vw", "dropped", "denoise_fallback"]

#: every symbol include/rh_gpu.h declares
EXPORTS = ["rh_create", "rh_destroy", "rh_last_error", "rh_add_species", "rh_upload_field",
           "rh_download_field", "rh_upload_flags", "rh_upload_particles", "rh_num_particles",
           "rh_download_particles", "rh_setup_get_fields", "rh_propagate", "rh_wait_particle_sends",
           "rh_clear_accumulation_arrays", "rh_accumulate", "rh_add_remote_updates",
           "rh_apply_boundary_conditions", "rh_propagate_B", "rh_step", "rh_diagnostics",
           "rh_counters", "rh_synchronize", "rh_set_timing", "rh_set_timing_only", "rh_kernel_times",
           "rh_reset_timing",
           "rh_layout_info", "rh_denoise_moments", "rh_set_denoiser", "rh_set_injector",
           "rh_inject_particles", "rh_set_timestep", "rh_get_timestep", "rh_injected", "rh_log_values",
           "rh_log_begin", "rh_log_open", "rh_log_write", "rh_write_vlsv", "rh_set_temporal_averages",
           "rh_set_species_streams"]


class GridDesc(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("periodic", C.c_int32 * 3), ("dx", C.c_double),
                ("x0", C.c_double), ("y0", C.c_double), ("z0", C.c_double),
                ("tile", C.c_int32 * 3)]


class Params(C.Structure):
    _fields_ = [("dt", C.c_double), ("use_hall", C.c_int32), ("use_e_pressure", C.c_int32),
                ("isothermal", C.c_int32), ("e_pressure_coeff", C.c_double),
                ("max_ion_speed2", C.c_double), ("max_electron_speed2", C.c_double),
                ("min_rho", C.c_double), ("max_E2", C.c_double), ("max_whistler", C.c_double),
                ("n_linear_smoothings", C.c_int32), ("use_gaussian", C.c_int32),
                ("gauss_coeffs", C.c_double * 4), ("imf", C.c_double * 3),
                ("imf_cell", C.c_int32 * 6), ("imf_face", C.c_int32 * 6),
                ("obstacle_ue", C.c_double * 3), ("initial_flow_through", C.c_int32),
                ("use_gravity", C.c_int32), ("GMdt", C.c_double)]


class SpeciesDesc(C.Structure):
    _fields_ = [("q", C.c_double), ("m", C.c_double), ("w", C.c_double),
                ("R2_obstacle", C.c_double), ("accumulate", C.c_int32),
                ("accelerate", C.c_int32), ("capacity", C.c_int64)]


class InjectorDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("seed", C.c_uint64), ("temperature", C.c_double),
                ("velocity", C.c_double * 3), ("density", C.c_double), ("ppc", C.c_double),
                ("cell_rate", C.c_void_p), ("exobase_radius", C.c_double), ("box", C.c_double * 6)]


class RhError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__("rh error %d: %s" % (code, msg))
        self.code = code


_lib = None


def load_library(path: str = LIB_PATH):
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError("HIP library %s is missing: build it with __graft_entry__.build()" % path)
    # One HIP runtime per process: torch's build carries its own
    # libamdhip64 (loaded by the unversioned name), the library needs
    # libamdhip64.so.7.  With torch loaded first, the library binds to the
    # runtime already in the process; loaded the other way round, the process
    # would hold two runtimes and torch would see no device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(path)
    vp, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
    L.rh_create.argtypes = [C.POINTER(GridDesc), C.POINTER(Params), i32, C.POINTER(vp)]
    L.rh_destroy.argtypes = [vp]
    L.rh_last_error.restype = C.c_char_p
    L.rh_last_error.argtypes = [vp]
    L.rh_add_species.argtypes = [vp, C.POINTER(SpeciesDesc), C.POINTER(i32)]
    L.rh_upload_field.argtypes = [vp, i32, vp]
    L.rh_download_field.argtypes = [vp, i32, vp]
    L.rh_upload_flags.argtypes = [vp, i32, vp]
    L.rh_upload_particles.argtypes = [vp, i32, i64, vp, vp]
    L.rh_num_particles.argtypes = [vp, i32, C.POINTER(i64)]
    L.rh_download_particles.argtypes = [vp, i32, i64, vp, vp]
    for name in ["rh_setup_get_fields", "rh_clear_accumulation_arrays", "rh_add_remote_updates",
                 "rh_propagate_B", "rh_synchronize", "rh_reset_timing"]:
        getattr(L, name).argtypes = [vp]
    for name in ["rh_propagate", "rh_wait_particle_sends", "rh_accumulate",
                 "rh_apply_boundary_conditions", "rh_step", "rh_set_timing", "rh_set_species_streams"]:
        getattr(L, name).argtypes = [vp, i32]
    L.rh_diagnostics.argtypes = [vp, i32, vp]
    L.rh_counters.argtypes = [vp, vp]
    L.rh_denoise_moments.argtypes = [vp, vp, C.c_double]
    L.rh_set_denoiser.argtypes = [vp, vp, C.c_double]
    L.rh_kernel_times.argtypes = [vp, C.c_char_p, C.POINTER(C.c_double), C.POINTER(i64)]
    L.rh_set_timing_only.argtypes = [vp, C.c_char_p]
    L.rh_layout_info.argtypes = [vp, vp]
    L.rh_set_injector.argtypes = [vp, i32, C.POINTER(InjectorDesc)]
    L.rh_inject_particles.argtypes = [vp, i32]
    L.rh_set_timestep.argtypes = [vp, i64]
    L.rh_get_timestep.argtypes = [vp, C.POINTER(i64)]
    L.rh_injected.argtypes = [vp, i32, C.POINTER(i64)]
    L.rh_log_values.argtypes = [vp, C.c_double, i32, vp, vp]
    L.rh_log_begin.argtypes = [vp, C.c_double]
    L.rh_log_open.argtypes = [vp, C.c_char_p, C.POINTER(C.c_char_p), i32, i32, C.c_double]
    L.rh_log_write.argtypes = [vp, C.c_double]
    L.rh_write_vlsv.argtypes = [vp, C.c_char_p, C.c_double, i64, C.POINTER(C.c_char_p), i32, C.POINTER(C.c_char_p)]
    L.rh_set_temporal_averages.argtypes = [vp, i32]
    for name in EXPORTS:
        if name not in ("rh_last_error",):
            getattr(L, name).restype = C.c_int
    _lib = L
    return L


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class Context:
    """One simulation domain on the current GPU (rh_ctx)."""

    def __init__(self, problem: Problem, precision: int = RH_FP32, tile=(0, 0, 0)):
        L = load_library()
        g, p = problem.grid, problem.params
        gd = GridDesc(g.nx, g.ny, g.nz, (C.c_int32 * 3)(*g.periodic), g.dx, g.x0, g.y0, g.z0,
                      (C.c_int32 * 3)(*tile))
        cp = Params()
        for name, _ in Params._fields_:
            v = getattr(p, name)
            if isinstance(v, (tuple, list)):
                arr = getattr(cp, name)
                for i, x in enumerate(v):
                    arr[i] = x
            else:
                setattr(cp, name, v)
        h = C.c_void_p()
        rc = L.rh_create(C.byref(gd), C.byref(cp), precision, C.byref(h))
        if rc != RH_OK:
            raise RhError(rc, "rh_create failed (%s)" % ("no GPU" if rc == RH_ERR_NO_DEVICE else "see code"))
        self._h = h
        self.problem = problem
        self.grid = g
        self.precision = precision
        self.species: List[int] = []
        for sp in problem.species:
            sid = C.c_int32()
            d = SpeciesDesc(sp.q, sp.m, sp.w, sp.R2_obstacle, sp.accumulate, sp.accelerate,
                            int(getattr(sp, "capacity", 0)))
            self._check(L.rh_add_species(self._h, C.byref(d), C.byref(sid)))
            self.species.append(sid.value)
        for name, arr in problem.fields.items():
            self.upload_field(name, arr)
        for name, arr in problem.flags.items():
            self._check(L.rh_upload_flags(self._h, FLAG_NAMES.index(name),
                                          _ptr(np.ascontiguousarray(arr, dtype=np.uint8))))
        for s, parts in enumerate(problem.particles):
            self.upload_particles(s, parts.cell, parts.state)
        for s, inj in enumerate(problem.injectors or []):
            if inj is not None:
                self.set_injector(s, inj)

    def _check(self, rc: int) -> None:
        if rc != RH_OK:
            msg = load_library().rh_last_error(self._h)
            raise RhError(rc, msg.decode() if msg else "")

    def close(self) -> None:
        if getattr(self, "_h", None):
            load_library().rh_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- data movement --------------------------------------------------
    def upload_field(self, name: str, arr: np.ndarray) -> None:
        a = np.ascontiguousarray(arr, dtype=np.float64)
        self._check(load_library().rh_upload_field(self._h, FIELD_NAMES.index(name), _ptr(a)))

    def field(self, name: str) -> np.ndarray:
        fid = FIELD_NAMES.index(name)
        out = np.empty(self.grid.nc * FIELD_DIM[fid], dtype=np.float64)
        self._check(load_library().rh_download_field(self._h, fid, _ptr(out)))
        return out

    def upload_particles(self, s: int, cell: np.ndarray, state: np.ndarray) -> None:
        cell = np.ascontiguousarray(cell, dtype=np.int64)
        state = np.ascontiguousarray(state, dtype=np.float64)
        self._check(load_library().rh_upload_particles(self._h, s, cell.shape[0], _ptr(cell), _ptr(state)))

    def num_particles(self, s: int) -> int:
        n = C.c_int64()
        self._check(load_library().rh_num_particles(self._h, s, C.byref(n)))
        return n.value

    def particles(self, s: int):
        n = self.num_particles(s)
        cell = np.empty(n, dtype=np.int64)
        st = np.empty((n, 7), dtype=np.float64)
        self._check(load_library().rh_download_particles(self._h, s, n, _ptr(cell), _ptr(st)))
        return cell, st

    # -- operators (reference names) ----------------------------------------
    def setup_get_fields(self) -> None: self._check(load_library().rh_setup_get_fields(self._h))
    def propagate(self, s: int) -> None: self._check(load_library().rh_propagate(self._h, s))
    def wait_particle_sends(self, s: int) -> None: self._check(load_library().rh_wait_particle_sends(self._h, s))
    def clear_accumulation_arrays(self) -> None: self._check(load_library().rh_clear_accumulation_arrays(self._h))
    def accumulate(self, s: int) -> None: self._check(load_library().rh_accumulate(self._h, s))
    def add_remote_updates(self) -> None: self._check(load_library().rh_add_remote_updates(self._h))
    def apply_boundary_conditions(self, s: int) -> None:
        self._check(load_library().rh_apply_boundary_conditions(self._h, s))
    def propagate_B(self) -> None: self._check(load_library().rh_propagate_B(self._h))
    def step(self, n: int = 1) -> None: self._check(load_library().rh_step(self._h, n))
    def synchronize(self) -> None: self._check(load_library().rh_synchronize(self._h))

    def deposit(self) -> None:
        """clear + accumulate every species + finalise: moments of the current particles."""
        self.clear_accumulation_arrays()
        for s in range(len(self.species)):
            self.accumulate(s)
        self.add_remote_updates()

    def diagnostics(self, s: int) -> np.ndarray:
        out = np.zeros(5)
        self._check(load_library().rh_diagnostics(self._h, s, _ptr(out)))
        return out

    def denoise_moments(self, denoiser, max_rel_change: float) -> None:
        """Learned denoiser on the deposited moments (no counterpart in the
        reference); call between add_remote_updates and propagate_B."""
        self._check(load_library().rh_denoise_moments(self._h, denoiser._h, float(max_rel_change)))

    def set_denoiser(self, denoiser, max_rel_change: float = 0.5) -> None:
        """Apply `denoiser` in every step before the field update (None detaches)."""
        self._denoiser = denoiser  # kept alive while attached
        h = denoiser._h if denoiser is not None else None
        self._check(load_library().rh_set_denoiser(self._h, h, float(max_rel_change)))

    # -- injection -------------------------------------------------------------
    def set_injector(self, s: int, inj) -> None:
        """Attach a problem.Injector to species s (None detaches)."""
        d = InjectorDesc()
        if inj is not None:
            d.kind, d.seed, d.temperature = inj.kind, inj.seed, inj.temperature
            for l in range(3):
                d.velocity[l] = inj.velocity[l]
            d.density, d.ppc, d.exobase_radius = inj.density, inj.ppc, inj.exobase_radius
            for l in range(6):
                d.box[l] = inj.box[l]
            rate = None
            if inj.cell_rate is not None:
                rate = np.ascontiguousarray(inj.cell_rate, dtype=np.float64)
                if rate.shape != (self.grid.nc,):
                    raise ValueError("cell_rate needs one value per cell")
                d.cell_rate = rate.ctypes.data
        self._check(load_library().rh_set_injector(self._h, s, C.byref(d)))

    def inject(self, s: int) -> None: self._check(load_library().rh_inject_particles(self._h, s))

    # places of one injection launch (32-bit place indices)
    MAX_PLACES = (1 << 31) - 1

    def _slabs(self, inj):
        """A uniform injector whose fixed places (ceil(ppc) per cell) would
        exceed MAX_PLACES, split into z slabs of its box that hold every cell
        centre exactly once (slab bounds on cell faces).  The particles drawn
        do not depend on the split: each one's generator counter is its cell,
        the timestep and its index in the cell."""
        import dataclasses
        import math
        g = self.grid
        if inj is None or inj.kind != P.INJ_UNIFORM:
            return [inj]
        per_plane = math.ceil(inj.ppc) * g.nx * g.ny
        planes = max(1, self.MAX_PLACES // max(per_plane, 1))
        if per_plane * g.nz <= self.MAX_PLACES:
            return [inj]
        out = []
        for k0 in range(0, g.nz, planes):
            k1 = min(g.nz, k0 + planes)
            b = list(inj.box)
            b[4] = max(b[4], g.z0 + k0 * g.dx)
            b[5] = min(b[5], g.z0 + k1 * g.dx)
            if b[4] < b[5]:
                out.append(dataclasses.replace(inj, box=tuple(b)))
        return out

    def fill(self, s: int, injector, timesteps=(1,)) -> None:
        """Run `injector` once per timestep in `timesteps` for species s (e.g.
        the reference's InjectorUniform at timestep 1), then restore the
        species' own injector and the timestep."""
        t0 = self.timestep()
        for part in self._slabs(injector):
            self.set_injector(s, part)
            for t in timesteps:
                self.set_timestep(t)
                self.inject(s)
        own = (self.problem.injectors or [None] * (s + 1))
        self.set_injector(s, own[s] if s < len(own) else None)
        self.set_timestep(t0)
    def set_timestep(self, t: int) -> None: self._check(load_library().rh_set_timestep(self._h, int(t)))

    def timestep(self) -> int:
        t = C.c_int64()
        self._check(load_library().rh_get_timestep(self._h, C.byref(t)))
        return t.value

    def injected(self, s: int) -> int:
        n = C.c_int64()
        self._check(load_library().rh_injected(self._h, s, C.byref(n)))
        return n.value

    # -- logs (the reference's field.log and pop*.log) ----------------------
    def log_values(self, t: float, include_inner: bool = False):
        """(field[41], pop[nspecies][17]) of one log line at time t; the log
        counters restart."""
        field = np.zeros(41)
        pop = np.zeros((len(self.species), 17))
        self._check(load_library().rh_log_values(self._h, float(t), int(include_inner), _ptr(field), _ptr(pop)))
        return field, pop

    def log_begin(self, t: float) -> None:
        self._check(load_library().rh_log_begin(self._h, float(t)))

    def log_open(self, directory: str, names=None, precision: int = 10, include_inner: bool = False,
                 max_B_stop: float = 0.0) -> None:
        names = names or [sp.name for sp in self.problem.species]
        arr = (C.c_char_p * len(names))(*[n.encode() for n in names])
        self._check(load_library().rh_log_open(self._h, directory.encode(), arr, int(precision),
                                               int(include_inner), float(max_B_stop)))

    def log_write(self, t: float) -> None:
        self._check(load_library().rh_log_write(self._h, float(t)))

    # -- output ------------------------------------------------------------------
    def set_temporal_averages(self, on: bool) -> None:
        """WRITE_GRID_TEMPORAL_AVERAGES: running sums of cellB and of each
        species' charge and current, written (and restarted) by write_vlsv's
        "cellBAverage", "n_ave" and "v_ave"."""
        self._check(load_library().rh_set_temporal_averages(self._h, int(on)))

    def write_vlsv(self, filename: str, t: float, timestep: int, variables, populations=None) -> None:
        """VLSV file of the selected cell variables (UserDataOP::writeData)."""
        pops = populations or [sp.name for sp in self.problem.species]
        v = (C.c_char_p * max(len(variables), 1))(*[x.encode() for x in variables])
        p = (C.c_char_p * len(pops))(*[x.encode() for x in pops])
        self._check(load_library().rh_write_vlsv(self._h, filename.encode(), float(t), int(timestep), v,
                                                 len(variables), p))

    def counters(self) -> Dict[str, float]:
        out = np.zeros(len(COUNTER_NAMES))
        self._check(load_library().rh_counters(self._h, _ptr(out)))
        return dict(zip(COUNTER_NAMES, out.tolist()))

    # -- measurement --------------------------------------------------------
    def set_timing(self, on: bool) -> None: self._check(load_library().rh_set_timing(self._h, int(on)))

    def set_timing_only(self, name) -> None:
        """Time only the launches of kernel `name` (None: every launch)."""
        self._check(load_library().rh_set_timing_only(self._h, name.encode() if name else None))

    def reset_timing(self) -> None: self._check(load_library().rh_reset_timing(self._h))

    def set_species_streams(self, on: bool) -> None:
        """step(): each species' push, migration and deposit on a stream of its
        own, side by side (the same results; for species crowded into few
        tiles)."""
        self._check(load_library().rh_set_species_streams(self._h, int(on)))

    def kernel_time(self, name: str):
        ms, n = C.c_double(), C.c_int64()
        self._check(load_library().rh_kernel_times(self._h, name.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def layout(self) -> Dict[str, int]:
        out = np.zeros(6, dtype=np.int64)
        self._check(load_library().rh_layout_info(self._h, _ptr(out)))
        return dict(zip(["ntiles", "tx", "ty", "tz", "ncells", "bytes_per_particle"], out.tolist()))


# ---------------------------------------------------------------------------
# The reference's operator objects, bound to a Context.

class BorisBuneman:
    """Particle propagator "BorisBuneman" (src/particle_propagator_boris_buneman.h:36-48).
    propagateCell runs for every cell of the species in one launch."""

    def __init__(self, ctx: Context, species: int):
        self.ctx, self.species = ctx, species

    def propagateCell(self) -> bool:
        self.ctx.propagate(self.species)
        return True


class Accumulator:
    """Particle accumulator "HybridAccumulator" (src/particle_accumulator.h:31-67).
    As in the reference, the first accumulator clears the arrays and the last
    one folds the remote (here: cross-tile) updates; the count of accumulators
    is kept per context (the reference keeps it in a static,
    src/particle_accumulator.cpp:25-32)."""

    def __init__(self, ctx: Context, species: int):
        self.ctx, self.species = ctx, species
        self.myOrderNumber = getattr(ctx, "_n_accumulators", 0)
        ctx._n_accumulators = self.myOrderNumber + 1

    def clearAccumulationArrays(self) -> bool:
        if self.myOrderNumber == 0:
            self.ctx.clear_accumulation_arrays()
        return True

    def accumulateBoundaryCells(self) -> bool:
        # boundary and inner cells are one launch on one GPU
        return True

    def accumulateInnerCells(self) -> bool:
        self.ctx.accumulate(self.species)
        return True

    def sendUpdates(self) -> bool:
        return True

    def wait(self) -> bool:
        return True

    def addRemoteUpdates(self) -> bool:
        if self.myOrderNumber == self.ctx._n_accumulators - 1:
            self.ctx.add_remote_updates()
        return True


class ParticleInjector:
    """A species' injector (ParticleInjectorBase, src/particle_injector.h):
    inject() adds the new particles of the context's current timestep."""

    def __init__(self, ctx: Context, species: int, injector=None):
        self.ctx, self.species = ctx, species
        if injector is not None:
            ctx.set_injector(species, injector)

    def inject(self) -> bool:
        self.ctx.inject(self.species)
        return True


def setupGetFields(ctx: Context) -> None:
    """src/hybrid_propagator.cpp:1861-1876."""
    ctx.setup_get_fields()


def propagateB(ctx: Context) -> bool:
    """src/hybrid_propagator.cpp:42-440."""
    ctx.propagate_B()
    return True


def propagate(ctx: Context, propagators: List[BorisBuneman], accumulators: List[Accumulator],
              injectors: Optional[List[ParticleInjector]] = None) -> bool:
    """One time step in the order of propagate() (src/user.cpp:118-134),
    without logging.  The caller advances the timestep first (Corsair's
    main loop does, ctx.set_timestep)."""
    setupGetFields(ctx)
    for p in propagators:
        p.propagateCell()
    for a in accumulators:
        a.clearAccumulationArrays()
    for s in range(len(propagators)):
        ctx.wait_particle_sends(s)
    for a in accumulators:
        a.accumulateBoundaryCells()
        a.accumulateInnerCells()
        a.sendUpdates()
    for a in accumulators:
        a.wait()
        a.addRemoteUpdates()
    for s in range(len(propagators)):
        ctx.apply_boundary_conditions(s)
    for inj in injectors or []:
        inj.inject()
    return propagateB(ctx)
