"""Training pairs for the learned moment denoiser.

A pair is two deposits of the same particle state: the target is the moment
field (n, J) of a high-ppc run; the input is the deposit of a random 1/k
subsample of its particles with k times the weight (the noise of a run at
ppc/k), together with the run's cellB.  Both deposits use the library's own
kernels; the subsample is drawn on the host.
"""
from __future__ import annotations

import dataclasses
from typing import List, Tuple

import numpy as np

from . import problem as P
from .gpu import RH_FP32, Context


def channel_scales(problem: P.Problem) -> List[float]:
    """Normalisation of the 7 input channels: upstream charge density,
    charge density times flow speed, upstream |B|."""
    meta = problem.meta
    rho = float(meta.get("density", 1e6)) * P.QE
    U = float(np.linalg.norm(meta.get("U", (4e5, 0.0, 0.0)))) or 4e5
    B = float(np.linalg.norm(problem.fields["faceB"][:3])) or 1e-9
    return [rho, rho * U, rho * U, rho * U, B, B, B]


def _moments(ctx: Context, g: P.Grid) -> np.ndarray:
    rho = ctx.field("cellRhoQi")
    ji = ctx.field("cellJi").reshape(-1, 3)
    return np.concatenate([rho[:, None], ji], 1).reshape(g.nz, g.ny, g.nx, 4)


def make_pairs(problem: P.Problem, subsample: int, snapshots: int, steps_between: int,
               seed: int = 1) -> Tuple[np.ndarray, np.ndarray]:
    """Run `problem` (high ppc) on the GPU and return (inputs [S,Z,Y,X,7],
    targets [S,Z,Y,X,4]) for `snapshots` states `steps_between` steps apart."""
    g = problem.grid
    hi = Context(problem, precision=RH_FP32)
    lo_problem = dataclasses.replace(
        problem,
        species=[dataclasses.replace(sp, w=sp.w * subsample) for sp in problem.species],
        particles=[P.Particles(p.cell[:0].copy(), p.state[:0].copy()) for p in problem.particles])
    lo = Context(lo_problem, precision=RH_FP32)
    rng = np.random.default_rng(seed)
    xs, ts = [], []
    for _ in range(snapshots):
        hi.step(steps_between)
        hi.deposit()
        target = _moments(hi, g)
        B = hi.field("cellB").reshape(g.nz, g.ny, g.nx, 3)
        for s in range(len(problem.species)):
            cell, state = hi.particles(s)
            keep = rng.random(cell.shape[0]) < 1.0 / subsample
            st = state[keep].copy()
            st[:, 6] *= subsample
            lo.upload_particles(s, cell[keep], st)
        lo.deposit()
        xs.append(np.concatenate([_moments(lo, g), B], -1))
        ts.append(target)
    hi.close()
    lo.close()
    return np.stack(xs).astype(np.float32), np.stack(ts).astype(np.float32)


def rel_error(a: np.ndarray, b: np.ndarray, scales) -> Tuple[float, float]:
    """Relative L2 error of n and of J (normalised channels) of a against b."""
    s = np.asarray(scales[:4], dtype=np.float64)
    d = (a.astype(np.float64) - b) / s
    r = b / s
    en = np.linalg.norm(d[..., 0]) / np.linalg.norm(r[..., 0])
    eJ = np.linalg.norm(d[..., 1:]) / np.linalg.norm(r[..., 1:])
    return float(en), float(eJ)
