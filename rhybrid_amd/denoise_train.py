"""Training pairs from independent runs, patch sets and training of the
learned moment denoiser at full size.

A pair is the state of two runs of one seeded distribution at the same
step: a high-ppc run (the target: its deposited n and J) and a low-ppc run
of the same initial distribution, drawn independently with its own seed
(the input: its deposited n and J and its cellB).  Both runs are filled on
the device (the reference's InjectorUniform at timestep 1) and fed by their
solar-wind injectors, so a run at hundreds of particles per cell never
crosses the host.  Patches of 32^3 voxels are cut at random offsets from the
snapshots and kept on the device; training runs the library's own BF16-MFMA
forward, backward and Adam kernels on batches of patches.
"""
from __future__ import annotations

import time
from typing import Callable, Dict, List, Optional, Tuple

import numpy as np

from . import denoise as D
from . import problem as P
from .gpu import RH_FP32, Context

PATCH = 32


def filled_context(pr: P.Problem, precision: int = RH_FP32, tile=(0, 0, 0)) -> Context:
    """A context of `pr` with the device fill of meta["fill"] done."""
    ctx = Context(pr, precision=precision, tile=tile)
    for s, inj in enumerate(pr.meta.get("fill") or []):
        if inj is not None:
            ctx.fill(s, inj)
    return ctx


def _volume(ctx: Context, g: P.Grid, with_B: bool):
    """[Z][Y][X][4 or 7] float32 torch tensor on the device: n, J (and cellB)."""
    import torch
    rho = ctx.field("cellRhoQi")
    ji = ctx.field("cellJi").reshape(-1, 3)
    parts = [rho[:, None], ji]
    if with_B:
        parts.append(ctx.field("cellB").reshape(-1, 3))
    v = np.concatenate(parts, 1).astype(np.float32).reshape(g.nz, g.ny, g.nx, -1)
    return torch.from_numpy(v).cuda()


def paired_runs(make: Callable[[int, int], P.Problem], hi_ppc: int, lo_ppc: int, seed_hi: int, seed_lo: int,
                warm: int, snapshots: int, every: int, crops_per_snapshot: int, crop_seed: int,
                log=None) -> Tuple[object, object, Dict[str, float]]:
    """Step a high-ppc and a low-ppc run of make(ppc, seed) side by side:
    `warm` steps, then `snapshots` snapshots `every` steps apart; from each,
    `crops_per_snapshot` 32^3 crops at random offsets.  Returns device
    tensors (inputs [P][32][32][32][7], targets [P][32][32][32][4]) and
    timings."""
    import torch
    t0 = time.perf_counter()
    pr_hi, pr_lo = make(hi_ppc, seed_hi), make(lo_ppc, seed_lo)
    g = pr_hi.grid
    hi, lo = filled_context(pr_hi), filled_context(pr_lo)
    del pr_hi, pr_lo
    hi.step(warm)
    lo.step(warm)
    hi.synchronize()
    t_warm = time.perf_counter() - t0
    rng = np.random.default_rng(crop_seed)
    X = torch.empty((snapshots * crops_per_snapshot, PATCH, PATCH, PATCH, 7), dtype=torch.float32, device="cuda")
    T = torch.empty((snapshots * crops_per_snapshot, PATCH, PATCH, PATCH, 4), dtype=torch.float32, device="cuda")
    q = 0
    for s in range(snapshots):
        hi.step(every)
        lo.step(every)
        hi.deposit()
        lo.deposit()
        vt = _volume(hi, g, False)
        vx = _volume(lo, g, True)
        for _ in range(crops_per_snapshot):
            z, y, x = (int(rng.integers(0, n - PATCH + 1)) for n in (g.nz, g.ny, g.nx))
            X[q] = vx[z:z + PATCH, y:y + PATCH, x:x + PATCH]
            T[q] = vt[z:z + PATCH, y:y + PATCH, x:x + PATCH]
            q += 1
        if log is not None and (s + 1) % max(1, snapshots // 10) == 0:
            log("pairs: snapshot %d/%d (%.1f s)" % (s + 1, snapshots, time.perf_counter() - t0))
    hi.close()
    lo.close()
    return X, T, {"seconds_warm": t_warm, "seconds_total": time.perf_counter() - t0}


def patch_errors(dn: Optional[D.Denoiser], X, T, scales, batch: int = 16) -> Tuple[float, float]:
    """Relative L2 error of n and J (normalised channels) of the denoised
    inputs (dn None: the raw inputs) against the targets over a patch set."""
    import torch
    s = torch.tensor(scales[:4], dtype=torch.float64, device=X.device)
    dn_, dJ, rn, rJ = 0.0, 0.0, 0.0, 0.0
    for b in range(0, X.shape[0], batch):
        x, t = X[b:b + batch].contiguous(), T[b:b + batch]
        y = x[..., :4] if dn is None else dn.forward(x)
        d = (y.double() - t.double()) / s
        r = t.double() / s
        dn_ += float((d[..., 0] ** 2).sum())
        dJ += float((d[..., 1:] ** 2).sum())
        rn += float((r[..., 0] ** 2).sum())
        rJ += float((r[..., 1:] ** 2).sum())
    return float(np.sqrt(dn_ / rn)), float(np.sqrt(dJ / rJ))


def train(dn: D.Denoiser, X, T, epochs: int, batch: int, lr: float, lr_final: float, seed: int = 0,
          log=None) -> Dict[str, object]:
    """Adam over shuffled batches of the patch set, the learning rate
    decaying geometrically from lr to lr_final; returns the per-epoch mean
    loss and the wall time."""
    import torch
    n = X.shape[0]
    steps = epochs * (n // batch)
    rng = np.random.default_rng(seed)
    losses: List[float] = []
    t0 = time.perf_counter()
    it = 0
    for e in range(epochs):
        order = torch.from_numpy(rng.permutation(n)).cuda()
        acc = 0.0
        for b in range(n // batch):
            idx = order[b * batch:(b + 1) * batch]
            x, t = X[idx].contiguous(), T[idx].contiguous()
            r = lr * (lr_final / lr) ** (it / max(steps - 1, 1))
            acc += dn.train_step(x, t, lr=r)
            it += 1
        losses.append(acc / max(n // batch, 1))
        if log is not None:
            log("train: epoch %d/%d loss %.4e (%.1f s)" % (e + 1, epochs, losses[-1], time.perf_counter() - t0))
    dn.synchronize()
    return {"epoch_loss": losses, "seconds": time.perf_counter() - t0, "steps": it, "batch": batch}
