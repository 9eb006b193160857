"""The learned denoiser in the loop, against the plain method.

On one obstacle-box geometry (config_obstacle, smaller than the bench box):

  1. training pairs from a 256-ppc FP32 run (inputs: 1/32 subsamples, i.e.
     8 ppc, with the run's cellB; targets: the 256-ppc deposit) and a few
     hundred Adam steps on the GPU;
  2. from one fresh seed: an FP64 256-ppc run (the reference), plain FP32
     runs at 8, 64 and 256 ppc, and an FP32 8-ppc run with the denoiser
     attached (applied before every field update, fallback bound against the
     binomially filtered deposit);
  3. per run: wall time per step (HIP-synchronised), and the relative L2
     error of n and J against the reference over the plasma cells (outside
     the obstacle, not on the exterior layer) after the same number of steps.

    python -m rhybrid_amd.denoise_loop [--steps 40] [--nx 64 --ny 32 --nz 32]
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import denoise as D
from . import denoise_data as DD
from . import problem as P
from .gpu import RH_FP32, RH_FP64, Context


def _timed_run(pr, precision, steps, denoiser=None, bound=0.5):
    ctx = Context(pr, precision=precision)
    if denoiser is not None:
        ctx.set_denoiser(denoiser, bound)
    ctx.step(1)            # warm-up (first launches)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.step(steps - 1)
    ctx.synchronize()
    dt = (time.perf_counter() - t0) / max(steps - 1, 1)
    ctx.deposit()          # moments of the final particles (the denoised run: raw deposit)
    if denoiser is not None:
        ctx.denoise_moments(denoiser, bound)
    out = {"rho": ctx.field("cellRhoQi"), "J": ctx.field("cellJi").reshape(-1, 3),
           "ms_per_step": dt * 1e3, "fallbacks": ctx.counters()["denoise_fallback"]}
    if denoiser is not None:
        ctx.set_denoiser(None)
    ctx.close()
    return out


def compare(nx=64, ny=32, nz=32, R_cells=5.0, steps=40, train_snapshots=4, train_steps_between=6,
            train_iters=300, lr=2e-3, bound=0.5, seed_train=11, seed_eval=23):
    geom = dict(nx=nx, ny=ny, nz=nz, R_cells=R_cells)
    # 1. training
    pr_train = P.config_obstacle(ppc=256, seed=seed_train, **geom)
    scales = DD.channel_scales(pr_train)
    t0 = time.perf_counter()
    X, T = DD.make_pairs(pr_train, subsample=32, snapshots=train_snapshots,
                         steps_between=train_steps_between, seed=seed_train)
    t_pairs = time.perf_counter() - t0
    del pr_train
    import torch
    V = nx * ny * nz
    dn = D.Denoiser(train_snapshots * V, scales=scales, weights=D.init_weights(seed=2, gain=0.1))
    xt, tt = torch.from_numpy(X).cuda(), torch.from_numpy(T).cuda()
    t0 = time.perf_counter()
    losses = [dn.train_step(xt, tt, lr=lr) for _ in range(train_iters)]
    t_train = time.perf_counter() - t0
    raw_err = DD.rel_error(X[..., :4], T, scales)
    del xt, tt, X, T
    # 2. runs from one fresh seed
    runs = {}
    runs["fp64_256"] = _timed_run(P.config_obstacle(ppc=256, seed=seed_eval, **geom), RH_FP64, steps)
    for ppc in (8, 64, 256):
        runs["fp32_%d" % ppc] = _timed_run(P.config_obstacle(ppc=ppc, seed=seed_eval, **geom), RH_FP32, steps)
    runs["fp32_8_denoised"] = _timed_run(P.config_obstacle(ppc=8, seed=seed_eval, **geom), RH_FP32, steps,
                                         denoiser=dn, bound=bound)
    dn.close()
    # 3. errors over the plasma cells
    g = P.config_obstacle(ppc=1, seed=seed_eval, **geom)
    plasma = g.grid.all_neighbours() & (g.flags["innerFlagField"] == 0)
    ref = runs["fp64_256"]
    out = {"geometry": "%dx%dx%d obstacle box, R=%g cells, H+ and He++ (4%%), %d steps" % (nx, ny, nz, R_cells, steps),
           "training": {"pairs": "%d snapshots of a 256-ppc FP32 run, 1/32 subsamples as 8-ppc inputs" % train_snapshots,
                        "iters": train_iters, "loss_first": losses[0], "loss_last": losses[-1],
                        "seconds_pairs": t_pairs, "seconds_train": t_train,
                        "raw_8ppc_error_on_pairs": {"n": raw_err[0], "J": raw_err[1]}},
           "fallback_bound": bound, "runs": {}}
    for name, r in runs.items():
        en = np.linalg.norm((r["rho"] - ref["rho"])[plasma]) / np.linalg.norm(ref["rho"][plasma])
        eJ = np.linalg.norm((r["J"] - ref["J"])[plasma]) / np.linalg.norm(ref["J"][plasma])
        entry = {"ms_per_step": round(r["ms_per_step"], 4), "n_error": float(en), "J_error": float(eJ)}
        if name == "fp32_8_denoised":
            entry["fallback_fraction"] = r["fallbacks"] / float(steps * g.grid.nc)
        out["runs"][name] = entry
    d8 = out["runs"]["fp32_8_denoised"]
    p64 = out["runs"]["fp32_64"]
    out["denoised_8_vs_plain_64"] = {"time_ratio": p64["ms_per_step"] / d8["ms_per_step"],
                                     "n_error_ratio": d8["n_error"] / p64["n_error"],
                                     "J_error_ratio": d8["J_error"] / p64["J_error"]}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--nx", type=int, default=64)
    ap.add_argument("--ny", type=int, default=32)
    ap.add_argument("--nz", type=int, default=32)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--iters", type=int, default=300)
    ap.add_argument("--bound", type=float, default=0.5)
    a = ap.parse_args()
    print(json.dumps(compare(nx=a.nx, ny=a.ny, nz=a.nz, steps=a.steps, train_iters=a.iters, bound=a.bound)),
          flush=True)


if __name__ == "__main__":
    main()


def _checkpointed_run(pr, precision, steps, every, dn=None, bound=0.5, log=None, name=""):
    """Run `steps` full steps of a device-filled problem, `every` steps a
    checkpoint of n and J (float32, host); the time per step excludes the
    checkpoints' downloads."""
    from .denoise_train import filled_context
    ctx = filled_context(pr, precision=precision)
    if dn is not None:
        ctx.set_denoiser(dn, bound)
    ctx.synchronize()
    cps, t_steps, done = [], 0.0, 0
    while done < steps:
        k = min(every, steps - done)
        t0 = time.perf_counter()
        ctx.step(k)
        ctx.synchronize()
        t_steps += time.perf_counter() - t0
        done += k
        cps.append((done, ctx.field("cellRhoQi").astype(np.float32),
                    ctx.field("cellJi").reshape(-1, 3).astype(np.float32)))
        if log is not None:
            log("%s: step %d, %.3f ms/step" % (name, done, t_steps / done * 1e3))
    out = {"ms_per_step": t_steps / steps * 1e3, "checkpoints": cps,
           "particles": [ctx.num_particles(s) for s in range(len(pr.species))],
           "fallbacks": ctx.counters()["denoise_fallback"]}
    if dn is not None:
        ctx.set_denoiser(None)
    ctx.close()
    return out


def compare_uneven(dn, nx=256, ny=192, nz=192, R_cells=12.0, H_cells=5.0, steps=2000, every=250,
                   ppcs=(8, 64, 256), bound=0.5, seed=23, log=None):
    """The denoiser in the loop on the uneven-load geometry (config_exosphere:
    H+ solar wind at ppc, O+ photoions in proportion): plain FP32 runs at each
    of `ppcs` and a denoised FP32 run at the lowest, from the same seeded
    distribution (device fill), against the highest-ppc run as the truth;
    n and J errors over the plasma cells every `every` steps, time per step,
    fallback fraction."""
    make = lambda ppc: P.config_exosphere(nx=nx, ny=ny, nz=nz, ppc=ppc, R_cells=R_cells, H_cells=H_cells,
                                          seed=seed, o_per_ppc=True, steps=steps)
    runs = {}
    truth_ppc, low = max(ppcs), min(ppcs)
    for ppc in sorted(ppcs, reverse=True):
        runs["fp32_%d" % ppc] = _checkpointed_run(make(ppc), RH_FP32, steps, every, log=log, name="fp32_%d" % ppc)
    runs["fp32_%d_denoised" % low] = _checkpointed_run(make(low), RH_FP32, steps, every, dn=dn, bound=bound,
                                                       log=log, name="denoised")
    g = make(1)
    plasma = g.grid.all_neighbours() & (g.flags["innerFlagField"] == 0)
    truth = runs["fp32_%d" % truth_ppc]["checkpoints"]
    out = {"geometry": "%dx%dx%d cells, exospheric O+ around an absorbing sphere of %g cells, H+ solar wind; "
                       "%d steps" % (nx, ny, nz, R_cells, steps),
           "truth": "plain FP32 at %d ppc" % truth_ppc, "fallback_bound": bound, "runs": {}}
    for name, r in runs.items():
        rows = []
        for (st, rho, J), (_, rho_j, J_j) in zip(r["checkpoints"], truth):
            en = np.linalg.norm((rho - rho_j)[plasma]) / np.linalg.norm(rho_j[plasma])
            eJ = np.linalg.norm((J - J_j)[plasma]) / np.linalg.norm(J_j[plasma])
            rows.append({"step": st, "n_error": float(en), "J_error": float(eJ)})
        entry = {"ms_per_step": round(r["ms_per_step"], 4), "errors": rows, "particles": r["particles"]}
        if "denoised" in name:
            entry["fallback_fraction"] = r["fallbacks"] / float(steps * g.grid.nc)
        out["runs"][name] = entry
    d = out["runs"]["fp32_%d_denoised" % low]
    p64 = out["runs"].get("fp32_64")
    if p64 is not None:
        last_d, last_p = d["errors"][-1], p64["errors"][-1]
        out["denoised_vs_plain_64"] = {
            "time_ratio": p64["ms_per_step"] / d["ms_per_step"],
            "n_error_ratio_last": last_d["n_error"] / last_p["n_error"],
            "J_error_ratio_last": last_d["J_error"] / last_p["J_error"],
            "n_error_ratio_mean": float(np.mean([a["n_error"] / b["n_error"] for a, b in zip(d["errors"], p64["errors"])])),
            "J_error_ratio_mean": float(np.mean([a["J_error"] / b["J_error"] for a, b in zip(d["errors"], p64["errors"])]))}
    return out


def truth_precision(nx=64, ny=48, nz=48, R_cells=6.0, H_cells=2.5, ppc=256, steps=100, seed=23):
    """The truth run's precision: FP32 against FP64 at the truth run's ppc on a
    sub-box of the uneven-load geometry, from the same device fill."""
    from .denoise_train import filled_context
    pr = P.config_exosphere(nx=nx, ny=ny, nz=nz, ppc=ppc, R_cells=R_cells, H_cells=H_cells, seed=seed,
                            o_per_ppc=True)
    res = {}
    for prec in (RH_FP32, RH_FP64):
        ctx = filled_context(pr, precision=prec)
        ctx.step(steps)
        res[prec] = (ctx.field("cellRhoQi"), ctx.field("cellJi"), ctx.field("faceB"), ctx.field("nodeE"))
        ctx.close()
    names = ("n", "J", "B", "E")
    return {"box": "%dx%dx%d, %d ppc, %d steps" % (nx, ny, nz, ppc, steps),
            **{k: float(np.linalg.norm(a - b) / np.linalg.norm(b)) for k, a, b in zip(names, res[RH_FP32], res[RH_FP64])}}
