"""Training the denoiser on pairs from full-size obstacle-box runs, then
running it in the loop on the uneven-load geometry.

  1. pairs (the training configuration): a high-ppc and an independent
     low-ppc run of the inflow/outflow obstacle box (128x96x96 by default,
     both filled on the device from their own seeds), 32^3 crops of their
     moment fields at the same steps; a held-out set from a second pair of
     runs with other seeds;
  2. training on the GPU by the library's kernels (Adam, BF16 MFMA);
  3. the denoiser in the loop on the uneven-load geometry against the plain
     method at 8, 64 and 256 ppc (denoise_loop.compare_uneven), and the
     truth run's FP32-vs-FP64 precision on a sub-box.

    python -m rhybrid_amd.denoise_experiment [--patches 20000] [--steps 2000] [--out file.json]
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import denoise as D
from . import denoise_data as DD
from . import denoise_loop as DL
from . import denoise_train as DT
from . import problem as P


def log(msg: str) -> None:
    print("# " + msg, file=sys.stderr, flush=True)


def run(train_box=(128, 96, 96), R_train=10.0, hi_ppc=256, lo_ppc=8, patches=20000, crops=36, warm=300, every=2,
        heldout_patches=1152, epochs=8, batch=8, lr=2e-3, lr_final=1e-4, eval_box=(256, 192, 192),
        R_eval=12.0, H_eval=5.0, steps=2000, every_eval=250, bound=0.5, truth_box=(64, 48, 48),
        truth_steps=100, weights_out=None, train_geometry="obstacle"):
    t_all = time.perf_counter()
    nx, ny, nz = train_box
    if train_geometry == "obstacle":
        # config 4: pairs from the inflow/outflow obstacle box (H+ and He++)
        make = lambda ppc, seed: P.config_obstacle_filled(nx=nx, ny=ny, nz=nz, ppc=ppc, seed=seed, R_cells=R_train)
    else:
        # pairs from the uneven-load geometry itself (H+ and exospheric O+),
        # scaled to the training box
        make = lambda ppc, seed: P.config_exosphere(nx=nx, ny=ny, nz=nz, ppc=ppc, R_cells=R_train,
                                                    H_cells=H_eval * R_train / R_eval, seed=seed, o_per_ppc=True,
                                                    steps=warm + 4 * patches // crops)
    scales = DD.channel_scales(make(1, 1))
    snapshots = -(-patches // crops)
    X, T, tp = DT.paired_runs(make, hi_ppc, lo_ppc, seed_hi=101, seed_lo=202, warm=warm, snapshots=snapshots,
                              every=every, crops_per_snapshot=crops, crop_seed=7, log=log)
    X, T = X[:patches], T[:patches]
    hs = -(-heldout_patches // crops)
    Xh, Th, _ = DT.paired_runs(make, hi_ppc, lo_ppc, seed_hi=303, seed_lo=404, warm=warm, snapshots=hs,
                               every=every, crops_per_snapshot=crops, crop_seed=8, log=log)
    log("pairs done: %d training, %d held-out patches" % (X.shape[0], Xh.shape[0]))
    dn = D.Denoiser(batch * DT.PATCH ** 3, scales=scales, weights=D.init_weights(seed=2, gain=0.1))
    raw = DT.patch_errors(None, Xh, Th, scales, batch=batch)
    tr = DT.train(dn, X, T, epochs=epochs, batch=batch, lr=lr, lr_final=lr_final, log=log)
    den = DT.patch_errors(dn, Xh, Th, scales, batch=batch)
    w = dn.weights()
    dn.close()
    if weights_out:
        np.save(weights_out, w)
    del X, T, Xh, Th
    import torch
    torch.cuda.empty_cache()
    flops = 3 * D.forward_flops(DT.PATCH ** 3) * patches * epochs
    training = {"pairs": "%dx%dx%d %s box (R=%g cells), %d-ppc target and independent %d-ppc input runs, "
                         "%d warm-up steps, snapshots every %d steps, %d crops of 32^3 each"
                         % (nx, ny, nz, train_geometry, R_train, hi_ppc, lo_ppc, warm, every, crops),
                "patches": patches, "heldout_patches": heldout_patches, "epochs": epochs, "batch": batch,
                "epoch_loss": tr["epoch_loss"], "seconds_pairs": tp["seconds_total"], "seconds_train": tr["seconds"],
                "train_TFLOPs": flops / tr["seconds"] / 1e12,
                "heldout_error_raw": {"n": raw[0], "J": raw[1]}, "heldout_error_denoised": {"n": den[0], "J": den[1]}}
    log("training: held-out n %.4f -> %.4f, J %.4f -> %.4f" % (raw[0], den[0], raw[1], den[1]))
    ex, ey, ez = eval_box
    dn = D.Denoiser(ex * ey * ez, scales=scales, weights=w)
    loop = DL.compare_uneven(dn, nx=ex, ny=ey, nz=ez, R_cells=R_eval, H_cells=H_eval, steps=steps,
                             every=every_eval, bound=bound, log=log)
    dn.close()
    jx, jy, jz = truth_box
    truth = DL.truth_precision(nx=jx, ny=jy, nz=jz, R_cells=R_eval * jx / ex, H_cells=H_eval * jx / ex,
                               steps=truth_steps)
    return {"training": training, "in_loop": loop, "truth_fp32_vs_fp64": truth,
            "seconds_total": time.perf_counter() - t_all}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train-box", type=int, nargs=3, default=[128, 96, 96])
    ap.add_argument("--R-train", type=float, default=10.0)
    ap.add_argument("--patches", type=int, default=20000)
    ap.add_argument("--heldout", type=int, default=1152)
    ap.add_argument("--warm", type=int, default=300)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--eval-box", type=int, nargs=3, default=[256, 192, 192])
    ap.add_argument("--R-eval", type=float, default=12.0)
    ap.add_argument("--H-eval", type=float, default=5.0)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--every", type=int, default=250)
    ap.add_argument("--bound", type=float, default=0.5)
    ap.add_argument("--truth-box", type=int, nargs=3, default=[64, 48, 48])
    ap.add_argument("--weights-out", default=None)
    ap.add_argument("--out", default=None)
    ap.add_argument("--train-geometry", choices=("obstacle", "exosphere"), default="obstacle")
    a = ap.parse_args()
    res = run(train_box=tuple(a.train_box), R_train=a.R_train, patches=a.patches, heldout_patches=a.heldout,
              warm=a.warm, epochs=a.epochs, eval_box=tuple(a.eval_box), R_eval=a.R_eval, H_eval=a.H_eval,
              steps=a.steps, every_eval=a.every, bound=a.bound, truth_box=tuple(a.truth_box),
              weights_out=a.weights_out, train_geometry=a.train_geometry)
    line = json.dumps(res)
    if a.out:
        with open(a.out, "w") as f:
            f.write(line + "\n")
    print(line, flush=True)


if __name__ == "__main__":
    main()
