"""Benchmark of the hot path on one MI355X.

Workload: 128x96x96-cell inflow/outflow box (dx=300 km) with an absorbing
sphere of radius 10 cells, H+ and He++ (4% by number) at 32 particles per
cell each (~71 M macroparticles), drifting Maxwellians at 450 km/s, IMF
(-1.5, 2.5, 0) nT, FP32 cell-local particles; uniform preload plus
solar-wind inflow of both species through the +x wall.  One step = push +
cell sort + CIC deposit + FP64 ghost fold + particle boundary conditions +
injection + field update (propagate() of the reference, src/user.cpp:118-134).

value = particle-steps per second over the whole timed region (all of the
step, not only push+deposit); roofline = the kernel with the largest total
time, algorithmic bytes per launch over its mean launch time from HIP events
on the library's stream; cpu_baseline = the FP64 oracle on the host cores on
a smaller box of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0      # MI355X published HBM3E bandwidth
BENCH_TILE = (4, 4, 4)     # the tile shape the context picks for the headline box
BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 matrix peak


def workload(scale: int = 1, seed: int = 6461674):
    from rhybrid_amd import problem as P
    return P.config_obstacle(nx=128 // scale, ny=96 // scale, nz=96 // scale, ppc=32,
                             R_cells=10.0 / scale, seed=seed, inflow=True)


def kernel_bytes(name: str, ctx, np_total: float, nspecies: int, npp: float = None) -> float:
    """Algorithmic HBM bytes of one launch of a kernel (FP32 store: 6 floats
    per particle), for the mean species population (or `npp` particles of
    the launch's species)."""
    lay = ctx.layout()
    nc, ntiles = lay["ncells"], lay["ntiles"]
    next_ = (lay["tx"] + 2) * (lay["ty"] + 2) * (lay["tz"] + 2)
    extk = next_ * 8
    nkeys = ntiles * lay["tx"] * lay["ty"] * lay["tz"] * 8
    if npp is None:
        npp = np_total / nspecies  # particles per species launch (mean)
    nh = (lay["tx"] + 1) * (lay["ty"] + 1) * (lay["tz"] + 1)
    tc = lay["tx"] * lay["ty"] * lay["tz"]
    fblock = 4.0 * (3 * nh + 7 * tc) * ntiles   # per-tile field blocks (FP32 + neighbourhood masks)
    if name == "push_fields":  # read 9 f32 fields per cell, write the per-tile blocks
        return 36.0 * nc + fblock
    if name == "push":       # pass 0: read 6 f32 + local cell byte and the field blocks; write 16-bit slot counts
        return 25.0 * npp + fblock + 2.0 * extk * ntiles + 8.0 * ntiles
    if name == "move":       # pass 1: read and write 6 f32 + local cell byte; field blocks; 16-bit slot
        # offsets and the 32-bit starts of the extended region's subcells
        return 50.0 * npp + fblock + 6.0 * extk * ntiles + 8.0 * ntiles
    if name == "deposit":    # read 6 f32 per particle and the subcell starts; write the f32 tile ghosts
        return 24.0 * npp + 4.0 * nkeys + 16.0 * next_ * ntiles
    if name == "fold":       # read the f32 tile ghosts once, write rho + J (f32)
        return 16.0 * next_ * ntiles + 16.0 * nc
    if name == "sort_slots":  # read and write the 16-bit slot table, write the totals
        return 4.0 * extk * ntiles + 4.0 * nkeys
    if name == "field_update":  # fused propagateB: read faceB, rho, Ji, eta (32 B), flags (3 B); write faceB,
        # cellUe, cellEp (36 B) and the per-tile push-field blocks (the last step of a call also writes the
        # reference's other arrays; not counted)
        return 71.0 * nc + fblock
    return 0.0


PROFILE_KERNELS = {"push": "rh::k_push_move<float>_[pass_0:_push]", "move": "rh::k_push_move<float>_[pass_1:_move]",
                   "deposit": "rh::k_deposit<float>", "fold": "rh::k_fold<float>",
                   "sort_slots": "rh::k_slot_offsets"}


def measured_traffic(kernel: str, profile: str = "hbm_traffic_obstacle128x96x96_ppc32_inflow.txt",
                     species: str = ""):
    """HBM bytes per launch of a kernel from a committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE, each scaled by the factor a 1 GiB copy
    gives; `species` selects a per-species line "<kernel>:<species>"); None if
    the summary does not list the kernel."""
    path = os.path.join(ROOT, "profiles", profile)
    name = PROFILE_KERNELS.get(kernel)
    if name is None or not os.path.exists(path):
        return None
    if species:
        name += ":" + species
    for line in open(path):
        if line.startswith("#"):
            continue
        f = line.rstrip("\n").split("\t")
        if f[0] == name:
            return float(f[6])
    return None


def step_bound(np_mean: float, ncells: int, nspec: int, ms_per_step: float) -> dict:
    """The whole step against the path's per-unit bytes: 56 B per
    particle-step (push + deposit, FP32), 52 B per cell-step per species
    (slot and ghost traffic) and 68 B per cell-step (the fused field update);
    the time those bytes take at the HBM peak over the measured step."""
    b = 56.0 * np_mean + 52.0 * ncells * nspec + 68.0 * ncells
    bound_ms = b / (HBM_PEAK_GBS * 1e9) * 1e3
    return {"per_unit": "56 B/particle-step + 52 B/cell-step per species + 68 B/cell-step",
            "bytes_per_step": b, "bound_ms_at_peak": round(bound_ms, 4),
            "frac": round(bound_ms / ms_per_step, 4) if ms_per_step > 0 else None}


def gpu_bench(args, pr):
    from rhybrid_amd.gpu import RH_FP32, Context
    n0 = pr.n_particles
    t0 = time.time()
    ctx = Context(pr, precision=RH_FP32)
    t_up = time.time() - t0
    print("# workload %s: %d cells, %d particles (upload %.1fs)" %
          (pr.name, pr.grid.nc, n0, t_up), file=sys.stderr, flush=True)
    ctx.step(args.warmup)
    ctx.synchronize()
    nspec = len(ctx.species)
    np_before = sum(ctx.num_particles(s) for s in range(nspec))
    # the timed region records events around the dominant kernel's launches
    # only (events around every launch add about 0.3 ms per step)
    dominant = "move"
    ctx.reset_timing()
    ctx.set_timing_only(dominant)
    ctx.set_timing(True)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.step(args.steps)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    ctx.set_timing(False)
    dom_ms, dom_n = ctx.kernel_time(dominant)
    np_after = sum(ctx.num_particles(s) for s in range(nspec))
    np_mean = 0.5 * (np_before + np_after)
    # per-kernel breakdown: a further run with every launch timed
    nb = min(args.steps, 10)
    ctx.reset_timing()
    ctx.set_timing_only(None)
    ctx.set_timing(True)
    ctx.step(nb)
    ctx.synchronize()
    ctx.set_timing(False)
    names = ["push_fields", "push", "sort_slots", "sort_count", "scan", "sort_base", "move", "deposit",
             "deposit_dense", "fold", "boundary", "inject_slots", "inject_count", "inject_fill", "sort_still",
             "cell_B_bc", "neumann", "imf", "cell2node", "node_J", "node2cell", "cell_J_Ue", "node_Ue_B_E",
             "node_avg", "face_curl_Ep", "field_update", "fill", "diag"]
    ktimes = {}
    for nm in names:
        ms, n = ctx.kernel_time(nm)
        if n:
            ktimes[nm] = (ms, n)
    total_k = sum(v[0] for v in ktimes.values())
    pd_names = ("push_fields", "push", "sort_slots", "sort_count", "scan", "sort_base", "move", "deposit",
                "deposit_dense", "fold")
    t_pd = sum(ktimes[k][0] for k in pd_names if k in ktimes) / 1e3
    dom = max(ktimes, key=lambda k: ktimes[k][0])
    per_kernel = {}
    for nm, (ms, n) in ktimes.items():
        b = kernel_bytes(nm, ctx, np_mean, nspec)
        avg_s = ms / 1e3 / n
        per_kernel[nm] = {"ms_total": round(ms, 3), "launches": n, "us_avg": round(avg_s * 1e6, 2),
                          "GBps": round(b / avg_s / 1e9, 1) if b > 0 else None}
    if dom == dominant and dom_n:
        dom_avg = dom_ms / 1e3 / dom_n   # measured in the timed region
    else:                                # (another kernel dominates: its breakdown figure)
        dom_avg = ktimes[dom][0] / 1e3 / ktimes[dom][1]
    dom_bytes = kernel_bytes(dom, ctx, np_mean, nspec)
    achieved = dom_bytes / dom_avg / 1e9
    lay = ctx.layout()
    counters = ctx.counters()
    ctx.close()
    # the push+deposit part of the step against its algorithmic bytes
    # (56 B per particle-step + 52 B per cell-step per species, FP32)
    pd_bytes = 56.0 * np_mean + 52.0 * lay["ncells"] * nspec
    res = {
        "value": np_mean * args.steps / wall,
        "ms_per_step": wall / args.steps * 1e3,
        "np": np_mean,
        "wall_s": wall,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(dom),
                     "algorithmic_bytes": dom_bytes},
        "step_bound": step_bound(np_mean, lay["ncells"], nspec, wall / args.steps * 1e3),
        "push_deposit": {"particle_steps_per_s": np_mean * nb / t_pd if t_pd > 0 else None,
                         "ms_per_step": t_pd / nb * 1e3,
                         "effective_GBps_vs_56B_per_particle": round(pd_bytes * nb / t_pd / 1e9, 1)},
        "kernels": per_kernel,
        "kernels_from": "%d further steps with every launch timed" % nb,
        "kernel_ms_sum_per_step": total_k / nb,
        "tiles": [lay["tx"], lay["ty"], lay["tz"]],
        "dropped_particles": counters.get("dropped", 0),
    }
    return res


def denoiser_bench(patches: int = 8, reps: int = 10):
    """The learned moment denoiser on `patches` volumes of 32^3 voxels x 7
    channels (the training patch shape): forward passes and training steps
    (forward, MSE loss, backward-data, backward-weights, Adam), timed with
    HIP events; FLOPs count the 3x3x3 convolutions only."""
    import torch
    from rhybrid_amd import denoise as D
    P, S = patches, 32
    V = P * S ** 3
    g = torch.Generator(device="cuda").manual_seed(11)
    x = torch.randn((P, S, S, S, 7), device="cuda", generator=g)
    t = x[..., :4] + 0.1 * torch.randn((P, S, S, S, 4), device="cuda", generator=g)
    d = D.Denoiser(V, scales=[1.0] * 7, weights=D.init_weights(seed=0))
    units_fwd = sum(ci * co for ci, co in D.LAYERS)                     # 2400
    units_train = units_fwd + sum(ci * co for ci, co in D.LAYERS[1:]) + units_fwd  # fwd + dgrad + wgrad
    flop_unit = 2.0 * V * 27
    for _ in range(2):
        d.forward(x)
        d.train_step(x, t, lr=1e-4)
    d.synchronize()
    d.reset_timing()
    d.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(reps):
        d.forward(x)
    d.synchronize()
    t_fwd = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        d.train_step(x, t, lr=1e-4)
    d.synchronize()
    t_train = (time.perf_counter() - t0) / reps
    d.set_timing(False)
    kern = {}
    for nm in ("pack_input", "conv1", "conv2", "conv3", "conv4", "loss", "dgrad4", "dgrad3", "dgrad2",
               "wgrad4", "wgrad3", "wgrad2", "wgrad1", "wgrad_reduce", "adam", "pack"):
        ms, n = d.kernel_time(nm)
        if n:
            kern[nm] = {"us_avg": round(ms * 1e3 / n, 2), "launches": n}
    d.close()
    conv_units = {"conv1": 7 * 32, "conv2": 1024, "conv3": 1024, "conv4": 128, "dgrad4": 128, "dgrad3": 1024,
                  "dgrad2": 1024, "wgrad4": 128, "wgrad3": 1024, "wgrad2": 1024, "wgrad1": 7 * 32}
    for nm, u in conv_units.items():
        if nm in kern:
            kern[nm]["TFLOPs"] = round(u * flop_unit / (kern[nm]["us_avg"] * 1e-6) / 1e12, 1)
    dom = max(conv_units, key=lambda nm: kern.get(nm, {}).get("us_avg", 0) * kern.get(nm, {}).get("launches", 0))
    ach = kern[dom]["TFLOPs"]
    return {"workload": "%d volumes of 32^3 voxels x 7 channels, conv3d 7-32-32-32-4 (3^3), BF16 MFMA, "
                        "FP32 accumulation" % P,
            "forward_tflops": units_fwd * flop_unit / t_fwd / 1e12,
            "train_step_tflops": units_train * flop_unit / t_train / 1e12,
            "ms_forward": t_fwd * 1e3, "ms_train_step": t_train * 1e3,
            "roofline": {"bound": "mfma", "kernel": dom, "achieved": ach, "peak": BF16_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac": round(ach / BF16_PEAK_TFLOPS, 4)},
            "kernels": kern}


def denoiser_forward_volume(shape=(256, 192, 192), reps: int = 5):
    """The denoiser's forward pass on one volume of the uneven-load box's
    size (the in-loop call), per-layer times from HIP events."""
    import torch
    from rhybrid_amd import denoise as D
    X, Y, Z = shape
    V = X * Y * Z
    g = torch.Generator(device="cuda").manual_seed(12)
    x = torch.randn((1, Z, Y, X, 7), device="cuda", generator=g)
    d = D.Denoiser(V, scales=[1.0] * 7, weights=D.init_weights(seed=0))
    d.forward(x)
    d.synchronize()
    d.reset_timing()
    d.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(reps):
        d.forward(x)
    d.synchronize()
    t = (time.perf_counter() - t0) / reps
    d.set_timing(False)
    flop_unit = 2.0 * V * 27
    kern = {}
    for nm, u in (("pack_input", 0), ("conv1", 7 * 32), ("conv2", 1024), ("conv3", 1024), ("conv4", 128)):
        ms, n = d.kernel_time(nm)
        if n:
            us = ms * 1e3 / n
            kern[nm] = {"us_avg": round(us, 1)}
            if u:
                kern[nm]["TFLOPs"] = round(u * flop_unit / (us * 1e-6) / 1e12, 1)
    d.close()
    return {"volume": list(shape), "voxels": V, "ms_forward": t * 1e3,
            "forward_tflops": D.forward_flops(V) / t / 1e12, "ns_per_voxel": t / V * 1e9, "kernels": kern}


DENOISER_DATA = os.path.join(ROOT, "rhybrid_amd", "data")


def denoiser_config5(steps: int = 10, warmup: int = 2, ppcs=(8, 64)):
    """The denoiser in the loop at the uneven-load geometry (256x192x192
    cells, exospheric O+ in proportion to the H+ ppc): time per step of the
    plain FP32 method at each of `ppcs` and of the lowest-ppc run with the
    denoiser attached (gather, forward over the 9.4 M-voxel volume, fallback
    bound, scatter, every step), each device-filled and timed over `steps`
    steps after `warmup`.  The trained weights and the 2000-step error curves
    against a 256-ppc truth run come from rhybrid_amd/denoise_experiment.py
    (its committed result, rhybrid_amd/data/denoiser_experiment.json)."""
    from rhybrid_amd import denoise as D
    from rhybrid_amd import denoise_data as DD
    from rhybrid_amd import problem as P
    from rhybrid_amd.denoise_train import filled_context
    make = lambda ppc: P.config_exosphere(ppc=ppc, o_per_ppc=True, seed=23)
    wpath = os.path.join(DENOISER_DATA, "denoiser_weights.npy")
    w = np.load(wpath) if os.path.exists(wpath) else D.init_weights(seed=0)
    out = {"geometry": "256x192x192 uneven-load box (H+ solar wind, O+ in proportion to the H+ ppc), FP32",
           "weights": os.path.relpath(wpath, ROOT) if os.path.exists(wpath) else "untrained (timing only)",
           "runs": {}}

    def timed(ppc, dn=None):
        pr = make(ppc)
        ctx = filled_context(pr)
        if dn is not None:
            ctx.set_denoiser(dn, 0.5)
        ctx.step(warmup)
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.step(steps)
        ctx.synchronize()
        ms = (time.perf_counter() - t0) / steps * 1e3
        n = [ctx.num_particles(s) for s in range(2)]
        fb = ctx.counters()["denoise_fallback"] if dn is not None else None
        if dn is not None:
            ctx.set_denoiser(None)
        ctx.close()
        return ms, n, fb, pr

    for ppc in sorted(ppcs):
        ms, n, _, pr = timed(ppc)
        out["runs"]["plain_%d" % ppc] = {"ms_per_step": round(ms, 3), "particles": n}
        print("# config-5 timing: plain %d ppc %.2f ms/step" % (ppc, ms), file=sys.stderr, flush=True)
    low = min(ppcs)
    pr = make(low)
    g = pr.grid
    dn = D.Denoiser(g.nc, scales=DD.channel_scales(pr), weights=w)
    ms, n, fb, _ = timed(low, dn)
    dn.close()
    out["runs"]["denoised_%d" % low] = {"ms_per_step": round(ms, 3), "particles": n,
                                        "fallback_fraction": fb / float((warmup + steps) * g.nc)}
    print("# config-5 timing: denoised %d ppc %.2f ms/step" % (low, ms), file=sys.stderr, flush=True)
    if 64 in ppcs:
        out["time_ratio_plain64_over_denoised8"] = out["runs"]["plain_64"]["ms_per_step"] / ms
    rpath = os.path.join(DENOISER_DATA, "denoiser_experiment.json")
    if os.path.exists(rpath):
        exp = json.load(open(rpath))
        loop = exp.get("in_loop", {})
        out["experiment"] = {"file": os.path.relpath(rpath, ROOT),
                             "command": "python -m rhybrid_amd.denoise_experiment",
                             "training": {k: exp["training"].get(k) for k in
                                          ("pairs", "patches", "epochs", "heldout_error_raw", "heldout_error_denoised")},
                             "truth": loop.get("truth"), "steps": loop.get("geometry"),
                             "denoised_vs_plain_64": loop.get("denoised_vs_plain_64"),
                             "final_errors": {k: v["errors"][-1] for k, v in loop.get("runs", {}).items()},
                             "ms_per_step": {k: v["ms_per_step"] for k, v in loop.get("runs", {}).items()},
                             "truth_fp32_vs_fp64": exp.get("truth_fp32_vs_fp64")}
    return out


def _rel_errors(x_gpu: np.ndarray, x_orc: np.ndarray, dim: int) -> dict:
    a = x_gpu.reshape(-1, dim)
    b = x_orc.reshape(-1, dim)
    nb = np.linalg.norm(b, axis=1)
    big = nb > 1e-3 * nb.max()
    d = np.linalg.norm(a - b, axis=1)
    return {"rel_l2": float(np.linalg.norm(d) / np.linalg.norm(nb)),
            "max_rel": float((d[big] / nb[big]).max())}


def host_cpu() -> dict:
    """The host's CPU as the process sees it (nproc, the CPUs it may run
    on, the model and socket count from /proc/cpuinfo)."""
    info = {"nproc": os.cpu_count()}
    try:
        info["affinity"] = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    try:
        models, phys = set(), set()
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                models.add(line.split(":", 1)[1].strip())
            elif line.startswith("physical id"):
                phys.add(line.split(":", 1)[1].strip())
        info["model"] = "; ".join(sorted(models))
        info["sockets"] = len(phys) or None
    except OSError:
        pass
    return info


def cpu_baseline(threads: int, pr, steps: int = 4):
    """The FP64 oracle (the reference algorithm in the reference's data
    layout, one slab of cells per thread) on the headline workload itself:
    `steps` full steps after one warm-up step on `threads` cores, and one
    step of the 64x48x48 box of the same workload on one core.  As the
    checker it also gives the matched-error tables of the FP32 GPU path from
    the same upload: the headline box after the oracle's 1 + `steps` steps,
    and the 32^3 periodic box (64 ppc, 2.1 M particles) after 100 steps."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from rhybrid_amd.gpu import RH_FP32, Context
    O.build()
    o = O.Oracle(pr, threads=threads)
    o.step(1)
    n = sum(o.num_particles(s) for s in range(o.nspecies))
    print("# cpu baseline: %d particles, %d threads" % (n, threads), file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    ctx = Context(pr, precision=RH_FP32, tile=BENCH_TILE)
    ctx.step(1 + steps)
    err_full = {"workload": "the headline box (128x96x96, ~71 M particles), %d full steps from the same upload, "
                            "FP32 GPU vs FP64 oracle" % (1 + steps), "tile": list(BENCH_TILE)}
    for key, nm, dim in (("n", "cellRhoQi", 1), ("J", "cellJi", 3), ("E", "nodeE", 3), ("B", "faceB", 3)):
        err_full[key] = _rel_errors(ctx.field(nm), o.field(nm).copy(), dim)
    err_full["particles_gpu"] = [ctx.num_particles(s) for s in range(len(pr.species))]
    err_full["particles_oracle"] = [o.num_particles(s) for s in range(o.nspecies)]
    ctx.close()
    del o
    small = workload(scale=2)
    o1 = O.Oracle(small, threads=1)
    n1 = sum(o1.num_particles(s) for s in range(o1.nspecies))
    t0 = time.perf_counter()
    o1.step(1)
    dt1 = time.perf_counter() - t0
    del o1, small
    cpu = {"value": n * steps / dt, "unit": "particle-steps/s", "cores": threads, "kind": "port",
           "sample": "FP64 oracle on the headline workload itself (%d particles), %d full steps incl. the field "
                     "update after one warm-up step, %d OpenMP threads (one slab of cells each); value_1core: one "
                     "step of the 64x48x48 box of the same workload on one core" % (n, steps, threads),
           "value_1core": n1 / dt1, "host": host_cpu()}
    # matched error of the FP32 GPU path on the 32^3 periodic box
    from rhybrid_amd import problem as P
    err_steps = 100
    pr = P.config_periodic(n=32, ppc=64)
    ctx = Context(pr, precision=RH_FP32, tile=BENCH_TILE)
    ctx.step(err_steps)
    g = {nm: ctx.field(nm) for nm in ("cellRhoQi", "cellJi", "nodeE", "faceB")}
    ctx.close()
    o = O.Oracle(pr, threads=threads)
    o.step(err_steps)
    err = {"workload": "32^3 periodic box, 64 ppc H+, %d full steps, FP32 GPU vs FP64 oracle; rel_l2 = "
                       "|X_gpu - X_oracle|_2 / |X_oracle|_2, max_rel over cells with |X| > 1e-3 max|X|" % err_steps,
           "tile": list(BENCH_TILE)}
    for key, nm, dim in (("n", "cellRhoQi", 1), ("J", "cellJi", 3), ("E", "nodeE", 3), ("B", "faceB", 3)):
        err[key] = _rel_errors(g[nm], o.field(nm).copy(), dim)
    err["headline_box"] = err_full
    return cpu, err


def error_curve(pr, steps: int = 500, every: int = 50, order_tile=(2, 4, 8)):
    """FP32 against FP64 on the headline box over `steps` full steps, both
    on the GPU from the same upload (the FP64 instantiation runs the same
    kernels in double and matches the FP64 oracle to 1e-9 in the tests): the
    relative errors of n, J, E and B every `every` steps.  Beside it, FP64
    against FP64 with another tile shape (`order_tile`): the same arithmetic
    in double, only the order of the deposit's and the ghost fold's sums
    differs, so the two start apart by rounding in the last bits.  Where that
    difference grows at the same rate as the FP32 one, the growth is the
    flow's amplification of any perturbation, not an inaccuracy of FP32."""
    from rhybrid_amd.gpu import RH_FP32, RH_FP64, Context
    c32 = Context(pr, precision=RH_FP32, tile=BENCH_TILE)
    c64 = Context(pr, precision=RH_FP64, tile=BENCH_TILE)
    c64o = Context(pr, precision=RH_FP64, tile=order_tile)
    names = (("n", "cellRhoQi", 1), ("J", "cellJi", 3), ("E", "nodeE", 3), ("B", "faceB", 3))
    rows = []
    for done in range(every, steps + 1, every):
        c32.step(every)
        c64.step(every)
        c64o.step(every)
        row = {"step": done}
        ref = {nm: c64.field(nm) for _, nm, _ in names}
        for key, nm, dim in names:
            row[key] = _rel_errors(c32.field(nm), ref[nm], dim)
        row["fp64_other_order"] = {key: _rel_errors(c64o.field(nm), ref[nm], dim) for key, nm, dim in names}
        row["particles_fp32"] = [c32.num_particles(s) for s in range(len(pr.species))]
        row["particles_fp64"] = [c64.num_particles(s) for s in range(len(pr.species))]
        rows.append(row)
        o = row["fp64_other_order"]
        print("# error curve step %d: n %.2e J %.2e E %.2e B %.2e | fp64 other order: n %.2e B %.2e"
              % (done, row["n"]["rel_l2"], row["J"]["rel_l2"], row["E"]["rel_l2"], row["B"]["rel_l2"],
                 o["n"]["rel_l2"], o["B"]["rel_l2"]), file=sys.stderr, flush=True)
    c32.close()
    c64.close()
    c64o.close()
    return {"workload": "the headline box, FP32 vs FP64 GPU paths from the same upload, %d full steps; "
                        "fp64_other_order: FP64 at tile %s vs FP64 at tile %s (summation order only)"
                        % (steps, "x".join(map(str, order_tile)), "x".join(map(str, BENCH_TILE))),
            "tile": list(BENCH_TILE), "rows": rows}


def exosphere_bench(steps: int = 10, warmup: int = 2, prefill: int = 600, tune_shapes: int = 4, tile=None,
                    species_streams: bool = True):
    """The uneven-load box (256x192x192 cells, dx=200 km, H+ at 24 ppc with
    +x-wall inflow, O+ photoions of an exponential corona around an absorbing
    sphere of 12 cells): tile shape from the autotuner, H+ filled on the device
    (InjectorUniform at timestep 1), O+ accumulated over `prefill` injection
    steps at their creation points, then `steps` timed full steps with the
    species' particle phases on streams of their own (the O+ launches, bound
    by a few crowded tiles, run beside the H+ ones), then `steps` more in the
    one-stream order for comparison, and the per-kernel table from further
    one-stream steps (launches that overlap have no time of their own)."""
    from rhybrid_amd import autotune
    from rhybrid_amd import problem as P
    from rhybrid_amd.gpu import RH_FP32, Context
    t0 = time.time()
    pr = P.config_exosphere()
    t_gen = time.time() - t0

    def prepare(ctx):
        ctx.set_species_streams(species_streams)  # the autotuner times the shapes as the box then runs
        ctx.fill(0, pr.meta["fill"][0])
        for t in range(1, prefill + 1, 100):  # in parts, each reported (long under a profiler)
            ctx.fill(1, pr.injectors[1], timesteps=range(t, min(t + 100, prefill + 1)))
            ctx.synchronize()
            print("# exosphere prefill: O+ injection steps %d-%d done" % (t, min(t + 99, prefill)),
                  file=sys.stderr, flush=True)

    if tile is None:
        shapes = autotune.candidates(pr.grid)[:tune_shapes]
        tuned = autotune.tune(pr, steps=6, shapes=shapes, prepare=prepare)
    else:  # a given tile shape (profiled runs): no autotuning
        tuned = {"best": tuple(tile), "ms_per_step": None}
    ctx = Context(pr, precision=RH_FP32, tile=tuned["best"])
    t0 = time.time()
    prepare(ctx)
    ctx.synchronize()
    t_fill = time.time() - t0
    ctx.step(warmup)
    ctx.synchronize()
    nO_cells = np.bincount(ctx.particles(1)[0], minlength=pr.grid.nc)
    n0 = [ctx.num_particles(s) for s in range(2)]
    ctx.set_species_streams(species_streams)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.step(steps)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    n1 = [ctx.num_particles(s) for s in range(2)]
    ctx.set_species_streams(False)
    wall_one = None
    if species_streams:
        t0 = time.perf_counter()
        ctx.step(steps)
        ctx.synchronize()
        wall_one = time.perf_counter() - t0
    # kernel breakdown per species from further steps with every launch timed
    nb = min(steps, 5)
    ctx.reset_timing()
    ctx.set_timing(True)
    ctx.step(nb)
    ctx.synchronize()
    ctx.set_timing(False)
    kt = {}
    per_species = {}
    for s, sp in enumerate(("H+", "O+")):
        npp = 0.5 * (n0[s] + n1[s])
        row = {}
        for nm in ("push", "sort_slots", "move", "deposit", "boundary", "inject_count", "inject_fill",
                   "inject_slots"):
            ms, n = ctx.kernel_time("%s:%d" % (nm, s))
            if n:
                b = kernel_bytes(nm, ctx, 0.0, 1, npp=npp)
                avg = ms / 1e3 / n
                row[nm] = {"us_avg": round(avg * 1e6, 2), "GBps": round(b / avg / 1e9, 1) if b > 0 else None,
                           "ms_per_step": round(ms / nb, 4)}
        per_species[sp] = {"particles": npp, "kernels": row}
    for nm in ("fold", "field_update", "push_fields"):
        ms, n = ctx.kernel_time(nm)
        if n:
            kt[nm] = {"us_avg": round(ms / n * 1e3, 2), "ms_per_step": round(ms / nb, 4)}
    # the dominant kernel (largest time per step over both species) against HBM
    dom_sp, dom_nm = max(((sp, nm) for sp in per_species for nm in per_species[sp]["kernels"]),
                         key=lambda k: per_species[k[0]]["kernels"][k[1]]["ms_per_step"])
    dk = per_species[dom_sp]["kernels"][dom_nm]
    dom_bytes = kernel_bytes(dom_nm, ctx, 0.0, 1, npp=per_species[dom_sp]["particles"])
    roof = {"bound": "hbm", "kernel": "%s (%s)" % (dom_nm, dom_sp), "achieved": dk["GBps"],
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(dk["GBps"] / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(dom_nm, "hbm_traffic_exosphere256x192x192.txt", dom_sp),
            "algorithmic_bytes": dom_bytes}
    ctx.close()
    np_mean = 0.5 * (sum(n0) + sum(n1))
    occ = nO_cells[nO_cells > 0]
    return {"workload": "256x192x192 cells, dx=200 km, H+ 24 ppc + inflow, O+ photoions (corona n0=1e12 m^-3 at "
                        "r0=12 cells, H=5 cells, 1e-6 1/s, 4x the solar wind's macroparticles per step), "
                        "absorbing sphere R=12 cells, FP32",
            "tile": list(tuned["best"]), "autotune_ms_per_step": tuned["ms_per_step"],
            "ms_per_step": wall / steps * 1e3, "particle_steps_per_s": np_mean * steps / wall,
            "species_streams": species_streams,
            "ms_per_step_one_stream": round(wall_one / steps * 1e3, 4) if wall_one else None,
            "particles": {"H+": n1[0], "O+": n1[1]},
            "O+_ppc": {"max": int(occ.max()) if occ.size else 0, "mean_occupied": float(occ.mean()) if occ.size else 0.0,
                       "cells_occupied": int(occ.size), "prefill_steps": prefill},
            "O+_production_per_s": pr.meta["O_total_rate"],
            "roofline": roof, "per_species": per_species, "grid_kernels": kt,
            "kernels_from": "%d further steps with every launch timed" % nb,
            "seconds_setup": round(t_gen, 1), "seconds_fill": round(t_fill, 2)}


def calibration_copy(gib: float) -> None:
    """A plain device-to-device copy of `gib` GiB (one copy dispatch of known
    size): hardware traffic counters (FETCH_SIZE, WRITE_SIZE) of the other
    dispatches are scaled by what the counters read for it."""
    import torch
    x = torch.ones(int(gib * (1 << 28)), dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    del x, y


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-denoiser", action="store_true")
    ap.add_argument("--no-in-loop", action="store_true", help="skip the denoiser's in-loop comparison runs")
    ap.add_argument("--only-denoiser", action="store_true", help="the denoiser section alone (its kernels' profile)")
    ap.add_argument("--no-exosphere", action="store_true", help="skip the uneven-load exosphere box")
    ap.add_argument("--no-error-curve", action="store_true", help="skip the 500-step FP32-vs-FP64 error curve")
    ap.add_argument("--only-exosphere", action="store_true", help="the uneven-load box alone (its kernels' profile)")
    ap.add_argument("--exosphere-tile", type=int, nargs=3, default=None, help="skip the autotuner, use this tile")
    ap.add_argument("--exosphere-one-stream", action="store_true",
                    help="time the uneven-load box in the one-stream order only (profiled runs: every launch "
                         "then has a time of its own and the species' launches a fixed order)")
    ap.add_argument("--calibration-copy-gib", type=float, default=0.0,
                    help="first copy this many GiB device to device (a traffic counter calibration dispatch)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (one domain per GPU; "
              "the halo and migration exchange between GPUs is not built).", file=sys.stderr)
        sys.exit(2)
    if args.only_exosphere:
        if args.calibration_copy_gib > 0:
            calibration_copy(args.calibration_copy_gib)
        print(json.dumps({"exosphere_box": exosphere_bench(steps=args.steps, warmup=args.warmup,
                                                           tile=args.exosphere_tile,
                                                           species_streams=not args.exosphere_one_stream)}),
              flush=True)
        return
    if args.only_denoiser:
        print(json.dumps({"denoiser": denoiser_bench(), "forward_volume": denoiser_forward_volume()}), flush=True)
        return
    if args.calibration_copy_gib > 0:
        calibration_copy(args.calibration_copy_gib)
    t0 = time.time()
    pr = workload()
    print("# workload generated in %.1fs" % (time.time() - t0), file=sys.stderr, flush=True)
    res = gpu_bench(args, pr)
    dn = None if args.no_denoiser else denoiser_bench()
    if dn is not None:
        dn["forward_volume"] = denoiser_forward_volume()
    if dn is not None and not args.no_in_loop:
        from rhybrid_amd import denoise_loop
        dn["in_loop"] = denoise_loop.compare()
        dn["in_loop_config5"] = denoiser_config5()
    exo = None
    if not args.no_exosphere:
        try:
            exo = exosphere_bench(tile=args.exosphere_tile, species_streams=not args.exosphere_one_stream)
        except Exception as e:  # reported in the line, the headline number stands
            exo = {"error": repr(e)}
    cpu, err = None, None
    if not args.no_cpu_baseline:
        # the process's CPU share (OMP_NUM_THREADS where the host sets it)
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
        cpu, err = cpu_baseline(threads, pr)
    curve = None if args.no_error_curve else error_curve(pr)
    del pr
    line = {
        "metric": "particle-steps/s (push+deposit) & step ms at matched n/J error vs FP64 oracle",
        "value": res["value"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "128x96x96-cell inflow/outflow box, dx=300 km, absorbing sphere R=10 cells, "
                               "H+ and He++ (4%) at 32 ppc each, 450 km/s, T=1e5 K, IMF (-1.5,2.5,0) nT; "
                               "full step (push, cell sort, CIC deposit, FP64 ghost fold, particle BCs, "
                               "solar-wind injection, field update); uniform preload plus +x-wall inflow",
                   "particles": res["np"], "cells": 128 * 96 * 96, "tile": res["tiles"],
                   "parallelism": "single GPU"},
        "roofline": res["roofline"],
        "cpu_baseline": cpu,
        "error_vs_oracle": err,
        "error_curve_fp32_vs_fp64": curve,
        "denoiser": dn,
        "exosphere_box": exo,
        "step_bound": res["step_bound"],
        "push_deposit": res["push_deposit"],
        "kernels": res["kernels"],
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
