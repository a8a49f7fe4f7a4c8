#!/usr/bin/env python
"""Throughput of the gfx950 angular-spectrum march (one JSON line on stdout).

Workload: the 256x256x1024 FP32 focused nonlinear beam (focused piston r=30 mm,
F=50 mm, 1 MHz 5-cycle burst at 1 MPa, split-step KT, power-law attenuation
with KK dispersion, Wendland + frequency-weighted + super-absorbing boundaries,
dZ = dX).  A step is one z-step of the march over the whole field; the field
stays resident in HBM.  value = nX*nY*nT*steps / wall time of the timed steps.

Also reported:
  roofline      - the dominant kernel's algorithmic HBM bytes per launch divided
                  by its average launch time (HIP events on the launch stream);
  step_roofline - the step's algorithmic bytes (the sum over the kernels launched
                  per step) at the measured step rate, as a fraction of 8 TB/s,
                  and the same with the measured HBM bytes per step (rocprofv3
                  FETCH_SIZE + WRITE_SIZE, calibrated) when the profile is present;
  cpu_baseline  - the FP32 NumPy/SciPy restatement of the reference march timed
                  on this host's cores: the 64x64x256 march in full, and 5 steps of
                  this workload after one warm-up step;
  batch         - 32 seeded transmit configurations of uneven size (nX=nY in
                  {128,192,256,384}, nT in {512,1024,2048}) solved as one batch
                  with autotuned launches, against the same solves one by one;
  learned       - the same march with the learned nonlinear step (bf16 conv net
                  on the matrix cores, KT where a column's max|dp/dt| exceeds a
                  threshold): for a sweep of thresholds, rel-L2 of the full run
                  against plain KT and the share of columns left to KT; at the
                  largest threshold within 1e-3 on the field, its step time, the
                  ratio to the plain step, and rel-L2 of the full run against
                  plain KT for field, pnp, ppp, pI and pIloss.  The network is
                  trained first, on the GPU, on 200k KT (before, after) pairs.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HBM_PEAK_GBS = 8000.0
METRIC = "voxel-z-steps/s at 256x256x1024 FP32; learned vs plain KT at rel-L2<=1e-3"


def algorithmic_bytes_per_launch(kind: str, vox: int, ncols: int, nT: int, nX: int = 0, band_rows: int = 0) -> float:
    """HBM bytes a kernel must move: one read + one write of the field state
    (4 B + 4 B per voxel: the complex64 half spectrum has the real field's size)
    for every sweep; the final column kernel only reads it and writes the
    per-column reductions (4 x 4 B) and the FP64 intensity carry (8 B r+w); the
    super-absorbing y pass reads and writes only its two boundary bands
    (band_rows of the nY rows of every x row)."""
    if kind == "column_final":
        return 4.0 * vox + ncols * (16 + 16) + 4 * nT
    if kind == "ysa" and band_rows:
        return 8.0 * vox * band_rows / (ncols // nX)
    return 8.0 * vox


def cpu_baseline(steps: int):
    """The FP32 NumPy/SciPy restatement of the reference march on this host: the
    64x64x256 march in full, then `steps` steady-state steps of the 256x256x1024
    march after one warm-up step (operator setup timed separately)."""
    import platform
    import numpy as np
    from oracle import asm_oracle as O
    from angularspectrum_heterogeneous_amd import workloads as W
    cores = min(16, len(os.sched_getaffinity(0)))
    O.set_workers(cores)
    f1, kw1 = W.config1()
    tim1 = {}
    t0 = time.perf_counter()
    res1 = O.angular_spectrum_solve(f1, O.OracleParams(**kw1), dtype=np.float32, timings=tim1)
    wall1 = time.perf_counter() - t0
    n1 = len(res1[5])
    field, kw = W.config2()
    tim = {}
    O.angular_spectrum_solve(field, O.OracleParams(**kw), dtype=np.float32, max_steps_run=steps + 1, timings=tim)
    t = sum(tim["step_s"][1:])
    vox = field.size
    cpu = platform.processor() or platform.machine()
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": vox * steps / t, "unit": "voxel-z-steps/s", "cores": cores, "kind": "port",
            "sample": f"{steps} z-steps of the same 256x256x1024 march after one warm-up step, by the FP32 "
                      f"NumPy/SciPy restatement of the reference (scipy.fft, {cores} workers); operator "
                      f"setup ({tim['setup_s']:.1f} s) excluded",
            "cpu": cpu,
            "step_s": [round(x, 3) for x in tim["step_s"]],
            "config1_full": {"steps": n1, "wall_s": round(wall1, 3), "setup_s": round(tim1["setup_s"], 3),
                             "value": f1.size * n1 / sum(tim1["step_s"]), "unit": "voxel-z-steps/s"}}


def timed_march(field_np, params, dz, K, Wm, P, net=None, dpdt_max=0.0, tune=None, launch=None):
    """W warm-up steps, K timed steps (synchronised on both sides), then P steps with
    HIP events around every launch.  Returns (seconds, per-kernel dict, max |p| seen)."""
    import ctypes
    import torch
    from angularspectrum_heterogeneous_amd import _lib
    from angularspectrum_heterogeneous_amd.plan import MarchPlan
    nX, nY, nT = field_np.shape
    plan = MarchPlan(field_np.shape, params)
    plan.set_dz(dz)
    if launch is not None:
        plan.set_launch(*launch)
    elif tune is not None:
        # the launch configuration (radix orders, column grid) chosen by the autotuner
        tune.update(plan.autotune())
    if net is not None:
        plan.set_learned(net, dpdt_max)
    rows = K + Wm + P + 1
    dev = torch.device("cuda")
    field = torch.from_numpy(field_np).to(dev)
    outs = torch.zeros((4, rows, nX, nY), dtype=torch.float32, device=dev)
    pax = torch.zeros((rows, nT), dtype=torch.float32, device=dev)
    stats = torch.zeros(rows + 1, dtype=torch.int32, device=dev)
    io = _lib.AshStepIO()
    io.max_rows = rows
    io.pnp_dev, io.ppp_dev, io.pI_dev, io.pIloss_dev = (outs[k].data_ptr() for k in range(4))
    io.pax_dev, io.stats_dev = pax.data_ptr(), stats.data_ptr()
    io.dz = dz
    plan.load(field, stats)
    row = 0
    for _ in range(Wm):
        io.row = row
        plan.step(io)
        row += 1
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(K):
        io.row = row
        plan.step(io)
        row += 1
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    # per-kernel launch times with HIP events on the launch stream
    msv = (ctypes.c_float * len(_lib.KERNEL_KINDS))()
    cnt = (ctypes.c_int32 * len(_lib.KERNEL_KINDS))()
    io.row = row
    _lib.check(plan.lib.ash_profile_step(plan.h, ctypes.byref(io), P, msv, cnt, _lib.current_stream()),
               "ash_profile_step")
    row += P
    kernels = {}
    for k, name in enumerate(_lib.KERNEL_KINDS):
        if cnt[k]:
            kernels[name] = {"avg_ms": round(float(msv[k]), 5), "launches_per_step": int(cnt[k])}
    fin = float(stats[: row + 1].view(torch.float32).max().item())
    plan.close()
    return elapsed, kernels, fin


def learned_leg(field_np, kw, dz, K, Wm, P, plain_ms, n_pairs, epochs):
    """The learned nonlinear step on the same workload: KT training pairs and training on
    the GPU, the timed march with the network, and rel-L2 of the full run against plain KT."""
    import numpy as np
    import torch
    from angularspectrum_heterogeneous_amd import learned as L
    from angularspectrum_heterogeneous_amd.solver import SolverParams, angular_spectrum_solve
    nT = field_np.shape[2]
    sp = SolverParams(**kw)
    c = dz * L.burgers_coefficient(sp) / sp.dT
    t0 = time.perf_counter()
    before, after = L.generate_pairs(sp, nT, dz, n_pairs=n_pairs, stride=8, pairs_per_column=50, seed=0)
    torch.cuda.synchronize()
    t_gen = time.perf_counter() - t0
    net = L.LearnedStep(nT, max_cols=256, seed=0)
    t0 = time.perf_counter()
    hist = L.train(net, before, after, c, epochs=epochs, batch=256, lr=2e-3)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    progress(f"learned leg: trained in {t_train:.1f} s; plain reference march and threshold sweep")
    del before, after
    torch.cuda.empty_cache()
    ref = angular_spectrum_solve(field_np, SolverParams(**kw), verbose=False)
    names = ("field", "pnp", "ppp", "pI", "pIloss")

    def rel(a, b):
        a = np.asarray(a, np.float64)
        b = np.asarray(b, np.float64)
        return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))

    sweep = []
    for thr in (1e30, 1.2e13, 6e12, 3e12, 1.5e12):
        progress(f"learned leg: march at max|dp/dt| threshold {thr:.3g} Pa/s")
        shares = []

        def cb(cc, z, f, thr=thr):
            u = f.reshape(-1, nT)
            A = u.abs().amax(1)
            live = A > 1e-3 * A.max()
            D = (u[:, 1:] - u[:, :-1]).abs().amax(1) / sp.dT
            shares.append(float((D[live] > thr).float().mean()))

        got = angular_spectrum_solve(field_np, SolverParams(**kw, ash_learned=net, ash_dpdt_max=thr),
                                     verbose=False, per_step_callback=cb)
        sweep.append({"dpdt_max": thr, "rel_l2_vs_kt": {n: rel(got[i], ref[i]) for i, n in enumerate(names)},
                      "kt_share_of_beam_columns": float(np.mean(shares))})
    ok = [e for e in sweep if e["rel_l2_vs_kt"]["field"] <= 1e-3]
    pick = ok[0] if ok else sweep[-1]
    elapsed, kernels, _ = timed_march(field_np, SolverParams(**kw), dz, K, Wm, P, net=net,
                                      dpdt_max=pick["dpdt_max"])
    ms = elapsed / K * 1e3
    return {"ms_per_step": ms, "value": field_np.size * K / elapsed, "speedup_vs_plain": plain_ms / ms,
            "dpdt_max": pick["dpdt_max"], "rel_l2_vs_kt": pick["rel_l2_vs_kt"],
            "meets_1e-3": bool(ok), "steps_compared": int(len(ref[5])), "sweep": sweep,
            "kernels": kernels, "train": {"pairs": n_pairs, "epochs": epochs, "batch": 256,
                                          "loss_first_epoch": hist[0], "loss_last_epoch": hist[-1],
                                          "pairs_s": round(t_gen, 2), "train_s": round(t_train, 2)},
            "net": "4 x conv1d(k=9, 32 ch) over t, bf16 weights/activations on v_mfma_f32_16x16x32_bf16, "
                   "fp32 accumulate; trained on the GPU by the library's own kernels"}


def batch_leg(n, seed, prop, sizes=(128, 192, 256, 384), lengths=(512, 1024, 2048), branches=(1, 4)):
    """Seeded transmit configurations of uneven size solved as one batch (autotuned
    launches), with the instances in 1 and in 4 concurrent graph branches, against the
    same instances solved one by one."""
    import numpy as np
    import torch
    from angularspectrum_heterogeneous_amd import workloads as W
    from angularspectrum_heterogeneous_amd.batch import BatchSolve
    from angularspectrum_heterogeneous_amd.solver import SolverParams, angular_spectrum_solve
    spec = W.config4(n=n, seed=seed, propDist=prop, sizes=sizes, lengths=lengths)
    inst = [(f, SolverParams(**kw)) for f, kw, _ in spec]
    runs = {}
    res = None
    tuned = None
    for nb in branches:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b = BatchSolve(inst, autotune=True, branches=nb)
        t_setup = time.perf_counter() - t0
        if tuned is None:
            tuned = [dict(it.plan.tuned) for it in b.items]
        t0 = time.perf_counter()
        r = b.run()
        torch.cuda.synchronize()
        runs[nb] = {"batch_s": round(time.perf_counter() - t0, 3), "setup_and_autotune_s": round(t_setup, 3)}
        b.close()
        if res is None:
            res = r
        else:
            del r
    work = sum(int(np.prod(f.shape)) * len(r_[5]) for (f, _), r_ in zip(inst, res))
    errs = []
    t_single = 0.0
    for (f, p), r in zip(inst, res):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ref = angular_spectrum_solve(f, p, verbose=False)
        torch.cuda.synchronize()
        t_single += time.perf_counter() - t0
        e = {}
        for name, a, c in zip(("field", "pnp", "ppp", "pI", "pIloss"), r, ref):
            a = np.asarray(a, np.float64)
            c = np.asarray(c, np.float64)
            e[name] = float(np.linalg.norm(a - c) / max(np.linalg.norm(c), 1e-300))
        errs.append(e)
        del ref
    worst = {k: max(e[k] for e in errs) for k in errs[0]}
    shapes = {}
    for f, _, m in spec:
        key = f"{m['nX']}x{m['nX']}x{m['nT']}"
        shapes[key] = shapes.get(key, 0) + 1
    for nb, rr in runs.items():
        rr["value"] = work / rr["batch_s"]
    best = min(runs, key=lambda k: runs[k]["batch_s"])
    return {"instances": n, "seed": seed, "propDist_m": prop, "shapes": shapes,
            "voxel_z_steps": work, "batch_s": runs[best]["batch_s"], "branches": best,
            "value": runs[best]["value"], "unit": "voxel-z-steps/s", "by_branches": runs,
            "one_by_one_s": round(t_single, 3), "one_by_one_value": work / t_single,
            "speedup_vs_one_by_one": t_single / runs[best]["batch_s"],
            "max_rel_l2_vs_single": worst,
            "autotune_choices": {k: sorted({str((t["x_radix_order"], t["y_radix_order"], t["column_grid"]))
                                            for t, (_, _, m) in zip(tuned, spec)
                                            if f"{m['nX']}x{m['nX']}x{m['nT']}" == k}) for k in shapes},
            "note": "wall time of BatchSolve.run() (marching, stability checks and the transfer of every "
                    "instance's outputs to the host), one graph launch per z-step for all active instances"}


def progress(msg: str) -> None:
    """One progress line on stderr (stdout carries only the result line)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--profile-iters", type=int, default=5)
    ap.add_argument("--cpu-steps", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned nonlinear step leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-solve leg")
    ap.add_argument("--no-autotune", action="store_true",
                    help="march with the default launch configuration instead of the autotuned one")
    ap.add_argument("--batch-instances", type=int, default=32)
    ap.add_argument("--learned-pairs", type=int, default=200_000)
    ap.add_argument("--learned-epochs", type=int, default=3)
    ap.add_argument("--launch", type=str, default=None,
                    help="pin the launch configuration 'x_radix,y_radix,column_grid' instead of autotuning "
                         "(a profile then holds only the march's own dispatches)")
    ap.add_argument("--copy-calib", action="store_true",
                    help="also run 3 elementwise passes over 512 MiB (known HBM bytes) to scale PMC byte counters")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(--gpus must be 1)", file=sys.stderr)
        sys.exit(2)

    progress("importing torch and the library")
    import numpy as np
    import torch
    from angularspectrum_heterogeneous_amd import _lib, workloads as W
    from angularspectrum_heterogeneous_amd.solver import SolverParams

    field_np, kw = W.config2()
    params = SolverParams(**kw)
    nX, nY, nT = field_np.shape
    vox = nX * nY * nT
    dz = params.dZmin
    K, Wm, P = args.steps, args.warmup, args.profile_iters
    launch = tuple(int(v) for v in args.launch.split(",")) if args.launch else None
    tune = None if (args.no_autotune or launch) else {}
    progress(f"timed march: {nX}x{nY}x{nT}, {Wm} warm-up + {K} timed steps")
    elapsed, kernels, fin = timed_march(field_np, params, dz, K, Wm, P, tune=tune, launch=launch)
    progress(f"timed march done: {elapsed / K * 1e3:.3f} ms per step")
    if launch:
        tune = {"x_radix_order": launch[0], "y_radix_order": launch[1], "column_grid": launch[2]}
    value = vox * K / elapsed
    ms_step = elapsed / K * 1e3
    # rows of every x row the super-absorbing y pass touches: the two boundary bands
    # (rows 0..b and nY-1-b..nY-1, b = round(nY boundaryFactor))
    band = max(round(nY * params.boundaryFactor), 1)
    band_rows = min(2 * (band + 1), nY)
    dom = max(kernels, key=lambda n: kernels[n]["avg_ms"] * kernels[n]["launches_per_step"])
    dom_s = kernels[dom]["avg_ms"] * 1e-3
    ab = algorithmic_bytes_per_launch(dom, vox, nX * nY, nT, nX, band_rows)
    achieved = ab / dom_s / 1e9
    step_bytes = 0.0
    for name, kv in kernels.items():
        b = algorithmic_bytes_per_launch(name, vox, nX * nY, nT, nX, band_rows)
        kv["alg_bytes_per_launch"] = b
        kv["alg_GBps"] = round(b / (kv["avg_ms"] * 1e-3) / 1e9, 1)
        step_bytes += b * kv["launches_per_step"]
    # measured HBM bytes per launch (rocprofv3 FETCH_SIZE/WRITE_SIZE, calibrated
    # on a kernel of known traffic), recorded next to the profile summaries
    traffic = None
    measured_step_bytes = None
    tp = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "traffic_256x256x1024.json")
    if os.path.exists(tp):
        tk = json.load(open(tp))["kernels"]
        t = tk.get(dom)
        if t:
            traffic = t["traffic_bytes"]
        if all(n in tk for n in kernels):
            measured_step_bytes = sum(tk[n]["traffic_bytes"] * kv["launches_per_step"] for n, kv in kernels.items())
    step_frac = step_bytes / (elapsed / K) / 1e9 / HBM_PEAK_GBS

    dev = torch.device("cuda")
    # practical streaming ceiling beside the 8 TB/s peak: an elementwise pass that
    # reads 512 MiB and writes 512 MiB (torch's add), timed with events
    a = torch.ones(128 * 1024 * 1024, dtype=torch.float32, device=dev)
    bb = torch.empty_like(a)
    torch.add(a, 1.0, out=bb)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(10):
        torch.add(a, 1.0, out=bb)
    ev1.record()
    ev1.synchronize()
    stream_gbs = 2 * a.numel() * 4 * 10 / (ev0.elapsed_time(ev1) * 1e-3) / 1e9
    del a, bb
    if args.copy_calib:
        # an elementwise kernel of known traffic: reads 512 MiB, writes 512 MiB
        a = torch.ones(128 * 1024 * 1024, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        for _ in range(3):
            torch.add(a, 1.0, out=b)
        torch.cuda.synchronize()
        del a, b
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "voxel-z-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": Wm,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "focused nonlinear beam 256x256x1024 (focused piston r=30 mm F=50 mm, "
                               "1 MHz 5 cycles 1 MPa, split-step KT, power-law atten y=1.1, Wendland+"
                               "freq-weighted+super-absorbing boundaries, dZ=dX)",
                   "nX": nX, "nY": nY, "nT": nT, "dZ": dz, "z_steps_per_s": K / elapsed,
                   "max_abs_p_seen": fin, "launch": tune or "default"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "alg_bytes_per_launch": ab},
        "step_roofline": {"alg_bytes_per_voxel_step": round(step_bytes / vox, 2), "frac": round(step_frac, 4),
                          "measured_bytes_per_voxel_step": (round(measured_step_bytes / vox, 2)
                                                            if measured_step_bytes else None),
                          "measured_frac": (round(measured_step_bytes / (elapsed / K) / 1e9 / HBM_PEAK_GBS, 4)
                                            if measured_step_bytes else None)},
        "stream_reference": {"kernel": "elementwise add, 512 MiB in + 512 MiB out", "GBps": round(stream_gbs, 1),
                             "frac_of_peak": round(stream_gbs / HBM_PEAK_GBS, 4)},
        "kernels": kernels,
    }
    if not args.no_batch:
        progress(f"batch leg: {args.batch_instances} instances")
        result["batch"] = batch_leg(args.batch_instances, 0, 0.02)
        progress("batch leg: 16 small instances")
        # the small grids, where one instance leaves most of the GPU idle
        result["batch_small"] = batch_leg(16, 1, 0.02, sizes=(128, 192), lengths=(512, 1024))
    if not args.no_learned:
        progress("learned leg: training data, training, threshold sweep, timed march")
        result["learned"] = learned_leg(field_np, kw, dz, K, Wm, P, ms_step, args.learned_pairs,
                                        args.learned_epochs)
    if not args.no_cpu_baseline:
        progress("CPU baseline")
        result["cpu_baseline"] = cpu_baseline(args.cpu_steps)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
