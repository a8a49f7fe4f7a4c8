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
  step_roofline - the step's algorithmic bytes (104 B per voxel-z-step) at the
                  measured step rate, as a fraction of 8 TB/s;
  cpu_baseline  - the FP32 NumPy/SciPy restatement of the reference march timed
                  on this host's cores over a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HBM_PEAK_GBS = 8000.0
METRIC = "voxel-z-steps/s at 256x256x1024 FP32; learned vs plain KT at rel-L2<=1e-3"


def algorithmic_bytes_per_launch(kind: str, vox: int, ncols: int, nT: int) -> float:
    """HBM bytes a kernel must move: one read + one write of the field state
    (4 B + 4 B per voxel: the complex64 half spectrum has the real field's size)
    for every sweep; the final column kernel only reads it and writes the
    per-column reductions (4 x 4 B) and the FP64 intensity carry (8 B r+w)."""
    if kind == "column_final":
        return 4.0 * vox + ncols * (16 + 16) + 4 * nT
    return 8.0 * vox


def cpu_baseline(steps: int):
    import numpy as np
    from oracle import asm_oracle as O
    from angularspectrum_heterogeneous_amd import workloads as W
    cores = min(16, len(os.sched_getaffinity(0)))
    O.set_workers(cores)
    field, kw = W.config2()
    p = O.OracleParams(**kw)
    tim = {}
    O.angular_spectrum_solve(field, p, dtype=np.float32, max_steps_run=steps, timings=tim)
    t = sum(tim["step_s"])
    vox = field.size
    return {"value": vox * steps / t, "unit": "voxel-z-steps/s", "cores": cores, "kind": "port",
            "sample": f"{steps} z-steps of the same 256x256x1024 march by the FP32 NumPy/SciPy "
                      f"restatement of the reference (scipy.fft, {cores} workers); operator "
                      f"setup ({tim['setup_s']:.1f} s) excluded",
            "step_s": [round(s, 3) for s in tim["step_s"]]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--profile-iters", type=int, default=5)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calib", action="store_true",
                    help="also run 3 elementwise passes over 512 MiB (known HBM bytes) to scale PMC byte counters")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(--gpus must be 1)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from angularspectrum_heterogeneous_amd import _lib, workloads as W
    from angularspectrum_heterogeneous_amd.plan import MarchPlan
    from angularspectrum_heterogeneous_amd.solver import SolverParams

    field_np, kw = W.config2()
    params = SolverParams(**kw)
    nX, nY, nT = field_np.shape
    vox = nX * nY * nT
    dz = params.dZmin
    plan = MarchPlan(field_np.shape, params)
    plan.set_dz(dz)
    K, Wm, P = args.steps, args.warmup, args.profile_iters
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
    t1 = time.perf_counter()
    elapsed = t1 - t0
    value = vox * K / elapsed
    ms_step = elapsed / K * 1e3

    # per-kernel launch times with HIP events on the launch stream
    import ctypes
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
    dom = max(kernels, key=lambda n: kernels[n]["avg_ms"] * kernels[n]["launches_per_step"])
    dom_s = kernels[dom]["avg_ms"] * 1e-3
    ab = algorithmic_bytes_per_launch(dom, vox, nX * nY, nT)
    achieved = ab / dom_s / 1e9
    for name, kv in kernels.items():
        b = algorithmic_bytes_per_launch(name, vox, nX * nY, nT)
        kv["alg_GBps"] = round(b / (kv["avg_ms"] * 1e-3) / 1e9, 1)
    # measured HBM bytes per launch (rocprofv3 FETCH_SIZE/WRITE_SIZE, calibrated
    # on a kernel of known traffic), recorded next to the profile summaries
    traffic = None
    tp = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "traffic_256x256x1024.json")
    if os.path.exists(tp):
        t = json.load(open(tp))["kernels"].get(dom)
        if t:
            traffic = t["traffic_bytes"]
    step_bytes = 104.0 * vox  # 13 sweeps x 8 B at this configuration
    step_frac = step_bytes / (elapsed / K) / 1e9 / HBM_PEAK_GBS

    fin = float(stats[: row + 1].view(torch.float32).max().item())
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
                   "max_abs_p_seen": fin},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "alg_bytes_per_launch": ab},
        "step_roofline": {"alg_bytes_per_voxel_step": 104, "frac": round(step_frac, 4)},
        "kernels": kernels,
    }
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(args.cpu_steps)
    plan.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
