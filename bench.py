"""Benchmark of the D3Q19 lid-driven cavity step on one MI355X.

Workload: lid-driven cavity, D3Q19, 256^3 cells, Re = 1000, U0 = 0.1,
tau = 3 U0 N / Re + 1/2 = 0.5768, MRT, reference initial state.  A step is
the six-face + twelve-edge halo pack and unpack looped back onto the same
GPU followed by one fused pull-scheme stream + bounce-back + macro + collide
pass over the lattice (A-B buffers).  `value` is the plain FP64 path's MLUPS
for that step with the state resident in HBM; the FP32- and FP16-stored
variants are timed the same way and their relative L2 velocity error and
mass drift against the FP64 run (same initial state, same step count) are
reported beside them.  `roofline` is for the step kernel alone, timed over the
same number of steps without the exchange.

Prints ONE JSON line.  --gpus above 1 is refused: this build runs on one GPU.
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

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--n", type=int, default=256, help="cells per edge")
    ap.add_argument("--collision", default="mrt", choices=["mrt", "bgk"])
    ap.add_argument("--accuracy-steps", type=int, default=5000,
                    help="total steps after which the reduced-storage fields are compared with FP64")
    ap.add_argument("--variants", default="f32,f16", help="comma list of reduced storages to run ('' = none)")
    ap.add_argument("--cpu-baseline", type=int, default=1, help="time the CPU oracle on a bounded sample")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--halo-loopback", type=int, default=1,
                    help="1: pack/unpack the six faces and twelve edges every step, looped back")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run device copies of this many MiB (counter calibration) and exit")
    return ap.parse_args()


def timed_steps(cav, stream, args):
    """W untimed steps, then K steps between HIP events on the library's stream."""
    from mglc_amd.hip import Event

    cav.step(args.warmup)
    stream.synchronize()
    e0, e1 = Event(), Event()
    t0 = time.perf_counter()
    e0.record(stream)
    cav.step(args.steps)
    e1.record(stream)
    e1.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_ms(e1)
    e0.close()
    e1.close()
    return ms, wall


def run_storage(args, storage, stream):
    from mglc_amd.lbm import Cavity, CavityConfig

    n = args.n
    base = dict(Reynolds=1000.0, U0=0.1, collision=args.collision, storage=storage)
    with Cavity(CavityConfig(n, n, n, halo_loopback=args.halo_loopback, **base), stream=stream.handle.value) as cav:
        cav.initial()
        ms, wall = timed_steps(cav, stream, args)
        info = cav.info()
        done = args.warmup + args.steps
        if args.accuracy_steps > done:
            cav.step(args.accuracy_steps - done)
        fields = cav.macro()
        steps_total = max(done, args.accuracy_steps)
    if args.halo_loopback:
        # the step kernel alone, for its roofline
        with Cavity(CavityConfig(n, n, n, **base), stream=stream.handle.value) as cav:
            cav.initial()
            ms_kernel, _ = timed_steps(cav, stream, args)
    else:
        ms_kernel = ms
    cells = n ** 3
    return dict(storage=storage, ms=ms, ms_kernel=ms_kernel, wall_s=wall,
                mlups=cells * args.steps / (ms * 1e-3) / 1e6, bytes_per_lup=info.bytes_per_lup, fields=fields,
                steps_total=steps_total)


def cpu_baseline(args):
    """The oracle (CPU restatement, OpenMP) on a bounded sample of the same
    workload: 256^3 cells from the reference initial state, a few steps."""
    from oracle import oracle as O
    n = args.n
    threads = max(1, min(args.cpu_threads, os.cpu_count() or 1))
    tau = 0.1 * n / 1000.0 * 3.0 + 0.5
    coll = O.MRT if args.collision == "mrt" else O.BGK
    p = O.cavity_params(n, n, n, tau, 0.1, collision=coll)
    steps = 1
    O.cavity_run(p, steps, nthreads=threads)
    per_step = max(O.last_loop_seconds(), 1e-3)
    steps = int(max(1, min(200, 15.0 / per_step)))
    O.cavity_run(p, steps, nthreads=threads)
    sec = O.last_loop_seconds()
    return {"value": n ** 3 * steps / sec / 1e6, "unit": "MLUPS", "cores": threads, "kind": "port",
            "sample": f"{n}^3 {args.collision.upper()} FP64 cavity, {steps} steps from the reference "
                      f"initial state, oracle/lbm_oracle.c with {threads} OpenMP threads, loop time only"}


def traffic_from_profiles(workload_key):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as fh:
            data = json.load(fh)
        return data.get(workload_key)
    except (OSError, ValueError):
        return None


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} refused: this build runs the cavity on one GPU "
              "(the multi-GPU halo path is design only)", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        from mglc_amd.hip import device_copy_calibration
        device_copy_calibration(args.copy_calibration << 20, 5)
        print(json.dumps({"copy_calibration_bytes": args.copy_calibration << 20, "reps": 5}))
        return
    from mglc_amd.hip import Stream

    stream = Stream()
    primary = run_storage(args, "f64", stream)
    variants = {}
    for st in [s for s in args.variants.split(",") if s]:
        r = run_storage(args, st, stream)
        vel64 = np.stack(primary["fields"][1:])
        vel = np.stack(r["fields"][1:])
        l2 = float(np.linalg.norm((vel - vel64).ravel()) / np.linalg.norm(vel64.ravel()))
        m64 = float(primary["fields"][0].sum())
        drift = abs(float(r["fields"][0].sum()) - m64) / m64
        variants[st] = {"mlups": round(r["mlups"], 1), "ms_per_step": round(r["ms"] / args.steps, 4),
                        "bytes_per_lup": r["bytes_per_lup"],
                        "kernel_hbm_frac": round(r["bytes_per_lup"] * args.n ** 3 / (r["ms_kernel"] / args.steps * 1e-3)
                                                 / 1e9 / HBM_PEAK_GBS, 4),
                        "speedup_vs_f64": round(r["mlups"] / primary["mlups"], 3),
                        "rel_l2_velocity_vs_f64": l2, "mass_drift_vs_f64": drift,
                        "compared_at_step": r["steps_total"]}
    stream.close()

    ms_step = primary["ms"] / args.steps
    bpl = primary["bytes_per_lup"]
    achieved = bpl * args.n ** 3 / (primary["ms_kernel"] / args.steps * 1e-3) / 1e9
    m0 = float(primary["fields"][0].size)
    workload = f"lid-driven cavity D3Q19 {args.n}^3 Re=1000 U0=0.1 {args.collision.upper()}"
    line = {
        "metric": "D3Q19 MLUPS + %HBM peak; FP16/FP32-store vs FP64 at stated L2 error",
        "value": round(primary["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (reference initial state: rho=1, u=0, lid layer u=U0)",
        "config": {"workload": workload, "cells": args.n ** 3, "collision": args.collision,
                   "storage": "f64", "pattern": "A-B pull", "kernel": "k_cavity_step",
                   "halo": "6 faces x 5 pops + 12 edges, pack/unpack looped back every step"
                   if args.halo_loopback else "none",
                   "kernel_ms_per_step": round(primary["ms_kernel"] / args.steps, 4)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic_from_profiles(f"{args.n}^3 {args.collision} f64")},
        "mass_drift_f64": abs(float(primary["fields"][0].sum()) - m0) / m0,
        "variants": variants,
    }
    if args.cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(args)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
