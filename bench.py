"""Benchmark of the D3Q19 lattice-Boltzmann hot path on one MI355X.

Default workload (the headline line): lid-driven cavity, D3Q19, 256^3 cells,
Re = 1000, U0 = 0.1, tau = 3 U0 N / Re + 1/2 = 0.5768, MRT, reference initial
state.  A step is the six-face + twelve-edge halo pack and unpack looped back
onto the same GPU followed by one fused pull-scheme stream + bounce-back +
macro + collide pass over the lattice (A-B buffers).  `value` is the plain
FP64 path's MLUPS for that step with the state resident in HBM; the FP32- and
FP16-stored variants are timed the same way and their relative L2 velocity
error and mass drift against the FP64 run (same initial state, same step
count) are reported beside them.  `roofline` is for the step kernel alone,
timed over the same number of steps without the exchange.

Other workloads (--workload), one JSON line each in the same format:
  batch      64 lid-driven cavities of N in {32,48,64,80,96}^3, Re in
             [100,1000], U0 in [0.05,0.1] (seeded), FP32 storage, one
             persistent kernel replayed from 100-step hipGraphs, autotuned;
             aggregate MLUPS beside launching the instances one at a time
  particles  192^3 with 2000 rigid spheres of radius 4-8 at seeded random
             non-overlapping positions, rho_s / rho_f = 1.5, g = 1e-5 along
             -z, tau = 0.65, FP32 storage with FP64 force/torque sums; the map
             is checked bit for bit and the trajectories against the FP64 path
  thermal    buoyancy-driven cavity 128^3, D3Q19 MRT + Guo force and D3Q7
             MRT temperature, Ra = 1e5, Pr = 0.71, walls at +-0.5, seeded
             noise; Nusselt numbers after --accuracy-steps

With the default workload, `other_workloads` in the same line holds short
runs of the other three (1000 timed steps each, no CPU legs, no probes):
their MLUPS, step time and roofline as measured on the same GPU in the same
process (--other-workloads 0 leaves them out).

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
METRIC = "D3Q19 MLUPS + %HBM peak; FP16/FP32-store vs FP64 at stated L2 error"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="cavity", choices=["cavity", "batch", "particles", "thermal"])
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default per workload)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps (default per workload)")
    ap.add_argument("--n", type=int, default=None, help="cells per edge (cavity 256, particles 192, thermal 128)")
    ap.add_argument("--collision", default="mrt", choices=["mrt", "bgk"])
    ap.add_argument("--accuracy-steps", type=int, default=None,
                    help="total steps after which reduced storage is compared with FP64 (default per workload)")
    ap.add_argument("--variants", default="f32,f16",
                    help="cavity / thermal: comma list of reduced storages ('' = none)")
    ap.add_argument("--cpu-baseline", type=int, default=1, help="time the CPU oracle on a bounded sample")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads of the CPU oracle leg (0 = every host core this process may use, at most "
                         "OMP_NUM_THREADS when that is set)")
    ap.add_argument("--waves-per-cu", type=int, default=0,
                    help="cavity / thermal: occupancy cap of the step kernels (0 = the library's choice)")
    ap.add_argument("--transform", default="auto", choices=["auto", "valu", "mfma"],
                    help="cavity: moment transforms of the FP16-stored MRT step")
    ap.add_argument("--halo-loopback", type=int, default=1,
                    help="cavity: 1 = the six faces and twelve edges exchanged every step, looped back "
                         "(written by the step kernel)")
    ap.add_argument("--overlap", type=int, default=1,
                    help="cavity: also time y/z faces looped back, fused step vs split step (shell, then "
                         "interior)")
    ap.add_argument("--attainable", type=int, default=1,
                    help="cavity / batch: also time the step's traffic with no arithmetic (lbm_bandwidth_probe)")
    ap.add_argument("--device-warmup", type=int, default=1,
                    help="cavity / thermal: before the W warmup steps, step a throwaway context of the same configuration "
                         "for ~150 ms of device time (steady clocks; the measured state is untouched)")
    ap.add_argument("--other-workloads", type=int, default=1,
                    help="with the cavity workload: also time the batch, particle and buoyancy workloads briefly "
                         "and report them under other_workloads")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run device copies of this many MiB (counter calibration) and exit")
    args = ap.parse_args()
    defaults = {  # steps, warmup, n, accuracy steps
        "cavity": (1000, 50, 256, 20000),
        "batch": (1000, 200, 0, 0),
        "particles": (1000, 20, 192, 5000),
        "thermal": (2000, 50, 128, 20000),
    }[args.workload]
    args.steps = defaults[0] if args.steps is None else args.steps
    args.warmup = defaults[1] if args.warmup is None else args.warmup
    args.n = defaults[2] if args.n is None else args.n
    args.accuracy_steps = defaults[3] if args.accuracy_steps is None else args.accuracy_steps
    return args


def timed(step, stream, args):
    """W untimed steps, then K steps between HIP events on `stream`."""
    from mglc_amd.hip import Event

    step(args.warmup)
    stream.synchronize()
    e0, e1 = Event(), Event()
    t0 = time.perf_counter()
    e0.record(stream)
    step(args.steps)
    e1.record(stream)
    e1.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_ms(e1)
    e0.close()
    e1.close()
    return ms, wall


def device_warmup(make_ctx, stream, ms=150.0, chunk=20):
    """Bring the device to its steady clocks before a measurement: step a
    throwaway context of the same configuration (its state is discarded, the
    measured context's is untouched) for about `ms` of device time.  A GPU
    that idled while the host worked (downloads, the previous variant's
    comparison) starts slow; the VALU-bound FP16 step is the most
    clock-sensitive."""
    from mglc_amd.hip import Event
    done = 0.0
    with make_ctx() as w:
        w.initial()
        while done < ms:
            e0, e1 = Event(), Event()
            e0.record(stream)
            w.step(chunk)
            e1.record(stream)
            e1.synchronize()
            done += e0.elapsed_ms(e1)
            e0.close()
            e1.close()
    return round(done, 1)


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


def roofline(bytes_per_launch, ms_per_launch, traffic=None):
    achieved = bytes_per_launch / (ms_per_launch * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic}


def attainable(storage, n, achieved, thermal=False):
    """What the memory system delivers for this traffic with no arithmetic,
    measured now on this device: a one-stream copy of the same bytes and the
    kernel's own pull pattern on an n^3 lattice in the library's layout (the
    D3Q19 pull: 19 upstream loads and 19 stores per cell; with `thermal` the
    buoyancy step's D3Q19 + D3Q7 pulls and force plane, 27 and 27); the
    kernel's achieved rate as a fraction of the latter."""
    from mglc_amd.lbm import bandwidth_probe
    copy = bandwidth_probe(0, storage, n)
    if thermal:
        pull = bandwidth_probe(2, storage, n)
        return {"copy_one_stream_GBs": round(copy, 1), "pull_d3q19_d3q7_force_no_arithmetic_GBs": round(pull, 1),
                "kernel_frac_of_pattern": round(achieved / pull, 4), "probe_n": n, "storage": storage}
    pull = bandwidth_probe(1, storage, n)
    return {"copy_one_stream_GBs": round(copy, 1), "pull19_no_arithmetic_GBs": round(pull, 1),
            "kernel_frac_of_pull19": round(achieved / pull, 4), "probe_n": n, "storage": storage}


def line_head(args, value, ms_step, dtype, data, config):
    return {"metric": METRIC, "value": round(value, 1), "unit": "MLUPS", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(ms_step, 4), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": dtype, "data": data, "config": config}


def host_cpu():
    """The host's CPU: model name, logical CPUs of the machine, CPUs this
    process may run on, and the thread count a CPU leg uses by default (the
    latter, capped by OMP_NUM_THREADS when set: a shared box's CPU share)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    threads = usable
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        threads = min(threads, int(omp))
    return {"model": model, "cpu_count": os.cpu_count(), "usable_cpus": usable, "threads": threads}


def oracle_rate(run, cells, budget_s=15.0, max_steps=200):
    """MLUPS of an oracle run callable run(steps) -> seconds, on a sample of
    about budget_s seconds."""
    per = max(run(1), 1e-3)
    steps = int(max(1, min(max_steps, budget_s / per)))
    sec = run(steps)
    return cells * steps / sec / 1e6, steps


# ---- lid-driven cavity (headline) ------------------------------------------
def run_storage(args, storage, stream):
    from mglc_amd.lbm import Cavity, CavityConfig

    n = args.n
    base = dict(Reynolds=1000.0, U0=0.1, collision=args.collision, storage=storage, waves_per_cu=args.waves_per_cu,
                transform=args.transform if storage == "f16" else "auto")
    def make():
        return Cavity(CavityConfig(n, n, n, halo_loopback=args.halo_loopback, **base), stream=stream.handle.value)

    with make() as cav:
        cav.initial()
        warm_ms = device_warmup(make, stream) if args.device_warmup else 0.0
        # with the loopback every timed step is one launch of the step kernel,
        # which also writes the outgoing halo (the first, untimed step call
        # exchanged the initial state's halo once)
        ms, wall = timed(cav.step, stream, args)
        # a step call writes the macroscopic fields on its last step only:
        # time single-step calls (each such a step) to split the timed
        # region into its one macro step and K - 1 plain steps
        ms_macro, _ = timed(lambda k: [cav.step(1) for _ in range(k)], stream,
                            argparse.Namespace(warmup=1, steps=5))
        ms_macro /= 5
        info = cav.info()
        halo_bytes = cav.halo_buffer_bytes() if args.halo_loopback else 0
        done = args.warmup + args.steps + 6
        if args.accuracy_steps > done:
            cav.step(args.accuracy_steps - done)
        fields = cav.macro()
        steps_total = max(done, args.accuracy_steps)
    ms_no_exchange = ms
    if args.halo_loopback:
        # the same step with no halo written, for comparison
        with Cavity(CavityConfig(n, n, n, **base), stream=stream.handle.value) as cav:
            cav.initial()
            ms_no_exchange, _ = timed(cav.step, stream, args)
    cells = n ** 3
    plain = (ms - ms_macro) / (args.steps - 1) if args.steps > 1 else None
    return dict(storage=storage, ms=ms, ms_kernel=ms, ms_no_exchange=ms_no_exchange, wall_s=wall, warm_ms=warm_ms,
                ms_macro_step=ms_macro, ms_plain_step=plain,
                mlups=cells * args.steps / (ms * 1e-3) / 1e6, bytes_per_lup=info.bytes_per_lup,
                # a step call writes rho, u, v, w (FP64) on its last step: the K
                # timed launches write them once
                bytes_per_launch=info.bytes_per_lup * cells + halo_bytes + 32.0 * cells / args.steps,
                halo_bytes=halo_bytes, fields=fields,
                steps_total=steps_total)


def overlap_compare(args, stream):
    """The y and z faces looped back (a periodic y-z box: the exchange of a
    subdomain with four neighbours, on one GPU), timed with the fused step
    (which writes the next step's halo itself) and with the split step
    (boundary shell first, writing the halo, then the interior)."""
    from mglc_amd.lbm import Cavity, CavityConfig

    n = args.n
    base = dict(Reynolds=1000.0, U0=0.1, collision=args.collision, storage="f64", halo_faces=0x3C, halo_loopback=1)
    steps = min(args.steps, 500)
    out = {"halo": "y and z faces looped back: 4 faces x 5 pops + 4 y-z edges per step", "storage": "f64",
           "steps": steps}
    for name, ov in (("fused", 0), ("split", 1)):
        with Cavity(CavityConfig(n, n, n, halo_overlap=ov, **base), stream=stream.handle.value) as cav:
            cav.initial()
            ms, _ = timed(cav.step, stream, argparse.Namespace(warmup=args.warmup, steps=steps))
            out[f"ms_per_step_{name}"] = round(ms / steps, 4)
    return out


def cpu_baseline_cavity(args):
    """The oracle (CPU restatement, OpenMP) on a bounded sample of the same
    workload: 256^3 cells from the reference initial state, a few steps."""
    from oracle import oracle as O
    n = args.n
    host = host_cpu()
    threads = args.cpu_threads if args.cpu_threads > 0 else host["threads"]
    tau = 0.1 * n / 1000.0 * 3.0 + 0.5
    coll = O.MRT if args.collision == "mrt" else O.BGK
    p = O.cavity_params(n, n, n, tau, 0.1, collision=coll)

    def run(steps):
        O.cavity_run(p, steps, nthreads=threads)
        return O.last_loop_seconds()

    rate, steps = oracle_rate(run, n ** 3)
    return {"value": rate, "unit": "MLUPS", "cores": threads, "kind": "port", "host": host,
            "sample": f"{n}^3 {args.collision.upper()} FP64 cavity, {steps} steps from the reference "
                      f"initial state, oracle/lbm_oracle.c with {threads} OpenMP threads, loop time only"}


def cpu_reference(kind, n, cells, threads, budget_s=15.0, max_steps=100):
    """The reference's own subroutines (oracle/_ref, built from the
    reference's Fortran sources) on the host cores: `threads` processes at
    once, each stepping one z slab of the n^3 grid (the subdomains of the
    reference's MPI decomposition, without its halo exchange, so an upper
    bound on that run's rate), timed on a sample of about budget_s seconds of
    loop time; None where that build is absent."""
    from oracle import oracle as O
    exe = O.REF_CAVITY if kind == "cavity" else O.REF_THERMAL
    if not os.path.exists(exe):
        return None
    procs = max(1, min(threads, n // 2))
    nzs = [n // procs + (1 if p < n % procs else 0) for p in range(procs)]
    per = max(max(O.reference_loop_seconds_slabs(kind, n, n, nzs, 1)), 1e-3)
    steps = int(max(1, min(max_steps, round(budget_s / per))))
    sec = max(O.reference_loop_seconds_slabs(kind, n, n, nzs, steps))
    what = ("MPI/Lid_driven_cavity/fortran/3d/mpi_3d_blocked collision, streaming, bounceback, macro"
            if kind == "cavity" else
            "MPI/Buoyancy_driven_cavity/fortran/3d/mpi_blocked collision, streaming, bounceback, collisionT, "
            "streamingT, bouncebackT, macro, macroT")
    return {"value": cells * steps / sec / 1e6, "unit": "MLUPS", "cores": procs, "kind": "reference",
            "host": host_cpu(), "seconds": round(sec, 2),
            "sample": f"{n}^3 as {procs} z slabs of {min(nzs)}-{max(nzs)} layers, one process per slab at once, "
                      f"{steps} steps of the reference's own {what} (amdflang -O3; loop time of the slowest "
                      f"process; no halo exchange between slabs; its compile-time tau and Ra differ from this "
                      f"workload's, which changes no operation count)"}


def bench_cavity(args, stream):
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
                        "kernel_hbm_frac": round(r["bytes_per_launch"] / (r["ms_kernel"] / args.steps * 1e-3)
                                                 / 1e9 / HBM_PEAK_GBS, 4),
                        "ms_per_step_no_exchange": round(r["ms_no_exchange"] / args.steps, 4),
                        "ms_macro_step": round(r["ms_macro_step"], 4),
                        "ms_per_step_without_macro": None if r["ms_plain_step"] is None
                        else round(r["ms_plain_step"], 4),
                        "speedup_vs_f64": round(r["mlups"] / primary["mlups"], 3),
                        "rel_l2_velocity_vs_f64": l2, "mass_drift_vs_f64": drift,
                        "compared_at_step": r["steps_total"], "device_warmup_ms": r["warm_ms"]}
    m0 = float(primary["fields"][0].size)
    workload = f"lid-driven cavity D3Q19 {args.n}^3 Re=1000 U0=0.1 {args.collision.upper()}"
    config = {"workload": workload, "cells": args.n ** 3, "collision": args.collision,
              "storage": "f64", "pattern": "A-B pull", "kernel": "k_cavity_step",
              "halo": "6 faces x 5 pops + 12 edges looped back every step (written by the step kernel)"
              if args.halo_loopback else "none",
              "kernel_ms_per_step": round(primary["ms_kernel"] / args.steps, 4),
              "ms_macro_step": round(primary["ms_macro_step"], 4),
              "ms_per_step_without_macro": None if primary["ms_plain_step"] is None
              else round(primary["ms_plain_step"], 4),
              "waves_per_cu": args.waves_per_cu or "library default (16)",
              "bytes_per_launch": {"lattice": int(primary["bytes_per_lup"] * args.n ** 3),
                                   "macro_fields_once_per_K_launches": 32 * args.n ** 3,
                                   "outgoing_halo": int(primary["halo_bytes"])},
              "ms_per_step_no_exchange": round(primary["ms_no_exchange"] / args.steps, 4),
              "device_warmup_ms": primary["warm_ms"]}
    line = line_head(args, primary["mlups"], primary["ms"] / args.steps, "f64",
                     "synthetic (reference initial state: rho=1, u=0, lid layer u=U0)", config)
    line["roofline"] = roofline(primary["bytes_per_launch"], primary["ms_kernel"] / args.steps,
                                traffic_from_profiles(f"{args.n}^3 {args.collision} f64"))
    line["mass_drift_f64"] = abs(float(primary["fields"][0].sum()) - m0) / m0
    line["variants"] = variants
    if args.overlap:
        ov = overlap_compare(args, stream)
        ov["ms_per_step_no_exchange"] = round(primary["ms_no_exchange"] / args.steps, 4)
        line["overlap"] = ov
    if args.attainable:
        line["roofline"]["attainable"] = attainable("f64", args.n, line["roofline"]["achieved"])
        for st, v in variants.items():
            v["attainable"] = attainable(st, args.n, v["kernel_hbm_frac"] * HBM_PEAK_GBS)
    if args.cpu_baseline:
        line["cpu_baseline"] = cpu_baseline_cavity(args)
        threads = args.cpu_threads if args.cpu_threads > 0 else host_cpu()["threads"]
        ref = cpu_reference("cavity", args.n, args.n ** 3, threads)
        if ref is not None:
            line["cpu_reference"] = ref
    return line


# ---- batched cavities --------------------------------------------------------
def bench_batch(args, stream_unused):
    from mglc_amd.hip import Stream
    from mglc_amd.lbm import Batch, batch_configs

    cfgs = batch_configs(64, seed=12345, storage="f32", collision=args.collision)
    cells = sum(c.nx * c.ny * c.nz for c in cfgs)
    args.steps = max(100, args.steps // 100 * 100)   # whole graph replays
    args.warmup = max(100, args.warmup // 100 * 100)
    each_steps = 100
    with Batch(cfgs) as b:
        b.initial()
        threads, blocks = b.autotune(trial_steps=4)
        info = b.info()
        s = Stream(borrowed=info.stream)
        ms, wall = timed(b.step, s, args)
        a2 = argparse.Namespace(warmup=2, steps=each_steps)
        ms_each, _ = timed(b.step_each, s, a2)
        info = b.info()
        s.close()
    mlups = cells * args.steps / (ms * 1e-3) / 1e6
    mlups_each = cells * each_steps / (ms_each * 1e-3) / 1e6
    sizes = sorted({c.nx for c in cfgs})
    config = {"workload": f"64 lid-driven cavities D3Q19, N in {{32,48,64,80,96}}^3 (seed 12345), Re U[100,1000], "
                          f"U0 U[0.05,0.1], {args.collision.upper()}, FP32 storage",
              "cells": cells, "instances": len(cfgs), "sizes_present": sizes, "storage": "f32",
              "kernel": "k_batch_step (persistent, 100-step hipGraph replays)", "threads": threads,
              "blocks": blocks, "cells_per_thread": info.cells_per_thread, "graph_steps": info.graph_steps}
    line = line_head(args, mlups, ms / args.steps, "f32", "synthetic (reference initial state per instance)", config)
    # kernel time per launch: the graph holds one step kernel per step and
    # one 1-thread step-counter kernel per 100 steps
    line["roofline"] = roofline(info.bytes_per_lup * cells, ms / args.steps, traffic_from_profiles("batch64 f32"))
    if args.attainable:
        # probed on one cube of about the batch's total cell count (the
        # instances together stream as much data), edge a multiple of 64 so
        # that whole waves cover its rows
        edge = max(64, int(cells ** (1.0 / 3.0)) // 64 * 64)
        line["roofline"]["attainable"] = attainable("f32", edge, line["roofline"]["achieved"])
    line["one_at_a_time"] = {"mlups": round(mlups_each, 1), "ms_per_step": round(ms_each / each_steps, 4),
                             "launches_per_step": len(cfgs), "speedup_batched": round(mlups / mlups_each, 3)}
    if args.cpu_baseline:
        from oracle import oracle as O
        threads_c = args.cpu_threads if args.cpu_threads > 0 else host_cpu()["threads"]
        params = [O.cavity_params(c.nx, c.ny, c.nz, c.tau(), c.U0, collision=O.MRT if args.collision == "mrt"
                                  else O.BGK) for c in cfgs]

        def run(steps):
            tot = 0.0
            for p in params:
                O.cavity_run(p, steps, nthreads=threads_c)
                tot += O.last_loop_seconds()
            return tot

        rate, steps = oracle_rate(run, cells, budget_s=15.0, max_steps=50)
        line["cpu_baseline"] = {"value": rate, "unit": "MLUPS", "cores": threads_c, "kind": "port",
                                "host": host_cpu(),
                                "sample": f"the 64 instances in FP64, {steps} steps each, one after another, "
                                          f"oracle/lbm_oracle.c with {threads_c} OpenMP threads, loop time only"}
    return line


# ---- particle-laden flow -----------------------------------------------------
def node_map(shape, xc, r):
    """Owner map by direct enumeration: node (i,j,k) at 1-based coordinates,
    squared distance summed as ((dx*dx + dy*dy) + dz*dz), the last sphere
    claiming a node keeps it."""
    nz, ny, nx = shape
    owner = -np.ones(shape, dtype=np.int32)
    for q in range(len(r)):
        lo = np.maximum(np.floor(xc[q] - r[q]).astype(int) - 2, 0)
        hi = np.minimum(np.ceil(xc[q] + r[q]).astype(int) + 1, np.array([nx, ny, nz]))
        k, j, i = np.meshgrid(np.arange(lo[2], hi[2]), np.arange(lo[1], hi[1]), np.arange(lo[0], hi[0]),
                              indexing="ij")
        dx, dy, dz = (i + 1.0) - xc[q, 0], (j + 1.0) - xc[q, 1], (k + 1.0) - xc[q, 2]
        sel = (dx * dx + dy * dy) + dz * dz <= r[q] * r[q]
        owner[k[sel], j[sel], i[sel]] = q
    return owner


def bench_particles(args, stream):
    from mglc_amd.lbm import Cavity, CavityConfig, random_spheres

    n, count, tau, rho_s, g = args.n, 2000, 0.65, 1.5, 1e-5
    xc0, r = random_spheres(n, n, n, count=count, rmin=4.0, rmax=8.0, seed=2025)

    def cfg(storage):
        return CavityConfig(n, n, n, U0=0.0, tauf=tau, meq12_rho=1, storage=storage, collision=args.collision)

    out = {}
    for storage in ("f32", "f64"):
        with Cavity(cfg(storage), stream=stream.handle.value) as cav:
            cav.particles_init(xc0, r, rho_solid=rho_s, gravity=g)
            fluid_nodes = int((cav.particle_map() < 0).sum())
            ms, _ = timed(cav.step, stream, args)
            done = args.warmup + args.steps
            if args.accuracy_steps > done:
                cav.step(args.accuracy_steps - done)
            st = cav.particles()
            owner = cav.particle_map()
            info = cav.info()
        out[storage] = dict(ms=ms, st=st, owner=owner, fluid0=fluid_nodes, bpl=info.bytes_per_lup)
    f32, f64 = out["f32"], out["f64"]
    steps_total = max(args.warmup + args.steps, args.accuracy_steps)
    disp = np.linalg.norm(f64["st"]["xc"] - xc0, axis=1)
    dev = np.linalg.norm(f32["st"]["xc"] - f64["st"]["xc"], axis=1)
    map_exact = bool(np.array_equal(f32["owner"], node_map(f32["owner"].shape, f32["st"]["xc"], r)) and
                     np.array_equal(f64["owner"], node_map(f64["owner"].shape, f64["st"]["xc"], r)))
    cells = n ** 3
    mlups = cells * args.steps / (f32["ms"] * 1e-3) / 1e6
    # algorithmic bytes of a step: every fluid node reads and writes its 19
    # populations and reads its owner entry (the macro fields, rho, u, v, w in
    # FP64, are written on the last step of a call only)
    step_bytes = f32["fluid0"] * (f32["bpl"] + 4)
    config = {"workload": f"particle-laden D3Q19 {n}^3, {count} spheres r in [4,8] (seed 2025), rho_s/rho_f=1.5, "
                          f"g=1e-5 along -z, tau={tau}, {args.collision.upper()}, FP32 storage",
              "cells": cells, "fluid_nodes": f32["fluid0"], "spheres": count, "storage": "f32",
              "kernel": "particle step: links + fluid + update + map + refill kernels",
              "fluid_mlups": round(f32["fluid0"] * args.steps / (f32["ms"] * 1e-3) / 1e6, 1)}
    line = line_head(args, mlups, f32["ms"] / args.steps, "f32",
                     "synthetic (seeded random sphere packing, fluid at rest)", config)
    line["roofline"] = roofline(step_bytes, f32["ms"] / args.steps, traffic_from_profiles(f"particles{n} f32"))
    line["fp64"] = {"mlups": round(cells * args.steps / (f64["ms"] * 1e-3) / 1e6, 1),
                    "ms_per_step": round(f64["ms"] / args.steps, 4)}
    line["accuracy"] = {"compared_at_step": steps_total, "map_bit_exact": map_exact,
                        "max_displacement_f64": float(disp.max()),
                        "max_trajectory_deviation_f32_vs_f64": float(dev.max()),
                        "rel_trajectory_deviation": float(dev.max() / max(disp.max(), 1e-300)),
                        "rel_velocity_deviation": float(np.abs(f32["st"]["U"] - f64["st"]["U"]).max() /
                                                        max(np.abs(f64["st"]["U"]).max(), 1e-300))}
    if args.cpu_baseline:
        from oracle import oracle as O
        p = O.particle_params(n, n, n, count, tau, rho_solid=rho_s, gravity=g,
                              collision=O.MRT if args.collision == "mrt" else O.BGK)

        threads = args.cpu_threads if args.cpu_threads > 0 else host_cpu()["threads"]
        O.set_threads(threads)

        def run(steps):
            t0 = time.perf_counter()
            O.particle_run(p, steps, xc0, r)
            return time.perf_counter() - t0

        t_init = run(0)
        steps = 4
        sec = run(steps) - t_init
        line["cpu_baseline"] = {"value": cells * steps / sec / 1e6, "unit": "MLUPS", "cores": threads,
                                "kind": "port", "host": host_cpu(),
                                "sample": f"the same {n}^3 / {count}-sphere case in FP64, {steps} steps, "
                                          f"oracle/particle_oracle.c with {threads} OpenMP threads (per-node "
                                          "loops; link sums, springs and refill serial), set-up time subtracted"}
    return line


# ---- buoyancy-driven cavity ----------------------------------------------------
def bench_thermal(args, stream):
    from mglc_amd.lbm import Cavity, ThermalConfig, buoyancy_initial_temperature

    n = args.n
    T0 = buoyancy_initial_temperature(n, n, n, seed=2024)
    z = np.zeros_like(T0)

    def make(storage):
        return ThermalConfig(n, n, n, Rayleigh=1e5, Prandtl=0.71, Thot=0.5, Tcold=-0.5, thermal_axis=1,
                             storage=storage, waves_per_cu=args.waves_per_cu)

    def run(storage):
        with Cavity(make(storage), stream=stream.handle.value) as cav:
            cav.load_fields(np.ones_like(T0), z, z, z, T0)
            warm = 0.0
            if args.device_warmup:
                warm = device_warmup(lambda: Cavity(make(storage), stream=stream.handle.value), stream)
            ms, _ = timed(cav.step, stream, args)
            done = args.warmup + args.steps
            if args.accuracy_steps > done:
                cav.step(args.accuracy_steps - done)
            out = dict(ms=ms, nu=cav.nusselt(), info=cav.info(), fields=cav.macro(), T=cav.temperature(),
                       warm=warm)
        return out

    cfg = make("f64")
    primary = run("f64")
    cells = n ** 3
    ms = primary["ms"]
    mlups = cells * args.steps / (ms * 1e-3) / 1e6
    # k_thermal_step: 19 + 7 FP64 populations read and written, Fz read and written
    bpl = primary["info"].bytes_per_lup
    nu_vol, nu_wall = primary["nu"]
    config = {"workload": f"buoyancy-driven cavity D3Q19+D3Q7 {n}^3 Ra=1e5 Pr=0.71 walls +-0.5, MRT, FP64",
              "cells": cells, "storage": "f64", "kernel": "k_thermal_step", "bytes_per_lup": bpl,
              "device_warmup_ms": primary["warm"]}
    line = line_head(args, mlups, ms / args.steps, "f64",
                     "synthetic (rho=1, u=0, T=0 plus seeded noise +-1e-4, held wall layers)", config)
    line["roofline"] = roofline(bpl * cells, ms / args.steps, traffic_from_profiles(f"thermal{n} f64"))
    if args.attainable:
        line["roofline"]["attainable"] = attainable("f64", n, line["roofline"]["achieved"], thermal=True)
    steps_total = max(args.warmup + args.steps, args.accuracy_steps)
    line["nusselt"] = {"at_step": steps_total, "nu_volume": nu_vol, "nu_wall": nu_wall}
    variants = {}
    vel64 = np.stack(primary["fields"][1:])
    for st in [s for s in args.variants.split(",") if s]:
        r = run(st)
        vel = np.stack(r["fields"][1:])
        m64 = float(primary["fields"][0].sum())
        b = r["info"].bytes_per_lup
        variants[st] = {
            "mlups": round(cells * args.steps / (r["ms"] * 1e-3) / 1e6, 1),
            "ms_per_step": round(r["ms"] / args.steps, 4), "bytes_per_lup": b,
            "kernel_hbm_frac": round(b * cells / (r["ms"] / args.steps * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "speedup_vs_f64": round(ms / r["ms"], 3),
            "rel_l2_velocity_vs_f64": float(np.linalg.norm((vel - vel64).ravel()) / np.linalg.norm(vel64.ravel())),
            "rel_l2_T_vs_f64": float(np.linalg.norm((r["T"] - primary["T"]).ravel()) /
                                     np.linalg.norm(primary["T"].ravel())),
            "mass_drift_vs_f64": abs(float(r["fields"][0].sum()) - m64) / m64,
            "nu_volume": r["nu"][0], "nu_wall": r["nu"][1], "compared_at_step": steps_total,
            "device_warmup_ms": r["warm"]}
    line["variants"] = variants
    if args.cpu_baseline:
        from oracle import oracle as O
        d = cfg.derived()
        threads = args.cpu_threads if args.cpu_threads > 0 else host_cpu()["threads"]
        p = O.thermal_params(n, n, n, d["tau"], d["para_a"], d["g_beta"], 0.0, 0.5, -0.5, 0.0, tconst_axis=1)
        fields = (np.ones_like(T0), z, z, z, T0)

        def run(steps):
            t0 = time.perf_counter()
            O.thermal_run(p, steps, fields=fields, nthreads=threads)
            return time.perf_counter() - t0

        t_init = run(0)
        rate, steps = oracle_rate(lambda s: max(run(s) - t_init, 1e-6), cells)
        line["cpu_baseline"] = {"value": rate, "unit": "MLUPS", "cores": threads, "kind": "port",
                                "host": host_cpu(),
                                "sample": f"the same {n}^3 case in FP64, {steps} steps, oracle/thermal_oracle.c "
                                          f"with {threads} OpenMP threads, set-up time subtracted"}
        ref = cpu_reference("thermal", n, cells, threads)
        if ref is not None:
            line["cpu_reference"] = ref
    return line


# the short runs of the other workloads beside the headline line: timed
# steps, warm-up steps, cells per edge, accuracy steps (0: compared at the
# end of the timed run)
OTHER_WORKLOADS = {
    "particles": (1000, 20, 192, 0),
    "thermal": (1000, 50, 128, 0),
    "batch": (1000, 200, 0, 0),
}


def other_workloads(args, stream):
    out = {}
    for name, (steps, warmup, n, acc) in OTHER_WORKLOADS.items():
        a = argparse.Namespace(**vars(args))
        a.workload, a.steps, a.warmup, a.n, a.accuracy_steps = name, steps, warmup, n, acc
        a.cpu_baseline, a.attainable = 0, 0
        t0 = time.perf_counter()
        line = {"particles": bench_particles, "thermal": bench_thermal, "batch": bench_batch}[name](a, stream)
        keep = {k: line[k] for k in ("value", "unit", "ms_per_step", "dtype", "steps", "warmup", "config",
                                     "roofline", "fp64", "accuracy", "nusselt", "variants", "one_at_a_time")
                if k in line}
        keep["wall_s"] = round(time.perf_counter() - t0, 1)
        out[name] = keep
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} refused: this build runs on one GPU "
              "(the multi-GPU halo path is design only)", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        from mglc_amd.hip import device_copy_calibration
        device_copy_calibration(args.copy_calibration << 20, 5)
        print(json.dumps({"copy_calibration_bytes": args.copy_calibration << 20, "reps": 5}))
        return
    from mglc_amd.hip import Stream

    stream = Stream()
    fn = {"cavity": bench_cavity, "batch": bench_batch, "particles": bench_particles,
          "thermal": bench_thermal}[args.workload]
    line = fn(args, stream)
    if args.workload == "cavity" and args.other_workloads:
        line["other_workloads"] = other_workloads(args, stream)
    stream.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
