"""Benchmark of the explicit finite-volume step on one MI355X.

Workload (the 2-D configuration the headline metric is quoted on): four-quadrant Riemann
problem (Lax-Liu configuration 3) on [0,1]^2 at 2048 x 2048 cells, limited linear
reconstruction (Barth-Jespersen) + HLLC, SSP-RK3, CFL 0.45, transmissive boundaries, FP64.
A step = one SSP-RK3 step (dt pass + three fused stages) over the whole grid, state resident
in HBM.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from turbostructred_amd import capi  # noqa: E402

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s


def lax_liu_state(cfg):
    import util
    return util.quadrants(cfg)


def gpu_run(n, steps, warmup, precision=capi.FP64):
    cfg = capi.make_config(ndims=2, nx=n, ny=n, cfl=0.45, riemann=capi.RIEMANN_HLLC,
                           recon=capi.RECON_LINEAR_LIM, limiter=capi.LIMITER_BJ, rk_stages=3,
                           precision=precision, bcs=[dict(type=capi.BC_NATURAL)])
    U = lax_liu_state(cfg)
    s = capi.Solver(cfg)
    s.set_state(U)
    if warmup:
        s.step(warmup)
    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:  # torch is only used for the contract's synchronisation bracket
        sync = lambda: None
    sync()
    t0 = time.perf_counter()
    info = s.step(steps)
    sync()
    t1 = time.perf_counter()
    assert info.steps_done == steps, info.steps_done
    elapsed = t1 - t0
    # per-kernel durations, measured with HIP events on the library's stream
    s.enable_kernel_timing(True)
    s.step(min(steps, 20))
    kern = {}
    names = {capi.KIND_DT: "dt_pass", capi.KIND_STAGE1: "stage1", capi.KIND_STAGE2: "stage2",
             capi.KIND_STAGE3: "stage3"}
    ncell = n * n
    for kind, name in names.items():
        ms, cnt = s.kernel_timing(kind)
        if cnt == 0:
            continue
        avg = ms / cnt * 1e-3
        b = s.algorithmic_bytes_per_cell(kind) * ncell
        kern[name] = dict(avg_us=avg * 1e6, launches=cnt, algorithmic_bytes=b, achieved_gbs=b / avg / 1e9,
                          frac=b / avg / HBM_PEAK)
    s.enable_kernel_timing(False)
    state = s.get_state()
    return cfg, elapsed, info, kern, state


def cpu_baseline(budget_s=15.0):
    """The FP64 oracle (a CPU port of the reference's step, restated) on this host's cores, on a
    bounded sample of the same workload."""
    from oracle import pyoracle
    threads = min(16, os.cpu_count() or 1)
    n = 512
    cfg = capi.make_config(ndims=2, nx=n, ny=n, cfl=0.45, riemann=capi.RIEMANN_HLLC,
                           recon=capi.RECON_LINEAR_LIM, limiter=capi.LIMITER_BJ, rk_stages=3,
                           bcs=[dict(type=capi.BC_NATURAL)])
    o = pyoracle.Oracle(cfg, threads=threads)
    o.set_state(lax_liu_state(cfg))
    t0 = time.perf_counter()
    o.step(1)
    one = time.perf_counter() - t0
    k = max(1, min(200, int(budget_s / max(one, 1e-6))))
    t0 = time.perf_counter()
    o.step(k)
    el = time.perf_counter() - t0
    return dict(value=n * n * k / el, unit="cell-updates/s", cores=threads, kind="port",
                sample=f"same problem at {n}x{n}, {k} SSP-RK3 steps (+1 warm-up), FP64 oracle, "
                       f"{threads} std::threads over grid lines")


def traffic_from_profiles(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary, if present."""
    path = os.path.join(ROOT, "profiles", "c2_2048_fp64_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(f"| {kernel} |"):
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            try:
                return float(cells[-1])
            except ValueError:
                return None
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)
    prec = capi.FP64 if args.precision == "fp64" else capi.FP32_RESIDUAL
    cfg, elapsed, info, kern, state = gpu_run(args.n, args.steps, args.warmup, prec)
    ncell = args.n * args.n
    value = ncell * args.steps / elapsed
    dom = max(kern.items(), key=lambda kv: kv[1]["avg_us"] * kv[1]["launches"])
    dname, d = dom
    traffic = traffic_from_profiles(dname)
    bytes_step = sum(s.get("algorithmic_bytes", 0) for s in kern.values()) / ncell
    out = {
        "metric": "cell-updates/s & % HBM roofline; served-solve p50/p99 at L1 err<=1e-5",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == capi.FP64 else "f32 residual / f64 state",
        "data": "synthetic (Lax-Liu configuration 3 initial state, no dataset)",
        "config": {"workload": f"2-D four-quadrant Riemann problem (Lax-Liu 3), {args.n}x{args.n} cells, "
                               "limited linear (Barth-Jespersen) + HLLC, SSP-RK3, CFL 0.45, transmissive",
                   "cells": ncell, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dname, "achieved": d["achieved_gbs"], "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": d["frac"], "traffic": traffic,
                     "algorithmic_bytes_per_launch": d["algorithmic_bytes"], "avg_launch_us": d["avg_us"]},
        "step_roofline_frac": value * bytes_step / HBM_PEAK,
        "algorithmic_bytes_per_cell_step": bytes_step,
        "kernels": kern,
        "final_time": info.time,
        "finite": bool(np.isfinite(state).all()),
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
