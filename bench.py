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


def gpu_run(n, steps, warmup, precision=capi.FP64, autotune=True):
    cfg = capi.make_config(ndims=2, nx=n, ny=n, cfl=0.45, riemann=capi.RIEMANN_HLLC,
                           recon=capi.RECON_LINEAR_LIM, limiter=capi.LIMITER_BJ, rk_stages=3,
                           precision=precision, bcs=[dict(type=capi.BC_NATURAL)])
    U = lax_liu_state(cfg)
    # tile height timed on this state (the library's autotuner); the context below uses the winner
    tuning = None
    if autotune:
        chosen, cands = capi.autotune_state(cfg, U, steps=5)
        tuning = {"chosen_rows_per_tile": chosen, "occupancy_model_rows_per_tile": cands[0][0] if cands else None,
                  "kernel_ms_per_step": {str(r): round(ms, 4) for r, ms in cands}}
    s = capi.Solver(cfg)
    t0 = time.perf_counter()
    s.set_state(U)
    t_set = time.perf_counter() - t0
    if warmup:
        s.step(warmup)
    # ts_step returns only after the library's stream has drained (it reads the step-control
    # block back), so the timed region starts and ends with the GPU idle: that is the
    # synchronisation bracket.  torch's GPU runtime is never initialised in this process: torch
    # ships its own HIP and HSA runtime libraries, and a second runtime opening the device beside
    # the library's fails to find it, or worse.
    t0 = time.perf_counter()
    info = s.step(steps)
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
    t0 = time.perf_counter()
    state = s.get_state()
    t_get = time.perf_counter() - t0
    # launch forms: a call of 256 steps replays the context's captured hipGraph of 256 steps (the
    # first such call builds it, untimed here), a call of 255 launches the kernels directly
    s.step(256)
    forms = {}
    for k, name in ((256, "graph_256_steps"), (255, "direct_255_steps")):
        t0 = time.perf_counter()
        s.step(k)
        forms[name + "_ms_per_step"] = (time.perf_counter() - t0) / k * 1e3
    tuning = {"autotune": tuning, "launch_forms": forms}
    return cfg, elapsed, info, kern, state, (t_set, t_get), tuning


def tgv_state(cfg, seed=2025):
    """Taylor-Green vortex on [0, 2 pi]^3: V0 = 1, M = 0.1, p0 = 1/(gamma M^2),
    p = p0 + (cos 2x + cos 2y)(cos 2z + 2)/16, rho = 1 + 1e-3 xi with xi ~ U(-1, 1) from Philox."""
    g = cfg.gamma
    n = cfg.nx
    c = (np.arange(n) + 0.5) * (2 * np.pi / n)
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    xi = np.random.Generator(np.random.Philox(seed)).uniform(-1, 1, (n, n, n))
    rho = 1 + 1e-3 * xi
    u = np.sin(x) * np.cos(y) * np.cos(z)
    v = -np.cos(x) * np.sin(y) * np.cos(z)
    p = 1 / (g * 0.01) + (np.cos(2 * x) + np.cos(2 * y)) * (np.cos(2 * z) + 2) / 16
    U = np.zeros((n + 2, n + 2, n + 2, 5))
    I = U[1:-1, 1:-1, 1:-1]
    I[..., 0] = rho
    I[..., 1] = rho * u
    I[..., 2] = rho * v
    I[..., 4] = p / (g - 1) + 0.5 * rho * (u * u + v * v)
    return U


def tgv_3d(n=256, steps=100):
    """3-D Taylor-Green vortex, periodic, limited linear + HLLC, SSP-RK3, CFL 0.45: FP64 and
    FP32-residual rates, and the FP32-residual density error against FP64 after the same steps."""
    out = {"workload": f"3-D Taylor-Green vortex {n}^3 periodic, limited linear (BJ) + HLLC, SSP-RK3, CFL 0.45",
           "steps": steps}
    U = None
    states = {}
    for prec, name in ((capi.FP64, "fp64"), (capi.FP32_RESIDUAL, "fp32_residual")):
        cfg = capi.make_config(ndims=3, nx=n, ny=n, nz=n, lx=2 * np.pi, ly=2 * np.pi, lz=2 * np.pi,
                               periodic=(1, 1, 1), cfl=0.45, riemann=capi.RIEMANN_HLLC,
                               recon=capi.RECON_LINEAR_LIM, rk_stages=3, precision=prec)
        if U is None:
            U = tgv_state(cfg)
        s = capi.Solver(cfg)
        s.set_state(U)
        s.step(2)
        t0 = time.perf_counter()
        info = s.step(steps)
        el = time.perf_counter() - t0
        bps = s.algorithmic_bytes_per_cell(capi.KIND_ALL)
        rate = n ** 3 * steps / el
        out[name] = {"cell_updates_per_s": rate, "ms_per_step": el / steps * 1e3,
                     "algorithmic_bytes_per_cell_step": bps, "step_roofline_frac": rate * bps / HBM_PEAK,
                     "final_time": info.time}
        states[name] = capi.interior(cfg, s.get_state())[..., 0].copy()
        s.close()
    d = np.abs(states["fp32_residual"] - states["fp64"])
    out["density_l1_fp32_vs_fp64"] = float(d.mean())
    out["density_linf_fp32_vs_fp64"] = float(d.max())
    return out


def viscous_channel(n=2048, steps=40):
    """Plane channel with viscous fluxes and the potential force Sigma (periodic x, walls at y = 0
    and y = 0.5, the layout of RunPoiseuille2D), n x n cells, Roe (Eps 0.05) + limited linear,
    FP64, forward Euler and SSP-RK3, started from the Poiseuille profile at u_c = 0.3."""
    mu, H, uc = 1e-3, 0.5, 0.3
    sigma = 8 * mu * uc / H**2
    out = {"workload": f"viscous channel {n}x{n}, periodic x, walls y, mu {mu}, Sigma ({sigma:.4g}, 0, 0), "
                       "Roe + limited linear (BJ), CFL 0.45, FP64", "steps": steps}
    for rk in (1, 3):
        cfg = capi.make_config(ndims=2, nx=n, ny=n, lx=1.0, ly=H, periodic=(1, 0, 0), cfl=0.45, eps=0.05,
                               riemann=capi.RIEMANN_ROE, recon=capi.RECON_LINEAR_LIM, rk_stages=rk,
                               bcs=[dict(type=capi.BC_WALL)], viscous=True, viscosity=mu, sigma=(sigma, 0, 0))
        y = (np.arange(n) + 0.5) * (H / n)
        u = 4 * uc * y * (H - y) / H**2
        U = np.zeros(capi.state_shape(cfg))
        U[..., 0] = 1.0
        U[1:-1, 1:-1, 1] = u[:, None]
        U[..., 4] = 1.0 / 0.4
        U[1:-1, 1:-1, 4] += 0.5 * u[:, None] ** 2
        s = capi.Solver(cfg)
        s.set_state(U)
        s.step(2)
        t0 = time.perf_counter()
        info = s.step(steps)
        el = time.perf_counter() - t0
        s.enable_kernel_timing(True)
        s.step(min(steps, 10))
        kern = {}
        for kind, name in ((capi.KIND_VISCOUS, "viscous_prepass"), (capi.KIND_DT, "dt_pass"),
                           (capi.KIND_STAGE1, "stage1"), (capi.KIND_STAGE2, "stage2"), (capi.KIND_STAGE3, "stage3")):
            ms, cnt = s.kernel_timing(kind)
            if cnt == 0:
                continue
            avg = ms / cnt * 1e-3
            b = s.algorithmic_bytes_per_cell(kind) * n * n
            kern[name] = dict(avg_us=avg * 1e6, launches=cnt, achieved_gbs=b / avg / 1e9, frac=b / avg / HBM_PEAK)
        bps = s.algorithmic_bytes_per_cell(capi.KIND_ALL)
        rate = n * n * steps / el
        out["euler" if rk == 1 else "ssp_rk3"] = {
            "cell_updates_per_s": rate, "ms_per_step": el / steps * 1e3, "algorithmic_bytes_per_cell_step": bps,
            "step_roofline_frac": rate * bps / HBM_PEAK, "kernels": kern, "final_time": info.time}
        s.close()
    return out


def sod_c1_config():
    """The reference's own CPU case: Sod, 1000 cells, piecewise constant + Roe (Eps 0), forward
    Euler, CFL 0.4, Natural boundaries, run until Time >= 0.2 (tests.h:13-80)."""
    return capi.make_config(ndims=1, nx=1000, cfl=0.4, riemann=capi.RIEMANN_ROE, recon=capi.RECON_PWC,
                            bcs=[dict(type=capi.BC_NATURAL)])


def sod_c1():
    """Time to solution of the Sod case on the GPU, state resident (the run is launch-bound:
    1000 cells per step, two launches per step replayed from hipGraphs)."""
    import util
    cfg = sod_c1_config()
    U = util.sod(cfg)
    s = capi.Solver(cfg)
    s.set_state(U)
    s.step(5)  # warm-up: graph capture and first launches
    s.set_state(U)
    s.set_time(0.0, 0)
    t0 = time.perf_counter()
    info = s.step(1 << 20, 0.2)
    el = time.perf_counter() - t0
    return {"workload": "Sod 1-D 1000 cells, PWC + Roe, forward Euler, CFL 0.4, to t = 0.2",
            "steps": info.steps_done, "final_time": info.time, "seconds": el,
            "steps_per_s": info.steps_done / el, "cell_updates_per_s": 1000 * info.steps_done / el}


def host_cores():
    """Logical CPUs of the host, the ones this process may run on, and physical cores (distinct
    (package, core) pairs of /proc/cpuinfo; None where that file does not say)."""
    logical = os.cpu_count()
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = logical
    phys = None
    try:
        pairs, pkg, core = set(), None, None
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k = k.strip()
            if k == "physical id":
                pkg = v.strip()
            elif k == "core id":
                core = v.strip()
            elif not k and pkg is not None and core is not None:
                pairs.add((pkg, core))
                pkg = core = None
        if pkg is not None and core is not None:
            pairs.add((pkg, core))
        phys = len(pairs) or None
    except OSError:
        pass
    return {"logical_cpus": logical, "usable_cpus": usable, "physical_cores": phys}


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
    # three timed runs of k steps each (continuing the same solve), the median reported
    k = max(1, min(100, int(budget_s / 3 / max(one, 1e-6))))
    rates = []
    for _ in range(3):
        t0 = time.perf_counter()
        o.step(k)
        rates.append(n * n * k / (time.perf_counter() - t0))
    out = dict(value=float(np.median(rates)), unit="cell-updates/s", cores=threads, kind="port",
               sample=f"same problem at {n}x{n}, median of 3 runs of {k} SSP-RK3 steps (+1 warm-up), FP64 "
                      f"oracle, {threads} std::threads over grid lines", runs=rates, host=host_cores())
    # the Sod case run whole on one core (the reference's configuration runs single-rank)
    import util
    c1 = sod_c1_config()
    o1 = pyoracle.Oracle(c1, threads=1)
    o1.set_state(util.sod(c1))
    t0 = time.perf_counter()
    i1 = o1.step(1 << 20, 0.2)
    el1 = time.perf_counter() - t0
    out["sod_c1"] = {"seconds": el1, "steps": i1.steps_done, "cores": 1,
                     "cell_updates_per_s": 1000 * i1.steps_done / el1}
    return out


def batched_tubes(count=4096, sample=48, seed=2024):
    """Batched form: `count` independent shock tubes (exact Godunov with FP64 Newton, FP32 bulk,
    SSP-RK3) in one persistent work-queue launch per stage, against one-at-a-time FP64 solves of
    a random sample of the same tubes (cell-updates/s of each), with density L1 error."""
    from turbostructred_amd import workloads
    tubes = workloads.shock_tubes(count, seed=seed, precision=capi.FP32_RESIDUAL)

    def whole_run(b):
        t0 = time.perf_counter()
        tickets = [b.submit(cfg, U, 10**6, te) for cfg, U, te in tubes]
        while b.run(32) > 0:
            pass
        el = time.perf_counter() - t0
        res = {}
        for t in tickets:
            done, st, info = b.poll(t)
            assert done
            res[t] = (st, info)
        return el, tickets, res

    rates = {}
    for auto in (True, False):
        b = capi.Batch(hbm_budget_bytes=4 << 30, autotune_tiles=auto)
        # warm-up: kernel loading and graph capture on a few tubes
        w = [b.submit(cfg, U, 20, te) for cfg, U, te in tubes[:8]]
        while w:
            b.run(8)
            w = [t for t in w if not b.poll(t)[0]]
        # with the autotuner: a first whole run that times the candidate tilings as the load
        # drains, then the measured run with the tilings it chose
        runs = 2 if auto else 1
        for r in range(runs):
            t_batch, tickets, results = whole_run(b)
            cu_batch = sum(cfg.nx * results[t][1].steps_done for (cfg, _, _), t in zip(tubes, tickets))
            rates[(auto, r)] = cu_batch / t_batch
        if auto:
            trials = b.tuning()
        else:
            keep = (t_batch, tickets, results, cu_batch)
        b.close()
    t_batch, tickets, results, cu_batch = keep
    chosen = {t["load_bucket"]: t["tile_param"] for t in trials if t["chosen"]}
    # one-at-a-time FP64 solves of a sample
    rng = np.random.default_rng(seed + 1)
    idx = rng.choice(count, size=sample, replace=False)
    cu_single, t_single, l1 = 0, 0.0, []
    for i in idx:
        cfg32, U, te = tubes[i]
        cfg = capi.make_config(ndims=1, nx=cfg32.nx, cfl=0.45, gamma=cfg32.gamma, riemann=capi.RIEMANN_GODUNOV,
                               recon=capi.RECON_LINEAR_LIM, rk_stages=3, precision=capi.FP64)
        s = capi.Solver(cfg)
        s.set_state(U)
        t1 = time.perf_counter()
        info = s.step(10**6, te, te)
        t_single += time.perf_counter() - t1
        cu_single += cfg.nx * info.steps_done
        ref = capi.interior(cfg, s.get_state())[..., 0]
        got = capi.interior(cfg32, results[tickets[i]][0])[..., 0]
        l1.append(float(np.abs(got - ref).mean()))
    rate_b, rate_s = cu_batch / t_batch, cu_single / t_single
    return {"workload": f"{count} shock tubes, N~logU[64,4096] ({sum(c.nx for c, _, _ in tubes)} cells), "
                        "exact Godunov (FP64 Newton) + limited linear, SSP-RK3, FP32 residual, t_end in-domain",
            "cell_updates_per_s": rate_b, "seconds": t_batch, "tiling": "load rule (the default)",
            "autotuned_first_run_while_tuning_cell_updates_per_s": rates[(True, 0)],
            "autotuned_second_run_cell_updates_per_s": rates[(True, 1)],
            "tuned_strips_per_work_item_by_log2_cells": {str(k): chosen[k] for k in sorted(chosen)},
            "fp64_one_at_a_time_cell_updates_per_s": rate_s, "fp64_sample_tubes": sample,
            "speedup_vs_fp64_one_at_a_time": rate_b / rate_s,
            "density_l1_vs_fp64_max": max(l1), "density_l1_vs_fp64_mean": float(np.mean(l1))}


def serving(rates=(50, 100, 200, 500, 1000, 2000), requests=200, seed=7):
    """Served form: 2-D Riemann problems arrive as a Poisson process; a background engine thread
    admits them by HBM footprint and batches them continuously (FP32 residual, FP64 state).
    Latency = arrival -> result on the host, `requests` arrivals per rate (p99 from 200 is the
    second-largest latency).  The baseline is the same request mix solved one at a time in FP64
    on the GPU, every request, timed around each solve."""
    import threading
    from turbostructred_amd import workloads
    out = {"workload": f"2-D four-quadrant Riemann problems, L~logU[64,1024], steps~U{{50..500}}, "
                       f"HLLC + limited linear + SSP-RK3, FP32 residual / FP64 state, {requests} requests per rate",
           "rates": []}
    reqs = workloads.riemann_2d(requests, seed=seed, precision=capi.FP32_RESIDUAL)
    b = capi.Batch(hbm_budget_bytes=16 << 30)
    w = [b.submit(cfg, U, 5) for cfg, U, _ in reqs[:4]]  # warm-up
    while w:
        b.run(5)
        w = [t for t in w if not b.poll(t)[0]]
    b.start_server()
    served = {}
    for lam in rates:
        rng = np.random.default_rng(seed + lam)
        arrivals = np.cumsum(rng.exponential(1.0 / lam, requests))
        lat = [None] * requests
        tick = [None] * requests
        t0 = time.perf_counter()
        done_count = [0]
        failed = set()

        def submitter():
            for i, (cfg, U, k) in enumerate(reqs):
                wait = t0 + arrivals[i] - time.perf_counter()
                if wait > 0:
                    time.sleep(wait)
                tick[i] = b.submit(cfg, U, k)

        th = threading.Thread(target=submitter)
        th.start()
        while done_count[0] < requests:
            for i in range(requests):
                if lat[i] is None and tick[i] is not None:
                    try:
                        d, st, _ = b.poll(tick[i])
                    except capi.TSError as e:  # a solve that turns unphysical (FP64 fails it too)
                        if e.code != capi.TS_ERR_UNPHYSICAL:
                            raise
                        d, st = True, None
                        failed.add(i)
                    if d:
                        lat[i] = time.perf_counter() - (t0 + arrivals[i])
                        done_count[0] += 1
                        if lam == rates[0] and st is not None:
                            served[i] = st
            time.sleep(0.0005)
        th.join()
        span = time.perf_counter() - t0
        ok = [i for i in range(requests) if i not in failed]
        la = np.array([lat[i] for i in ok]) * 1e3
        out["rates"].append({"rate_per_s": lam, "p50_ms": float(np.percentile(la, 50)),
                             "p99_ms": float(np.percentile(la, 99)),
                             "throughput_cell_updates_per_s": sum(reqs[i][0].nx * reqs[i][0].ny * reqs[i][2] for i in ok) / span,
                             "requests": requests, "unphysical": len(failed)})
        print(f"serving: {lam}/s done", file=sys.stderr, flush=True)
    b.stop_server()
    b.close()
    # the same mix one solve at a time in FP64 (every request), and the density error of the served
    # FP32-residual results against those FP64 solves
    l1 = []
    t_fp64, cu_fp64 = 0.0, 0
    fp64_failed = 0
    t_e2e = 0.0  # the same solves timed end to end: context, upload, steps, download
    for i, (cfg32, U, k) in enumerate(reqs):
        cfg = capi.make_config(ndims=2, nx=cfg32.nx, ny=cfg32.ny, cfl=0.45, riemann=capi.RIEMANN_HLLC,
                               recon=capi.RECON_LINEAR_LIM, rk_stages=3, precision=capi.FP64)
        t0 = time.perf_counter()
        s = capi.Solver(cfg)
        s.set_state(U)
        t1 = time.perf_counter()
        try:
            s.step(k)
        except capi.TSError as e:
            if e.code != capi.TS_ERR_UNPHYSICAL:
                raise
            fp64_failed += 1
            s.close()
            continue
        t2 = time.perf_counter()
        t_fp64 += t2 - t1
        cu_fp64 += cfg.nx * cfg.ny * k
        st = s.get_state()
        t_e2e += time.perf_counter() - t0
        ref = capi.interior(cfg, st)[..., 0]
        if i in served:
            l1.append(float(np.abs(capi.interior(cfg32, served[i])[..., 0] - ref).mean()))
        s.close()
    out["fp64_unphysical"] = fp64_failed
    rate64 = cu_fp64 / t_fp64
    sat = max(r["throughput_cell_updates_per_s"] for r in out["rates"])
    out["density_l1_vs_fp64_max"] = max(l1)
    out["density_l1_vs_fp64_mean"] = float(np.mean(l1))
    out["fp64_one_at_a_time_cell_updates_per_s"] = rate64
    out["fp64_one_at_a_time_requests"] = requests
    out["served_over_fp64_one_at_a_time"] = sat / rate64
    # the served throughput counts each request from its arrival to its result on the host, so
    # the like-for-like baseline also counts context creation, upload and download
    out["fp64_one_at_a_time_end_to_end_cell_updates_per_s"] = cu_fp64 / t_e2e
    out["served_over_fp64_one_at_a_time_end_to_end"] = sat / (cu_fp64 / t_e2e)
    return out


_COPY_PROBE = """
import torch
n = {nbytes}
x = torch.empty(n // 8, dtype=torch.float64, device="cuda")
y = torch.empty_like(x)
x.fill_(1.0)
for _ in range(3):
    y.copy_(x)
e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
e0.record()
for _ in range({reps}):
    y.copy_(x)
e1.record()
e1.synchronize()
print(2 * n / (e0.elapsed_time(e1) * 1e-3 / {reps}) / 1e9)
"""


def copy_bandwidth(nbytes=1 << 30, reps=20):
    """Device-to-device copy rate in GB/s (bytes read + bytes written per second) of two 1 GiB
    buffers, four times the 256 MB Infinity Cache, so the copy runs at the rate HBM sustains; it
    is printed next to the vendor peak.  torch times the copies with events on its own stream, in
    a child process started before this process makes any HIP call: torch brings its own HIP and
    HSA runtime libraries, which must not open the device beside the library's in one process.
    None if the child fails."""
    import subprocess
    try:
        r = subprocess.run([sys.executable, "-c", _COPY_PROBE.format(nbytes=nbytes, reps=reps)],
                           capture_output=True, text=True, timeout=300)
        return float(r.stdout.strip().splitlines()[-1]) if r.returncode == 0 else None
    except (subprocess.SubprocessError, ValueError, IndexError):
        return None


def traffic_from_profiles(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary, if present."""
    path = os.path.join(ROOT, "profiles", "c2_2048_fp64_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    found, in_table = None, False
    for line in open(path):  # the table whose last column is "HBM bytes per launch"
        if line.startswith("| kernel |"):
            in_table = line.rstrip().rstrip("|").strip().endswith("HBM bytes per launch")
        elif in_table and line.startswith(f"| {kernel} |"):
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            try:
                found = float(cells[-1])
            except ValueError:
                pass
    return found


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-3d", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-viscous", action="store_true")
    ap.add_argument("--no-sod", action="store_true")
    # the occupancy model's tiling without the autotuner's trial launches (for profiler runs, whose
    # per-kernel averages the trials would otherwise enter)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--tgv-steps", type=int, default=100)  # timed steps of the 3-D leg
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (the library has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)
    prec = capi.FP64 if args.precision == "fp64" else capi.FP32_RESIDUAL
    # first, while this process has made no HIP call: the child is then not a fork of a process
    # holding the device
    copy_gbs = copy_bandwidth()
    cfg, elapsed, info, kern, state, (t_set, t_get), tuning = gpu_run(args.n, args.steps, args.warmup, prec,
                                                                      not args.no_autotune)
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
                     "algorithmic_bytes_per_launch": d["algorithmic_bytes"], "avg_launch_us": d["avg_us"],
                     "measured_copy_gbs": copy_gbs,
                     "frac_of_measured_copy": d["achieved_gbs"] / copy_gbs if copy_gbs else None},
        "step_roofline_frac": value * bytes_step / HBM_PEAK,
        "algorithmic_bytes_per_cell_step": bytes_step,
        "kernels": kern,
        "autotune": tuning["autotune"],
        "launch_forms": tuning["launch_forms"],
        # the same K steps with the host-buffer hand-over of the boundary included (AoS state with
        # ghosts in via ts_set_state, out via ts_get_state: layout conversion + PCIe copies)
        "pcie_inclusive_cell_updates_per_s": ncell * args.steps / (t_set + elapsed + t_get),
        "set_state_s": t_set,
        "get_state_s": t_get,
        "final_time": info.time,
        "finite": bool(np.isfinite(state).all()),
    }
    if args.precision == "fp64":
        # the same run with FP32 residual arithmetic (FP64 state): rate and density error
        _, el32, _, _, st32, _, _ = gpu_run(args.n, args.steps, args.warmup, capi.FP32_RESIDUAL,
                                            not args.no_autotune)
        d = np.abs(capi.interior(cfg, st32)[..., 0] - capi.interior(cfg, state)[..., 0])
        out["fp32_residual"] = {"cell_updates_per_s": ncell * args.steps / el32,
                                "density_l1_vs_fp64": float(d.mean()), "density_linf_vs_fp64": float(d.max())}
    if not args.no_3d:
        out["tgv_3d"] = tgv_3d(steps=args.tgv_steps)
    if not args.no_viscous:
        out["viscous_channel"] = viscous_channel()
    if not args.no_sod:
        out["sod_c1"] = sod_c1()
    if not args.no_batch:
        out["batched"] = batched_tubes()
    if not args.no_serve:
        out["serving"] = serving()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
