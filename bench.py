#!/usr/bin/env python3
"""Benchmark of the MI355X compact-Helmholtz direct solver.

Workload (the configuration the headline metric is quoted on): 2D compact
Helmholtz on the unit square, Nx = Ny = 1023, k = 50, Dirichlet, complex
FP64, 4th- and 6th-order schemes, 64 right-hand sides per order built by the
GPU right-hand-side assembly from manufactured solutions
u = e^{i phi} sin(a pi x) sin(b pi y), a, b ~ U{1..20}, phi ~ U[0, 2 pi)
(seeded).  One step = one batched solve of the 64 right-hand sides for each
order (2 x 64 x 1023^2 complex DOF), inputs already resident in HBM.

Prints one JSON line: DOF/s, p50/p99 batched-solve latency, the relative
error of sampled results against the CPU FP64 oracle, the roofline of the
dominant kernel (HIP events on the launch stream, algorithmic bytes) and the
CPU baseline (the oracle's restatement of the reference solver() timed on
this host's cores).

    python bench.py [--gpus 1] [--steps 10] [--warmup 2]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
CATS = ["dst_fft", "dst_dense", "tridiagonal", "residual_update", "rhs"]


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--N", type=int, default=1023)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--k", type=float, default=50.0)
    ap.add_argument("--orders", default="4,6")
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="budget of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true", help="skip the oracle comparison (profiling runs)")
    ap.add_argument("--no-sections", action="store_true", help="headline only: no C1 / C3 sections")
    ap.add_argument("--check-rhs", type=int, default=2, help="right-hand sides per order compared with the oracle")
    ap.add_argument("--workload", default="c2", choices=["c2", "3d", "uneven", "serve"],
                    help="c2: the headline 2D batch (default); 3d: 255^3 FP64 and mixed precision; uneven: 512 "
                         "uneven 2D problems in one HIP graph; serve: Poisson requests through the server")
    ap.add_argument("--qps", type=float, default=100.0, help="serve: mean request rate")
    ap.add_argument("--requests", type=int, default=300, help="serve: requests in the timed stream")
    ap.add_argument("--lanes", type=int, default=4, help="serve: server lanes (streams)")
    ap.add_argument("--problems", type=int, default=512, help="uneven: number of problems")
    ap.add_argument("--sweeps", type=int, default=2, help="uneven: fixed refinement sweeps")
    ap.add_argument("--by-order", action="store_true",
                    help="uneven: one plan per (N, order) group instead of one per N (orders 4 and 6 of one N "
                         "share a multi-operator plan by default)")
    ap.add_argument("--tune", action="store_true",
                    help="uneven: autotune each group's tridiagonal shape and spectrum layout before capture, "
                         "timing each candidate as a graph replay of the group's solve")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="after the run, copy MIB MiB three times with an elementwise kernel (y = x * 1), the "
                         "known-size reference for scaling HBM counters")
    return ap.parse_args()


def manufactured(torch, N, B, kk, seed, dev):
    """f, u and their boundary data on the extended grid for B sine solutions."""
    rng = np.random.default_rng(seed)
    a = rng.integers(1, 21, size=B)
    b = rng.integers(1, 21, size=B)
    phi = rng.uniform(0.0, 2.0 * np.pi, size=B)
    h = 1.0 / (N + 1)
    x = torch.arange(N + 2, dtype=torch.float64, device=dev) * h
    A = torch.tensor(a, dtype=torch.float64, device=dev)[:, None]
    Bb = torch.tensor(b, dtype=torch.float64, device=dev)[:, None]
    sx = torch.sin(A * math.pi * x[None, :])           # (B, N+2)
    sy = torch.sin(Bb * math.pi * x[None, :])
    amp = torch.tensor(np.exp(1j * phi), dtype=torch.complex128, device=dev)
    fac = torch.tensor(kk * kk - np.pi ** 2 * (a ** 2 + b ** 2), dtype=torch.complex128, device=dev) * amp
    u_e = (amp[:, None, None] * (sy[:, :, None] * sx[:, None, :])).contiguous()     # [b][j][i]
    f_e = (fac[:, None, None] * (sy[:, :, None] * sx[:, None, :])).contiguous()
    cy = torch.cos(Bb * math.pi * x[None, :])
    fy_lo = fac[:, None] * sx[:, 1:-1] * (Bb * math.pi) * cy[:, :1]
    fy_hi = fac[:, None] * sx[:, 1:-1] * (Bb * math.pi) * cy[:, -1:]
    fy_b = torch.cat([fy_lo, fy_hi], dim=1).contiguous()
    ax = (A * math.pi) ** 2
    by = (Bb * math.pi) ** 2

    def ring(g, s):
        return torch.cat([-s * g[:, 0, :], -s * g[:, -1, :], -s * g[:, :, 0], -s * g[:, :, -1]], dim=1).contiguous()

    fxx_b = ring(f_e, ax.to(torch.complex128))
    fyy_b = ring(f_e, by.to(torch.complex128))
    return dict(a=a, b=b, phi=phi, f_e=f_e, u_e=u_e, fy_b=fy_b, fxx_b=fxx_b, fyy_b=fyy_b)


def host_threads():
    """Cores this process may use (capped at 16, the share of one GPU's host)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return max(1, min(n, 16))


def host_cpus():
    return {"online": os.cpu_count(), "usable": host_threads()}


def cpu_baseline(O, N, coef, F_host, seconds, threads=None):
    """Oracle restatement of the reference 2D solver() on this host's cores
    (threads=1: serial, as the reference runs without -DUSE_OMP)."""
    threads = host_threads() if threads is None else threads
    LUU = O.lu2d(N, N, coef, threads)
    done, t_used = 0, 0.0
    while t_used < seconds:   # cycle over the right-hand sides until the budget is spent
        t0 = time.perf_counter()
        O.solver2d(N, N, F_host[done % F_host.shape[0]], LUU, threads)
        t_used += time.perf_counter() - t0
        done += 1
    return {"value": done * N * N / t_used, "unit": "complex DOF/s", "cores": threads, "kind": "port",
            "sample": f"oracle solver() (own FFT DST, plan creation inside the call, "
                      f"{'serial' if threads == 1 else 'OpenMP over rows/modes'}) on "
                      f"{done} 6th-order right-hand sides of the same N={N} k=50 workload, {t_used:.1f} s"}


def cpu_steady(O, N, coef, F_host, seconds, threads=None):
    """The same solver() with its DST plan created once outside the timed
    calls (steady state, plans reused)."""
    threads = host_threads() if threads is None else threads
    LUU = O.lu2d(N, N, coef, threads)
    plan = O.DstPlan(N)
    done, t_used = 0, 0.0
    while t_used < seconds:
        t0 = time.perf_counter()
        O.solver2d_planned(plan, N, N, F_host[done % F_host.shape[0]], LUU, threads)
        t_used += time.perf_counter() - t0
        done += 1
    return {"value": done * N * N / t_used, "unit": "complex DOF/s", "cores": threads, "kind": "port",
            "sample": f"oracle solver() with its DST plan reused across calls (steady state), "
                      f"{'serial' if threads == 1 else 'OpenMP over rows/modes'}, {done} right-hand sides of "
                      f"the same N={N} workload, {t_used:.1f} s"}


def _events(torch, n):
    return [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(n)]


def run_c1(torch, H, O, args):
    """C1: 2D order 6, N = 100 (N+1 = 101: the dense DST path), k = 10,
    u = e^{i pi/4} sin(3 pi x) sin(2 pi y) (the configuration's (1+i) up to the
    factor sqrt 2, which changes no relative figure).  GPU latency of one
    solve, the oracle check, and the reference's own CPU configuration
    (./Direct_Solver 6 100 2: two OpenMP threads) as it runs (DST plans
    created inside the call) and in steady state."""
    N, k, order = 100, 10.0, 6
    s = O.solve2d(order, N, O.SINE, [k, 0.0, 3, 2, math.pi / 4], threads=2)
    ctx = H.Context(0)
    plan = H.Plan2D(ctx, N, N, s["coef"])
    rhs = torch.from_numpy(s["rhs"]).cuda()
    out = torch.empty_like(rhs)
    for _ in range(5):
        plan.solve(rhs, out=out)
    torch.cuda.synchronize()
    ev = _events(torch, 50)
    t0 = time.perf_counter()
    for a, b in ev:
        a.record()
        plan.solve(rhs, out=out)
        b.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    lat = [a.elapsed_time(b) for a, b in ev]
    g = out.cpu().numpy()
    # the same solve recorded once into a HIP graph and replayed: the launch
    # overhead of the solve's kernels is paid at capture, not per solve
    gout = torch.empty_like(rhs)
    graph = H.Graph(ctx, [(plan, rhs, gout)])
    try:
        for _ in range(5):
            graph.launch()
        torch.cuda.synchronize()
        gev = _events(torch, 50)
        for a, b in gev:
            a.record()
            graph.launch()
            b.record()
        torch.cuda.synchronize()
        glat = [a.elapsed_time(b) for a, b in gev]
        graph_same = bool(np.array_equal(gout.cpu().numpy(), g))
    finally:
        graph.close()
    res = {"workload": "2D order 6, N=100, k=10, u=e^{i pi/4} sin(3 pi x) sin(2 pi y), one right-hand side",
           "gpu_dof_per_s": N * N * len(ev) / wall, "p50_ms": float(np.percentile(lat, 50)),
           "p99_ms": float(np.percentile(lat, 99)),
           "graph_p50_ms": float(np.percentile(glat, 50)), "graph_p99_ms": float(np.percentile(glat, 99)),
           "graph_same_bits": graph_same,
           "max_rel_err_vs_oracle": float(np.abs(g - s["sol"]).max() / np.abs(s["sol"]).max()),
           "disc_err_vs_exact": float(np.abs(g - s["u"]).max() / np.abs(s["u"]).max())}
    if not args.no_cpu:
        F = s["rhs"][None, :]
        cpu = cpu_baseline(O, N, s["coef"], F, 2.0, threads=2)
        cpu["sample"] = cpu["sample"].replace(f"6th-order right-hand sides of the same N={N} k=50 workload",
                                              "solves of the C1 right-hand side")
        res["cpu_baseline_2threads"] = cpu
        res["cpu_baseline_2threads_steady"] = cpu_steady(O, N, s["coef"], F, 2.0, threads=2)
    return res


def run_c3_brief(torch, H, O, args):
    """C3: 3D order 6, 255^3, k = 30, u = (1+i) sin(a pi x) sin(b pi y) sin(c pi z)
    with (a, b, c) drawn from the seed: FP64 and mixed precision (two fixed
    sweeps) on the GPU, the oracle check, and the reference's 3D CPU solver
    (serial, as it ships) beside the OpenMP restatement."""
    N, k, order = 255, 30.0, 6
    rng = np.random.default_rng((args.seed, 3))
    a, b, c = (int(v) for v in rng.integers(1, 21, size=3))
    s = O.setup3d(order, N, N, N, O.SINE, [k, 0.0, a, b, c, 1.0, 1.0], threads=host_threads())
    LUU = O.lu3d(N, N, N, s["coef"])
    ref = O.solver3d(N, N, N, s["rhs"], LUU, host_threads())
    ctx = H.Context(0)
    rhs = torch.from_numpy(s["rhs"]).cuda()
    out = torch.empty_like(rhs)
    res = {"workload": f"3D order 6, N=255^3, k=30, u=(1+i) sin(a pi x) sin(b pi y) sin(c pi z), "
                       f"(a,b,c)=({a},{b},{c}) from seed ({args.seed}, 3)"}
    for name, prec in (("fp64", H.PREC_F64), ("mixed_2_sweeps", H.PREC_MIXED)):
        plan = H.Plan3D(ctx, N, N, N, s["coef"].reshape(12, N), precision=prec)
        if prec == H.PREC_MIXED:
            plan.set_refine(0.0, 2)
        for _ in range(2):
            plan.solve(rhs, out=out)
        torch.cuda.synchronize()
        ev = _events(torch, 10)
        t0 = time.perf_counter()
        for e0, e1 in ev:
            e0.record()
            plan.solve(rhs, out=out)
            e1.record()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        lat = [e0.elapsed_time(e1) for e0, e1 in ev]
        g = out.cpu().numpy()
        res[name] = {"dof_per_s": N ** 3 * len(ev) / wall, "p50_ms": float(np.percentile(lat, 50)),
                     "max_rel_err_vs_oracle": float(np.abs(g - ref).max() / np.abs(ref).max())}
        plan.close()
    res["mixed_over_fp64"] = res["mixed_2_sweeps"]["dof_per_s"] / res["fp64"]["dof_per_s"]
    if not args.no_cpu:
        def sample(nthreads, seconds):
            done, t_used = 0, 0.0
            while t_used < seconds:
                t0 = time.perf_counter()
                O.solver3d(N, N, N, s["rhs"], LUU, nthreads)
                t_used += time.perf_counter() - t0
                done += 1
            how = "serial, as the reference's 3D CPU solver ships" if nthreads == 1 else "OpenMP over layers/modes"
            return {"value": done * N ** 3 / t_used, "unit": "complex DOF/s", "cores": nthreads, "kind": "port",
                    "sample": f"oracle 3D solver() ({how}), {done} solves of the C3 right-hand side, {t_used:.1f} s"}
        res["cpu_baseline"] = sample(host_threads(), 3.0)
        res["cpu_baseline_serial"] = sample(1, 1.0)
    return res


# ---------------------------------------------------------------- other workloads
def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _dof_line(metric_extra, value, unit, steps, warmup, ms, config, **kw):
    out = {"metric": "complex DOF/s at <=1e-10 rel err vs FP64 oracle; p50/p99 solve latency" + metric_extra,
           "value": value, "unit": unit, "n_gpus": 1, "steps": steps, "warmup": warmup, "ms_per_step": ms,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": kw.pop("dtype", "f64"),
           "data": "synthetic", "config": config}
    out.update(kw)
    return out


def _timing(H, ctx):
    import ctypes
    lib = H.lib()
    ms = (ctypes.c_double * 5)()
    nl = (ctypes.c_long * 5)()
    by = (ctypes.c_double * 5)()
    lib.hh_context_read_timing(ctx.h, ms, nl, by)
    return {CATS[c]: {"ms": round(ms[c], 4), "launches": int(nl[c]),
                      "avg_us": round(1e3 * ms[c] / nl[c], 2) if nl[c] else None,
                      "GBps": round(by[c] / (ms[c] * 1e-3) / 1e9, 1) if ms[c] > 0 else None}
            for c in range(5) if nl[c]}


def manufactured3d(torch, N, kk, abc, dev, amp=1 + 1j):
    """hz^2 f and u on the extended grid for u = amp sin(a pi x) sin(b pi y) sin(c pi z)."""
    h = 1.0 / (N + 1)
    x = torch.arange(N + 2, dtype=torch.float64, device=dev) * h
    a, b, c = abc
    sx, sy, sz = (torch.sin(m * math.pi * x) for m in (a, b, c))
    u = (sz[:, None, None] * sy[None, :, None] * sx[None, None, :]).to(torch.complex128) * amp
    fac = kk * kk - math.pi ** 2 * (a * a + b * b + c * c)
    return (h * h * fac) * u, u


def run_3d(args, torch, H):
    """3D order 6, N^3 = 255^3, k = 30: plain FP64 against FP32 bulk + FP64 refinement, adaptive
    (stopping on the residual) and with two fixed sweeps."""
    dev = torch.device("cuda:0")
    N, kk, order = 255 if args.N == 1023 else args.N, 30.0, 6
    h = 1.0 / (N + 1)
    rng = np.random.default_rng((args.seed, 3))
    abc = [int(v) for v in rng.integers(1, 21, size=3)]
    ctx = H.Context(0)
    k2e = np.full(N + 2, kk * kk, np.complex128)
    coef = H.coefficients3d(order, N, N, N, h, h, h, k2e)
    p64 = H.Plan3D(ctx, N, N, N, coef)
    pmx = H.Plan3D(ctx, N, N, N, coef, precision=H.PREC_MIXED)
    pmx.set_refine(1e-12, 6)            # adaptive: stops on the residual (a host sync per sweep)
    pfx = H.Plan3D(ctx, N, N, N, coef, precision=H.PREC_MIXED)
    pfx.set_refine(0.0, 2)              # two fixed sweeps, no host sync (graph-capturable)
    f_e, u_e = manufactured3d(torch, N, kk, abc, dev)
    zf = torch.zeros(2 * (N + 2) * (N + 2), dtype=torch.complex128, device=dev)
    F = p64.rhs(order, h, h, h, f_e, u_e, k2e, fxx_f=zf, fyy_f=zf, fzz_f=zf)
    u_exact = u_e[1:-1, 1:-1, 1:-1].contiguous()
    del f_e, u_e
    s64, smx, sfx = torch.empty_like(F), torch.empty_like(F), torch.empty_like(F)
    res = {}
    log(f"3d: N={N} (a,b,c)={abc} right-hand side assembled")
    paths = (("fp64", p64, s64), ("mixed", pmx, smx), ("mixed_fixed", pfx, sfx))
    for name, plan, out in paths:
        log(f"3d: timing {name}")
        for _ in range(args.warmup):
            plan.solve(F, out=out)
        torch.cuda.synchronize()
        lib = H.lib()
        lib.hh_context_reset_timing(ctx.h)
        lib.hh_context_enable_timing(ctx.h, 1)
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
        t0 = time.perf_counter()
        for a, b in ev:
            a.record()
            plan.solve(F, out=out)
            b.record()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        lib.hh_context_enable_timing(ctx.h, 0)
        lat = [a.elapsed_time(b) for a, b in ev]
        res[name] = {"dof_per_s": N ** 3 * args.steps / wall, "ms_per_solve": 1e3 * wall / args.steps,
                     "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                     "kernels": _timing(H, ctx)}
    sw, rel = pmx.last_refine()
    res["mixed"].update(sweeps=sw, rel_residual=rel)
    res["mixed_fixed"].update(sweeps=2)
    # exact solution check on the GPU (norms kernel), oracle check on the host
    for name, plan, out in paths:
        m, d = ctx.norms(out, u_exact)
        res[name]["disc_err_vs_exact"] = d / max(m, 1e-300)
    cpu = cpu_serial = None
    if not (args.no_check and args.no_cpu):
        from oracle import oracle as O
        log("3d: oracle solve on the host")
        Fh = F.cpu().numpy().ravel()
        threads = host_threads()
        t0 = time.perf_counter()
        LUU = O.lu3d(N, N, N, np.concatenate(coef))
        t_lu = time.perf_counter() - t0
        ref = O.solver3d(N, N, N, Fh, LUU, threads)
        for name, plan, out in paths:
            g = out.cpu().numpy().ravel()
            res[name]["max_rel_err_vs_oracle"] = float(np.abs(g - ref).max() / np.abs(ref).max())
        if not args.no_cpu:
            def sample(nthreads, seconds):
                done, t_used = 0, 0.0
                while t_used < seconds:   # solver() only, as the reference times it
                    t0 = time.perf_counter()
                    O.solver3d(N, N, N, Fh, LUU, nthreads)
                    t_used += time.perf_counter() - t0
                    done += 1
                how = "serial, as the reference's 3D CPU solver" if nthreads == 1 else "OpenMP over layers/modes"
                return {"value": done * N ** 3 / t_used, "unit": "complex DOF/s", "cores": nthreads, "kind": "port",
                        "sample": f"oracle solver() of the 3D reference CPU path ({how}), {done} solves of the same "
                                  f"{N}^3 k=30 right-hand side, {t_used:.1f} s (factorisation LU() {t_lu:.1f} s, "
                                  f"untimed like the reference)"}
            cpu = sample(threads, args.cpu_seconds)
            cpu_serial = sample(1, args.cpu_seconds / 2)
    names = [p[0] for p in paths]
    ok = {k: res[k].get("max_rel_err_vs_oracle", 0.0) <= 1e-10 for k in names}
    best = max((k for k in names if ok[k]), key=lambda k: res[k]["dof_per_s"], default="fp64")
    best_mixed = max((k for k in names[1:] if ok[k]), key=lambda k: res[k]["dof_per_s"], default="mixed")
    return _dof_line(" (3D)", res[best]["dof_per_s"], "complex DOF/s", args.steps, args.warmup,
                     res[best]["ms_per_solve"],
                     {"workload": f"3D compact Helmholtz N={N}^3, k=30, order 6, u=(1+i) sin(a pi x) sin(b pi y) "
                                  f"sin(c pi z), (a,b,c)={tuple(abc)} from seed ({args.seed}, 3), GPU RHS assembly",
                      "grid": [N, N, N], "orders": [order]},
                     dtype="f32+f64" if best != "fp64" else "f64", value_path=best,
                     p50_ms=res[best]["p50_ms"], p99_ms=res[best]["p99_ms"],
                     mixed_over_fp64=res[best_mixed]["dof_per_s"] / res["fp64"]["dof_per_s"], results=res,
                     cpu_baseline=cpu, cpu_baseline_serial=cpu_serial, host_cpus=host_cpus())


def run_uneven(args, torch, H):
    """512 uneven 2D problems (N log-uniform in 63..2047, k in [1,80], order 4/6), FP32 bulk + FP64
    refinement with fixed sweeps, grouped by N into multi-operator plans (one operator per problem, so
    both orders share a plan; --by-order: one plan per (N, order)) and replayed as one HIP graph."""
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(args.seed)
    Ns = np.array([63, 127, 255, 511, 1023, 2047])
    P = args.problems
    Nsel = rng.choice(Ns, size=P)
    ksel = rng.uniform(1.0, 80.0, size=P)
    osel = rng.choice([4, 6], size=P)
    ctx = H.Context(0)
    groups = {}
    for i in range(P):
        groups.setdefault((int(Nsel[i]), int(osel[i])), []).append(i)
    solves64, solvesmx, info = [], [], []
    total_dof = 0
    for (N, order), idx in sorted(groups.items()):
        h = 1.0 / (N + 1)
        ks = ksel[idx]
        k2es = [np.full(N + 2, k * k, np.complex128) for k in ks]
        ops = [H.coefficients2d(order, N, N, h, h, k2e) for k2e in k2es]
        p64 = H.Plan2D(ctx, N, N, ops)
        pmx = H.Plan2D(ctx, N, N, ops, precision=H.PREC_MIXED)
        pmx.set_refine(0.0, args.sweeps)
        B = len(idx)
        m = manufactured(torch, N, B, 0.0, args.seed + N + order, dev)
        # per-problem k: f = (k^2 - pi^2 (a^2 + b^2)) u, rebuilt from u
        fac = torch.tensor(ks ** 2 - np.pi ** 2 * (m["a"] ** 2 + m["b"] ** 2), dtype=torch.complex128, device=dev)
        f_e = m["u_e"] * fac[:, None, None]
        x = torch.arange(N + 2, dtype=torch.float64, device=dev) * h
        A = torch.tensor(m["a"], dtype=torch.float64, device=dev)[:, None]
        Bb = torch.tensor(m["b"], dtype=torch.float64, device=dev)[:, None]
        amp = torch.tensor(np.exp(1j * m["phi"]), dtype=torch.complex128, device=dev)
        sx = torch.sin(A * math.pi * x[None, :])
        cy = torch.cos(Bb * math.pi * x[None, :])
        fy_b = torch.cat([(fac * amp)[:, None] * sx[:, 1:-1] * (Bb * math.pi) * cy[:, :1],
                          (fac * amp)[:, None] * sx[:, 1:-1] * (Bb * math.pi) * cy[:, -1:]], dim=1).contiguous()

        def ring(g, s):
            return torch.cat([-s * g[:, 0, :], -s * g[:, -1, :], -s * g[:, :, 0], -s * g[:, :, -1]],
                             dim=1).contiguous()

        fxx_b = ring(f_e, ((A * math.pi) ** 2).to(torch.complex128))
        fyy_b = ring(f_e, ((Bb * math.pi) ** 2).to(torch.complex128))
        F = p64.rhs(order, h, h, f_e.contiguous(), m["u_e"], np.concatenate(k2es), fy_b=fy_b, fxx_b=fxx_b,
                    fyy_b=fyy_b, k2yb=np.zeros(2 * B))
        u_ex = m["u_e"][:, 1:-1, 1:-1].contiguous()
        del m, f_e
        solves64.append((p64, F, torch.empty_like(F)))
        solvesmx.append((pmx, F, torch.empty_like(F)))
        shapes = {}
        if args.tune and args.by_order:
            # the autotuner picks each group's tridiagonal shape and spectrum layout, timing every
            # candidate (the built-in one included) as a graph replay, before capture
            from hackathon_amd import autotune
            t_tune = time.perf_counter()
            for name, plan, buf in (("fp64", p64, solves64[-1][2]), ("mixed", pmx, solvesmx[-1][2])):
                best = autotune.tune(plan, F, buf, reps=3, tiles=(0, -1), graph=True)
                shapes[name] = {"W": best["W"], "rows": best["rows"], "rhs": best["rhs"], "tile": best["tile"],
                                "us": round(best["us"], 1)}
            log(f"uneven: tuned N={N} order={order} in {time.perf_counter() - t_tune:.1f} s: {shapes}")
        info.append({"N": N, "order": order, "problems": B, "k": ks, "ops": ops, "u": u_ex, "shapes": shapes,
                     "solve": len(solves64) - 1, "off": 0})
        total_dof += B * N * N
        log(f"uneven: group N={N} order={order}: {B} problems")
    if not args.by_order:
        # both orders of one N in one multi-operator plan: half the launches, twice the batch
        byN = {}
        for gi in info:
            byN.setdefault(gi["N"], []).append(gi)
        m64, mmx, uu = [], [], []
        for N, gl in sorted(byN.items()):
            ops = [op for gi in gl for op in gi["ops"]]
            F = torch.cat([solves64[gi["solve"]][1] for gi in gl])
            off = 0
            for gi in gl:
                gi["solve"], gi["off"] = len(m64), off
                off += gi["problems"]
            p64 = H.Plan2D(ctx, N, N, ops)
            pmx = H.Plan2D(ctx, N, N, ops, precision=H.PREC_MIXED)
            pmx.set_refine(0.0, args.sweeps)
            m64.append((p64, F, torch.empty_like(F)))
            mmx.append((pmx, F, torch.empty_like(F)))
            uu.append(torch.cat([gi["u"] for gi in gl]))
        solves64, solvesmx = m64, mmx
        exact = uu
        if args.tune:   # each merged plan's shape and layout, timed as graph replays, before capture
            from hackathon_amd import autotune
            for (p64, F, b64), (pmx, _, bmx), (N, gl) in zip(m64, mmx, sorted(byN.items())):
                t_tune = time.perf_counter()
                shapes = {}
                for name, plan, buf in (("fp64", p64, b64), ("mixed", pmx, bmx)):
                    best = autotune.tune(plan, F, buf, reps=3, tiles=(0, -1), graph=True)
                    shapes[name] = {"W": best["W"], "rows": best["rows"], "rhs": best["rhs"], "tile": best["tile"],
                                    "us": round(best["us"], 1)}
                for gi in gl:
                    gi["shapes"] = shapes
                log(f"uneven: tuned N={N} in {time.perf_counter() - t_tune:.1f} s: {shapes}")
    else:
        exact = [gi["u"] for gi in info]
    torch.cuda.synchronize()
    out = {}
    for name, solves in (("fp64", solves64), ("mixed", solvesmx)):
        log(f"uneven: recording and timing the {name} graph")
        g = H.Graph(ctx, solves)
        for _ in range(args.warmup):
            g.launch()
        torch.cuda.synchronize()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
        t0 = time.perf_counter()
        for a, b in ev:
            a.record()
            g.launch()
            b.record()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        lat = [a.elapsed_time(b) for a, b in ev]
        out[name] = {"dof_per_s": total_dof * args.steps / wall, "ms_per_launch": 1e3 * wall / args.steps,
                     "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                     "graph_nodes": g.num_nodes}
        # discretisation error of every problem against its exact solution
        worst = 0.0
        for (plan, F, sol), u in zip(solves, exact):
            mx, d = ctx.norms(sol, u)
            worst = max(worst, d / mx)
        out[name]["max_disc_err_vs_exact"] = worst
        g.close()
    check = {"fp64": {}, "mixed": {}}
    if not args.no_check:
        from oracle import oracle as O
        log("uneven: oracle checks")
        for gi in info:
            if gi["N"] > 1023:
                continue
            N, i, o = gi["N"], gi["solve"], gi["off"]
            F, sol64, solmx = solves64[i][1], solves64[i][2], solvesmx[i][2]
            LUU = O.lu2d(N, N, gi["ops"][0], 16)
            ref = O.solver2d(N, N, F[o].cpu().numpy().ravel(), LUU, 16)
            for name, sol in (("fp64", sol64), ("mixed", solmx)):
                g = sol[o].cpu().numpy().ravel()
                check[name][f"N{N}_order{gi['order']}"] = float(np.abs(g - ref).max() / np.abs(ref).max())
        for name in ("fp64", "mixed"):
            out[name]["max_rel_err_vs_oracle"] = max(check[name].values()) if check[name] else None
    ok = out["mixed"].get("max_rel_err_vs_oracle") is not None and out["mixed"]["max_rel_err_vs_oracle"] <= 1e-10
    best = "mixed" if ok and out["mixed"]["dof_per_s"] >= out["fp64"]["dof_per_s"] else "fp64"
    sizes = {str(k): len(v) for k, v in sorted(groups.items())}
    return _dof_line(" (uneven batch)", out[best]["dof_per_s"], "complex DOF/s", args.steps, args.warmup,
                     out[best]["ms_per_launch"],
                     {"workload": f"{P} uneven 2D problems, N log-uniform in {{63..2047}}, k~U[1,80], order~U{{4,6}}, "
                                  f"manufactured sines (seed {args.seed}); one HIP graph launch per step, "
                                  f"{'one plan per (N, order)' if args.by_order else 'one multi-operator plan per N'}",
                      "problems": P, "total_dof": total_dof, "groups": sizes,
                      "tridiagonal_shapes": "built-in" if not args.tune else
                      {f"N{gi['N']}_order{gi['order']}": gi["shapes"] for gi in info}},
                     dtype="f32+f64" if best == "mixed" else "f64", value_path=best,
                     mixed_over_fp64=out["mixed"]["dof_per_s"] / out["fp64"]["dof_per_s"], results=out,
                     oracle_check=check)


def run_serve(args, torch, H):
    """Poisson request stream of mixed 2D/3D solves through the continuous-batching server."""
    import threading
    rng = np.random.default_rng(args.seed)
    shapes2 = [127, 255, 511, 1023, 2047]
    shapes3 = [63, 127, 255]
    # right-hand sides and solutions live in pinned host memory (as a client
    # streaming to the GPU would keep them), so the copies run at PCIe speed
    def pinned(n):
        return torch.empty(n, dtype=torch.complex128, pin_memory=True).numpy()

    pools = {}
    for d, Ns in ((2, shapes2), (3, shapes3)):
        for N in Ns:
            n = N ** d
            pools[(d, N)] = []
            for _ in range(2):
                a = pinned(n)
                a[:] = rng.standard_normal(n) + 1j * rng.standard_normal(n)
                pools[(d, N)].append(a)
    reqs = []
    for _ in range(args.requests):
        if rng.uniform() < 0.7:
            d, N = 2, int(rng.choice(shapes2))
        else:
            d, N = 3, int(rng.choice(shapes3))
        reqs.append((d, N, int(rng.choice([4, 6])), float(rng.uniform(1.0, 60.0)), int(rng.integers(0, 2))))
    gaps = rng.exponential(1.0 / args.qps, size=len(reqs))
    srv = H.Server(0, hbm_budget=96 << 30, lanes=args.lanes, max_batch=8)
    # warm-up: every grid once per lane (plans, tables, code objects)
    warm = [srv.submit(d, N, 4, 10.0, pools[(d, N)][0]) for (d, N) in pools for _ in range(args.lanes)]
    for t in warm:
        srv.wait(t)
    log(f"serve: warm-up of {len(warm)} requests done; streaming {len(reqs)} at {args.qps:g}/s")
    st0 = srv.stats()
    results = [None] * len(reqs)
    tickets = [None] * len(reqs)
    ready = [threading.Event() for _ in reqs]

    keep = set()   # solutions kept for the oracle check; the others return to the pinned pool at once
    for i, (d, N, order, k, which) in enumerate(reqs):
        if d == 2 and N <= 1023 and len(keep) < 8:
            keep.add(i)

    def waiter():
        for i in range(len(reqs)):
            ready[i].wait()
            sol, info = srv.wait(tickets[i])
            results[i] = (sol if i in keep else None, info, time.perf_counter())

    th = threading.Thread(target=waiter)
    th.start()
    t_start = time.perf_counter()
    t_next = t_start
    for i, (d, N, order, k, which) in enumerate(reqs):
        t_next += gaps[i]
        delay = t_next - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        tickets[i] = srv.submit(d, N, order, k, pools[(d, N)][which], out=pinned(N ** d))
        ready[i].set()
    th.join()
    t_end = max(r[2] for r in results)
    wall = t_end - t_start
    st1 = srv.stats()
    lat = np.array([r[1]["latency_s"] for r in results]) * 1e3
    q = np.array([r[1]["queue_s"] for r in results]) * 1e3
    dof = sum(N ** d for d, N, _, _, _ in reqs)
    check = {}
    if not args.no_check:
        from oracle import oracle as O
        seen = set()
        for i, (d, N, order, k, which) in enumerate(reqs):
            if i not in keep or (N, order) in seen or len(seen) >= 4:
                continue
            seen.add((N, order))
            h = 1.0 / (N + 1)
            coef = H.coefficients2d(order, N, N, h, h, np.full(N + 2, k * k, np.complex128))
            ref = O.solver2d(N, N, pools[(d, N)][which], O.lu2d(N, N, coef, 16), 16)
            g = results[i][0].ravel()
            check[f"2D N{N} order{order} k{k:.1f}"] = float(np.abs(g - ref).max() / np.abs(ref).max())
    srv.close()
    return _dof_line(" (serving)", dof / wall, "complex DOF/s", len(reqs), len(warm), 1e3 * wall / len(reqs),
                     {"workload": f"Poisson stream, {args.qps:g} req/s, {len(reqs)} requests: 70% 2D N in "
                                  f"{{127..2047}}, 30% 3D N in {{63,127,255}}, k~U[1,60], order~U{{4,6}}, FP64, "
                                  f"right-hand sides and solutions in pinned host memory (seed {args.seed}); "
                                  f"{args.lanes} lanes, max batch 8",
                      "qps": args.qps, "lanes": args.lanes},
                     p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)),
                     p50_queue_ms=float(np.percentile(q, 50)), p99_queue_ms=float(np.percentile(q, 99)),
                     latency_unit="arrival to solution in host memory (includes PCIe copies)",
                     requests_per_s=len(reqs) / wall, batches=st1["batches"] - st0["batches"],
                     oracle_check=check,
                     max_rel_err_vs_oracle=max(check.values()) if check else None)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); multi-GPU runs are not part of "
              "this build", file=sys.stderr)
        sys.exit(2)
    import torch

    from hackathon_amd import helmholtz as H

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.workload != "c2":
        out = {"3d": run_3d, "uneven": run_uneven, "serve": run_serve}[args.workload](args, torch, H)
        print(json.dumps(out), flush=True)
        return
    dev = torch.device("cuda:0")
    N, B, kk = args.N, args.batch, args.k
    orders = [int(o) for o in args.orders.split(",")]
    h = 1.0 / (N + 1)
    ctx = H.Context(0)
    m = manufactured(torch, N, B, kk, args.seed, dev)
    k2e = np.full(N + 2, kk * kk, np.complex128)
    plans, F, sol = {}, {}, {}
    for o in orders:
        coef = H.coefficients2d(o, N, N, h, h, k2e)
        plans[o] = H.Plan2D(ctx, N, N, coef)
        F[o] = plans[o].rhs(o, h, h, m["f_e"], m["u_e"], k2e, fy_b=m["fy_b"], fxx_b=m["fxx_b"],
                            fyy_b=m["fyy_b"], k2yb=np.zeros(2))
        sol[o] = torch.empty_like(F[o])
    del m["f_e"]
    torch.cuda.synchronize()

    def step(ev=None):
        for i, o in enumerate(orders):
            if ev is not None:
                ev[2 * i].record()
            plans[o].solve(F[o], out=sol[o])
            if ev is not None:
                ev[2 * i + 1].record()

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    lib = H.lib()
    lib.hh_context_reset_timing(ctx.h)
    lib.hh_context_enable_timing(ctx.h, 1)
    events = [[torch.cuda.Event(enable_timing=True) for _ in range(2 * len(orders))] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        step(events[s])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    lib.hh_context_enable_timing(ctx.h, 0)
    import ctypes
    ms = (ctypes.c_double * 5)()
    nl = (ctypes.c_long * 5)()
    by = (ctypes.c_double * 5)()
    lib.hh_context_read_timing(ctx.h, ms, nl, by)
    lat = [events[s][2 * i].elapsed_time(events[s][2 * i + 1]) for s in range(args.steps) for i in range(len(orders))]

    dof_step = len(orders) * B * N * N
    value = dof_step * args.steps / wall
    # dominant kernel of the solve
    solve_cats = [0, 1, 2]
    dom = max(solve_cats, key=lambda c: ms[c])
    achieved = by[dom] / (ms[dom] * 1e-3) / 1e9 if ms[dom] > 0 else 0.0
    kernels = {CATS[c]: {"ms": round(ms[c], 4), "launches": int(nl[c]),
                         "avg_us": round(1e3 * ms[c] / nl[c], 2) if nl[c] else None,
                         "GBps": round(by[c] / (ms[c] * 1e-3) / 1e9, 1) if ms[c] > 0 else None}
               for c in range(5) if nl[c]}
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if os.path.exists(tfile):
        try:
            t = json.load(open(tfile))
            traffic = t.get(CATS[dom])
        except Exception:
            traffic = None

    # correctness: sampled results against the CPU FP64 oracle
    check = {}
    cpu = None
    cpu_serial = None
    cpu_steady16 = None
    if not (args.no_check and args.no_cpu):
        from oracle import oracle as O
        for o in orders:
            coef = plans[o].coef
            Fh = F[o][: max(args.check_rhs, 1)].cpu().numpy().reshape(-1, N * N)
            if not args.no_check:
                LUU = O.lu2d(N, N, coef, 16)
                errs, derr = [], []
                for r in range(min(args.check_rhs, B)):
                    ref = O.solver2d(N, N, Fh[r], LUU, 16)
                    g = sol[o][r].cpu().numpy().ravel()
                    errs.append(float(np.abs(g - ref).max() / np.abs(ref).max()))
                    u = m["u_e"][r, 1:-1, 1:-1].cpu().numpy().ravel()
                    derr.append(float(np.abs(g - u).max() / np.abs(u).max()))
                check[f"order{o}"] = {"max_rel_err_vs_oracle": max(errs), "disc_err_vs_exact": max(derr),
                                      "rhs_checked": len(errs)}
                # white noise puts energy in the nearly singular low modes,
                # which the manufactured sines leave out
                wn = torch.randn(N * N, dtype=torch.complex128, device=dev,
                                 generator=torch.Generator(device=dev).manual_seed(args.seed + o))
                wsol = plans[o].solve(wn).cpu().numpy()
                wref = O.solver2d(N, N, wn.cpu().numpy(), LUU, 16)
                check[f"order{o}"]["white_noise_rel_err_vs_oracle"] = float(
                    np.abs(wsol - wref).max() / np.abs(wref).max())
                check[f"order{o}"]["near_resonant_modes"] = plans[o].mode_info()[0]
            if not args.no_cpu and o == orders[-1]:
                Fall = F[o].cpu().numpy().reshape(B, N * N)
                cpu = cpu_baseline(O, N, coef, Fall, args.cpu_seconds)
                cpu_serial = cpu_baseline(O, N, coef, Fall, args.cpu_seconds / 2, threads=1)
                cpu_steady16 = cpu_steady(O, N, coef, Fall, args.cpu_seconds / 2)

    out = {
        "metric": "complex DOF/s at <=1e-10 rel err vs FP64 oracle; p50/p99 solve latency",
        "value": value,
        "unit": "complex DOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"2D compact Helmholtz N={N}x{N}, k={kk:g}, orders {args.orders}, "
                               f"{B} manufactured-sine right-hand sides per order (seed {args.seed}), "
                               "plain FP64 path, batched solve",
                   "global_batch": B * len(orders), "grid": [N, N], "orders": orders, "precision": "fp64"},
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "latency_unit": f"one batched solve of {B} right-hand sides",
        "check": check,
        "kernels": kernels,
        "roofline": {"bound": "hbm", "kernel": CATS[dom], "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "bytes_per_dof_per_pass": 32},
        "solve_level_frac": value * 32 / (HBM_PEAK_GBS * 1e9),
        "cpu_baseline": cpu,
        "cpu_baseline_serial": cpu_serial,
        "cpu_baseline_steady": cpu_steady16,
        "host_cpus": host_cpus(),
    }
    if not args.no_sections:
        # the other single-GPU configurations, beside the headline
        from oracle import oracle as O
        log("C1 section")
        out["c1"] = run_c1(torch, H, O, args)
        log("C3 section")
        out["c3"] = run_c3_brief(torch, H, O, args)
    print(json.dumps(out), flush=True)
    if args.copy_probe > 0:
        n = args.copy_probe * (1 << 20) // 8
        x = torch.ones(n, dtype=torch.float64, device=dev)
        y = torch.empty_like(x)
        for _ in range(3):
            torch.mul(x, 1.0, out=y)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": n * 8, "copy_probe_bytes_written": n * 8}), file=sys.stderr)


if __name__ == "__main__":
    main()
