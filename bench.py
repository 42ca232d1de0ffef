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
    ap.add_argument("--check-rhs", type=int, default=2, help="right-hand sides per order compared with the oracle")
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


def cpu_baseline(O, N, coef, F_host, seconds):
    """Oracle restatement of the reference 2D solver() on this host's cores."""
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    threads = max(1, min(threads, 16))
    LUU = O.lu2d(N, N, coef, threads)
    done, t_used = 0, 0.0
    while t_used < seconds:   # cycle over the right-hand sides until the budget is spent
        t0 = time.perf_counter()
        O.solver2d(N, N, F_host[done % F_host.shape[0]], LUU, threads)
        t_used += time.perf_counter() - t0
        done += 1
    return {"value": done * N * N / t_used, "unit": "complex DOF/s", "cores": threads, "kind": "port",
            "sample": f"oracle solver() (own FFT DST, plan creation inside the call, OpenMP over rows/modes) on "
                      f"{done} 6th-order right-hand sides of the same N={N} k=50 workload, {t_used:.1f} s"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); multi-GPU runs are not part of "
              "this build", file=sys.stderr)
        sys.exit(2)
    import torch

    from hackathon_amd import helmholtz as H

    assert torch.cuda.is_available(), "bench.py needs a GPU"
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
            if not args.no_cpu and o == orders[-1]:
                Fall = F[o].cpu().numpy().reshape(B, N * N)
                cpu = cpu_baseline(O, N, coef, Fall, args.cpu_seconds)

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
    }
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
