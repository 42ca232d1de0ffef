"""Autotuner of the kernel launch shapes.

The tridiagonal kernel's block shape (W neighbouring modes per block, rows
per thread, hence chunks per mode and the depth of the reduced system that
parallel cyclic reduction solves) is chosen by timing the candidates on the
device with the library's own kernel timers (HIP events on the launch
stream).  Results are kept in a JSON table keyed by
(dimension, system length, modes, batch bucket, precision) so that a shape
does not depend on what else a batch contains; `apply_tuned(plan, batch)`
sets the tabulated shape on a plan.  With graph=True each candidate is
instead recorded into a HIP graph of the solve and its replays are timed
(HIP events around the launches), which is what a graph-captured batch
runs: launch gaps and the order of the kernels in the replay count, and the
eager per-kernel timers do not see them.  The DST variants (half-length or
odd-extension transform, 256-, 64- or, for strided lines, 512-thread
blocks, and the automatic choice) are timed as well.

    python -m hackathon_amd.autotune --dim 2 --N 1023 --batch 64
"""
import argparse
import ctypes
import json
import os

import numpy as np

from . import helmholtz as H

TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")
CANDIDATES = [(w, m, r) for m in (8, 16) for w in (1, 2, 4, 8, 16) for r in (1, 2)]


def _bucket(batch):
    b = 1
    while b < batch:
        b *= 2
    return b


def key(dim, n, nmodes, batch, precision):
    return f"{dim}:{n}:{nmodes}:{_bucket(batch)}:{precision}"


def load_table(path=TABLE):
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def save_table(table, path=TABLE):
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)


def time_tri(plan, rhs, out, reps=3):
    """(kernel microseconds of one solve of rhs, average milliseconds per
    launch of each kernel category), from the library's event timers."""
    lib = H.lib()
    ctx = plan.ctx
    plan.solve(rhs, out=out)  # warm-up (twiddles, scratch, code objects)
    lib.hh_context_reset_timing(ctx.h)
    lib.hh_context_enable_timing(ctx.h, 1)
    for _ in range(reps):
        plan.solve(rhs, out=out)
    lib.hh_context_enable_timing(ctx.h, 0)
    ms = (ctypes.c_double * 5)()
    nl = (ctypes.c_long * 5)()
    lib.hh_context_read_timing(ctx.h, ms, nl, None)
    return 1e3 * sum(ms) / reps, [ms[k] / max(nl[k], 1) for k in range(5)]


def time_graph(plan, rhs, out, reps=3):
    """(microseconds of one replay of a graph holding the solve of rhs,
    [None] * 5): the solve recorded once, replayed after one warm-up."""
    import torch
    g = H.Graph(plan.ctx, [(plan, rhs, out)])
    try:
        g.launch()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        for _ in range(reps):
            g.launch()
        e.record()
        e.synchronize()
        return 1e3 * s.elapsed_time(e) / reps, [None] * 5
    finally:
        g.close()


def tune(plan, rhs, out, candidates=CANDIDATES, reps=3, verbose=False, tiles=None, graph=False):
    """Try the built-in shape and every candidate shape (and, in 2D,
    spectrum layout, or only the given tile widths) that fits; leave the
    fastest set on the plan: by kernel time of the whole solve, or, with
    graph=True, by the time of a graph replay of it."""
    lib = H.lib()
    results = []
    batch = plan._batch_of(rhs)
    if tiles is None:
        tiles = (0, -1, 8, 16, 32) if plan.dim == 2 else (-1,)
    for w, m, r in [(0, 0, 1)] + list(candidates):
        if r > 1 and batch < 2 * r:
            continue
        if lib.hh_plan_set_tri_shape(plan.h, w, m, r) != 0:
            continue
        for tw in tiles:
            if tw > 0 and w > 0 and (tw < w or tw % w):
                continue
            H._check(lib.hh_plan_set_tile_width(plan.h, tw))
            try:
                t, per = (time_graph if graph else time_tri)(plan, rhs, out, reps)
            except H.HHError:
                continue
            dst_us = per[0] * 1e3 if per[0] is not None else None
            tri_us = per[2] * 1e3 if per[2] is not None else None
            results.append((t, w, m, r, tw, dst_us, tri_us))
            if verbose:
                detail = "graph replay" if graph else f"DST {dst_us:7.1f}, tridiagonal {tri_us:7.1f} us per launch"
                print(f"  W={w:3d} rows={m:3d} rhs={r} tile={tw:3d}: solve {t:9.1f} us ({detail})", flush=True)
    results.sort()
    best = results[0]
    lib.hh_plan_set_tri_shape(plan.h, best[1], best[2], best[3])
    lib.hh_plan_set_tile_width(plan.h, best[4])
    return {"W": best[1], "rows": best[2], "rhs": best[3], "tile": best[4], "us": best[0],
            "all": [{"W": w, "rows": m, "rhs": r, "tile": tw, "us": t, "dst_us": d, "tri_us": q}
                    for t, w, m, r, tw, d, q in results]}


def apply_tuned(plan, batch, table=None):
    """Set the tabulated shape on a plan, if the table has one."""
    table = load_table() if table is None else table
    n = plan.Ny if plan.dim == 2 else plan.Nz
    nmodes = plan.Nx if plan.dim == 2 else plan.Nx * plan.Ny
    e = table.get(key(plan.dim, n, nmodes, batch, plan.precision))
    if e:
        H._check(H.lib().hh_plan_set_tri_shape(plan.h, e["W"], e["rows"], e.get("rhs", 1)))
        H._check(H.lib().hh_plan_set_tile_width(plan.h, e.get("tile", -1)))
    return e


def main():
    import torch
    ap = argparse.ArgumentParser()
    ap.add_argument("--dim", type=int, default=2)
    ap.add_argument("--N", type=int, default=1023)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--precision", type=int, default=H.PREC_F64)
    ap.add_argument("--k", type=float, default=50.0)
    ap.add_argument("--save", action="store_true")
    ap.add_argument("--graph", action="store_true", help="time each candidate as a graph replay of the solve")
    args = ap.parse_args()
    ctx = H.Context(0)
    N = args.N
    h = 1.0 / (N + 1)
    if args.dim == 2:
        coef = H.coefficients2d(6, N, N, h, h, np.full(N + 2, args.k ** 2, np.complex128))
        plan = H.Plan2D(ctx, N, N, coef, precision=args.precision)
        shape = (args.batch, N, N)
    else:
        coef = H.coefficients3d(6, N, N, N, h, h, h, np.full(N + 2, args.k ** 2, np.complex128))
        plan = H.Plan3D(ctx, N, N, N, coef, precision=args.precision)
        shape = (args.batch, N, N, N)
    g = torch.Generator(device="cuda").manual_seed(0)
    rhs = torch.randn(shape, dtype=torch.complex128, device="cuda", generator=g)
    out = torch.empty_like(rhs)
    print(f"tridiagonal shapes, dim={args.dim} N={N} batch={args.batch} precision={args.precision}", flush=True)
    if args.precision == H.PREC_MIXED:
        plan.set_refine(0.0, 2)   # fixed sweeps: graph-capturable, and the same work per timing
    res = tune(plan, rhs, out, verbose=True, graph=args.graph)
    print(json.dumps({k: v for k, v in res.items() if k != "all"}), flush=True)
    # -1 automatic; bit 0 odd extension, bit 1 one-wave blocks, bit 2 512-thread
    # blocks for strided lines (N+1 in 128..256)
    for algo in (-1, 0, 1, 2, 3, 4):
        ctx.set_dst_algorithm(algo)
        _, per = time_tri(plan, rhs, out)
        print(f"DST algorithm {algo}: {1e3 * per[0]:9.1f} us per launch", flush=True)
    ctx.set_dst_algorithm(-1)
    if args.save:
        table = load_table()
        n = N
        nmodes = N if args.dim == 2 else N * N
        table[key(args.dim, n, nmodes, args.batch, args.precision)] = {"W": res["W"], "rows": res["rows"],
                                                                       "rhs": res["rhs"], "tile": res["tile"],
                                                                       "us": round(res["us"], 1)}
        save_table(table)


if __name__ == "__main__":
    main()
