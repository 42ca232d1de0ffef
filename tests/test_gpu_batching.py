"""GPU tests of the batching forms: HIP-graph replay of uneven solves,
fixed-sweep mixed precision, and operator replacement in a resident plan.

Bars: graph replays are bit-identical to the eager solves they recorded
(same kernels, same shapes); mixed precision <= 1e-10 relative to the FP64
oracle; a plan whose operators were replaced gives the same bits as a plan
created with those operators.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def H():
    import torch
    assert torch.cuda.is_available(), "GPU test run without a GPU"
    from hackathon_amd import helmholtz
    return helmholtz


@pytest.fixture(scope="module")
def ctx(H):
    return H.Context(0)


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.complex128)).cuda()


def relerr(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


# (order, N, sine problem [kr, ki, a, b, phi], precision)
UNEVEN = [(4, 63, [20.0, 0.0, 3, 5, 0.4], 1), (6, 127, [45.0, 0.0, 7, 2, 1.0], 0),
          (6, 255, [70.0, 0.0, 11, 4, 2.0], 1), (4, 31, [5.0, 0.0, 1, 2, 0.0], 0)]


def test_graph_replays_uneven_batch(H, ctx, oracle):
    import torch
    solves, refs, eager = [], [], []
    for order, N, p, prec in UNEVEN:
        s = oracle.solve2d(order, N, 1, p, threads=8)
        plan = H.Plan2D(ctx, N, N, s["coef"], precision=prec)
        if prec == H.PREC_MIXED:
            plan.set_refine(0.0, 2)     # fixed sweeps: capturable
        rhs = dev(np.stack([s["rhs"], 0.5 * s["rhs"]]))
        out = torch.empty_like(rhs)
        e = plan.solve(rhs).cpu().numpy()
        solves.append((plan, rhs, out))
        refs.append(s["sol"])
        eager.append(e)
        err = relerr(e[0].ravel(), s["sol"])
        say(f"eager order={order} N={N} prec={prec} rel err vs oracle {err:.2e}")
        assert err <= (1e-10 if prec == H.PREC_MIXED else 1e-12)
    g = H.Graph(ctx, solves)
    say(f"graph of {len(solves)} solves, {g.num_nodes} nodes")
    scratch = torch.randn(1 << 20, dtype=torch.float64, device="cuda")
    for rep in range(3):
        for _, _, out in solves:
            out.zero_()
        g.launch()
        torch.cuda.synchronize()
        for (plan, rhs, out), e in zip(solves, eager):
            assert np.array_equal(out.cpu().numpy(), e), f"replay {rep} differs from the eager solve"
        # eager work of the program between replays
        scratch.mul_(1.0001)
        solves[1][0].solve(solves[1][1])
        torch.cuda.synchronize()
    g.close()


def test_fixed_sweeps_3d_mixed(H, ctx, oracle):
    s = oracle.solve3d(6, 31, 1, [15.0, 0.0, 2, 3, 1, 1.0, 1.0], threads=8)
    plan = H.Plan3D(ctx, 31, 31, 31, s["coef"], precision=H.PREC_MIXED)
    plan.set_refine(0.0, 2)
    sol = plan.solve(dev(s["rhs"])).cpu().numpy().ravel()
    err = relerr(sol, s["sol"])
    sweeps, rel = plan.last_refine()
    say(f"3D mixed fixed sweeps={sweeps} rel={rel} err={err:.2e}")
    assert sweeps == 2 and rel == -1.0
    assert err <= 1e-10


def test_mixed_3d_255_row_shapes(H, ctx):
    """3D 255^3 in mixed precision: the built-in shape of its FP32
    elimination (W = 8 modes x 32 chunks of 8 rows), 16 rows per thread and
    W = 8 x 8 rows all land within 1e-10 of the FP64 solve of the same
    right-hand side (the FP64 solver is pinned to the oracle by the parity
    tests)."""
    import ctypes
    import torch
    N = 255
    h = 1.0 / (N + 1)
    coef = H.coefficients3d(6, N, N, N, h, h, h, np.full(N + 2, 900.0 + 0j))
    g = torch.Generator(device="cuda").manual_seed(21)
    rhs = torch.randn((1, N, N, N), dtype=torch.complex128, device="cuda", generator=g)
    ref = H.Plan3D(ctx, N, N, N, coef).solve(rhs).cpu().numpy().ravel()
    plan = H.Plan3D(ctx, N, N, N, coef, precision=H.PREC_MIXED)
    plan.set_refine(0.0, 2)
    v = [ctypes.c_int() for _ in range(4)]
    H._check(H.lib().hh_plan_get_tri_shape(plan.h, *[ctypes.byref(x) for x in v]))
    assert [x.value for x in v] == [8, 32, 8, 1]
    for shape in ((0, 0, 0), (16, 16, 1), (8, 8, 1)):
        H._check(H.lib().hh_plan_set_tri_shape(plan.h, *shape))
        sol = plan.solve(rhs).cpu().numpy().ravel()
        err = relerr(sol, ref)
        say(f"3D 255^3 mixed shape={shape}: rel err vs FP64 {err:.2e}")
        assert err <= 1e-10


def test_real_band_shape_and_complex_fallback(H, ctx, oracle):
    """A shape whose reduced system fits the 160 KiB of LDS only with real
    bands (W = 8 modes x 128 chunks of 8 rows at N = 1023: 1024 threads,
    128 KiB) is accepted for a real operator and solves to the oracle; once
    hh_plan_set_operators makes the operator complex the plan reports and
    runs its built-in shape instead."""
    import ctypes
    N = 1023
    sr = oracle.solve2d(6, N, 1, [50.0, 0.0, 7, 11, 1.3], threads=8)
    sc = oracle.solve2d(6, N, 1, [50.0, 0.5, 7, 11, 1.3], threads=8)
    plan = H.Plan2D(ctx, N, N, sr["coef"])

    def shape():
        v = [ctypes.c_int() for _ in range(4)]
        H._check(H.lib().hh_plan_get_tri_shape(plan.h, *[ctypes.byref(x) for x in v]))
        return [x.value for x in v]

    H._check(H.lib().hh_plan_set_tri_shape(plan.h, 8, 8, 1))
    assert shape() == [8, 128, 8, 1]
    e = relerr(plan.solve(dev(sr["rhs"])).cpu().numpy().ravel(), sr["sol"])
    say(f"real operator, W=8 x 8 rows: err {e:.2e}")
    assert e <= 1e-12
    plan.set_operators([sc["coef"]])
    assert H.lib().hh_plan_is_real_operator(plan.h) == 0
    assert shape()[0] != 8
    e = relerr(plan.solve(dev(sc["rhs"])).cpu().numpy().ravel(), sc["sol"])
    say(f"complex operator, built-in shape {shape()}: err {e:.2e}")
    assert e <= 1e-12
    # a complex plan refuses the shape outright
    assert H.lib().hh_plan_set_tri_shape(plan.h, 8, 8, 1) == H.HH_ERR_UNSUPPORTED


def test_set_operators_matches_new_plan(H, ctx, oracle):
    N = 127
    probs = [oracle.solve2d(o, N, 1, [k, 0.0, 3, 2, 0.3], threads=8) for o, k in
             [(4, 10.0), (6, 30.0), (4, 55.0), (6, 70.0)]]
    ops = [p["coef"] for p in probs]
    rhs = dev(np.stack([p["rhs"] for p in probs]))
    plan = H.Plan2D(ctx, N, N, ops)          # four operator slots
    first = plan.solve(rhs).cpu().numpy()
    for b in range(4):
        assert relerr(first[b].ravel(), probs[b]["sol"]) <= 1e-12
    plan.set_operators([ops[3], ops[2]])     # two operators: rhs b uses op b % 2
    rhs2 = dev(np.stack([probs[3]["rhs"], probs[2]["rhs"]]))
    swapped = plan.solve(rhs2).cpu().numpy()
    fresh = H.Plan2D(ctx, N, N, [ops[3], ops[2]]).solve(rhs2).cpu().numpy()
    say("set_operators vs fresh plan")
    assert np.array_equal(swapped, fresh)
    assert np.array_equal(swapped[0], first[3]) and np.array_equal(swapped[1], first[2])
    assert relerr(swapped[0].ravel(), probs[3]["sol"]) <= 1e-12
    with pytest.raises(H.HHError):
        plan.set_operators(ops + ops)        # more operators than slots


@pytest.mark.parametrize("nop,batch", [(1, 7), (2, 7), (3, 8)])
def test_rhs_per_thread_same_bits(H, ctx, oracle, nop, batch):
    """Several right-hand sides per thread (sharing an operator) give the
    same bits as one per thread, for batches that do not fill the groups."""
    import torch
    N = 255
    probs = [oracle.solve2d(o, N, 1, [k, 0.0, 3, 2, 0.3], threads=8) for o, k in [(4, 12.0), (6, 40.0), (6, 75.0)]]
    ops = [p["coef"] for p in probs[:nop]]
    plan = H.Plan2D(ctx, N, N, ops)
    g = torch.Generator(device="cuda").manual_seed(nop * 10 + batch)
    rhs = torch.randn((batch, N, N), dtype=torch.complex128, device="cuda", generator=g)
    ref = None
    for W, rows in ((2, 8), (4, 4)):
        for r in (1, 2, 4):
            H._check(H.lib().hh_plan_set_tri_shape(plan.h, W, rows, r))
            out = plan.solve(rhs).cpu().numpy()
            say(f"nop={nop} batch={batch} W={W} rows={rows} rhs/thread={r}")
            if r == 1:
                ref = out
            else:
                assert np.array_equal(out, ref)
    # and the solution of right-hand side b matches the oracle for its operator b % nop
    rhs_o = dev(np.stack([probs[b % nop]["rhs"] for b in range(batch)]))
    H._check(H.lib().hh_plan_set_tri_shape(plan.h, 2, 8, 4))
    sol = plan.solve(rhs_o).cpu().numpy()
    for b in range(batch):
        assert relerr(sol[b].ravel(), probs[b % nop]["sol"]) <= 1e-12


def test_mixed_rhs_per_thread_same_bits(H, ctx, oracle):
    """FP32 storage with FP64 elimination (mixed precision): pairing two
    right-hand sides per thread, as the built-in shape does for a batch with
    two or more per operator, gives the same bits as one per thread."""
    import torch
    N = 1023
    s = oracle.solve2d(6, N, 1, [50.0, 0.0, 4, 9, 0.7], threads=8)
    plan = H.Plan2D(ctx, N, N, s["coef"], precision=H.PREC_MIXED)
    plan.set_refine(0.0, 2)
    g = torch.Generator(device="cuda").manual_seed(11)
    rhs = torch.randn((5, N, N), dtype=torch.complex128, device="cuda", generator=g)
    rhs[0] = dev(s["rhs"]).reshape(N, N)
    outs = []
    for shape in ((0, 0, 0), (4, 16, 1), (4, 16, 2)):
        H._check(H.lib().hh_plan_set_tri_shape(plan.h, *shape))
        outs.append(plan.solve(rhs).cpu().numpy())
        say(f"mixed N={N} batch=5 shape={shape}")
    assert np.array_equal(outs[0], outs[2]) and np.array_equal(outs[1], outs[2])
    assert relerr(outs[0][0].ravel(), s["sol"]) <= 1e-10


def test_spectrum_layouts_same_bits(H, ctx, oracle):
    """Row-major and mode-tiled spectra (any tile width) give the same bits."""
    import torch
    N = 127
    s = oracle.solve2d(6, N, 1, [33.0, 0.0, 5, 3, 0.4], threads=8)
    plan = H.Plan2D(ctx, N, N, s["coef"])
    H._check(H.lib().hh_plan_set_tri_shape(plan.h, 4, 8, 1))   # tiles must hold whole blocks of 4 modes
    g = torch.Generator(device="cuda").manual_seed(5)
    rhs = torch.randn((3, N, N), dtype=torch.complex128, device="cuda", generator=g)
    rhs[0] = dev(s["rhs"]).reshape(N, N)
    outs = []
    for tw in (0, -1, 8, 16, 64):
        H._check(H.lib().hh_plan_set_tile_width(plan.h, tw))
        outs.append(plan.solve(rhs).cpu().numpy())
        say(f"tile width {tw}")
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])
    assert relerr(outs[0][0].ravel(), s["sol"]) <= 1e-12
    with pytest.raises(H.HHError):   # a tile narrower than the block is refused at launch
        H._check(H.lib().hh_plan_set_tile_width(plan.h, 2))
        plan.solve(rhs)


def test_graph_keeps_own_workspace(H, ctx, oracle):
    """A graph owns its workspace: an eager solve that grows (and moves) the
    context's scratch afterwards must not disturb later replays."""
    import torch
    N = 127
    s = oracle.solve2d(6, N, 1, [30.0, 0.0, 3, 2, 0.5], threads=8)
    plan = H.Plan2D(ctx, N, N, s["coef"], precision=H.PREC_MIXED)
    plan.set_refine(0.0, 2)
    rhs = dev(np.stack([s["rhs"], 0.25 * s["rhs"]]))
    out = torch.empty_like(rhs)
    eager = plan.solve(rhs).cpu().numpy()
    g = H.Graph(ctx, [(plan, rhs, out)])
    big = H.Plan2D(ctx, 255, 255, oracle.solve2d(6, 255, 1, [30.0, 0.0, 3, 2, 0.5], threads=8)["coef"],
                   precision=H.PREC_MIXED)
    big_rhs = dev(np.ones(16 * 255 * 255))
    for rep in range(3):
        out.zero_()
        g.launch()
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy(), eager), f"replay {rep}"
        # eager solves that need more workspace than the graph's solve
        big.solve(big_rhs)
        torch.cuda.synchronize()
    say(f"graph replays unchanged after the context scratch grew to {ctx.bytes_in_use()} bytes in use")
    g.close()


def test_graph_refuses_changed_plan(H, ctx, oracle):
    """A plan whose band kind changed since the recording refuses to replay;
    new operator values of the same kind reach the replay."""
    import torch
    N = 63
    s1 = oracle.solve2d(4, N, 1, [2.0, 0.0, 1, 1, 0.0], threads=8)
    s2 = oracle.solve2d(4, N, 1, [2.5, 0.0, 2, 1, 0.0], threads=8)
    plan = H.Plan2D(ctx, N, N, [s1["coef"]])
    # k < pi: every mode diagonally dominant; with no condition threshold
    # neither operator has a near-resonant mode, so only values change
    plan.set_resonance_threshold(1e300)
    rhs = dev(s2["rhs"])
    out = torch.empty_like(rhs)
    g = H.Graph(ctx, [(plan, rhs, out)])
    assert plan.mode_info()[0] == 0
    plan.set_operators([s2["coef"]])          # same band kind, no near-resonant mode either
    assert plan.mode_info()[0] == 0
    g.launch()
    torch.cuda.synchronize()
    err = relerr(out.cpu().numpy(), s2["sol"])
    say(f"replay after set_operators (same kind): rel err {err:.2e}")
    assert err <= 1e-12
    cplx = [c * (1 + 0.01j) for c in s2["coef"]]
    plan.set_operators([cplx])                # complex bands now: the graph is stale
    with pytest.raises(H.HHError) as ei:
        g.launch()
    assert ei.value.status == H.HH_ERR_ARG
    g.close()
