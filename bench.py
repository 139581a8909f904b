#!/usr/bin/env python3
"""Throughput of the doubly periodic stresslet spectral Ewald sum on one MI355X.

Workload (default): the stresslet double-layer potential on 200 non-overlapping
ellipses x 500 nodes (N = 1e5 sources = targets, unit box), trapezoid-weighted
smooth densities, Ewald parameters from the reference driver's formulas at
tol 1e-10 with the grid fixed at 512^2 (P = 24).  One step = the real-space sum
plus the Fourier-space sum (the two kernels the reference driver calls,
mex_stokes_dlp_real + mex_stokes_dlp_kspace) over all targets, inputs resident
in HBM.  The targets are the sources and are passed as the same array.  The
two calls are independent and run as one concurrent group
(se2p_group_begin / se2p_group_end; --sequential times them one after the
other, and the line carries both step times).

Prints ONE JSON line: target evaluations per second (one evaluation = one
2-vector u(x_t) for one density), the roofline of the dominant kernel (the
real-space pair sum, FP64 vector ALU bound) and the CPU oracle timed on this
host's cores on the same workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6      # MI355X FP64 vector (and matrix) peak
PEAK_HBM_GBS = 8000.0
# flop convention per real-space pair: exp = 20, divide = 1 (x2), FMA = 2;
# an evaluated stresslet pair costs 80, a pair rejected by the cutoff 6.
F_PAIR_IN = 80
F_PAIR_OUT = 6


def _pmc_traffic(kernel_tag: str):
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated),
    or None when no summary is committed."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        with open(path) as fh:
            data = json.load(fh)
        return data.get(kernel_tag, {}).get("hbm_bytes_per_launch")
    except (OSError, ValueError):
        return None


def _host_cpu():
    """Logical CPUs, physical cores and threads per core of this host (from
    /proc/cpuinfo), or None where it cannot be read."""
    try:
        phys, cores, logical = set(), set(), 0
        cur = {}
        with open("/proc/cpuinfo") as fh:
            for line in list(fh) + [""]:
                if not line.strip():
                    if cur:
                        logical += 1
                        cores.add((cur.get("physical id"), cur.get("core id")))
                        phys.add(cur.get("physical id"))
                    cur = {}
                    continue
                k, _, v = line.partition(":")
                cur[k.strip()] = v.strip()
        return {"logical_cpus": logical, "physical_cores": len(cores), "sockets": len(phys),
                "smt": logical > len(cores)}
    except OSError:
        return None


def _stage_model(N, Nt, G, P, pairs_in, pairs_out, stage_ms):
    """Algorithmic flops and HBM bytes per stage of one stresslet step, each
    stage's roofline time max(bytes / HBM peak, flops / FP64 peak) and its
    fraction of the measured stage time.  K = (P+1)^2 nodes per point, c_g = 3
    grids, c_in = 4 input components per source (f, n)."""
    import math

    K, cg, cin = (P + 1) ** 2, 3, 4
    lg = math.log2(G)
    model = {
        "real": (pairs_in * F_PAIR_IN + pairs_out * F_PAIR_OUT, N * 8 * (2 + cin) + 32 * Nt,
                 ["real_call.real"]),
        "spread": (N * K * (2 + 2 * cg), N * 8 * (2 + cin) + 8 * G * cg, ["kspace_call.spread"]),
        "fft_fwd": (cg * 2.5 * G * lg, cg * 24 * G, ["kspace_call.fft_fwd"]),
        "filter": (40 * G, 16 * G * (cg + 2), ["kspace_call.filter"]),
        "fft_inv": (2 * 2.5 * G * lg, 2 * 24 * G, ["kspace_call.fft_inv"]),
        "interp": (Nt * K * 6, 16 * G + 32 * Nt, ["kspace_call.interp"]),
        "sort": (0.0, 0.0, ["real_call.sort", "kspace_call.sort"]),
    }
    out, t_model = {}, 0.0
    for name, (fl, by, keys) in model.items():
        ms = sum(stage_ms.get(k, 0.0) for k in keys)
        tm = max(by / (PEAK_HBM_GBS * 1e9), fl / (PEAK_FP64_TFLOPS * 1e12)) * 1e3
        t_model += tm
        out[name] = {"flops": fl, "hbm_bytes": by, "measured_ms": ms, "model_ms": tm,
                     "frac": tm / ms if ms > 0 else None,
                     "tflops": fl / (ms * 1e-3) / 1e12 if ms > 0 else None,
                     "hbm_gbs": by / (ms * 1e-3) / 1e9 if ms > 0 else None}
    return out, t_model


def _direct_ewald_error(np, ur, uk):
    """Relative max-norm error of the step's output against the committed FP64
    direct Ewald sum on 256 targets (tests/golden/truth_dlp_ellipses_N1e5.json),
    per output; None if the fixture is absent."""
    path = os.path.join(ROOT, "tests", "golden", "truth_dlp_ellipses_N1e5.json")
    try:
        with open(path) as fh:
            d = json.load(fh)
    except (OSError, ValueError):
        return None
    sub = np.array(d["targets"])
    tr, tk = np.array(d["u_real"]), np.array(d["u_kspace"])
    tt = tr + tk
    u = ur[sub] + uk[sub]
    scale = np.abs(tt).max()
    return {"u": float(np.abs(u - tt).max() / scale),
            "u1": float(np.abs(u[:, 0] - tt[:, 0]).max() / np.abs(tt[:, 0]).max()),
            "u2": float(np.abs(u[:, 1] - tt[:, 1]).max() / np.abs(tt[:, 1]).max()),
            "real_part": float(np.abs(ur[sub] - tr).max() / scale),
            "fourier_part": float(np.abs(uk[sub] - tk).max() / scale),
            "targets": len(sub),
            "truth": "3x3-image real sum + direct Fourier sum, j in (-256,256]^2 (FP64)"}


def _batched_variant(args, pe, W, np, torch, dev, nrhs=64):
    """Per-RHS throughput of 64 densities on the 1000-point unit-box geometry
    (grid 64^2, tol 1e-10), Stokeslet + stresslet, as 64 plain calls of the
    four entry points versus one call of each batched entry point."""
    from fasttools2d_amd import params as Pm

    w = W.uniform_box(1000, seed=1)
    pr = Pm.select_params("slp", 1000, 1000, w.f, 1.0, 1.0, tol=1e-10, M=64)
    pd = Pm.select_params("dlp", 1000, 1000, w.f, 1.0, 1.0, tol=1e-10, M=64)
    rng = np.random.default_rng(100)
    F = torch.from_numpy(rng.standard_normal((nrhs, 1000, 2))).to(dev)
    ps = torch.from_numpy(w.psrc).to(dev)
    nv = torch.from_numpy(w.n).to(dev)
    ctx = pe.Context(0)
    stream = torch.cuda.Stream(dev)
    outs = [torch.empty((1000, 2), dtype=torch.float64, device=dev) for _ in range(4)]

    def plain():
        for r in range(nrhs):
            pe.mex_stokes_slp_real(ps, ps, F[r], pr.xi, pr.nside_x, pr.nside_y, 1.0, 1.0, ctx=ctx, out=outs[0])
            pe.mex_stokes_slp_kspace(ps, ps, pr.xi, pr.eta, F[r], pr.Mx, pr.My, 1.0, 1.0, pr.w, pr.P, ctx=ctx, out=outs[1])
            pe.mex_stokes_dlp_real(ps, ps, F[r], nv, pd.xi, pd.nside_x, pd.nside_y, 1.0, 1.0, ctx=ctx, out=outs[2])
            pe.mex_stokes_dlp_kspace(ps, ps, pd.xi, pd.eta, F[r], nv, pd.Mx, pd.My, 1.0, 1.0, pd.w, pd.P, ctx=ctx, out=outs[3])

    def batched():
        pe.stokes_slp_real_batched(ps, ps, F, pr.xi, pr.nside_x, pr.nside_y, 1.0, 1.0, ctx=ctx)
        pe.stokes_slp_kspace_batched(ps, ps, pr.xi, pr.eta, F, pr.Mx, pr.My, 1.0, 1.0, pr.w, pr.P, ctx=ctx)
        pe.stokes_dlp_real_batched(ps, ps, F, nv, pd.xi, pd.nside_x, pd.nside_y, 1.0, 1.0, ctx=ctx)
        pe.stokes_dlp_kspace_batched(ps, ps, pd.xi, pd.eta, F, nv, pd.Mx, pd.My, 1.0, 1.0, pd.w, pd.P, ctx=ctx)

    res = {}
    with torch.cuda.stream(stream):
        ctx.bind_torch_stream(stream)
        ctx.set_async(True)
        for name, fn in (("plain", plain), ("batched", batched)):
            fn()
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            reps = 3
            e0.record(stream)
            for _ in range(reps):
                fn()
            e1.record(stream)
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / reps
            # one evaluation = one target x one density, Stokeslet and stresslet counted once each
            res[name] = {"ms_per_call_set": ms, "target_evals_per_s": 2 * 1000 * nrhs / (ms * 1e-3)}
        ctx.set_async(False)
    res["batched_over_plain"] = res["batched"]["target_evals_per_s"] / res["plain"]["target_evals_per_s"]
    res["workload"] = (f"1000 uniform points (unit box, targets = sources), grid 64^2, tol 1e-10, "
                       f"{nrhs} densities, Stokeslet + stresslet (real + Fourier each)")
    print(f"# batched variant: plain {res['plain']['ms_per_call_set']:.3f} ms, batched "
          f"{res['batched']['ms_per_call_set']:.3f} ms", file=sys.stderr, flush=True)
    return res


def _timed_ms(torch, stream, fn, reps=3):
    """Average milliseconds of fn() on `stream` (HIP events), after one warm-up."""
    fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def _batched_ellipses_variant(args, pe, W, np, torch, dev, nrhs=64):
    """Per-RHS throughput of 64 densities on the benchmark's stresslet
    geometry (200 ellipses x 500 nodes, N = 1e5, grid 512^2, tol 1e-10), where
    a plain call is compute-bound rather than launch-bound: 64 plain calls of
    the two stresslet entry points against one call of each batched entry
    point.  Densities: N(0,1) times the trapezoid weights."""
    w, p = W.dlp_ellipses_config(seed=2, n_ell=200, n_nodes=500, tol=1e-10, M=512)
    N = len(w.psrc)
    wt = np.linalg.norm(w.f, axis=1, keepdims=True)
    rng = np.random.default_rng(101)
    F = torch.from_numpy(rng.standard_normal((nrhs, N, 2)) * wt[None]).to(dev)
    ps = torch.from_numpy(np.ascontiguousarray(w.psrc)).to(dev)
    nv = torch.from_numpy(np.ascontiguousarray(w.n)).to(dev)
    ctx = pe.Context(0)
    stream = torch.cuda.Stream(dev)
    ur = torch.empty((N, 2), dtype=torch.float64, device=dev)
    uk = torch.empty_like(ur)

    def plain():
        for r in range(nrhs):
            pe.mex_stokes_dlp_real(ps, ps, F[r], nv, p.xi, p.nside_x, p.nside_y, 1.0, 1.0,
                                   ctx=ctx, out=ur)
            pe.mex_stokes_dlp_kspace(ps, ps, p.xi, p.eta, F[r], nv, p.Mx, p.My, 1.0, 1.0, p.w,
                                     p.P, ctx=ctx, out=uk)

    def batched():
        pe.stokes_dlp_real_batched(ps, ps, F, nv, p.xi, p.nside_x, p.nside_y, 1.0, 1.0, ctx=ctx)
        pe.stokes_dlp_kspace_batched(ps, ps, p.xi, p.eta, F, nv, p.Mx, p.My, 1.0, 1.0, p.w, p.P,
                                     ctx=ctx)

    res = {}
    with torch.cuda.stream(stream):
        ctx.bind_torch_stream(stream)
        ctx.set_async(True)
        for name, fn in (("plain", plain), ("batched", batched)):
            ms = _timed_ms(torch, stream, fn)
            res[name] = {"ms_per_call_set": ms, "target_evals_per_s": N * nrhs / (ms * 1e-3)}
        ctx.set_async(False)
    torch.cuda.synchronize()
    ctx.close()
    res["batched_over_plain"] = res["batched"]["target_evals_per_s"] / res["plain"]["target_evals_per_s"]
    res["workload"] = (f"stresslet, 200 ellipses x 500 nodes (N=1e5), grid 512^2, tol 1e-10, "
                       f"{nrhs} densities (real + Fourier sums)")
    print(f"# batched on the ellipses: plain {res['plain']['ms_per_call_set']:.2f} ms, batched "
          f"{res['batched']['ms_per_call_set']:.2f} ms ({res['batched_over_plain']:.2f}x per density)",
          file=sys.stderr, flush=True)
    return res


def _fixed_grid_variant(args, pe, np, torch, w, ps, f, nv, stream, fwd_ms):
    """The benchmark workload on its stated 512^2 grid at tol 1e-10 with the
    fixed-grid parameters: xi from the drivers' k-space estimate at
    k_inf = 256 and the cells from their real-space estimate (the headline
    uses the drivers' cell size with the grid raised to 512^2, which their
    own rule would size at 368^2).  Step time and error against the FP64
    direct Ewald truth on 256 targets (the total is independent of xi).  The
    step is formed as the headline step."""
    from fasttools2d_amd import params as Pm

    N = len(w.psrc)
    p = Pm.select_params_grid("dlp", N, N, w.f, 1.0, 1.0, M=512, tol=1e-10)
    ctx = pe.Context(0)
    ur = torch.empty((N, 2), dtype=torch.float64, device=ps.device)
    uk = torch.empty_like(ur)

    def real():
        pe.mex_stokes_dlp_real(ps, ps, f, nv, p.xi, p.nside_x, p.nside_y, 1.0, 1.0, ctx=ctx, out=ur)

    def fourier():
        pe.mex_stokes_dlp_kspace(ps, ps, p.xi, p.eta, f, nv, p.Mx, p.My, 1.0, 1.0, p.w, p.P,
                                 ctx=ctx, out=uk)

    def step():
        # the headline's step form
        if args.sequential:
            real()
            fourier()
        else:
            with ctx.group():
                fourier()
                real()

    with torch.cuda.stream(stream):
        ctx.bind_torch_stream(stream)
        ctx.set_async(True)
        ms = _timed_ms(torch, stream, step, reps=10)
        ctx.set_async(False)
    torch.cuda.synchronize()
    err = _direct_ewald_error(np, ur.cpu().numpy(), uk.cpu().numpy())
    ctx.close()
    print(f"# fixed-grid parameters on 512^2: nside {p.nside_x} xi {p.xi:.2f}: {ms:.3f} ms "
          f"({fwd_ms / ms:.2f}x the headline step), error vs direct Ewald "
          f"{err['u'] if err else float('nan'):.2e}", file=sys.stderr, flush=True)
    return {"ms_per_step": ms, "target_evals_per_s": N / (ms * 1e-3),
            "over_headline": fwd_ms / ms,
            "error_vs_direct_ewald": err["u"] if err else None,
            "params": {"nside": p.nside_x, "xi": p.xi, "eta": p.eta, "Mx": p.Mx, "tol": 1e-10,
                       "mode": "fixed grid"}}


def _multibox_variant(args, pe, W, np, torch, dev, tune=False):
    """The many-box workload: 32 independent periodic boxes, L ~ U[0.5, 4],
    N ~ log-uniform in [1e3, 2e5] uniform points each (1.41e6 in all; seed 4),
    N(0,1) forces, per-box parameters from the driver formulas at tol 1e-10.
    One step is the real-space sums of every box (one launch set) and their
    Fourier sums (packed spreading and interpolation launches; per-box sorts
    and transforms on four streams).  Timed as eager async calls and as a
    replayed HIP graph of the same step; the Stokeslet (the workload's
    forces) first, then the stresslet on the same boxes with its own driver
    parameters.  With ``tune``, also the same step at the parameters and
    spreading split from se2p_autotune_multibox (its tuning takes minutes on
    a fresh process: every candidate grid size builds rocFFT plans)."""
    ctx = pe.Context(0)
    stream = torch.cuda.Stream(dev)

    def timed(fn, reps=5):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    def eager_and_graph(fn):
        fn()
        fn()
        torch.cuda.synchronize()
        eager = timed(fn)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=stream):
            fn()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        return eager, timed(g.replay)

    out = {}
    for kernel in ("slp", "dlp"):
        boxes = W.multibox_config(seed=4, nbox=32, kind=kernel)
        tb = []
        for b in boxes:
            d = dict(b)
            for k in ("psrc", "f", "n"):
                d[k] = torch.from_numpy(np.ascontiguousarray(b[k])).to(dev)
            d["ptar"] = d["psrc"]
            if kernel == "slp":
                d.pop("n")
            tb.append(d)
        ntot = sum(len(b["psrc"]) for b in boxes)
        real = pe.stokes_slp_real_multibox if kernel == "slp" else pe.stokes_dlp_real_multibox
        fourier = pe.stokes_slp_kspace_multibox if kernel == "slp" else pe.stokes_dlp_kspace_multibox

        def step():
            return real(tb, ctx=ctx) + fourier(tb, ctx=ctx)

        with torch.cuda.stream(stream):
            ctx.bind_torch_stream(stream)
            ctx.set_async(True)
            eager_ms, graph_ms = eager_and_graph(step)
            parts = {name: dict(zip(("eager_ms", "graph_ms"), eager_and_graph(fn))) for name, fn in (
                ("real", lambda: real(tb, ctx=ctx)), ("fourier", lambda: fourier(tb, ctx=ctx)))}
            ctx.set_async(False)
        name = "Stokeslet" if kernel == "slp" else "stresslet"
        autotuned = None
        if tune:
            # the same step at parameters from se2p_autotune_multibox (own
            # context: the spreading split is a context setting)
            tctx = pe.Context(0)
            tuned = [dict(d) for d in tb]
            t0 = time.perf_counter()
            # each box's cell size first (se2p_autotune at P = 24, cached per
            # process, which the multi-box tuner then reuses), one line per box
            for i, d in enumerate(tuned):
                prm, _ = pe.autotune(kernel, d["psrc"], d["psrc"], d["f"], d.get("n"), d["Lx"],
                                     d["Ly"], tol=1e-10, P=24, ctx=tctx)
                print(f"# multibox {name} tuning box {i + 1}/{len(tuned)}: N {d['psrc'].shape[0]}, "
                      f"Nb {prm.Nb:.0f}, grid {prm.Mx}, {time.perf_counter() - t0:.1f} s",
                      file=sys.stderr, flush=True)
            tinfo = pe.autotune_multibox(kernel, tuned, tol=1e-10, ctx=tctx)
            tune_s = time.perf_counter() - t0
            with torch.cuda.stream(stream):
                tctx.bind_torch_stream(stream)
                tctx.set_async(True)
                t_eager, t_graph = eager_and_graph(lambda: real(tuned, ctx=tctx) + fourier(tuned, ctx=tctx))
                tctx.set_async(False)
            print(f"# multibox variant {name} autotuned ({tune_s:.1f} s): eager {t_eager:.3f} ms, "
                  f"graph {t_graph:.3f} ms", file=sys.stderr, flush=True)
            autotuned = {"eager_ms": t_eager, "graph_ms": t_graph,
                         "target_evals_per_s": ntot / (min(t_eager, t_graph) * 1e-3),
                         "tuning_seconds": tune_s, "tuner": tinfo}
        print(f"# multibox variant {name}: eager {eager_ms:.3f} ms, graph {graph_ms:.3f} ms (real "
              f"{parts['real']['graph_ms']:.3f}, Fourier {parts['fourier']['graph_ms']:.3f} ms replayed)",
              file=sys.stderr, flush=True)
        out[kernel] = {
            "eager": {"ms_per_step": eager_ms, "target_evals_per_s": ntot / (eager_ms * 1e-3)},
            "graph": {"ms_per_step": graph_ms, "target_evals_per_s": ntot / (graph_ms * 1e-3)},
            "parts": parts,
            "autotuned": autotuned,
            "grids": sorted({int(b["Mx"]) for b in boxes}),
            "workload": f"32 boxes, {ntot} points in all (targets = sources), {name} real + Fourier "
                        "sums, per-box parameters from the {kernel} driver formulas at tol 1e-10"
                        .replace("{kernel}", kernel)}
    return out


def _graph_variant(pe, torch, step, ctx, stream, N, reps=10):
    """The headline step captured once in a HIP graph and replayed, against
    the same step as async calls."""
    with torch.cuda.stream(stream):
        ctx.bind_torch_stream(stream)
        ctx.set_async(True)
        step()
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=stream):
            step()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        res = {}
        for name, fn in (("async", step), ("graph", g.replay)):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(reps):
                fn()
            e1.record(stream)
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / reps
            res[name] = {"ms_per_step": ms, "target_evals_per_s": N / (ms * 1e-3)}
        ctx.set_async(False)
    print(f"# graph variant: async {res['async']['ms_per_step']:.3f} ms, graph "
          f"{res['graph']['ms_per_step']:.3f} ms", file=sys.stderr, flush=True)
    return res


def _adjoint_variant(args, pe, np, torch, w, p, ps, pt, f, nv, ctx, stream, fwd_ms):
    """Forward + adjoint on the headline workload: one step is the two forward
    sums plus the two backward entry points with every gradient (dL/dpsrc,
    dL/dptar, dL/df, dL/dn) for a seeded du.  Measured with the forward calls
    keeping their sorted state for the backward calls (se2p_set_keep, the
    training pattern; the reported step) and with the backward calls
    rebuilding it; each as the headline step is formed (the forward pair
    and the backward pair each one concurrent group, or, with --sequential,
    one call after the other) and in the other form.  ``fwd_ms`` maps each
    form to its forward step time."""
    N = ps.shape[0]
    du = torch.from_numpy(np.random.default_rng(77).standard_normal((N, 2))).to(ps.device)
    ur = torch.empty_like(du)
    uk = torch.empty_like(du)
    form = "sequential" if args.sequential else "grouped"
    other = "grouped" if args.sequential else "sequential"

    def pair(first, second, grouped):
        if grouped:
            with ctx.group():
                first()
                second()
        else:
            second()
            first()

    def make_step(grouped):
        def fwd_k():
            pe.mex_stokes_dlp_kspace(ps, pt, p.xi, p.eta, f, nv, p.Mx, p.My, w.Lx, w.Ly, p.w,
                                     p.P, ctx=ctx, out=uk)

        def fwd_r():
            pe.mex_stokes_dlp_real(ps, pt, f, nv, p.xi, p.nside_x, p.nside_y, w.Lx, w.Ly,
                                   ctx=ctx, out=ur)

        def bwd_k():
            pe.mex_stokes_dlp_kspace_backward(ps, pt, p.xi, p.eta, f, nv, p.Mx, p.My, w.Lx,
                                              w.Ly, p.w, p.P, du, ctx=ctx)

        def bwd_r():
            pe.mex_stokes_dlp_real_backward(ps, pt, f, nv, p.xi, p.nside_x, p.nside_y, w.Lx,
                                            w.Ly, du, ctx=ctx)

        def step():
            pair(fwd_k, fwd_r, grouped)
            pair(bwd_k, bwd_r, grouped)
        return step

    step = make_step(not args.sequential)

    def timed_steps(step=step):
        for _ in range(max(1, args.warmup)):
            step()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = max(3, args.steps // 2)
        e0.record(stream)
        for _ in range(reps):
            step()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    def stage_split():
        # forward calls (keeping state when enabled), then the backward calls timed
        ctx.set_async(False)
        stages = {}
        for _ in range(3):
            pe.mex_stokes_dlp_real(ps, pt, f, nv, p.xi, p.nside_x, p.nside_y, w.Lx, w.Ly, ctx=ctx,
                                   out=ur)
            pe.mex_stokes_dlp_kspace(ps, pt, p.xi, p.eta, f, nv, p.Mx, p.My, w.Lx, w.Ly, p.w, p.P,
                                     ctx=ctx, out=uk)
            ctx.set_timing(True)
            pe.mex_stokes_dlp_real_backward(ps, pt, f, nv, p.xi, p.nside_x, p.nside_y, w.Lx, w.Ly,
                                            du, ctx=ctx)
            for k, v in ctx.stage_times().items():
                if v >= 0:
                    stages.setdefault("real_backward." + k, []).append(v)
            pe.mex_stokes_dlp_kspace_backward(ps, pt, p.xi, p.eta, f, nv, p.Mx, p.My, w.Lx, w.Ly,
                                              p.w, p.P, du, ctx=ctx)
            for k, v in ctx.stage_times().items():
                if v >= 0:
                    stages.setdefault("kspace_backward." + k, []).append(v)
            ctx.set_timing(False)
        return {k: statistics.mean(v) for k, v in stages.items()}

    other_step = make_step(args.sequential)
    with torch.cuda.stream(stream):
        ctx.bind_torch_stream(stream)
        ctx.set_async(True)
        ms_rebuild = timed_steps()
        ms_rebuild_other = timed_steps(other_step)
        st_rebuild = stage_split()
        ctx.set_keep(True)
        ctx.set_async(True)
        ms = timed_steps()
        ms_other = timed_steps(other_step)
        st_keep = stage_split()
        ctx.set_keep(False)
        ctx.set_async(False)
    torch.cuda.synchronize()
    f0, f1 = fwd_ms[form], fwd_ms[other]
    print(f"# forward+adjoint ({form}): {ms:.3f} ms per step (kept forward state, "
          f"{ms / f0:.2f}x the forward), {ms_rebuild:.3f} ms (rebuilt, {ms_rebuild / f0:.2f}x); "
          f"{other}: {ms_other:.3f} ms ({ms_other / f1:.2f}x), {ms_rebuild_other:.3f} ms "
          f"({ms_rebuild_other / f1:.2f}x)", file=sys.stderr, flush=True)
    return {"ms_per_step": ms, "target_evals_per_s": N / (ms * 1e-3),
            "over_forward": ms / f0,
            "adjoint_over_forward_time": (ms - f0) / f0,
            "step_form": form,
            "stage_ms": st_keep,
            "rebuilt": {"ms_per_step": ms_rebuild, "over_forward": ms_rebuild / f0,
                        "stage_ms": st_rebuild},
            other: {"ms_per_step": ms_other, "over_forward": ms_other / f1,
                    "forward_ms": f1,
                    "rebuilt": {"ms_per_step": ms_rebuild_other,
                                "over_forward": ms_rebuild_other / f1}},
            "workload": "same as the headline; step = forward real + Fourier sums, then both "
                        "backward entry points with all four gradients; the forward calls keep "
                        "their sorted state for the backward calls (se2p_set_keep); "
                        "over_forward divides by the forward step of the same form"}


def _autotune_variant(args, pe, np, torch, w, ps, f, nv, stream, fwd_ms):
    """The headline geometry with parameters chosen by se2p_autotune at the
    same tolerance (1e-10): points per cell, spreading work split and
    support/grid (P 16, 20 or the drivers' 24, on oversampled grids, checked
    against P = 24) by timing, on a context of its own (the split is a
    context setting).  One step is the same two MEX-equivalent calls at the
    tuned parameters, formed as the headline step; the output is checked against the FP64 direct Ewald sum on
    the 256 truth targets (the real + Fourier total does not depend on the
    split, so the truth at the headline's xi serves any xi)."""
    N = ps.shape[0]
    ctx = pe.Context(0)
    t0 = time.perf_counter()
    prm, info = pe.autotune("dlp", ps, ps, f, nv, w.Lx, w.Ly, tol=1e-10, ctx=ctx)
    tune_s = time.perf_counter() - t0
    ur = torch.empty((N, 2), dtype=torch.float64, device=ps.device)
    uk = torch.empty_like(ur)

    def real():
        pe.mex_stokes_dlp_real(ps, ps, f, nv, prm.xi, prm.nside_x, prm.nside_y, w.Lx, w.Ly,
                               ctx=ctx, out=ur)

    def fourier():
        pe.mex_stokes_dlp_kspace(ps, ps, prm.xi, prm.eta, f, nv, prm.Mx, prm.My, w.Lx, w.Ly,
                                 prm.w, prm.P, ctx=ctx, out=uk)

    def step():
        # the headline's step form; the tuner times the two sums side by side
        if args.sequential:
            real()
            fourier()
        else:
            with ctx.group():
                fourier()
                real()

    with torch.cuda.stream(stream):
        ctx.bind_torch_stream(stream)
        ctx.set_async(True)
        for _ in range(max(1, args.warmup)):
            step()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(args.steps):
            step()
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / args.steps
        ctx.set_async(False)
    acc = _direct_ewald_error(np, ur.cpu().numpy(), uk.cpu().numpy())
    print(f"# autotuned: P {prm.P}, nside {prm.nside_x}, grid {prm.Mx}^2, spread chunk "
          f"{info['spread_chunk'] or 'default'}: {ms:.3f} ms per step, "
          f"error vs direct Ewald {acc['u'] if acc else float('nan'):.2e}", file=sys.stderr,
          flush=True)
    return {"ms_per_step": ms, "target_evals_per_s": N / (ms * 1e-3),
            "speedup_vs_headline": fwd_ms / ms,
            "params": {"P": prm.P, "Nb": prm.Nb, "nside": [prm.nside_x, prm.nside_y],
                       "Mx": prm.Mx, "My": prm.My, "xi": prm.xi, "tol": 1e-10},
            "tuning_seconds": tune_s, "candidates": info["candidates"],
            "spread_chunk": info["spread_chunk"], "spread_chunks": info["chunks"],
            "supports": info["supports"],
            "error_vs_direct_ewald": acc,
            "workload": "headline geometry (stresslet, 200 ellipses x 500 nodes), tol 1e-10, "
                        "parameters from se2p_autotune"}


def _at_scale_variant(args, pe, W, np, torch, dev):
    """The at-scale workload: N = 4e6 uniform points in the unit box (targets
    = sources, seed 5), tol 1e-12 on a 2048^2 grid for both kernels: the
    grid is below the drivers' 2 k_inf at this tolerance, so xi and the cells
    come from the fixed-grid mode (the k-space estimate inverted at
    k_inf = 1024, the real-space estimate for the cells).  Forward step (real
    + Fourier sums) and forward + adjoint step (both backward entries, every
    gradient), with the forward state kept for the backward calls and
    rebuilt by them, each timed over a few steps after a warm-up.  The error
    of the forward total on 256 targets is measured against a tighter run
    (tests/golden/truth_uniform_N4e6.json: the oracle at tol 1e-14, 4096^2
    grid, the drivers' 121 x 121 cells)."""
    from fasttools2d_amd import params as Pm

    N = 4_000_000
    w = W.uniform_box(N, seed=5)
    truth = json.load(open(os.path.join(ROOT, "tests", "golden", "truth_uniform_N4e6.json")))
    sub = np.array(truth["targets"])
    assert np.array_equal(sub, np.random.default_rng(5).choice(N, 256, replace=False))
    du = torch.from_numpy(np.random.default_rng(78).standard_normal((N, 2))).to(dev)
    ps = torch.from_numpy(np.ascontiguousarray(w.psrc)).to(dev)
    f = torch.from_numpy(np.ascontiguousarray(w.f)).to(dev)
    nv = torch.from_numpy(np.ascontiguousarray(w.n)).to(dev)
    ur = torch.empty((N, 2), dtype=torch.float64, device=dev)
    uk = torch.empty_like(ur)
    out = {}
    for kind in ("slp", "dlp"):
        p = Pm.select_params_grid(kind, N, N, w.f, 1.0, 1.0, M=2048, tol=1e-12)
        ctx = pe.Context(0)
        stream = torch.cuda.Stream(dev)

        def pair(real, fourier):
            # as the headline step: one concurrent group, Fourier-space call first
            if args.sequential:
                real()
                fourier()
            else:
                with ctx.group():
                    fourier()
                    real()

        def fwd():
            if kind == "slp":
                pair(lambda: pe.mex_stokes_slp_real(ps, ps, f, p.xi, p.nside_x, p.nside_y, 1.0, 1.0,
                                                    ctx=ctx, out=ur),
                     lambda: pe.mex_stokes_slp_kspace(ps, ps, p.xi, p.eta, f, p.Mx, p.My, 1.0, 1.0,
                                                      p.w, p.P, ctx=ctx, out=uk))
            else:
                pair(lambda: pe.mex_stokes_dlp_real(ps, ps, f, nv, p.xi, p.nside_x, p.nside_y, 1.0,
                                                    1.0, ctx=ctx, out=ur),
                     lambda: pe.mex_stokes_dlp_kspace(ps, ps, p.xi, p.eta, f, nv, p.Mx, p.My, 1.0,
                                                      1.0, p.w, p.P, ctx=ctx, out=uk))

        def fwd_adj():
            fwd()
            if kind == "slp":
                pair(lambda: pe.mex_stokes_slp_real_backward(ps, ps, f, p.xi, p.nside_x, p.nside_y,
                                                             1.0, 1.0, du, ctx=ctx),
                     lambda: pe.mex_stokes_slp_kspace_backward(ps, ps, p.xi, p.eta, f, p.Mx, p.My,
                                                               1.0, 1.0, p.w, p.P, du, ctx=ctx))
            else:
                pair(lambda: pe.mex_stokes_dlp_real_backward(ps, ps, f, nv, p.xi, p.nside_x,
                                                             p.nside_y, 1.0, 1.0, du, ctx=ctx),
                     lambda: pe.mex_stokes_dlp_kspace_backward(ps, ps, p.xi, p.eta, f, nv, p.Mx,
                                                               p.My, 1.0, 1.0, p.w, p.P, du,
                                                               ctx=ctx))

        def timed(fn, reps=3):
            fn()
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(reps):
                fn()
            e1.record(stream)
            torch.cuda.synchronize()
            return e0.elapsed_time(e1) / reps

        with torch.cuda.stream(stream):
            ctx.bind_torch_stream(stream)
            ctx.set_async(True)
            f_ms = timed(fwd)
            fr_ms = timed(fwd_adj)      # the backward calls rebuild the forward state
            ctx.set_keep(True)
            fa_ms = timed(fwd_adj)      # the backward calls reuse the kept state
            ctx.set_keep(False)
            ctx.set_async(False)
        torch.cuda.synchronize()
        fwd()
        u = (ur + uk).cpu().numpy()[sub]
        torch.cuda.synchronize()
        ctx.close()
        ut = np.array(truth[kind]["u"])
        err = float(np.abs(u - ut).max() / np.abs(ut).max())
        print(f"# at scale {kind}: N=4e6 grid {p.Mx}^2 nside {p.nside_x} xi {p.xi:.2f}: forward "
              f"{f_ms:.2f} ms, forward+adjoint {fa_ms:.2f} ms kept / {fr_ms:.2f} ms rebuilt; "
              f"error vs tighter run {err:.2e}", file=sys.stderr, flush=True)
        out[kind] = {"forward_ms": f_ms, "forward_evals_per_s": N / (f_ms * 1e-3),
                     "forward_plus_adjoint_ms": fa_ms, "over_forward": fa_ms / f_ms,
                     "forward_plus_adjoint_rebuilt_ms": fr_ms,
                     "over_forward_rebuilt": fr_ms / f_ms,
                     "error_vs_tighter_run": err,
                     "params": {"Mx": p.Mx, "nside": p.nside_x, "xi": p.xi, "eta": p.eta,
                                "P": p.P, "tol": 1e-12, "mode": "fixed grid"}}
    out["workload"] = ("N = 4e6 uniform points, unit box, targets = sources, tol 1e-12, "
                       "grid 2048^2 for both kernels (fixed-grid parameters)")
    out["step_form"] = "sequential" if args.sequential else "grouped"
    out["error_reference"] = ("forward total (real + Fourier) on 256 targets against the oracle "
                              "at tol 1e-14 on a 4096^2 grid with the drivers' 121 x 121 cells "
                              "(tests/golden/truth_uniform_N4e6.json)")
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--cpu-runs", type=int, default=5)
    ap.add_argument("--no-variants", action="store_true",
                    help="skip the plain-vs-batched right-hand-side comparison")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the forward+adjoint variant")
    ap.add_argument("--no-scale", action="store_true",
                    help="skip the N = 4e6 forward / forward+adjoint variant")
    ap.add_argument("--sequential", action="store_true",
                    help="time the two sums one after the other instead of as one concurrent "
                         "group (se2p_group_begin / se2p_group_end)")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="also run a plain device copy of this many MiB (scales HBM counters)")
    ap.add_argument("--variant", choices=["batched", "batched_ellipses", "multibox", "at_scale"],
                    default=None,
                    help="run only this variant and print its JSON line (for profiling)")
    ap.add_argument("--no-tune", action="store_true",
                    help="--variant multibox: skip the multi-box tuner (it takes minutes)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "the spectral Ewald path runs on one GPU; run with --gpus 1",
                          "n_gpus": args.gpus}), flush=True)
        return 2

    import numpy as np
    import torch

    from fasttools2d_amd import periodic_ewald as pe
    from fasttools2d_amd import workloads as W

    if args.variant is not None:
        dev = torch.device("cuda", 0)
        torch.cuda.set_device(dev)
        run = {"batched": _batched_variant, "batched_ellipses": _batched_ellipses_variant,
               "multibox": lambda *a: _multibox_variant(*a, tune=not args.no_tune),
               "at_scale": _at_scale_variant}[args.variant]
        print(json.dumps({"variant": args.variant, "result": run(args, pe, W, np, torch, dev)}),
              flush=True)
        return 0

    w, p = W.dlp_ellipses_config(seed=2, n_ell=200, n_nodes=500, tol=1e-10, M=512)
    N = len(w.psrc)
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    ps, f, nv = t(w.psrc), t(w.f), t(w.n)
    # the targets are the sources: one array, as a caller evaluating the
    # potential on its own points passes it (the library then builds one set
    # of cell lists and bins for both)
    assert np.array_equal(w.ptar, w.psrc)
    pt = ps
    ur = torch.empty((N, 2), dtype=torch.float64, device=dev)
    uk = torch.empty((N, 2), dtype=torch.float64, device=dev)

    ctx = pe.Context(0)
    stream = torch.cuda.Stream(dev)
    ctx.bind_torch_stream(stream)

    def call_real():
        pe.mex_stokes_dlp_real(ps, pt, f, nv, p.xi, p.nside_x, p.nside_y, w.Lx, w.Ly, ctx=ctx, out=ur)

    def call_kspace():
        pe.mex_stokes_dlp_kspace(ps, pt, p.xi, p.eta, f, nv, p.Mx, p.My, w.Lx, w.Ly, p.w, p.P,
                                 ctx=ctx, out=uk)

    def step_sequential():
        call_real()
        call_kspace()

    def step_grouped():
        # neither sum reads the other's result: one concurrent group runs them
        # side by side (the Fourier-space call on the context stream, the
        # real-space call on a stream slot)
        with ctx.group():
            call_kspace()
            call_real()

    step, other_step = ((step_sequential, step_grouped) if args.sequential
                        else (step_grouped, step_sequential))

    torch.cuda.synchronize()
    with torch.cuda.stream(stream):
        ctx.bind_torch_stream(stream)
        ctx.set_async(True)
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        print(f"# warmup done ({args.warmup} steps)", file=sys.stderr, flush=True)
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        wall0 = time.perf_counter()
        ev0.record(stream)
        for _ in range(args.steps):
            step()
        ev1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - wall0
        gpu_ms = ev0.elapsed_time(ev1)
        print(f"# timed {args.steps} steps: {gpu_ms:.3f} ms (wall {wall * 1e3:.3f} ms)",
              file=sys.stderr, flush=True)
        # the other step form, for comparison (same outputs, bitwise)
        other_ms = _timed_ms(torch, stream, other_step, reps=args.steps)
        print(f"# {'grouped' if args.sequential else 'sequential'} step: {other_ms:.3f} ms",
              file=sys.stderr, flush=True)

        # per-stage kernel durations (HIP events on the context stream)
        ctx.set_timing(True)
        stages = {}
        nstage = max(3, min(args.steps, 10))
        for _ in range(nstage):
            pe.mex_stokes_dlp_real(ps, pt, f, nv, p.xi, p.nside_x, p.nside_y, w.Lx, w.Ly, ctx=ctx, out=ur)
            for k, v in ctx.stage_times().items():
                if v >= 0:
                    stages.setdefault("real_call." + k, []).append(v)
            pe.mex_stokes_dlp_kspace(ps, pt, p.xi, p.eta, f, nv, p.Mx, p.My, w.Lx, w.Ly, p.w, p.P,
                                     ctx=ctx, out=uk)
            for k, v in ctx.stage_times().items():
                if v >= 0:
                    stages.setdefault("kspace_call." + k, []).append(v)
        ctx.set_timing(False)
        ctx.set_async(False)
    torch.cuda.synchronize()
    stage_ms = {k: statistics.mean(v) for k, v in stages.items()}
    if args.calib_copy_mb > 0:
        # reference kernel for HBM counters: reads and writes exactly this many bytes
        nbytes = args.calib_copy_mb << 20
        a = torch.empty(nbytes // 8, dtype=torch.float64, device=dev).fill_(1.0)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    pairs_in, pairs_out = pe.count_pairs(ps, pt, p.nside_x, p.nside_y, w.Lx, w.Ly, ctx=ctx)
    flops_real = pairs_in * F_PAIR_IN + pairs_out * F_PAIR_OUT
    # the dominant kernel alone (k_real_win<1>), and the whole real-space
    # stage around it (chunk planning, work items, split-chunk sums)
    real_ms = stage_ms.get("real_call.pair", stage_ms["real_call.real"])
    real_stage_ms = stage_ms["real_call.real"]
    achieved_tflops = flops_real / (real_ms * 1e-3) / 1e12

    ms_per_step = gpu_ms / args.steps
    value = N * args.steps / (gpu_ms * 1e-3)

    ur_h, uk_h = ur.cpu().numpy(), uk.cpu().numpy()
    accuracy = _direct_ewald_error(np, ur_h, uk_h)
    stages_model, t_model = _stage_model(N, N, p.Mx * p.My, p.P, pairs_in, pairs_out, stage_ms)

    # CPU baseline leg: the oracle (the CPU restatement of the reference) on
    # this host, which also checks the measured result on a target subset
    cpu = None
    parity = None
    if not args.no_cpu:
        # threads spread over the cores, bound before the oracle's OpenMP
        # runtime starts; the oracle build for the host's instruction set
        # (all builds give bitwise the same results)
        os.environ.setdefault("OMP_PROC_BIND", "spread")
        os.environ.setdefault("OMP_PLACES", "cores")
        from oracle import cref

        isa = cref.host_isa()
        cref.use_isa(isa)
        sub = np.random.default_rng(55).choice(N, 256, replace=False)
        c = (cref.dlp_real(w.psrc, w.ptar[sub], w.f, w.n, p.xi, p.nside_x, p.nside_y, w.Lx, w.Ly) +
             cref.dlp_kspace(w.psrc, w.ptar[sub], p.xi, p.eta, w.f, w.n, p.Mx, p.My, w.Lx, w.Ly,
                             p.w, p.P))
        parity = float(np.abs((ur_h + uk_h)[sub] - c).max() / np.abs(c).max())
        threads = cref.get_threads()

        def cpu_real():
            cref.dlp_real(w.psrc, w.ptar, w.f, w.n, p.xi, p.nside_x, p.nside_y, w.Lx, w.Ly)

        def cpu_kspace():
            cref.dlp_kspace(w.psrc, w.ptar, p.xi, p.eta, w.f, w.n, p.Mx, p.My, w.Lx, w.Ly, p.w, p.P)

        def timed(fn, runs, label):
            ts = []
            for _ in range(runs):
                t0 = time.perf_counter()
                fn()
                ts.append(time.perf_counter() - t0)
            print(f"# cpu {label}: {', '.join(f'{x:.3f}' for x in ts)} s", file=sys.stderr, flush=True)
            return statistics.median(ts)

        cpu_real()   # warm-up
        cpu_kspace()
        real_s = timed(cpu_real, args.cpu_runs, f"real, {threads} threads")
        k_s = timed(cpu_kspace, args.cpu_runs, f"Fourier, {threads} threads")
        med = real_s + k_s
        # one thread, one run of each sum (a few seconds)
        cref.set_threads(1)
        real1 = timed(cpu_real, 1, "real, 1 thread")
        k1 = timed(cpu_kspace, 1, "Fourier, 1 thread")
        cref.set_threads(threads)
        cpu = {"value": N / med, "unit": "target evals/s", "cores": threads, "kind": "port",
               "sample": f"full workload (N={N} targets, real + Fourier sums), median of "
                         f"{args.cpu_runs} runs of each sum after 1 warm-up, {threads} OpenMP threads",
               "build": f"gcc -O3 -march={isa} -fopenmp -ffp-contract=off",
               "binding": {"OMP_PROC_BIND": os.environ.get("OMP_PROC_BIND"),
                           "OMP_PLACES": os.environ.get("OMP_PLACES")},
               "seconds_per_step": med,
               "stage_seconds": {"real": real_s, "fourier": k_s},
               "one_thread": {"value": N / (real1 + k1), "seconds_per_step": real1 + k1,
                              "stage_seconds": {"real": real1, "fourier": k1}},
               "host_cpu": _host_cpu()}

    variants = None
    if not args.no_variants:
        variants = _batched_variant(args, pe, W, np, torch, dev)
        variants["batched_ellipses"] = _batched_ellipses_variant(args, pe, W, np, torch, dev)
        variants["fixed_grid"] = _fixed_grid_variant(args, pe, np, torch, w, ps, f, nv, stream,
                                                     ms_per_step)
        variants["multibox"] = _multibox_variant(args, pe, W, np, torch, dev)
        variants["graph"] = _graph_variant(pe, torch, step, ctx, stream, N)
        variants["autotuned"] = _autotune_variant(args, pe, np, torch, w, ps, f, nv, stream,
                                                  ms_per_step)
        if not args.no_scale:
            variants["at_scale"] = _at_scale_variant(args, pe, W, np, torch, dev)
    adjoint = None
    if not args.no_adjoint:
        fwd_forms = {"sequential" if args.sequential else "grouped": ms_per_step,
                     "grouped" if args.sequential else "sequential": other_ms}
        adjoint = _adjoint_variant(args, pe, np, torch, w, p, ps, pt, f, nv, ctx, stream, fwd_forms)

    traffic = _pmc_traffic("k_real_win_dlp")
    out = {
        "metric": "target evals/s at 1e-10 rel err, 1 GPU (stresslet DLP, plain forward)",
        "value": value,
        "unit": "target evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded ellipses, seed 2)",
        "config": {"workload": "stresslet DLP, 200 ellipses x 500 nodes (N=1e5), grid 512^2, "
                               "P=24, tol 1e-10",
                   "N": N, "Mx": p.Mx, "My": p.My, "P": p.P, "xi": p.xi,
                   "nside": [p.nside_x, p.nside_y], "tol": p.tol},
        "roofline": {"bound": "valu",
                     "kernel": "k_real_win<1> (real-space stresslet pair sum over chunk windows)",
                     "achieved": achieved_tflops, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved_tflops / PEAK_FP64_TFLOPS, "traffic": traffic,
                     "flops_per_launch": flops_real, "pairs_in": pairs_in, "pairs_out": pairs_out,
                     "avg_kernel_ms": real_ms,
                     "real_stage_ms": real_stage_ms,
                     "real_stage_frac": flops_real / (real_stage_ms * 1e-3) / 1e12 / PEAK_FP64_TFLOPS,
                     "note": "achieved = pair-count flops / the pair kernel's HIP-event time "
                             "in the stage-timed calls, where it runs alone on the GPU (in the "
                             "grouped step it shares the GPU with the Fourier-space call); "
                             "real_stage_* counts the whole real-space stage (chunk plan, "
                             "work-item scan, pair kernel, split-chunk sums)"},
        "step_form": ("sequential: real-space call, then Fourier-space call" if args.sequential
                      else "grouped: the Fourier-space and real-space calls in one concurrent "
                           "group (se2p_group_begin / se2p_group_end), side by side on the GPU"),
        "step_forms_ms": {"grouped": other_ms if args.sequential else ms_per_step,
                          "sequential": ms_per_step if args.sequential else other_ms},
        "stage_ms": stage_ms,
        "stage_roofline": {"stages": stages_model, "model_ms": t_model,
                           "frac": t_model / ms_per_step,
                           "note": "model = sum over stages of max(HBM bytes / 8 TB/s, "
                                   "flops / 78.6 TFLOP/s); sort stages have no model time"},
        "parity_vs_oracle": parity,
        "error_vs_direct_ewald": accuracy,
        "cpu_baseline": cpu,
        "variants": variants,
        "forward_plus_adjoint": adjoint,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
